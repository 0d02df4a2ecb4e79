"""Builds oracle/cpu/flusi_cpu, the C++ OpenMP restatement used as the CPU
baseline and as a second checker (test infrastructure, not the product).

    python oracle/cpu/build.py

No FMA contraction (-ffp-contract=off): the restatement then rounds as the
numpy oracle does, which matters where the reference's wing taper jumps
from 1/2 to 1 at the root plane and a grid point lies on that plane.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "flusi_cpu")


def build() -> str:
    tmp = BIN + ".tmp"
    cmd = ["g++", "-O3", "-std=c++17", "-fopenmp", "-march=x86-64-v3", "-ffp-contract=off",
           os.path.join(HERE, "flusi_cpu.cpp"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle/cpu build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, BIN)
    return BIN


if __name__ == "__main__":
    print(build())
