"""Builds libflusi_hip.so (gfx950) in-tree with hipcc.

    python -m flusi_amd.build [--force]

Every source is compiled to an object under _build/ (in parallel), then linked
into flusi_amd/libflusi_hip.so next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libflusi_hip.so")
SOURCES = ["passes.hip", "tiles_wave.hip", "zstep.hip", "xrhs_wave.hip", "fields.hip", "adjoint.hip", "solver.hip",
           "host_params.cpp", "h5bind.cc"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FLUSI_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         # packed FP32 instructions are kept out of all device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include")]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cc"):  # host-only C++ (dlopen of the HDF5 layer), no device pass
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        return obj
    cmd = [HIPCC, *FLAGS]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(force: bool = True, jobs: int = 4) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_h5io()
    build_fortran()
    build_probe()
    return LIB


PROBE_LIB = os.path.join(HERE, "libflusi_probe.so")


def build_probe() -> str:
    """libflusi_probe.so: the HBM stream probe bench.py compares the passes
    with (copy, read-only and write-only bandwidth of this device)."""
    tmp = PROBE_LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-shared", os.path.join(CSRC, "probe.hip"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"probe build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, PROBE_LIB)
    return PROBE_LIB


HDF5_DIR = os.environ.get("FLUSI_HDF5_DIR", "/opt/conda")
H5LIB = os.path.join(HERE, "libflusi_h5.so")


def build_h5io() -> str:
    """libflusi_h5.so: the HDF5 field-file layer (plain C against the image's
    libhdf5), loaded by libflusi_hip.so on the first backup or snapshot.
    Returns its path, or "" when no HDF5 installation is found (those entry
    points then report an error)."""
    inc = os.path.join(HDF5_DIR, "include")
    lib = os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"warning: no hdf5.h under {inc}: HDF5 snapshots and backups unavailable", file=sys.stderr)
        return ""
    tmp = H5LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c99", "-shared", "-fPIC", "-I", inc, os.path.join(CSRC, "h5io.c"), "-o", tmp,
           "-L", lib, "-lhdf5", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"HDF5 layer build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, H5LIB)
    return H5LIB


FLANG = os.environ.get("FLUSI_FC", "/opt/rocm/bin/amdflang")
FORTRAN_BIN = os.path.join(HERE, "bin", "flusi_amd")


def build_fortran() -> str:
    """The iso_c_binding module and the Fortran host driver (needs amdflang);
    returns the driver path, or "" when no Fortran compiler is present."""
    if not os.path.exists(FLANG):
        return ""
    fdir = os.path.join(HERE, "fortran")
    obj = os.path.join(BUILD, "flusi_hip_f.o")
    cmds = [
        [FLANG, "-O2", "-J", BUILD, "-c", os.path.join(fdir, "flusi_hip.f90"), "-o", obj],
        [FLANG, "-O2", "-J", BUILD, os.path.join(fdir, "flusi_amd_main.f90"), obj, "-L", HERE, "-lflusi_hip",
         "-Wl,-rpath,$ORIGIN/..", "-o", FORTRAN_BIN],
    ]
    os.makedirs(os.path.dirname(FORTRAN_BIN), exist_ok=True)
    for cmd in cmds:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return FORTRAN_BIN


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
