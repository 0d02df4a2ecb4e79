"""Writes the .ini files of a batched ensemble of small penalized simulations
(FLUSI .ini format, one file per member).

The default is the 48-member ensemble: 16 members each at 64^3, 96^3 and
128^3 in a periodic box of side 2*pi, eta = 1e-3, nu = 1e-2, RK2 with a fixed
time step.  In every grid size half the members are spheres in a uniform
x-flow (radius uniform in [0.3, 0.8]) and half are single rigid rectangular
plates flapping in quiescent fluid (BodyType=nobody, LeftWing=no) with
5-term Fourier kinematics.  All random numbers come from
numpy.random.default_rng(seed): first the sphere radii (all sizes, in order),
then the plates' kinematics coefficients.

    python -m flusi_amd.ensemble_configs OUT_DIR
"""
from __future__ import annotations

import os
import sys
from typing import List, Sequence

import numpy as np

TWO_PI = 2.0 * np.pi


def _kinematics(rng, nfft: int = 5):
    i = np.arange(1, nfft + 1)
    k = {}
    for ang, lead, amp in (("phi", ("ai", 60.0), 4.0), ("alpha", ("bi", 40.0), 4.0), ("theta", None, 2.0)):
        ai = rng.uniform(-amp, amp, nfft) / i
        bi = rng.uniform(-amp, amp, nfft) / i
        if lead is not None:
            (ai if lead[0] == "ai" else bi)[0] = lead[1] + rng.uniform(-10.0, 10.0)
        k[ang] = (ai, bi)
    return k


def _write_kinematics(path: str, k, nfft: int = 5):
    fmt = lambda a: " ".join("%.15e" % v for v in a)
    with open(path, "w") as fh:
        fh.write("[kinematics]\ntype=fourier;\nunits=degree;\nconvention=flusi;\n")
        for ang in ("phi", "alpha", "theta"):
            ai, bi = k[ang]
            fh.write(f"nfft_{ang}={nfft};\na0_{ang}=0.0;\nai_{ang}={fmt(ai)};\nbi_{ang}={fmt(bi)};\n")


def _common(n: int, dt: float, nt: int, method: str, geometry: str = "") -> str:
    return (f"[Resolution]\nnx={n};\nny={n};\nnz={n};\n"
            f"[Time]\nTmax=1.0e9;\nnt={nt};\niTimeMethodFluid={method};\nCFL=0.1;\ndt_fixed={dt!r};\n"
            "[ReynoldsNumber]\nnu=1.0e-2;\n"
            "[InitialCondition]\ninicond=quiescent;\n"
            f"[Geometry]\nxl={TWO_PI!r};\nyl={TWO_PI!r};\nzl={TWO_PI!r};\n{geometry}"
            "[Saving]\ntsave=1.0e9;\ntintegral=1.0e9;\nitdrag=1000000000;\n"
            "[Dealiasing]\niDealias=1;\n")


def write_ensemble(out_dir: str, sizes: Sequence[int] = (64, 96, 128), per_size: int = 16, seed: int = 5,
                   dt: float = 5.0e-4, nt: int = 500, method: str = "RK2") -> List[str]:
    """Writes the member .ini files (and the plates' kinematics files) into
    out_dir; returns the .ini paths, sizes in order, spheres before plates."""
    os.makedirs(out_dir, exist_ok=True)
    rng = np.random.default_rng(seed)
    nsph = per_size // 2
    nplate = per_size - nsph
    radii = rng.uniform(0.3, 0.8, nsph * len(sizes))
    paths = []
    c = TWO_PI / 2.0
    for si, n in enumerate(sizes):
        for j in range(nsph):
            r = float(radii[si * nsph + j])
            p = os.path.join(out_dir, f"sphere_{n}_{j:02d}.ini")
            with open(p, "w") as fh:
                geo = f"Size={2.0 * r!r};\nx0={c - 1.0!r};\ny0={c!r};\nz0={c!r};\n"
                fh.write(_common(n, dt, nt, method, geo))
                fh.write("[Penalization]\niPenalization=1;\niMoving=0;\niMask=sphere;\neps=1.0e-3;\n"
                         "iCavity=no;\ncompute_forces=1;\n")
                fh.write("[MeanFlow]\niMeanFlow_x=fixed;\niMeanFlow_y=fixed;\niMeanFlow_z=fixed;\n"
                         "ux=1.0;\nuy=0.0;\nuz=0.0;\n")
            paths.append(p)
    for si, n in enumerate(sizes):
        for j in range(nplate):
            kin = f"kin_plate_{n}_{j:02d}.ini"
            _write_kinematics(os.path.join(out_dir, kin), _kinematics(rng))
            span = 1.0 + 0.2 * rng.uniform(-1.0, 1.0)
            chord = 0.25 + 0.05 * rng.uniform(-1.0, 1.0)
            p = os.path.join(out_dir, f"plate_{n}_{j:02d}.ini")
            with open(p, "w") as fh:
                fh.write(_common(n, dt, nt, method))
                fh.write("[Penalization]\niPenalization=1;\niMoving=1;\niMask=Insect;\neps=1.0e-3;\n"
                         "iCavity=no;\ncompute_forces=1;\n")
                fh.write(f"[Insects]\nWingShape=rectangular;\nb_top={0.5 * chord!r};\nb_bot={0.5 * chord!r};\n"
                         f"L_span={span!r};\nWingThickness=4.0*dx;\nFlappingMotion_right=from_file::{kin};\n"
                         "FlappingMotion_left=none;\nLeftWing=no;\nRightWing=yes;\nBodyType=nobody;\n"
                         "BodyMotion=tethered;\nx_pivot_r=0.0 0.0 0.0;\nx_pivot_l=0.0 0.0 0.0;\n"
                         f"x0={c!r} {c!r} {c!r};\nyawpitchroll_0=0.0 0.0 0.0;\neta0=0.0;\n")
                fh.write("[MeanFlow]\niMeanFlow_x=free;\niMeanFlow_y=free;\niMeanFlow_z=free;\n"
                         "ux=0.0;\nuy=0.0;\nuz=0.0;\n")
            paths.append(p)
    return paths


if __name__ == "__main__":
    for q in write_ensemble(sys.argv[1] if len(sys.argv) > 1 else "ensemble_cfg"):
        print(q)
