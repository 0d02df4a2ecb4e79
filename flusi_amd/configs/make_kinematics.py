"""Writes kin_cfg3.ini and kin_cfg4.ini: seeded Fourier wing kinematics of
the tethered flapping-insect workload (stroke phi: a1 = 70 deg, pitch alpha:
b1 = 45 deg, deviation theta small), in FLUSI's from_file kinematics format
(wings_motion.f90:98-171: [kinematics] type=fourier, units=degree).
kin_cfg3.ini: 5 terms from default_rng(3); kin_cfg4.ini: 2 terms from
default_rng(4), whose 12 coefficients are the adjoint workload's controls.

    python flusi_amd/configs/make_kinematics.py
"""
import os

import numpy as np


def make(seed: int = 3, nfft: int = 5):
    rng = np.random.default_rng(seed)
    i = np.arange(1, nfft + 1)
    ai_phi = rng.uniform(-5, 5, nfft) / i
    bi_phi = rng.uniform(-5, 5, nfft) / i
    ai_phi[0] = 70.0
    ai_alpha = rng.uniform(-5, 5, nfft) / i
    bi_alpha = rng.uniform(-5, 5, nfft) / i
    bi_alpha[0] = 45.0
    ai_theta = rng.uniform(-3, 3, nfft) / i
    bi_theta = rng.uniform(-3, 3, nfft) / i
    return dict(ai_phi=ai_phi, bi_phi=bi_phi, ai_alpha=ai_alpha, bi_alpha=bi_alpha,
                ai_theta=ai_theta, bi_theta=bi_theta, nfft=nfft)


def write(path: str, seed: int = 3, nfft: int = 5):
    k = make(seed, nfft)
    fmt = lambda a: " ".join("%.15e" % v for v in a)
    with open(path, "w") as fh:
        fh.write("; seeded wing kinematics (numpy default_rng(%d)), FLUSI from_file format\n" % seed)
        fh.write("[kinematics]\ntype=fourier;\nunits=degree;\nconvention=flusi;\n")
        for ang in ("phi", "alpha", "theta"):
            fh.write(f"nfft_{ang}={k['nfft']};\n")
            fh.write(f"a0_{ang}=0.0;\n")
            fh.write(f"ai_{ang}={fmt(k['ai_' + ang])};\n")
            fh.write(f"bi_{ang}={fmt(k['bi_' + ang])};\n")


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    write(os.path.join(here, "kin_cfg3.ini"))
    write(os.path.join(here, "kin_cfg4.ini"), seed=4, nfft=2)
