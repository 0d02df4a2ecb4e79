"""The C++ OpenMP restatement (oracle/cpu/flusi_cpu.cpp, the benchmark's CPU
baseline) against the numpy oracle: RK4 steps of the benchmark's insect
configuration class from rest (the wings drive the flow) and from a random
solenoidal state, on grids whose FFT lengths take every radix the
restatement has (2, 3, 4, 5)."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import flusi_oracle as fo
from tests._util import CONFIGS, ROOT, rewrite_ini

BIN = os.path.join(ROOT, "oracle", "cpu", "flusi_cpu")


def _bin():
    if not os.path.exists(BIN):
        pytest.skip("oracle/cpu/flusi_cpu is not built (python -c 'import __graft_entry__ as g; g.build()')")
    return BIN


def _ini(tmp_path, n, **kw):
    nx, ny, nz = n
    over = dict(Resolution__nx=nx, Resolution__ny=ny, Resolution__nz=nz, Time__dt_fixed="2.0e-3",
                Penalization__eps="1.0e-2")
    over.update(kw)
    return rewrite_ini(os.path.join(CONFIGS, "cfg3_insect_512.ini"), str(tmp_path), **over)


def _run(ini, steps, uk_in=None, tmp_path=None):
    out = str(tmp_path / "uk_out.bin")
    args = [_bin(), ini, str(steps), "4", uk_in or "-", out]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    return info, out


@pytest.mark.parametrize("n", [(32, 24, 40), (48, 32, 30)])
def test_steps_from_rest_match_oracle(tmp_path, n):
    ini = _ini(tmp_path, n)
    sim = fo.Simulation(ini)
    sim.init_fields()
    for _ in range(3):
        sim.step()
    info, out = _run(ini, 3, tmp_path=tmp_path)
    g = sim.g
    got = np.fromfile(out, dtype=np.complex128).reshape(3, g.nxh, g.ny, g.nz)
    err = np.linalg.norm(got - sim.uk) / np.linalg.norm(sim.uk)
    print(f"grid {n}: rel L2 {err:.3e}, |uk| {np.linalg.norm(sim.uk):.3e}, {info}")
    assert np.linalg.norm(sim.uk) > 0.0 and err <= 1e-12


def test_step_from_random_state_matches_oracle(tmp_path):
    ini = _ini(tmp_path, (40, 32, 36))
    sim = fo.Simulation(ini)
    sim.init_fields()
    g = sim.g
    rng = np.random.default_rng(11)
    v = rng.standard_normal((3, g.nxh, g.ny, g.nz)) + 1j * rng.standard_normal((3, g.nxh, g.ny, g.nz))
    v *= np.exp(-0.5 * g.K2 / (4.0 * (2 * np.pi / g.p.xl) ** 2))[None]
    v = fo.add_grad_pressure(g, v)
    v[:, 0, 0, 0] = 0.0
    v = fo.fft3(g, fo.ifft3(g, v))  # Hermitian-consistent
    uk0 = np.ascontiguousarray(v)
    path = str(tmp_path / "uk_in.bin")
    uk0.tofile(path)
    sim.uk = uk0.copy()
    sim.step()
    sim.step()
    info, out = _run(ini, 2, uk_in=path, tmp_path=tmp_path)
    got = np.fromfile(out, dtype=np.complex128).reshape(3, g.nxh, g.ny, g.nz)
    err = np.linalg.norm(got - sim.uk) / np.linalg.norm(sim.uk)
    print(f"rel L2 {err:.3e}, {info}")
    assert err <= 1e-12


def test_out_of_scope_configurations_are_refused(tmp_path):
    """The restatement covers one configuration class and says so."""
    for kw in (dict(Time__iTimeMethodFluid="AB2"), dict(Insects__BodyType="jerry"),
               dict(Time__dt_fixed="0.0"), dict(MeanFlow__iMeanFlow_x="fixed")):
        ini = _ini(tmp_path, (16, 16, 16), **kw)
        r = subprocess.run([_bin(), ini, "1", "1"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and "out of scope" in r.stderr, (kw, r.stderr)
