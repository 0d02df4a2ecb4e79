"""The benchmark configurations other than the 512^3 insect, on the GPU
against the FP64 oracle:

* the penalised sphere in the 4 x 2 x 2 box with the outlet_x vorticity
  sponge, fixed mean flow and CFL-adapted RK2 (flusi_amd/configs/
  cfg2_sphere_256x128x128.ini): steps at reduced resolution in both
  precisions, one right-hand side at the full 256 x 128 x 128, and one on a
  64 x 64 x 512 grid (the six-field z step at nz = 512);
* the 64^3 Gaussian-core vortex ring with AB2 (cfg1_vortex_ring_64.ini).
"""
import os

import numpy as np
import pytest

import flusi_amd
from oracle import flusi_oracle as fo
from tests._util import CONFIGS, GOLDEN, check_err, rel_l2, rewrite_ini

pytestmark = pytest.mark.gpu

PRECS = [(flusi_amd.FP64, 1e-12), (flusi_amd.FP32_MIXED, 2e-6)]


def _perturbed(sim, seed=3, amp=0.2):
    """The oracle's initial uk plus a smooth solenoidal field of a real velocity."""
    g = sim.g
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((3, g.nxh, g.ny, g.nz)) + 1j * rng.standard_normal((3, g.nxh, g.ny, g.nz))
    v *= np.exp(-0.5 * g.K2 / (4.0 * (2 * np.pi / g.p.xl) ** 2))[None]
    v = fo.add_grad_pressure(g, v)
    v[:, 0, 0, 0] = 0.0
    v = fo.fft3(g, fo.ifft3(g, v))
    v *= amp / max(1e-30, np.sqrt(fo.energy_k(g, v)))
    return sim.uk + v


@pytest.mark.parametrize("prec,tol", PRECS)
def test_sphere_outlet_sponge_rk2_steps(tmp_path, prec, tol):
    """3 CFL-adapted RK2 steps of the sphere workload at 64 x 32 x 32 (sponge
    over the last 7 of 64 x cells, the same 10 % as 26 of 256)."""
    ini = rewrite_ini(os.path.join(CONFIGS, "cfg2_sphere_256x128x128.ini"), str(tmp_path),
                      Resolution__nx=64, Resolution__ny=32, Resolution__nz=32, Sponge__sponge_thickness=7,
                      Saving__tintegral=100.0, Saving__itdrag=100000)
    sim = fo.Simulation(ini, workers=4)
    sim.init_fields()
    sim.uk = _perturbed(sim)
    with flusi_amd.Flusi(ini, prec) as s:
        s.init_fields()
        s.set_uk(sim.uk)
        m, _, _ = s.get_mask()
        assert np.max(np.abs(m - sim.mask)) < (1e-6 if prec else 1e-14)
        for _ in range(3):
            sim.step()
            s.step()
        t, it, _, dt1 = s.get_time()
        assert it == sim.it and abs(dt1 - sim.dt1) <= 1e-12 * sim.dt1
        err = rel_l2(s.get_uk(), sim.uk)
        check_err(err, 5e-11 if prec == flusi_amd.FP64 else 4e-8, "uk after 3 RK2 steps")


def test_sphere_full_size_rhs_fp64(tmp_path):
    """One right-hand side (penalisation, outlet_x sponge) of the sphere
    workload at its full 256 x 128 x 128."""
    ini = os.path.join(CONFIGS, "cfg2_sphere_256x128x128.ini")
    sim = fo.Simulation(ini, workers=8)
    sim.init_fields()
    uk = _perturbed(sim)
    nlk_ref, _ = sim.cal_nlk(0.0, uk)
    with flusi_amd.Flusi(ini, flusi_amd.FP64) as s:
        s.init_fields()
        s.set_uk(uk)
        err = rel_l2(s.cal_nlk(0.0), nlk_ref)
    print(f"sphere 256x128x128 cal_nlk FP64 rel L2 {err:.2e}", flush=True)
    assert err < 1e-11


@pytest.mark.parametrize("prec,tol", [(flusi_amd.FP64, 1e-11), (flusi_amd.FP32_MIXED, 1.1e-6)])
def test_sphere_sponge_rhs_nz512(tmp_path, prec, tol):
    """One right-hand side of the sphere workload with the outlet_x sponge on a
    64 x 64 x 512 grid: the z step's six-field (sponge) variant at nz = 512,
    which runs with spilled registers in both precisions."""
    ini = rewrite_ini(os.path.join(CONFIGS, "cfg2_sphere_256x128x128.ini"), str(tmp_path),
                      Resolution__nx=64, Resolution__ny=64, Resolution__nz=512, Sponge__sponge_thickness=7,
                      Saving__tintegral=100.0, Saving__itdrag=100000)
    sim = fo.Simulation(ini, workers=8)
    sim.init_fields()
    uk = _perturbed(sim)
    nlk_ref, _ = sim.cal_nlk(0.0, uk)
    with flusi_amd.Flusi(ini, prec) as s:
        s.init_fields()
        s.set_uk(uk)
        err = rel_l2(s.cal_nlk(0.0), nlk_ref)
    check_err(err, tol, "sponge nlk at nz = 512")


def test_vortex_ring_64_ab2_steps_fp64(tmp_path):
    """10 AB2 steps (euler_startup first) of the 64^3 Gaussian-core ring."""
    ini = rewrite_ini(os.path.join(CONFIGS, "cfg1_vortex_ring_64.ini"), str(tmp_path),
                      Saving__tintegral=100.0, Saving__itdrag=100000)
    sim = fo.Simulation(ini, workers=4)
    sim.init_fields()
    with flusi_amd.Flusi(ini, flusi_amd.FP64) as s:
        s.init_fields()
        assert rel_l2(s.get_uk(), sim.uk) < 1e-12
        for _ in range(10):
            sim.step()
            s.step()
        assert s.get_time()[1] == sim.it == 10
        err = rel_l2(s.get_uk(), sim.uk)
        e_gpu, e_ref = s.energy_k(), fo.energy_k(sim.g, sim.uk)
    assert err < 1e-10
    assert abs(e_gpu - e_ref) <= 1e-12 * abs(e_ref)


@pytest.mark.parametrize("dealias", [1, 0])
def test_cal_vis_and_energy_on_device_match_oracle(tmp_path, dealias):
    """flusi_cal_vis (vis.f90:9-66) and flusi_energy_k (integrals.f90:522-552)
    computed on the GPU: the integrating-factor table with and without the
    2/3 dealiasing, and the Parseval energy of the 32^3 ring."""
    ini = rewrite_ini(os.path.join(GOLDEN, "vortex_ring1_AB2", "vortexring1_AB2.ini"), str(tmp_path),
                      Dealiasing__iDealias=dealias, Geometry__xl=2.5)
    sim = fo.Simulation(ini)
    sim.init_fields()
    with flusi_amd.Flusi(ini, flusi_amd.FP64) as s:
        s.init_fields()
        for dt in (1e-3, 0.37):
            vis = s.cal_vis(dt)
            ref = fo.cal_vis(sim.g, dt, sim.p.nu)
            np.testing.assert_allclose(vis, ref, rtol=1e-14, atol=0.0)
            assert np.array_equal(vis == 0.0, ref == 0.0)
        e = s.energy_k()
    e_ref = fo.energy_k(sim.g, sim.uk)
    assert abs(e - e_ref) <= 1e-13 * abs(e_ref)
