"""The benchmarked workloads at their own sizes, on the GPU:

* the 512^3 tethered flapping insect (flusi_amd/configs/cfg3_insect_512.ini):
  one cal_nlk in FP64 and in FP32-mixed against the FP64 oracle, with the
  mask, u_s and colour the right-hand side drew;
* RK4 steps of the same insect on 64 x 64 x 512 and 512 x 64 x 64 grids in
  both precisions, which run the z step (nz = 512, the 8-line FP32 tile) and
  the x pass (nx = 512) that the 512^3 benchmark times, against the oracle;
* the adjoint workload (cfg4_insect_adjoint_128.ini: 128^3, 400 RK4 steps,
  checkpoints every 20, 12 controls) against central differences;
* the 48-member ensemble (16 each at 64^3, 96^3, 128^3): every way of
  advancing it gives the same fields, and members match the oracle.

Tolerances: FP64 within the round-off of reordered FFT sums; FP32-mixed
bounds are about five times the error measured on MI355X (printed by each
test).
"""
import os
import time

import numpy as np
import pytest

import flusi_amd
from oracle import flusi_oracle as fo
from tests._util import CONFIGS, rel_l2, rewrite_ini

pytestmark = pytest.mark.gpu

WORKERS = max(1, min(16, os.cpu_count() or 1))


def _lowmode_state(sim, seed, amp, kmax=6):
    """A solenoidal velocity of a real field whose spectrum fills the modes
    with |index| <= kmax in every direction (so that it is resolved on every
    grid used here), scaled to kinetic energy amp^2, added to the oracle's
    initial uk."""
    g = sim.g
    rng = np.random.default_rng(seed)
    iy = np.r_[0:kmax + 1, g.ny - kmax:g.ny]
    iz = np.r_[0:kmax + 1, g.nz - kmax:g.nz]
    shape = (3, kmax + 1, iy.size, iz.size)
    v = np.zeros((3, g.nxh, g.ny, g.nz), dtype=np.complex128)
    v[:, :kmax + 1][:, :, iy[:, None], iz[None, :]] = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    v = fo.add_grad_pressure(g, v)
    v[:, 0, 0, 0] = 0.0
    v = fo.fft3(g, fo.ifft3(g, v))
    v *= amp / np.sqrt(fo.energy_k(g, v))
    return sim.uk + v


# ------------------------------------------------------------- 512^3 insect


@pytest.fixture(scope="module")
def insect512(tmp_path_factory):
    d = tmp_path_factory.mktemp("cfg3")
    ini = rewrite_ini(os.path.join(CONFIGS, "cfg3_insect_512.ini"), str(d))
    print(f"\n[512^3 oracle] {WORKERS} workers", flush=True)
    t0 = time.perf_counter()
    sim = fo.Simulation(ini, workers=WORKERS)
    sim.init_fields()
    uk = _lowmode_state(sim, seed=7, amp=0.5)
    nlk, _ = sim.cal_nlk(0.0, uk)
    print(f"[512^3 oracle] cal_nlk done in {time.perf_counter() - t0:.1f} s", flush=True)
    return dict(ini=ini, uk=uk, nlk=nlk, mask=sim.mask, us=sim.us, color=sim.color)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("prec,tol", [(flusi_amd.FP64, 1e-11), (flusi_amd.FP32_MIXED, 1.5e-6)])
def test_insect_512_cal_nlk_matches_oracle(insect512, prec, tol):
    """rhs.f90:2-115 at the benchmark's 512^3 with the insect mask drawn at
    t = 0 (create_mask inside cal_nlk, iMoving = 1)."""
    ref = insect512
    print(f"[512^3 GPU] precision {prec}", flush=True)
    with flusi_amd.Flusi(ref["ini"], prec) as s:
        s.init_fields()
        s.set_uk(ref["uk"])
        nlk = s.cal_nlk(0.0)
        m, us, col = s.get_mask()
    err = rel_l2(nlk, ref["nlk"])
    em = float(np.max(np.abs(m - ref["mask"])))
    eu = float(np.max(np.abs(us - ref["us"]))) / max(1e-300, float(np.max(np.abs(ref["us"]))))
    ncol = int(np.count_nonzero(col != ref["color"].astype(np.int16)))
    print(f"insect 512^3 cal_nlk precision {prec}: rel L2 {err:.3e}, mask max err {em:.2e}, "
          f"us rel max err {eu:.2e}, colour mismatches {ncol}", flush=True)
    assert err < tol
    if prec == flusi_amd.FP64:
        assert em < 1e-12 and eu < 1e-12 and ncol == 0
    else:
        # mask and us are stored in FP32; a wing drawn over the body compares
        # with the stored (rounded) body value, which can flip the colour of
        # a point where the two are equal to FP32 precision
        assert em < 1e-6 and eu < 1e-6 and ncol <= 1e-6 * m.size


def _insect_ini(tmp_path, nx, ny, nz, **kw):
    over = dict(Resolution__nx=nx, Resolution__ny=ny, Resolution__nz=nz, Time__nt=3,
                Time__dt_fixed="4.0e-4", Penalization__eps="1.0e-2", Saving__tintegral=100.0,
                Saving__itdrag=100000)
    over.update(kw)
    return rewrite_ini(os.path.join(CONFIGS, "cfg3_insect_512.ini"), str(tmp_path), **over)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("shape", [(64, 64, 512), (512, 64, 64), (64, 64, 64)])
@pytest.mark.parametrize("prec,tol", [(flusi_amd.FP64, 1e-12), (flusi_amd.FP32_MIXED, 6e-7)])
def test_insect_rk4_steps_match_oracle(tmp_path, shape, prec, tol):
    """Three RK4 steps (fluid_time_step.f90:602-671: explicit diffusion,
    dealiasing of every stage, the mask redrawn at t, t+dt/2, t+dt/2, t+dt)
    of the flapping insect: nz = 512 runs the benchmark's z-step kernel, nx =
    512 its x pass."""
    ini = _insect_ini(tmp_path, *shape)
    print(f"[RK4 insect] grid {shape} precision {prec}", flush=True)
    sim = fo.Simulation(ini, workers=WORKERS)
    sim.init_fields()
    sim.uk = _lowmode_state(sim, seed=11, amp=0.3)
    with flusi_amd.Flusi(ini, prec) as s:
        s.init_fields()
        s.set_uk(sim.uk)
        for _ in range(3):
            sim.step()
            s.step()
        t, it, _, dt1 = s.get_time()
        assert it == sim.it == 3 and t == pytest.approx(sim.time, rel=1e-14)
        err = rel_l2(s.get_uk(), sim.uk)
        m, us, col = s.get_mask()
    em = float(np.max(np.abs(m - sim.mask)))
    print(f"insect RK4 x3 grid {shape} precision {prec}: rel L2 {err:.3e}, mask max err {em:.2e}", flush=True)
    assert err < tol
    ncol = int(np.count_nonzero(col != sim.color.astype(np.int16)))
    assert em < (1e-12 if prec == flusi_amd.FP64 else 1e-6)
    assert ncol == 0 if prec == flusi_amd.FP64 else ncol <= 1e-6 * col.size + 2, ncol


@pytest.mark.timeout(600)
@pytest.mark.parametrize("shape", [(64, 64, 512), (512, 64, 64)])
@pytest.mark.parametrize("prec,tol", [(flusi_amd.FP64, 1e-16), (flusi_amd.FP32_MIXED, 2e-8)])
def test_insect_rk4_fixed_mode_and_line_flags_match_general_path(tmp_path, monkeypatch, shape, prec, tol):
    """The same three RK4 insect steps through the z step's fixed-mode kernels
    (one per RK4 stage, operands loaded one position ahead) and the x pass's
    mask line flags (the default), and through the run-time-mode z-step
    kernel (FLUSI_ZSTEP_EU=1 selects it at nz >= 256) with every mask line
    read (FLUSI_MASK_FLAGS=0): the same arithmetic.  With nz < 256 both runs
    take the same z-step kernels and only the line flags differ, which must
    not change a bit; at nz = 512 the two z-step kernels agree to round-off
    (measured 9.8e-18 FP64 and 4.1e-9 FP32, the bounds about five times
    that)."""
    ini = _insect_ini(tmp_path, *shape)
    print(f"[fixed vs general path] grid {shape} precision {prec}", flush=True)
    sim = fo.Simulation(ini, workers=WORKERS)
    sim.init_fields()
    u0 = _lowmode_state(sim, seed=23, amp=0.3)

    def run():
        with flusi_amd.Flusi(ini, prec) as s:
            s.init_fields()
            s.set_uk(u0)
            for _ in range(3):
                s.step()
            return s.get_uk(), s.get_mask()[0]

    uk_fast, m_fast = run()
    monkeypatch.setenv("FLUSI_ZSTEP_EU", "1")
    monkeypatch.setenv("FLUSI_MASK_FLAGS", "0")
    uk_gen, m_gen = run()
    err = rel_l2(uk_fast, uk_gen)
    print(f"fixed-mode + line flags vs general path, grid {shape} precision {prec}: rel L2 {err:.3e}", flush=True)
    assert np.array_equal(m_fast, m_gen)
    if shape[2] < 256:
        assert err == 0.0
    assert err <= tol


# ------------------------------------------------------------- adjoint at 128^3


@pytest.mark.timeout(600)
def test_adjoint_cfg4_full_size(tmp_path):
    """d(mean lift over 400 RK4 steps)/d(12 kinematics coefficients) at 128^3
    with a checkpoint every 20 steps: the two largest components against
    central differences (h = 1e-4) of the FP64 mean lift, and the FP32-mixed
    gradient against the FP64 one."""
    ini = rewrite_ini(os.path.join(CONFIGS, "cfg4_insect_adjoint_128.ini"), str(tmp_path))
    K, every, h = 400, 20, 1.0e-4
    print("[adjoint 128^3] FP64 forward/backward", flush=True)
    with flusi_amd.Flusi(ini, flusi_amd.FP64) as s:
        th = s.adjoint_controls()
        assert th.size == 12
        J = s.adjoint_forward(th, K, every)
        g = s.adjoint_backward()
        big = np.argsort(-np.abs(g))[:2]
        fd = {}
        for i in big:
            tp, tm = th.copy(), th.copy()
            tp[i] += h
            tm[i] -= h
            fd[int(i)] = (s.adjoint_forward(tp, K, every) - s.adjoint_forward(tm, K, every)) / (2 * h)
    print(f"[adjoint 128^3] J = {J:.12g}; adjoint {[(i, g[i]) for i in fd]}; central differences {fd}", flush=True)
    for i, v in fd.items():
        # measured 2.6e-10 and 5.3e-9: the O(h^2) error of the differences
        assert abs(g[i] - v) <= 3e-8 * abs(v), (i, g[i], v)
    print("[adjoint 128^3] FP32-mixed forward/backward", flush=True)
    with flusi_amd.Flusi(ini, flusi_amd.FP32_MIXED) as s32:
        J32 = s32.adjoint_forward(th, K, every)
        g32 = s32.adjoint_backward()
    eg = float(np.linalg.norm(g32 - g) / np.linalg.norm(g))
    print(f"[adjoint 128^3] FP32-mixed: J rel err {abs(J32 - J) / abs(J):.2e}, gradient rel L2 {eg:.2e}", flush=True)
    # measured 2.2e-7 and 2.1e-7
    assert abs(J32 - J) <= 1.5e-6 * abs(J)
    assert eg <= 1.5e-6


# ------------------------------------------------------------- 48-member ensemble


@pytest.mark.timeout(900)
def test_ensemble_48_members_full_size(tmp_path):
    """The benchmark's ensemble (flusi_amd.ensemble_configs defaults: 16
    members each at 64^3, 96^3 and 128^3, spheres then plates per size, RK2,
    FP32-mixed): member-by-member, batched and graph-replayed advances agree
    bit for bit after the eager first step, the capture and two replays; one
    96^3 sphere and one 128^3 plate match the oracle after the second step
    (the first RK2 step from a uniform flow changes only its zero mode, since
    at it = 0 the stale integrating factor drops N0 and N1 = N0,
    fluid_time_step.f90:554-597)."""
    from flusi_amd.ensemble_configs import write_ensemble
    paths = write_ensemble(str(tmp_path / "ens"))
    assert len(paths) == 48
    modes = (flusi_amd.ENSEMBLE_MEMBERS, flusi_amd.ENSEMBLE_BATCHED, flusi_amd.ENSEMBLE_GRAPH)
    ens = {}
    for mo in modes:
        e = flusi_amd.Ensemble(paths, flusi_amd.FP32_MIXED)
        e.init_fields()
        ens[mo] = e
    assert sorted(ens[modes[2]].batches()) == sorted([(8, n, n, n) for n in (64, 96, 128) for _ in range(2)])
    checks = {8: 96, 47: 128}  # spheres of every size first, then the plates
    for step in range(4):
        print(f"[ensemble 48] step {step}", flush=True)
        for mo in modes:
            ens[mo].advance(1, mo)
        ref = ens[flusi_amd.ENSEMBLE_MEMBERS]
        for i in range(len(paths)):
            u0 = ref.member(i).get_uk()
            t0 = ref.member(i).get_time()
            for mo in modes[1:]:
                m = ens[mo].member(i)
                assert m.get_time()[:2] == t0[:2], (step, i, mo)
                assert np.array_equal(m.get_uk(), u0), (step, i, mo)
        if step == 1:
            for i, n in checks.items():
                sim = fo.Simulation(paths[i], workers=WORKERS)
                sim.init_fields()
                assert sim.g.nx == n
                sim.step()
                sim.step()
                err = rel_l2(ens[flusi_amd.ENSEMBLE_GRAPH].member(i).get_uk(), sim.uk)
                print(f"[ensemble 48] member {i} ({n}^3) after two steps: rel L2 vs oracle {err:.3e}", flush=True)
                assert err < 1e-5
    for e in ens.values():
        e.close()
