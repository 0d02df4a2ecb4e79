"""Independent checks of the oracle branches that no reference fixture runs
directly: the outlet_x vorticity sponge region, the ellipsoid_plain insect
body and the sinusoidal mean flow.  Each is tied to something the reference
fixes by other means (a fixture-pinned sibling branch, a closed form, a
limit), so that a plausible slip -- an off-by-one column, the wrong mask
threshold, b_body/L_body not taken from the file, the step's end time in
place of its start -- fails here.  CPU only."""
import math
import os

import numpy as np
import pytest

from oracle import flusi_oracle as fo
from tests._util import CONFIGS, GOLDEN, rewrite_ini

# ---------------------------------------------------------------- outlet_x


def _sponge_sim(tmp_path, kind, th=5, n=(32, 16, 16)):
    ini = rewrite_ini(os.path.join(CONFIGS, "cfg2_sphere_256x128x128.ini"), str(tmp_path), name=f"{kind}.ini",
                      Resolution__nx=n[0], Resolution__ny=n[1], Resolution__nz=n[2],
                      Sponge__iSpongeType=kind, Sponge__sponge_thickness=th)
    sim = fo.Simulation(ini)
    sim.mask = np.zeros((n[2], n[1], n[0]))
    return sim


def test_outlet_x_first_sponge_column(tmp_path):
    """sponge.f90:262-277: the sponge starts at ix = nx - sponge_thickness
    (0-based, 'ix >= nx-1-sponge_thickness+1').  Vorticity placed only in that
    column comes back as the full -w/eps_sponge; placed only in the column
    before it, the sponge term is exactly zero."""
    th = 5
    sim = _sponge_sim(tmp_path, "outlet_x", th)
    nx = sim.g.nx
    rng = np.random.default_rng(1)
    for col, inside in ((nx - th, True), (nx - th - 1, False), (nx - 1, True), (0, False)):
        w = np.zeros_like(sim.mask)
        w[:, :, col] = rng.standard_normal(w[:, :, col].shape)
        out = sim.penalize_vort(w)
        if inside:
            np.testing.assert_array_equal(out, -w * (1.0 / sim.p.eps_sponge))  # -vort*eps_inv
        else:
            assert not out.any(), col


def test_outlet_x_is_cavity_upper_x_slab(tmp_path):
    """Away from the y and z walls and in the upper half of x, the cavity
    sponge (pinned by the sphere_sponge_RK2 fixture) reduces to its
    'ix >= nx-1-th+1' clause, which is the whole outlet_x predicate: the two
    regions must agree there, over a mask that is solid in part of the
    sponge.  Everywhere else outlet_x is that clause alone."""
    th = 5
    so = _sponge_sim(tmp_path, "outlet_x", th)
    sc = _sponge_sim(tmp_path, "cavity", th)
    g = so.g
    rng = np.random.default_rng(2)
    m = np.where(rng.random(so.mask.shape) < 0.3, rng.random(so.mask.shape), 0.0)
    so.mask = m
    sc.mask = m.copy()
    w = np.ones_like(m)
    ro = so.penalize_vort(w) != 0.0
    rc = sc.penalize_vort(w) != 0.0
    ix = np.arange(g.nx)[None, None, :]
    iy = np.arange(g.ny)[None, :, None]
    iz = np.arange(g.nz)[:, None, None]
    interior = np.broadcast_to((iy >= th) & (iy <= g.ny - 1 - th) & (iz >= th) & (iz <= g.nz - 1 - th)
                               & (ix >= g.nx // 2), m.shape)
    assert interior.any() and ro[interior].any()
    np.testing.assert_array_equal(ro[interior], rc[interior])
    # outside that window outlet_x is the x clause only: the y/z walls and
    # the lower x slab of the cavity are not sponge
    lower = np.broadcast_to(ix < g.nx - th, m.shape)
    assert not ro[lower].any() and rc[lower].any()


def test_outlet_x_mask_threshold_is_single_precision(tmp_path):
    """outlet_x compares the FP64 mask with the literal 1e-12 written without
    a kind (sponge.f90:270), i.e. a default-real constant, 9.99999996e-13 after
    promotion; the cavity branch writes 1.0d-12.  A mask value between the two
    is fluid for the cavity sponge and solid for outlet_x."""
    th = 5
    lo = float(np.float32(1.0e-12))
    assert lo < 1.0e-12
    between = 0.5 * (lo + 1.0e-12)
    for kind, expect in (("outlet_x", False), ("cavity", True)):
        sim = _sponge_sim(tmp_path, kind, th)
        sim.mask[...] = between
        out = sim.penalize_vort(np.ones_like(sim.mask))
        assert bool(out[:, th + 1, sim.g.nx - 1].any()) == expect, kind
        sim.mask[...] = 0.5 * lo
        assert sim.penalize_vort(np.ones_like(sim.mask))[:, th + 1, sim.g.nx - 1].all(), kind


# ---------------------------------------------------------------- ellipsoid_plain


def test_ellipsoid_plain_reproduces_jerry_fixture(tmp_path, monkeypatch):
    """With b_body = 0.1, L_body = 1 read from the file and Jerry's head and
    eye spheres drawn after it, the ellipsoid_plain body reproduces the
    insects_mask_jerry fixture keys (mask and us at t = 0, 0.5, 1): the branch
    takes the ellipsoid from the file (body_geometry.f90:511-595 draws the
    same shape from hard-coded values)."""
    d = os.path.join(GOLDEN, "insects_mask_jerry")
    ini = rewrite_ini(os.path.join(d, "jerry.ini"), str(tmp_path), Insects__BodyType="ellipsoid_plain",
                      Insects__b_body=0.1, Insects__L_body=1.0)
    orig = fo.Insect._draw_body

    def body_then_head(self, g, mask, color, us):
        orig(self, g, mask, color, us)
        Mb, xc = self.st["M_body"], self.st["xc"]
        for xp, rr in ((self.x_head, self.R_head), (self.x_eye_l, self.R_eye), (self.x_eye_r, self.R_eye)):
            self._drawsphere(g, xc + Mb.T @ xp, rr, mask, color)

    monkeypatch.setattr(fo.Insect, "_draw_body", body_then_head)
    sim = fo.Simulation(ini)
    sim.insect = fo.Insect(sim.ini, sim.p, d)
    assert sim.insect.BodyType == "ellipsoid_plain" and sim.insect.R_head == 0.125 and sim.insect.R_eye == 0.0625
    for t in (0.0, 0.5, 1.0):
        sim.create_mask(t)
        tag = "%06d" % int(t * 1000)
        for nm, f in (("mask", sim.mask), ("usx", sim.us[0]), ("usy", sim.us[1]), ("usz", sim.us[2])):
            ok, errs = fo.compare_key(fo.keyvalues(f, t, sim.p), fo.read_key(os.path.join(d, f"{nm}_{tag}.ref")))
            assert ok, (nm, t, errs)


def _plain_body_volume(tmp_path, k, b=0.12, L=1.0):
    """Mask volume of the ellipsoid_plain body alone (no wings) on a 1.2 x 0.4 x
    0.4 box at dx = 0.0125 / k."""
    ini = rewrite_ini(os.path.join(CONFIGS, "cfg3_insect_512.ini"), str(tmp_path), name=f"vol{k}.ini",
                      Resolution__nx=96 * k, Resolution__ny=32 * k, Resolution__nz=32 * k,
                      Geometry__xl=1.2, Geometry__yl=0.4, Geometry__zl=0.4,
                      Insects__x0="0.6 0.2 0.2", Insects__LeftWing="no", Insects__RightWing="no",
                      Insects__b_body=b, Insects__L_body=L, Insects__yawpitchroll_0="0.0 0.0 0.0")
    sim = fo.Simulation(ini)
    sim.insect = fo.Insect(sim.ini, sim.p, os.path.dirname(ini))
    sim.create_mask(0.0)
    assert set(np.unique(sim.color)) <= {0, 1}
    return float(sim.mask.sum()) * sim.g.dx * sim.g.dy * sim.g.dz


def test_ellipsoid_plain_volume_converges(tmp_path):
    """The cosine-smoothed body's volume tends to the ellipsoid's
    (4/3) pi a b^2 with a = L_body/2, b = b_body as dx halves."""
    exact = 4.0 / 3.0 * math.pi * 0.5 * 0.12 ** 2
    errs = [abs(_plain_body_volume(tmp_path, k) - exact) / exact for k in (1, 2, 4)]
    print("ellipsoid_plain volume rel errors at dx = 0.0125, 0.00625, 0.003125:", errs)
    # second order: each halving of dx cuts the error about four times
    assert errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0
    assert errs[2] < 5.0e-4


# ---------------------------------------------------------------- sinusoidal mean flow


def _sinusoidal_sim(tmp_path, perturb):
    ini = rewrite_ini(os.path.join(GOLDEN, "vortex_ring1_AB2", "vortexring1_AB2.ini"), str(tmp_path),
                      name=f"sin{int(perturb)}.ini",
                      Resolution__nx=16, Resolution__ny=16, Resolution__nz=16,
                      Time__dt_fixed="1.0e-2", Time__iTimeMethodFluid="AB2",
                      InitialCondition__inicond="VortexRing" if perturb else "quiescent",
                      MeanFlow__iMeanFlow_x="sinusoidal", MeanFlow__iMeanFlow_y="free",
                      MeanFlow__iMeanFlow_z="sinusoidal", MeanFlow__ux=0.7, MeanFlow__uz=-0.3,
                      MeanFlow__umean_freq=2.5, Saving__tintegral=100.0, Saving__tsave=100.0)
    sim = fo.Simulation(ini)
    sim.init_fields()
    return sim


@pytest.mark.parametrize("perturb", [False, True])
def test_sinusoidal_mean_flow_zero_mode(tmp_path, perturb):
    """fluid_time_step.f90:122-125 sets the zero mode after each full step with
    the time the step started from (time_step.f90:57-80 advances time after
    FluidTimestep): after the step t_{n-1} -> t_n the x and z zero modes are
    U sin(2 pi f t_{n-1}) -- distinguishable from t_n since f dt = 0.025 --
    and the free y component keeps what the step left.  From rest with no
    obstacle the flow stays uniform (u x w = 0 for a uniform field), so every
    other mode stays exactly zero."""
    sim = _sinusoidal_sim(tmp_path, perturb)
    f = 2.5
    for n in range(6):
        t_start = sim.time
        sim.step()
        assert sim.time == pytest.approx(t_start + 1.0e-2, rel=1e-14)
        assert sim.uk[0, 0, 0, 0] == pytest.approx(0.7 * math.sin(2 * math.pi * f * t_start), abs=1e-15)
        assert sim.uk[2, 0, 0, 0] == pytest.approx(-0.3 * math.sin(2 * math.pi * f * t_start), abs=1e-15)
        if n > 0:
            assert abs(sim.uk[0, 0, 0, 0] - 0.7 * math.sin(2 * math.pi * f * sim.time)) > 1e-3
        if not perturb:
            rest = sim.uk.copy()
            rest[:, 0, 0, 0] = 0.0
            assert not rest.any()
            assert sim.uk[1, 0, 0, 0] == 0.0


# ------------------------------------------------- adjoint's smooth drawing
def _cfg4_draw(tmp_path, name, t, smooth, **over):
    """One drawing of the cfg4 insect (ellipsoid_plain body, two TwoEllipses
    wings, from_file kinematics) at 40^3 and time t."""
    ini = rewrite_ini(os.path.join(CONFIGS, "cfg4_insect_adjoint_128.ini"), str(tmp_path), name=name,
                      Resolution__nx=40, Resolution__ny=40, Resolution__nz=40, **over)
    sim = fo.Simulation(ini)
    sim.init_fields()
    sim.insect.smooth_mode = smooth
    sim.create_mask(t)
    return sim


@pytest.mark.parametrize("t", [0.0, 0.0375, 0.21])
def test_smooth_drawing_composes_single_parts(tmp_path, t):
    """The differentiable drawing against the reference's own drawing of each
    part alone.  Its mask is the largest part value, where each wing's value
    is the reference's times the root factor steps(-y, 0) in [1/2, 1] (1 off
    the root slab); its us is sum_w m_w us_w / (m_body + sum_w m_w) with us_w
    the reference's single-wing velocity (the body is tethered at rest, so
    the body-motion transform is the identity here).  A wrong weight, a
    missing body term or a taper on the wrong side fails one of these."""
    full = _cfg4_draw(tmp_path, "full.ini", t, True)
    body = _cfg4_draw(tmp_path, "body.ini", t, False, Insects__LeftWing="no", Insects__RightWing="no")
    right = _cfg4_draw(tmp_path, "right.ini", t, False, Insects__LeftWing="no", Insects__BodyType="nobody")
    left = _cfg4_draw(tmp_path, "left.ini", t, False, Insects__RightWing="no", Insects__BodyType="nobody")
    right_s = _cfg4_draw(tmp_path, "right_s.ini", t, True, Insects__LeftWing="no", Insects__BodyType="nobody")
    left_s = _cfg4_draw(tmp_path, "left_s.ini", t, True, Insects__RightWing="no", Insects__BodyType="nobody")
    ins = full.insect
    assert np.all(ins.st["M_body"] == np.eye(3)) and not np.any(ins.st["vc"]) and not np.any(ins.st["rot_body_b"])
    mb = body.mask
    for std, sm in ((right, right_s), (left, left_s)):
        # root factor: equal off the root slab, between 1/2 and 1 on it, and
        # the slab is thin (about one grid spacing of the span)
        r = np.where(std.mask > 0, sm.mask / np.where(std.mask > 0, std.mask, 1.0), 1.0)
        assert np.all(sm.mask[std.mask == 0] == 0)
        assert r.min() >= 0.5 - 1e-12 and r.max() <= 1.0 + 1e-12
        changed = (r < 1.0) & (std.mask > 0)
        assert changed.any() and changed.sum() < 0.2 * (std.mask > 0).sum()
        # a wing drawn alone moves its points with the reference's velocity
        # (to rounding: (m v) / m)
        on = sm.mask > 0
        np.testing.assert_allclose(sm.us[:, on], std.us[:, on], rtol=1e-14, atol=1e-15)
    mr, ml = right_s.mask, left_s.mask
    np.testing.assert_array_equal(full.mask, np.maximum(mb, np.maximum(mr, ml)))
    wsum = mr + ml
    on = wsum > 0
    assert on.any() and (on & (mb > 0)).any()
    want = np.zeros_like(full.us)
    for i in range(3):
        want[i][on] = (mr[on] * right.us[i][on] + ml[on] * left.us[i][on]) / (mb[on] + wsum[on])
    np.testing.assert_allclose(full.us, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # where no wing reaches, us is the tethered body's: zero
    assert not full.us[:, ~on].any()


def test_adjoint_objective_is_mean_of_cal_drag_lift(tmp_path):
    """adjoint_mean_lift with the file's own controls equals the mean over
    the steps of cal_drag's z force taken after each step by a plain stepping
    loop with the same differentiable drawing (the objective adds nothing to
    the fixture-pinned cal_drag and steppers), and the controls reach the
    kinematics: moving one coefficient moves J."""
    ini = rewrite_ini(os.path.join(CONFIGS, "cfg4_insect_adjoint_128.ini"), str(tmp_path),
                      Resolution__nx=24, Resolution__ny=24, Resolution__nz=24,
                      Time__dt_fixed="4.0e-4", Penalization__eps="1.0e-2")
    sim = fo.Simulation(ini)
    sim.init_fields()
    k = sim.insect.kine_r
    th = np.array([v for ang in ("phi", "alpha", "theta") for v in (k.ai[ang] + k.bi[ang])])
    assert th.size == 12
    J = fo.Simulation(ini).adjoint_mean_lift(th, 3)
    ref = fo.Simulation(ini)
    ref.init_fields()
    ref.insect.smooth_mode = True
    ref.create_mask(ref.time)
    fz = []
    for _ in range(3):
        ref.step()
        ref.cal_drag(ref.time, fo.ifft3(ref.g, ref.uk))
        fz.append(ref.last_integrals["force"][2])
    assert J == pytest.approx(sum(fz) / 3, rel=1e-14)
    th2 = th.copy()
    th2[0] += 5.0  # phi amplitude of the first harmonic, degrees
    assert abs(fo.Simulation(ini).adjoint_mean_lift(th2, 3) - J) > 1e-6 * abs(J)
