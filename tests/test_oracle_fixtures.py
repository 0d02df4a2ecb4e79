"""The oracle against the reference's own regression fixtures (golden data
from tests/tests_data.tar.gz of the reference, copied under tests/golden/),
compared with the reference's tolerance rules (postprocessing/compare_key.f90,
compare_timeseries.f90).  CPU only."""
import glob
import os

import pytest

from oracle import flusi_oracle as fo
from tests._util import series_ok


def _ref_keys(d):
    """6-value key files of the fixture, indexed by (prefix, stored time)."""
    out = {}
    for f in sorted(glob.glob(os.path.join(d, "*.ref"))):
        vals = open(f).read().split()
        if len(vals) < 6:
            continue
        pre = os.path.basename(f).split("_")[0]
        out[(pre, round(float(vals[0]), 6))] = fo.read_key(f)
    return out


def _check_run(d, ini, series):
    sim = fo.Simulation(os.path.join(d, ini), workers=4).run(save_keys=True)
    mine = {(k.split("_")[0], round(v[0], 6)): v for k, v in sim.keys.items()}
    refs = _ref_keys(d)
    assert refs
    for key, ref in refs.items():
        assert key in mine, key
        ok, errs = fo.compare_key(mine[key], ref)
        assert ok, (key, errs)
    for ts in series:
        ok, worst, msg = series_ok(sim.series.get(ts, []), os.path.join(d, ts))
        assert ok, (ts, msg)
    return sim


def test_insects_mask_jerry(golden):
    """Jerry body + TwoEllipses wings, Drosophila_hovering_fry / flapper_dickinson
    kinematics, dry run at t=0, 0.5, 1.0 (reference tests/insects_mask_jerry.sh)."""
    d = os.path.join(golden, "insects_mask_jerry")
    sim = fo.Simulation(os.path.join(d, "jerry.ini"))
    sim.insect = fo.Insect(sim.ini, sim.p, d)
    rows = []
    for t in (0.0, 0.5, 1.0):
        sim.create_mask(t)
        rows.append(sim.insect.kinematics_row(t))
        tag = "%06d" % int(t * 1000)
        for nm, f in (("mask", sim.mask), ("usx", sim.us[0]), ("usy", sim.us[1]), ("usz", sim.us[2])):
            ok, errs = fo.compare_key(fo.keyvalues(f, t, sim.p), fo.read_key(os.path.join(d, f"{nm}_{tag}.ref")))
            assert ok, (nm, t, errs)
    ok, worst = fo.compare_timeseries(rows, fo.read_timeseries(os.path.join(d, "kinematics.dry-run.t")))
    assert ok, worst


def test_insects_mask_BB_tethered(golden):
    """Bumblebee body with legs, antennae and proboscis, Fourier bumblebee
    wings, kinematics from file; dry run at t=0..1 (reference
    tests/insects_mask_BB_tethered.sh)."""
    d = os.path.join(golden, "insects_mask_BB_tethered")
    sim = fo.Simulation(os.path.join(d, "insect_test.ini"))
    sim.insect = fo.Insect(sim.ini, sim.p, d)
    rows = []
    for t in (0.0, 0.25, 0.5, 0.75, 1.0):
        sim.create_mask(t)
        rows.append(sim.insect.kinematics_row(t))
        tag = "%06d" % int(round(t * 1000))
        for nm, f in (("mask", sim.mask), ("usx", sim.us[0]), ("usy", sim.us[1]), ("usz", sim.us[2])):
            ok, errs = fo.compare_key(fo.keyvalues(f, t, sim.p), fo.read_key(os.path.join(d, f"{nm}_{tag}.ref")))
            assert ok, (nm, t, errs)
            assert max(errs) < 1e-8, (nm, t, errs)  # the keys agree to their printed digits
    ok, worst = fo.compare_timeseries(rows, fo.read_timeseries(os.path.join(d, "kinematics.dry-run.t")))
    assert ok, worst


@pytest.mark.slow
def test_insects_flow_BB_tethered_RK4(golden):
    """64^3 RK4 with a moving bumblebee, dynamic mean flow, CFL-adapted
    intelligent dt and a periodic insect: forces, per-part forces,
    kinematics, energy and dt series plus mask/us/|w| keys
    (insects_flow_BB_tethered_RK4.sh:80-100)."""
    _check_run(os.path.join(golden, "insects_flow_BB_tethered_RK4"), "insect_test.ini",
               ["forces.t", "forces_part1.t", "forces_part2.t", "forces_part3.t", "kinematics.t", "energy.t",
                "dt.t"])


@pytest.mark.slow
@pytest.mark.parametrize("case", ["insects_flow_BB_tethered_AB2", "insects_flow_BB_tethered_AB2_periodic"])
def test_insects_flow_BB_tethered_AB2(golden, case):
    """The bumblebee fixture with AB2 and CFL 0.1, and its variant with the
    insect near the corner of the periodic box (parts wrap around):
    forces, per-part forces, kinematics, energy and dt series plus
    mask/us/|w| keys (insects_flow_BB_tethered_AB2*.sh:80-100)."""
    _check_run(os.path.join(golden, case), "insect_test.ini",
               ["forces.t", "forces_part1.t", "forces_part2.t", "forces_part3.t", "kinematics.t", "energy.t",
                "dt.t"])


def test_vortex_ring4_sponge(golden):
    """AB2 with the cavity vorticity sponge: keys incl. pressure at t=0, 0.5
    (vortexring4_sponge.sh:21-23)."""
    _check_run(os.path.join(golden, "vortex_ring4_sponge"), "vortexring4_sponge.ini", ["dt.t", "energy.t",
                                                                                     "meanflow.t"])


def test_vortex_ring1_AB2(golden):
    _check_run(os.path.join(golden, "vortex_ring1_AB2"), "vortexring1_AB2.ini", ["dt.t", "energy.t", "meanflow.t"])


def test_vortex_ring2_RK2(golden):
    _check_run(os.path.join(golden, "vortex_ring2_RK2"), "vortexring2_RK2.ini", ["dt.t", "energy.t", "meanflow.t"])


@pytest.mark.slow
def test_sphere_sponge_RK2(golden):
    """48^3 RK2 + cavity vorticity sponge, forces.t compared (sphere_sponge_RK2.sh:75-88)."""
    _check_run(os.path.join(golden, "sphere_sponge_RK2"), "sphere_sponge_RK2.ini",
               ["forces.t", "dt.t", "energy.t", "meanflow.t"])


@pytest.mark.slow
def test_sphere_AB2(golden):
    _check_run(os.path.join(golden, "sphere"), "testing_sphere.ini", ["forces.t", "dt.t", "energy.t"])
