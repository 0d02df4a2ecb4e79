"""Batched ensembles and the kernel-variant autotuner on the GPU.

The three ways of advancing an ensemble (every member on its own, one launch
per pass and batch, and the batched step replayed from a captured hipGraph)
run the same kernels on the same data, so their fields must agree bit for
bit, step after step, with the host's own eager work between graph replays.
A plate member (rectangular wing, no body) and a sphere member are also
checked against the oracle in FP64.
"""
import os

import numpy as np
import pytest

import flusi_amd
from flusi_amd.ensemble_configs import write_ensemble
from oracle import flusi_oracle as fo
from tests._util import rel_l2

pytestmark = pytest.mark.gpu


def _paths(tmp_path, sizes=(32, 48, 64), per_size=4):
    return write_ensemble(str(tmp_path / "ens"), sizes=sizes, per_size=per_size, dt=5.0e-4, nt=50)


def _ens(paths, prec):
    e = flusi_amd.Ensemble(paths, prec)
    e.init_fields()
    return e


@pytest.mark.parametrize("prec", [flusi_amd.FP32_MIXED, flusi_amd.FP64])
def test_ensemble_modes_agree_every_replay(tmp_path, prec):
    paths = _paths(tmp_path)
    ens = {m: _ens(paths, prec) for m in (flusi_amd.ENSEMBLE_MEMBERS, flusi_amd.ENSEMBLE_BATCHED,
                                          flusi_amd.ENSEMBLE_GRAPH)}
    # spheres and plates of each grid form their own batches
    b = ens[flusi_amd.ENSEMBLE_GRAPH].batches()
    assert sorted(b) == sorted([(2, n, n, n) for n in (32, 48, 64) for _ in range(2)])
    for step in range(6):  # step 0 eager, the capture at step 1, then replays
        for mode, e in ens.items():
            e.advance(1, mode)
        ref = ens[flusi_amd.ENSEMBLE_MEMBERS]
        for i in range(len(paths)):
            u0 = ref.member(i).get_uk()
            t0 = ref.member(i).get_time()
            for mode in (flusi_amd.ENSEMBLE_BATCHED, flusi_amd.ENSEMBLE_GRAPH):
                m = ens[mode].member(i)
                assert m.get_time()[:2] == t0[:2], (step, i, mode)
                u = m.get_uk()
                assert np.array_equal(u, u0), (step, i, mode, float(np.max(np.abs(u - u0))))
        # the host's own eager work between replays: integrals (which redraw
        # the mask at the current time) on a plate and a sphere member
        for e in ens.values():
            for i in (0, len(paths) - 1):
                e.member(i).write_integrals(e.member(i).get_time()[0])
    for e in ens.values():
        e.close()


def test_ensemble_members_match_standalone_and_oracle(tmp_path):
    paths = _paths(tmp_path, sizes=(32,), per_size=2)  # one sphere, one plate
    e = _ens(paths, flusi_amd.FP64)
    e.advance(3, flusi_amd.ENSEMBLE_GRAPH)
    for i, p in enumerate(paths):
        s = flusi_amd.Flusi(p, flusi_amd.FP64)
        s.init_fields()
        for _ in range(3):
            s.step()
        assert np.array_equal(s.get_uk(), e.member(i).get_uk()), p
        sim = fo.Simulation(p)
        sim.init_fields()
        for _ in range(3):
            sim.step()
        assert rel_l2(e.member(i).get_uk(), sim.uk) < 1e-10, p
        m, us, col = e.member(i).get_mask()
        assert np.max(np.abs(m - sim.mask)) < 1e-12
        assert np.max(np.abs(us - sim.us)) < 1e-10
        assert np.array_equal(col, sim.color.astype(np.int16))
        s.close()
    e.close()


def test_ensemble_autotune_then_graph(tmp_path):
    paths = _paths(tmp_path, sizes=(64,), per_size=4)
    e = _ens(paths, flusi_amd.FP32_MIXED)
    f = _ens(paths, flusi_amd.FP32_MIXED)
    res = e.autotune(reps=2)
    assert len(res) == e.nbatches
    for v, ms in res:
        assert len(v) == 6 and all(x in (0, 1, 2, 3, 4) for x in v) and ms > 0.0
    e.advance(4, flusi_amd.ENSEMBLE_GRAPH)
    f.advance(4, flusi_amd.ENSEMBLE_MEMBERS)
    for i in range(len(paths)):
        # another kernel variant rounds differently in FP32
        assert rel_l2(e.member(i).get_uk(), f.member(i).get_uk()) < 1e-5
    e.close()
    f.close()


def test_single_context_autotune(tmp_path):
    paths = _paths(tmp_path, sizes=(64,), per_size=2)
    s = flusi_amd.Flusi(paths[1], flusi_amd.FP32_MIXED)
    s.init_fields()
    r = flusi_amd.Flusi(paths[1], flusi_amd.FP32_MIXED)
    r.init_fields()
    v, ms = s.autotune(reps=2)
    assert len(v) == 6 and ms > 0.0
    v2, ms2 = s.autotune(reps=2)  # cached
    assert v2 == v and ms2 == ms
    for _ in range(2):
        s.step()
        r.step()
    assert rel_l2(s.get_uk(), r.get_uk()) < 1e-5
    s.close()
    r.close()
