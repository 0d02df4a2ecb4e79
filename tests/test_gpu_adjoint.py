"""The discrete adjoint on the GPU: d(mean lift)/d(wing kinematics
coefficients) against central finite differences of the solver's own mean
lift, the objective and two gradient components against the oracle's
restatement of the same objective, checkpoint-interval independence,
FP32-mixed against FP64, and the torch.autograd binding.  The reference has
no adjoint; the oracle's objective is built from its fixture-pinned steppers
and cal_drag and the differentiable drawing pinned in
tests/test_oracle_pinning.py."""
import os

import numpy as np
import pytest

import flusi_amd
from oracle import flusi_oracle as fo
from tests._util import CONFIGS, rewrite_ini

pytestmark = pytest.mark.gpu

K, EVERY = 6, 2


def _ini(tmp_path, n=32):
    return rewrite_ini(os.path.join(CONFIGS, "cfg4_insect_adjoint_128.ini"), str(tmp_path),
                       Resolution__nx=n, Resolution__ny=n, Resolution__nz=n,
                       Time__dt_fixed="4.0e-4", Penalization__eps="1.0e-2")


def test_adjoint_gradient_matches_central_differences(tmp_path):
    s = flusi_amd.Flusi(_ini(tmp_path), flusi_amd.FP64)
    th0 = s.adjoint_controls()
    assert th0.size == 12
    J0 = s.adjoint_forward(th0, K, EVERY)
    g = s.adjoint_backward()
    h = 1.0e-4
    fd = np.zeros_like(g)
    for i in range(th0.size):
        tp, tm = th0.copy(), th0.copy()
        tp[i] += h
        tm[i] -= h
        fd[i] = (s.adjoint_forward(tp, K, EVERY) - s.adjoint_forward(tm, K, EVERY)) / (2 * h)
    print("J", J0, "\nadjoint", g, "\nfd     ", fd)
    assert np.max(np.abs(fd)) > 0.0
    assert np.max(np.abs(g - fd)) <= 1e-6 * np.max(np.abs(fd)), np.abs(g - fd) / np.max(np.abs(fd))
    s.close()


def test_adjoint_objective_and_gradient_match_oracle(tmp_path):
    """J of the GPU's forward sweep against the oracle's adjoint_mean_lift
    with the same controls, and the adjoint gradient's two largest
    components against central differences of the oracle's J."""
    ini = _ini(tmp_path)
    s = flusi_amd.Flusi(ini, flusi_amd.FP64)
    th = s.adjoint_controls()
    J = s.adjoint_forward(th, K, EVERY)
    g = s.adjoint_backward()
    s.close()
    Jo = fo.Simulation(ini).adjoint_mean_lift(th, K)
    print(f"J gpu {J!r} oracle {Jo!r} rel {abs(J - Jo) / abs(Jo):.3e}", flush=True)
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    h = 1.0e-4
    for i in np.argsort(-np.abs(g))[:2]:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (fo.Simulation(ini).adjoint_mean_lift(tp, K) - fo.Simulation(ini).adjoint_mean_lift(tm, K)) / (2 * h)
        print(f"control {i}: adjoint {g[i]!r} oracle central difference {fd!r}", flush=True)
        assert abs(g[i] - fd) <= 1e-6 * np.max(np.abs(g))


def test_adjoint_checkpoint_interval_and_precision(tmp_path):
    s = flusi_amd.Flusi(_ini(tmp_path), flusi_amd.FP64)
    th = s.adjoint_controls()
    grads = []
    for every in (1, 2, 4, K):
        J = s.adjoint_forward(th, K, every)
        grads.append(s.adjoint_backward())
    for g in grads[1:]:
        assert np.allclose(g, grads[0], rtol=1e-12, atol=1e-12 * np.max(np.abs(grads[0])))
    s32 = flusi_amd.Flusi(_ini(tmp_path), flusi_amd.FP32_MIXED)
    J32 = s32.adjoint_forward(th, K, EVERY)
    g32 = s32.adjoint_backward()
    assert abs(J32 - J) <= 1e-4 * abs(J)
    assert np.linalg.norm(g32 - grads[0]) <= 1e-3 * np.linalg.norm(grads[0])
    s.close()
    s32.close()


def test_adjoint_torch_autograd(tmp_path):
    torch = pytest.importorskip("torch")
    from flusi_amd.adjoint_torch import mean_lift
    s = flusi_amd.Flusi(_ini(tmp_path), flusi_amd.FP64)
    th = torch.tensor(s.adjoint_controls(), dtype=torch.float64, requires_grad=True)
    J = mean_lift(th, s, K, EVERY)
    J2 = mean_lift(th.detach() * 1.0001, s, K, EVERY)  # another run on the same solver in between
    (2.0 * J).backward()
    ref_J = s.adjoint_forward(th.detach().numpy(), K, EVERY)
    ref_g = s.adjoint_backward()
    assert J.detach().item() == ref_J and J2.detach().item() != ref_J
    assert np.allclose(th.grad.numpy(), 2.0 * ref_g, rtol=1e-12, atol=0.0)
    s.close()
