"""torch.autograd binding of the discrete adjoint: the mean lift of a tethered
flapping insect as a differentiable function of its wing kinematics' Fourier
coefficients.

    sim = flusi_amd.Flusi("cfg4_insect_adjoint_128.ini", flusi_amd.FP64)
    theta = torch.tensor(sim.adjoint_controls(), dtype=torch.float64, requires_grad=True)
    J = mean_lift(theta, sim, nsteps=400, every=20)
    J.backward()          # theta.grad = dJ/dtheta from the adjoint sweep

The forward pass runs the solver through the C ABI (flusi_adjoint_forward,
which keeps a checkpoint every `every` steps); the backward pass is one
adjoint sweep (flusi_adjoint_backward).  PyTorch is only the calling
convention here: the controls and gradient cross the boundary as host
float64 arrays.
"""
from __future__ import annotations

import itertools

import numpy as np
import torch

_run_ids = itertools.count(1)


class MeanLift(torch.autograd.Function):
    """J(theta) = (1/K) sum_n F_z(t_n) over K RK4 steps (see flusi_adjoint_forward)."""

    @staticmethod
    def forward(ctx, theta, sim, nsteps, every):
        th = theta.detach().to("cpu", torch.float64).numpy().copy()
        J = sim.adjoint_forward(th, int(nsteps), int(every))
        run = next(_run_ids)
        sim._adjoint_run = run
        ctx.sim, ctx.run, ctx.theta = sim, run, th
        ctx.nsteps, ctx.every = int(nsteps), int(every)
        return torch.tensor(J, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, gJ):
        sim = ctx.sim
        if getattr(sim, "_adjoint_run", None) != ctx.run:
            # another forward ran on this solver since: restore our checkpoints
            sim.adjoint_forward(ctx.theta, ctx.nsteps, ctx.every)
            sim._adjoint_run = ctx.run
        g = torch.from_numpy(np.asarray(sim.adjoint_backward(), dtype=np.float64))
        return (gJ.to("cpu", torch.float64) * g).to(gJ.dtype).to(gJ.device), None, None, None


def mean_lift(theta: torch.Tensor, sim, nsteps: int, every: int = 20) -> torch.Tensor:
    return MeanLift.apply(theta, sim, nsteps, every)
