"""torch.autograd.Function over the GPU loop and its hand-written adjoint.

    curve = MacroStressCurve.apply(params, job, n_increments, phase)

params: float64 tensor (tau0, tau1, theta0, theta1, n) of crystalline phase
`phase`; returns the macroscopic stress after each of n increments
(n x 6, components 11 22 33 23 13 12).  The forward resets the job to its
post-setup state with these parameters and records checkpoints every
`checkpoint_every` increments; backward runs evpgpu_adjoint_run.  The device
does the work: torch only carries the five numbers and the n x 6 curve.
"""
import numpy as np
import torch

from . import lib


class MacroStressCurve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, job, n_increments, phase=0, checkpoint_every=5, cg_tol=1e-10, cg_max=4000):
        q = params.detach().to(torch.float64).cpu().numpy()
        job.set_params(phase, q)
        curve, stats = job.adjoint_record(n_increments, checkpoint_every)
        ctx.job, ctx.phase, ctx.cg = job, phase, (cg_tol, cg_max)
        ctx.stats = stats
        return torch.from_numpy(curve)

    @staticmethod
    def backward(ctx, grad_curve):
        g, _ = ctx.job.adjoint_run(grad_curve.detach().to(torch.float64).cpu().numpy(), *ctx.cg)
        return torch.from_numpy(np.ascontiguousarray(g[ctx.phase])), None, None, None, None, None, None
