"""Differentiable effective diffusivity.

``effective_diffusivity(k)`` solves the per-voxel diffusivity field ``k``
([B,X,Y,Z] or [X,Y,Z] float32 on a HIP device) to convergence with the
gfx950 sweep kernels and returns D_eff per batch element.  Its backward pass
is one kernel over the converged field (libtauhip ``tf_grad_field``): the
dissipation identity D_rel = (X/(Y Z)) * W, with W minimised by the field,
makes dD_rel/dk the explicit derivative of W, so no iteration is stored or
replayed and no adjoint system is solved.
"""
from __future__ import annotations

import torch

from .solver import FieldSolver


class _EffectiveDiffusivity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, k, D_0, conv_crit, iter_limit, half_sweeps):
        s = FieldSolver(k.detach(), D_scaling=D_0, device=k.device)
        if half_sweeps:
            s.half_sweeps_per_pass = half_sweeps
        s.solve(verbose=False, conv_crit=conv_crit, iter_limit=iter_limit)
        ctx.solver = s
        ctx.shape = k.shape
        return torch.as_tensor(s.D_rel * D_0, dtype=torch.float64, device=k.device)

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.solver.grad_D_eff().to(torch.float64)
        g = g * grad_out.view(-1, 1, 1, 1)
        ctx.solver = None
        return g.to(torch.float32).reshape(ctx.shape), None, None, None, None


def effective_diffusivity(k: torch.Tensor, D_0: float = 1.0, conv_crit: float = 1e-3,
                          iter_limit: int = 10000, half_sweeps: int | None = None) -> torch.Tensor:
    """D_eff (float64 tensor [B]) of the voxel diffusivity field k, differentiable in k.

    The backward pass differentiates the fixed point, so its accuracy follows
    the forward's convergence: the default criterion is 1e-3, ten times
    tighter than the solvers' default 1e-2 (at 1e-2 the per-voxel gradient
    of the 24^3 sigmoid-k test volumes is off by up to a few percent of its
    largest value; tests/test_gpu_fullsize.py measures both)."""
    if k.dim() == 3:
        k = k.unsqueeze(0)
    return _EffectiveDiffusivity.apply(k, D_0, conv_crit, iter_limit, half_sweeps)
