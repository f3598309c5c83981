"""Differentiable BVE rollouts: the hand-written adjoint of the RK4 step
behind ``torch.autograd.Function`` (no reference counterpart).

``rollout(model, zeta0, dt, nsteps)`` advances ``zeta0`` by ``nsteps`` fused
RK4 steps and records the input state of every RK4 stage (stage
checkpointing: ``nsteps * 4`` states, 468 KB each at L = 170, kept in device
memory).  Its backward runs the reverse sweep of the library: per stage one
synthesis of the forward fields and the Jacobian cotangent, one Fourier
kernel with the transposed pointwise Jacobian, and one analysis with the
transposed spectral algebra and RK4 bookkeeping.

Gradients follow torch's convention for complex tensors (the real inner
product ``sum Re(conj(x) y)`` of the coefficient arrays).
"""
from __future__ import annotations

import torch

__all__ = ["BVERollout", "rollout", "sphere_misfit"]


class BVERollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, zeta0, model, dt, nsteps, forcing):
        y, ck = model.rk4_record(zeta0.detach(), dt, nsteps, forcing)
        ctx.model = model
        ctx.dt = float(dt)
        ctx.nsteps = int(nsteps)
        ctx.save_for_backward(ck)
        return y

    @staticmethod
    def backward(ctx, grad_y):
        (ck,) = ctx.saved_tensors
        g = ctx.model.rk4_adjoint(ck, ctx.dt, ctx.nsteps, grad_y.contiguous())
        return g, None, None, None, None


def rollout(model, zeta0: torch.Tensor, dt: float, nsteps: int, forcing=None) -> torch.Tensor:
    """``nsteps`` RK4 steps of a :class:`~aeolus_amd.physics.BarotropicVorticity`
    (or ``BarotropicBatch``) model, differentiable with respect to ``zeta0``."""
    return BVERollout.apply(zeta0, model, dt, nsteps, forcing)


def sphere_misfit(z: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """``1/2 ||z - target||^2`` in the sphere L2 product of real fields:
    ``sum_l [Re(d_l0)^2 + 2 sum_{0<m<=l} |d_lm|^2] / 2`` over the last two
    (l, m) axes, summed over any leading axes."""
    d = z - target
    n = d.shape[-1]
    l = torch.arange(n, device=d.device)[:, None]
    m = torch.arange(n, device=d.device)[None, :]
    valid = (m <= l).to(torch.float64)
    wr = valid * torch.where(m == 0, 1.0, 2.0).to(torch.float64)
    wi = valid * (m > 0).to(torch.float64) * 2.0
    return 0.5 * torch.sum(wr * d.real ** 2 + wi * d.imag ** 2)
