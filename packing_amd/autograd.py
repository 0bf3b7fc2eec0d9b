"""The overlap cost as a torch.autograd.Function.

Forward runs the cost-only kernel (no gradient accumulation); backward runs the
cost+gradient kernel and scales each instance's gradient by the incoming grad_output.
Both are the hand-written HIP kernels behind pk_overlap / pk_lut_overlap; the gradient is
the analytic one (shape derivative of the clipped area, SAT depth derivative, or the
trilinear table derivative).
"""
from __future__ import annotations

import torch

from . import pack_cuda, pack_cuda_lut


class OverlapCost(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyt, use_separation=False, lut=None):
        xyt = xyt.contiguous()
        cost = torch.empty(xyt.shape[0], dtype=xyt.dtype, device=xyt.device)
        if lut is not None:
            pack_cuda_lut.overlap_multi_ensemble(xyt, cost, lut)
        else:
            pack_cuda.overlap_multi_ensemble(xyt, xyt, bool(use_separation), out_cost=cost)
        ctx.save_for_backward(xyt)
        ctx.use_separation = bool(use_separation)
        ctx.lut = lut
        return cost

    @staticmethod
    def backward(ctx, grad_out):
        (xyt,) = ctx.saved_tensors
        cost = torch.empty(xyt.shape[0], dtype=xyt.dtype, device=xyt.device)
        g = torch.empty_like(xyt)
        if ctx.lut is not None:
            pack_cuda_lut.overlap_multi_ensemble(xyt, cost, ctx.lut, out_grads=g)
        else:
            pack_cuda.overlap_multi_ensemble(xyt, xyt, ctx.use_separation, out_cost=cost, out_grads=g)
        return g * grad_out.reshape(-1, 1, 1).to(g.dtype), None, None


def overlap_cost(xyt: torch.Tensor, use_separation: bool = False, lut=None) -> torch.Tensor:
    """Differentiable (n_ens,) overlap cost of poses (n_ens, N, 3)."""
    return OverlapCost.apply(xyt, use_separation, lut)
