"""The overlap cost as a torch.autograd.Function.

Forward runs the cost kernel once: with the gradient accumulated in the same launch when
the poses require a gradient (the analytic gradient is saved for backward), cost only
otherwise.  Backward scales each instance's saved gradient by the incoming grad_output;
it launches nothing.  The kernels are the hand-written HIP kernels behind pk_overlap /
pk_lut_overlap; the gradient is the analytic one (shape derivative of the clipped area,
SAT depth derivative, or the trilinear table derivative).
"""
from __future__ import annotations

import torch

from . import pack_cuda, pack_cuda_lut


class OverlapCost(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyt, use_separation=False, lut=None):
        xyt = xyt.contiguous()
        cost = torch.empty(xyt.shape[0], dtype=xyt.dtype, device=xyt.device)
        g = torch.empty_like(xyt) if ctx.needs_input_grad[0] else None
        if lut is not None:
            pack_cuda_lut.overlap_multi_ensemble(xyt, cost, lut, out_grads=g)
        else:
            pack_cuda.overlap_multi_ensemble(xyt, xyt, bool(use_separation), out_cost=cost, out_grads=g)
        if g is not None:
            ctx.save_for_backward(g)
        return cost

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return g * grad_out.reshape(-1, 1, 1).to(g.dtype), None, None


def overlap_cost(xyt: torch.Tensor, use_separation: bool = False, lut=None) -> torch.Tensor:
    """Differentiable (n_ens,) overlap cost of poses (n_ens, N, 3)."""
    return OverlapCost.apply(xyt, use_separation, lut)
