"""FP64 zero-overlap check of final packings on the GPU (pk_feasible).

A packing is feasible when no two trees have intersecting interiors (touching is
allowed), the collision rule of the Kaggle metric (core/pack_metric.py:130-141).
"""
from __future__ import annotations

import torch

from . import _native as N
from .context import device_geometry


def overlapping_pairs(xyt: torch.Tensor, inst_off: torch.Tensor | None = None, n_max: int | None = None,
                      stream=None, geometry=None) -> torch.Tensor:
    """int32 count of overlapping tree pairs per instance.

    xyt: (n_ens, N, 3), or (T, 3) with int32 inst_off (n_ens + 1) for a ragged batch."""
    if inst_off is None:
        if xyt.ndim != 3 or xyt.shape[2] != 3:
            raise ValueError("xyt must be (n_ens, N, 3)")
        n_ens, n_trees = xyt.shape[0], xyt.shape[1]
    else:
        n_ens = inst_off.numel() - 1
        n_trees = int(n_max) if n_max is not None else int((inst_off[1:] - inst_off[:-1]).max())
    xyt = xyt.contiguous()
    out = torch.empty(n_ens, dtype=torch.int32, device=xyt.device)
    if n_ens == 0:
        return out
    dg = device_geometry(xyt.device, geometry)
    N.check(N.lib().pk_feasible(dg.handle, N.dtype_code(xyt.dtype), xyt.data_ptr(), N.ptr(inst_off),
                                n_ens, n_trees, out.data_ptr(), N.stream_handle(stream)), "pk_feasible")
    return out


def feasible(xyt, inst_off=None, n_max=None, stream=None, geometry=None) -> torch.Tensor:
    return overlapping_pairs(xyt, inst_off, n_max, stream, geometry) == 0
