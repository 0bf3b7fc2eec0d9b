"""Drop-in for the reference module ``pack_cuda`` (core/pack_cuda.py) on MI355X.

Same entry points, argument order and meaning; arrays are torch tensors on the GPU
instead of CuPy arrays, and ``stream`` is a torch.cuda.Stream (None = current stream).
The work is done by hand-written HIP kernels in libpacking.so (pk_overlap,
pk_boundary); there is no other code path.
"""
from __future__ import annotations

import torch

from . import _native as N
from .context import device_geometry, get_geometry, set_geometry  # noqa: F401


def _check_xyt(name, t, dtype=None):
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    if t.ndim != 3 or t.shape[2] != 3:
        raise ValueError(f"{name} must be shape (n_ensembles, n_trees, 3), got {tuple(t.shape)}")
    if not t.is_cuda:
        raise ValueError(f"{name} must be on the GPU")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be C-contiguous")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name} must have dtype {dtype}, got {t.dtype}")


def _check_out(name, t, shape, dtype):
    if t is None:
        return
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if t.dtype != dtype:
        raise ValueError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if not t.is_contiguous() or not t.is_cuda:
        raise ValueError(f"{name} must be a C-contiguous GPU tensor")


def overlap_multi_ensemble(xyt1, xyt2, use_separation, out_cost=None, out_grads=None,
                           crystal_axes=None, only_self_interactions=False, N_periodic=2,
                           stream=None, block=0, stats=None):
    """Total overlap (area, or sum of squared SAT separations) per ensemble.

    Reference: core/pack_cuda.py:1040-1183.  Every reference call site passes
    ``xyt1 is xyt2`` (core/pack_cost.py:803-817,1113-1116); that is the supported form.
    ``out_cost`` (n_ens,) is required; ``out_grads`` (n_ens, N, 3) is optional.
    ``crystal_axes`` (n_ens, 4) enables periodic images (|cell| <= N_periodic).
    """
    if not isinstance(use_separation, bool):
        raise ValueError("use_separation must be a boolean")
    _check_xyt("xyt1", xyt1)
    if xyt2 is not xyt1 and (xyt2.data_ptr() != xyt1.data_ptr() or xyt2.shape != xyt1.shape):
        raise ValueError("xyt1 and xyt2 must be the same tensor (as at every reference call site)")
    n_ens, n_trees = xyt1.shape[0], xyt1.shape[1]
    if out_cost is None:
        raise ValueError("out_cost must be provided")
    dt = xyt1.dtype
    _check_out("out_cost", out_cost, (n_ens,), dt)
    _check_out("out_grads", out_grads, (n_ens, n_trees, 3), dt)
    if crystal_axes is not None:
        crystal_axes = crystal_axes.reshape(n_ens, 4) if crystal_axes.numel() == 4 * n_ens else crystal_axes
        _check_out("crystal_axes", crystal_axes, (n_ens, 4), dt)
    if n_ens == 0:
        return
    dg = device_geometry(xyt1.device)
    L = N.lib()
    rc = L.pk_overlap(dg.handle, N.dtype_code(dt), N.ptr(xyt1), None, n_ens, n_trees,
                      N.PK_MODE_SEP if use_separation else N.PK_MODE_AREA, N.ptr(crystal_axes),
                      int(N_periodic), int(bool(only_self_interactions)), N.ptr(out_cost),
                      N.ptr(out_grads), int(block), N.ptr(stats), N.stream_handle(stream))
    N.check(rc, "pk_overlap")


def overlap_ragged(xyt_flat, inst_off, n_max, use_separation, out_cost, out_grads=None,
                   stream=None, block=0, stats=None):
    """Ragged batch: trees of all instances back to back (T, 3), int32 offsets (n_ens + 1)."""
    n_ens = inst_off.numel() - 1
    if n_ens <= 0:
        return
    dg = device_geometry(xyt_flat.device)
    rc = N.lib().pk_overlap(dg.handle, N.dtype_code(xyt_flat.dtype), N.ptr(xyt_flat), N.ptr(inst_off),
                            n_ens, int(n_max), N.PK_MODE_SEP if use_separation else N.PK_MODE_AREA,
                            None, 0, 0, N.ptr(out_cost), N.ptr(out_grads), int(block), N.ptr(stats),
                            N.stream_handle(stream))
    N.check(rc, "pk_overlap")


def boundary_distance_multi_ensemble(xyt, h, out_cost=None, out_grads=None, out_grad_h=None,
                                     stream=None):
    """Boundary term per ensemble (core/pack_cuda.py:1186-1295); h is (n_ens, 3)."""
    _check_xyt("xyt", xyt)
    n_ens, n_trees = xyt.shape[0], xyt.shape[1]
    dt = xyt.dtype
    if out_cost is None:
        raise ValueError("out_cost must be provided")
    _check_out("h", h, (n_ens, 3), dt)
    _check_out("out_cost", out_cost, (n_ens,), dt)
    _check_out("out_grads", out_grads, (n_ens, n_trees, 3), dt)
    _check_out("out_grad_h", out_grad_h, (n_ens, 3), dt)
    if n_ens == 0:
        return
    dg = device_geometry(xyt.device)
    rc = N.lib().pk_boundary(dg.handle, N.dtype_code(dt), N.ptr(xyt), None, n_ens, n_trees, N.ptr(h),
                             N.ptr(out_cost), N.ptr(out_grads), N.ptr(out_grad_h),
                             N.stream_handle(stream))
    N.check(rc, "pk_boundary")
