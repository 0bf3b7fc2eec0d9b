"""Drop-in for the reference module ``lbfgs_torch_parallel`` on MI355X.

``lbfgs(func, x0, ...)`` has the reference's signature and semantics
(core/lbfgs_torch_parallel.py:411-729, line_search_fn=None): M independent systems,
two-loop recursion, first step lr * min(1, 1/sum|g|), max-step clamp, history kept
only from iteration 2 when y.s > 1e-10, the reference's convergence tests, inactive
systems that keep stepping while any system is active, and x0 updated in place.  The
optimizer arithmetic runs in HIP kernels (pk_lbfgs_*); only ``func`` runs between them.
The host never waits on the GPU except every ``sync_every`` iterations, to leave the
loop early once the device has marked the run as stopped.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import torch
from torch import Tensor

from . import _native as N

__all__ = ["lbfgs"]


def lbfgs(func, x0: Tensor, lr: float = 1.0, max_iter: int = 20, max_eval: Optional[int] = None,
          tolerance_grad: float = 1e-7, tolerance_change: float = 1e-9, tolerance_rel_change: float = 0.0,
          stop_on_cost_increase: bool = False, history_size: int = 100,
          line_search_fn: Optional[str] = None, max_step: Optional[float] = 0.5,
          sync_every: int = 32, stream=None):
    if line_search_fn is not None:
        raise NotImplementedError("line_search_fn: only None (no line search) is implemented on the GPU")
    if not x0.is_cuda:
        raise ValueError("x0 must be a GPU tensor")
    if x0.dtype not in (torch.float32, torch.float64):
        raise ValueError("x0 must be float32 or float64")
    if max_eval is None:
        max_eval = max_iter * 5 // 4
    M, n = x0.shape
    dev, dt = x0.device, x0.dtype
    x = x0.detach().clone()
    loss, grad = func(x, True)
    L = torch.empty(M, dtype=dt, device=dev)
    G = torch.empty(M, n, dtype=dt, device=dev)
    L.copy_(loss.detach().reshape(M))
    G.copy_(grad.detach().reshape(M, n))
    m = int(history_size)
    d = torch.zeros(M, n, dtype=dt, device=dev)
    gp = torch.zeros_like(d)
    S = torch.zeros(M, m, n, dtype=dt, device=dev)
    Y = torch.zeros_like(S)
    ro = torch.zeros(M, m, dtype=torch.float64, device=dev)
    al = torch.zeros_like(ro)
    sysr = torch.zeros(M * N.PK_LBFGS_SYS_BYTES, dtype=torch.uint8, device=dev)
    counters = torch.zeros(2 * max_iter + 4, dtype=torch.int32, device=dev)
    b = N.LbfgsBatch(M=M, n=n, history=m, dtype=N.dtype_code(dt), x=x.data_ptr(), g=G.data_ptr(),
                     loss=L.data_ptr(), d=d.data_ptr(), g_prev=gp.data_ptr(), s_hist=S.data_ptr(),
                     y_hist=Y.data_ptr(), ro=ro.data_ptr(), al=al.data_ptr(), sys=sysr.data_ptr(),
                     counters=counters.data_ptr(), lr=float(lr), tol_grad=float(tolerance_grad),
                     tol_change=float(tolerance_change), tol_rel=float(tolerance_rel_change),
                     max_step=0.0 if max_step is None else float(max_step), max_eval=int(max_eval),
                     stop_on_increase=int(bool(stop_on_cost_increase)))
    lib = N.lib()
    sh = N.stream_handle(stream)
    N.check(lib.pk_lbfgs_begin(C.byref(b), sh), "pk_lbfgs_begin")
    for it in range(1, max_iter + 1):
        if sync_every and it % sync_every == 0 and int(counters[0].item()):
            break
        N.check(lib.pk_lbfgs_direction(C.byref(b), it, sh), "pk_lbfgs_direction")
        N.check(lib.pk_lbfgs_step(C.byref(b), it, sh), "pk_lbfgs_step")
        if it != max_iter:
            loss, grad = func(x, True)
            L.copy_(loss.detach().reshape(M))
            G.copy_(grad.detach().reshape(M, n))
    x0.copy_(x)
    return x0
