"""Drop-in for the reference module ``lbfgs_torch_parallel`` on MI355X.

``lbfgs(func, x0, ...)`` has the reference's signature and semantics
(core/lbfgs_torch_parallel.py:411-729): M independent systems, two-loop recursion,
first step lr * min(1, 1/sum|g|), max-step clamp, history kept only from iteration 2
when y.s > 1e-10, the reference's convergence tests, and x0 updated in place.

* ``line_search_fn=None``: inactive systems keep stepping while any system is active,
  as in the reference.  The host never waits on the GPU except every ``sync_every``
  iterations, to leave the loop early once the device has marked the run as stopped.
* ``line_search_fn='strong_wolfe'``: the reference's batched strong-Wolfe search
  (:70-407) for the systems still active, with its batch-shared loop (25 search
  iterations in total per L-BFGS iteration, bracketing then zoom) and ``func`` called on
  the rows still searching only, as a smaller batch (so ``func`` must treat rows
  alike, as pack_dynamics' objective does).  Every search decision (Armijo and
  curvature tests, cubic steps, bracket updates) runs in device kernels
  (pk_lbfgs_ls_*); the host reads one count per search iteration to pick the rows to
  evaluate, as the reference's ``active.any()`` does.

The optimizer arithmetic runs in HIP kernels (pk_lbfgs_*); only ``func`` runs between them.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import torch
from torch import Tensor

from . import _native as N

__all__ = ["lbfgs"]

MAX_LS = 25   # search iterations per L-BFGS iteration (the reference's default)


def lbfgs(func, x0: Tensor, lr: float = 1.0, max_iter: int = 20, max_eval: Optional[int] = None,
          tolerance_grad: float = 1e-7, tolerance_change: float = 1e-9, tolerance_rel_change: float = 0.0,
          stop_on_cost_increase: bool = False, history_size: int = 100,
          line_search_fn: Optional[str] = None, max_step: Optional[float] = 0.5,
          sync_every: int = 32, stream=None):
    if line_search_fn not in (None, "strong_wolfe"):
        raise ValueError(f"line_search_fn: None or 'strong_wolfe', not {line_search_fn!r}")
    if not x0.is_cuda:
        raise ValueError("x0 must be a GPU tensor")
    if x0.dtype not in (torch.float32, torch.float64):
        raise ValueError("x0 must be float32 or float64")
    if max_eval is None:
        max_eval = max_iter * 5 // 4
    if stream is not None:
        # host reads of device counts below use the current stream
        with torch.cuda.stream(stream):
            return lbfgs(func, x0, lr, max_iter, max_eval, tolerance_grad, tolerance_change,
                         tolerance_rel_change, stop_on_cost_increase, history_size, line_search_fn,
                         max_step, sync_every, None)
    ls = line_search_fn == "strong_wolfe"
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
                     stop_on_increase=int(bool(stop_on_cost_increase)), line_search=int(ls))
    lib = N.lib()
    sh = N.stream_handle(stream)
    N.check(lib.pk_lbfgs_begin(C.byref(b), sh), "pk_lbfgs_begin")
    if ls:
        _run_line_search(func, b, x, d, L, G, counters, max_iter, sh, stream)
    else:
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


def _run_line_search(func, b, x, d, L, G, counters, max_iter, sh, stream):
    """L-BFGS iterations with the strong-Wolfe search (core/lbfgs_torch_parallel.py:
    606-622); the history update and convergence tests of each iteration run at the
    start of the next pk_lbfgs_direction, as without a search."""
    lib = N.lib()
    M, n = x.shape
    dev, dt = x.device, x.dtype
    f_t = torch.zeros(M, dtype=dt, device=dev)
    g_t = torch.zeros(M, n, dtype=dt, device=dev)
    x_t = torch.zeros(M, n, dtype=dt, device=dev)
    gprev = torch.zeros(M, n, dtype=dt, device=dev)
    bg = torch.zeros(M, 2, n, dtype=dt, device=dev)
    lsr = torch.zeros(M * N.PK_LBFGS_LS_BYTES, dtype=torch.uint8, device=dev)
    flag = torch.zeros(M, dtype=torch.int32, device=dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    q = N.LbfgsLs(f=f_t.data_ptr(), g=g_t.data_ptr(), x=x_t.data_ptr(), g_prev=gprev.data_ptr(),
                  bg=bg.data_ptr(), sys=lsr.data_ptr(), flag=flag.data_ptr(), count=count.data_ptr())
    bp, qp = C.byref(b), C.byref(q)

    def evaluate(main):
        """Objective on the flagged rows (ascending row order), written into f_t, g_t."""
        idx = torch.nonzero(flag, as_tuple=False).reshape(-1)
        f, g = func(x_t.index_select(0, idx), main)
        f_t.index_copy_(0, idx, f.detach().reshape(-1).to(dt))
        g_t.index_copy_(0, idx, g.detach().reshape(idx.numel(), n).to(dt))
        N.check(lib.pk_lbfgs_ls_gtd(bp, qp, sh), "pk_lbfgs_ls_gtd")

    for it in range(1, max_iter + 1):
        N.check(lib.pk_lbfgs_direction(bp, it, sh), "pk_lbfgs_direction")
        c = counters[[1 + 2 * (it - 1), 2 + 2 * it]].cpu()
        if int(c[0]) == 0 or int(c[1]) == 0:
            break                                  # no system active: the reference's loop exit / break
        N.check(lib.pk_lbfgs_ls_begin(bp, qp, sh), "pk_lbfgs_ls_begin")
        evaluate(True)
        ls_iter = 0
        while ls_iter < MAX_LS:
            N.check(lib.pk_lbfgs_ls_bracket(bp, qp, ls_iter, sh), "pk_lbfgs_ls_bracket")
            if int(count.item()) == 0:
                break
            evaluate(False)
            ls_iter += 1
        N.check(lib.pk_lbfgs_ls_zoom_begin(bp, qp, ls_iter, sh), "pk_lbfgs_ls_zoom_begin")
        active = int(count.item())
        while active > 0 and ls_iter < MAX_LS:
            N.check(lib.pk_lbfgs_ls_zoom_trial(bp, qp, sh), "pk_lbfgs_ls_zoom_trial")
            if int(count.item()) == 0:
                break
            evaluate(False)
            ls_iter += 1
            N.check(lib.pk_lbfgs_ls_zoom_update(bp, qp, sh), "pk_lbfgs_ls_zoom_update")
            active = int(count.item())
        N.check(lib.pk_lbfgs_ls_end(bp, qp, sh), "pk_lbfgs_ls_end")
    # the history update / tests after the last iteration change no iterate
    return x
