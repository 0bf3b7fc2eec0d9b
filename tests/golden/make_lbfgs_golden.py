"""Writes tests/golden/lbfgs_golden.npz and lbfgs_wolfe_golden.npz: iterates of the
reference's batched L-BFGS (core/lbfgs_torch_parallel.py, torch CPU, float64) on
closed-form objectives, without and with the strong-Wolfe line search.

Nothing else pins that function's exact iterate semantics (first-step scaling, max-step
clamp, history shift, batch-coupled stopping, inactive systems that keep stepping), so
its input/output vectors are recorded here.  Run in the build container only (needs the
reference checkout); the .npz it writes is the committed fixture.

    python tests/golden/make_lbfgs_golden.py /root/reference
"""
import os
import sys

import numpy as np
import torch

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
sys.path.insert(0, os.path.join(REF, "core"))
import lbfgs_torch_parallel as L  # noqa: E402

torch.set_default_dtype(torch.float64)


def quad_fn(A, b):
    """f_m(x) = 0.5 x^T A_m x - b_m . x per system"""
    def f(x, _main=True):
        Ax = torch.einsum("mij,mj->mi", A, x)
        return 0.5 * (x * Ax).sum(1) - (b * x).sum(1), Ax - b
    return f


def rosen_fn(x, _main=True):
    """Chained Rosenbrock per system."""
    a, b = x[:, :-1], x[:, 1:]
    f = (100 * (b - a ** 2) ** 2 + (1 - a) ** 2).sum(1)
    g = torch.zeros_like(x)
    g[:, :-1] += -400 * a * (b - a ** 2) - 2 * (1 - a)
    g[:, 1:] += 200 * (b - a ** 2)
    return f, g


def abs_fn(x, _main=True):
    """Nonsmooth: sum |x - c| + 0.1 x^2 (exercises the directional-derivative stop)."""
    c = torch.linspace(-1, 1, x.shape[1])
    f = (x - c).abs().sum(1) + 0.1 * (x ** 2).sum(1)
    g = torch.sign(x - c) + 0.2 * x
    return f, g


def main():
    rng = np.random.default_rng(20251203)
    out = {}
    M, n = 6, 7
    Q = rng.normal(size=(M, n, n))
    A = np.einsum("mij,mkj->mik", Q, Q) + 0.5 * np.eye(n)[None]
    b = rng.normal(size=(M, n))
    x0q = rng.normal(size=(M, n))
    x0r = rng.uniform(-1.5, 1.5, size=(M, 5))
    x0a = rng.normal(size=(3, 4))
    cases = []
    for name, fn, x0 in (("quad", quad_fn(torch.tensor(A), torch.tensor(b)), x0q),
                         ("rosen", rosen_fn, x0r), ("abs", abs_fn, x0a)):
        for opts in (
            dict(max_iter=1, history_size=3, max_step=0.5),
            dict(max_iter=2, history_size=3, max_step=0.5),
            dict(max_iter=5, history_size=3, max_step=0.5),
            dict(max_iter=40, history_size=3, max_step=0.5),
            dict(max_iter=40, history_size=8, max_step=0.1, tolerance_grad=0.0, tolerance_change=0.0),
            dict(max_iter=25, history_size=2, max_step=None, lr=0.5),
            dict(max_iter=30, history_size=4, max_step=0.3, stop_on_cost_increase=True),
            dict(max_iter=60, history_size=4, max_step=0.5, tolerance_rel_change=1e-3),
        ):
            cases.append((name, fn, x0, opts))
    meta = []
    for k, (name, fn, x0, opts) in enumerate(cases):
        x = torch.tensor(x0.copy())
        res = L.lbfgs(fn, x, **opts)
        out[f"x0_{k}"] = x0
        out[f"x_{k}"] = res.numpy()
        meta.append(f"{name}|{repr(opts)}")
    out["meta"] = np.array(meta)
    out["quad_A"], out["quad_b"] = A, b
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lbfgs_golden.npz")
    np.savez(path, **out)
    print(f"wrote {len(cases)} cases to {path}")

    # line_search_fn='strong_wolfe' (core/lbfgs_torch_parallel.py:70-407): bracketing and
    # zoom phases, batch-shared line-search iteration count, evaluation on the active rows
    # only; recorded with the number of objective calls and rows per call.
    # the line search evaluates only the rows still searching, as a smaller batch, so
    # the objective must not depend on the row index (as pack_dynamics' does not): the
    # quadratic shares system 0's A and b
    wout, wmeta = {}, []
    x0w = rng.normal(size=(5, 6))
    wcases = []
    A1 = np.broadcast_to(A[:1], A.shape).copy()
    b1 = np.broadcast_to(b[:1], b.shape).copy()

    def quad_shared(x, _main=True):
        At, bt = torch.tensor(A1[0]), torch.tensor(b1[0])
        Ax = x @ At.T
        return 0.5 * (x * Ax).sum(1) - (bt * x).sum(1), Ax - bt

    for name, fn, x0 in (("quad_shared", quad_shared, x0q),
                         ("rosen", rosen_fn, x0r), ("abs", abs_fn, x0a), ("rosen", rosen_fn, x0w)):
        for opts in (
            dict(max_iter=1, history_size=3, max_step=0.5),
            dict(max_iter=3, history_size=3, max_step=0.5),
            dict(max_iter=12, history_size=3, max_step=0.5),
            dict(max_iter=30, history_size=5, max_step=None, tolerance_grad=0.0, tolerance_change=0.0),
            dict(max_iter=25, history_size=4, max_step=0.2, lr=0.7),
            dict(max_iter=30, history_size=4, max_step=0.5, tolerance_rel_change=1e-4),
            dict(max_iter=20, history_size=2, max_step=1.0, stop_on_cost_increase=True),
        ):
            wcases.append((name, fn, x0, dict(opts, line_search_fn="strong_wolfe")))
    for k, (name, fn, x0, opts) in enumerate(wcases):
        calls = []

        def counted(x, main, fn=fn, calls=calls):
            calls.append(x.shape[0])
            return fn(x, main)

        x = torch.tensor(x0.copy())
        res = L.lbfgs(counted, x, **opts)
        wout[f"x0_{k}"] = x0
        wout[f"x_{k}"] = res.numpy()
        wout[f"calls_{k}"] = np.array(calls, dtype=np.int64)
        wmeta.append(f"{name}|{repr(opts)}")
    wout["meta"] = np.array(wmeta)
    wout["quad_A"], wout["quad_b"] = A1[0], b1[0]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lbfgs_wolfe_golden.npz")
    np.savez(path, **wout)
    print(f"wrote {len(wcases)} line-search cases to {path}")


if __name__ == "__main__":
    main()
