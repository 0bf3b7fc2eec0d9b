"""Drop-in lbfgs (device kernels) against the iterates recorded from the reference
implementation (tests/golden/lbfgs_golden.npz), float64 on the GPU."""
import ast
import os

import numpy as np
import pytest
import torch

from gpu_util import require_gpu

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "lbfgs_golden.npz"))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def torch_fns():
    A = torch.tensor(G["quad_A"], device="cuda")
    b = torch.tensor(G["quad_b"], device="cuda")

    def quad(x, _m=True):
        Ax = torch.einsum("mij,mj->mi", A[: x.shape[0]], x)
        return 0.5 * (x * Ax).sum(1) - (b[: x.shape[0]] * x).sum(1), Ax - b[: x.shape[0]]

    def rosen(x, _m=True):
        a, c = x[:, :-1], x[:, 1:]
        f = (100 * (c - a ** 2) ** 2 + (1 - a) ** 2).sum(1)
        g = torch.zeros_like(x)
        g[:, :-1] += -400 * a * (c - a ** 2) - 2 * (1 - a)
        g[:, 1:] += 200 * (c - a ** 2)
        return f, g

    def absf(x, _m=True):
        cc = torch.linspace(-1, 1, x.shape[1], device=x.device, dtype=x.dtype)
        return (x - cc).abs().sum(1) + 0.1 * (x ** 2).sum(1), torch.sign(x - cc) + 0.2 * x

    return {"quad": quad, "rosen": rosen, "abs": absf}


@pytest.mark.parametrize("k", range(len(G["meta"])))
def test_lbfgs_matches_reference(oracle_mod, k):
    from packing_amd.lbfgs_torch_parallel import lbfgs
    from test_oracle_lbfgs import FNS
    name, opts = str(G["meta"][k]).split("|", 1)
    opts = ast.literal_eval(opts)
    x0 = G[f"x0_{k}"]
    x = torch.tensor(x0, device="cuda")
    lbfgs(torch_fns()[name], x, **opts)
    ref = G[f"x_{k}"]
    a = oracle_mod.lbfgs(FNS[name], x0, **opts)
    b = oracle_mod.lbfgs(FNS[name], x0 * (1 + 1e-15), **opts)
    tol = max(1e-9, 10 * np.abs(a - b).max())
    assert np.abs(x.cpu().numpy() - ref).max() <= tol


def test_float32_converges_to_minimiser():
    """float32: a well-conditioned batched quadratic reaches A^-1 b."""
    from packing_amd.lbfgs_torch_parallel import lbfgs
    rng = np.random.default_rng(4)
    M, n = 64, 30
    d = torch.tensor(rng.uniform(1.0, 4.0, (M, n)), device="cuda", dtype=torch.float32)
    b = torch.tensor(rng.normal(size=(M, n)), device="cuda", dtype=torch.float32)

    def f(x, _m=True):
        return 0.5 * (d * x * x).sum(1) - (b * x).sum(1), d * x - b

    x = torch.zeros(M, n, device="cuda")
    lbfgs(f, x, max_iter=60, history_size=8, max_step=None, tolerance_grad=0, tolerance_change=0)
    np.testing.assert_allclose(x.cpu().numpy(), (b / d).cpu().numpy(), atol=1e-4)


W = np.load(os.path.join(HERE, "golden", "lbfgs_wolfe_golden.npz"))


def torch_wolfe_fns(calls):
    A = torch.tensor(W["quad_A"], device="cuda")
    b = torch.tensor(W["quad_b"], device="cuda")

    def quad_shared(x, _m=True):
        calls.append(x.shape[0])
        Ax = x @ A.T
        return 0.5 * (x * Ax).sum(1) - (b * x).sum(1), Ax - b

    fns = torch_fns()

    def counted(f):
        def g(x, _m=True):
            calls.append(x.shape[0])
            return f(x, _m)
        return g

    return {"quad_shared": quad_shared, "rosen": counted(fns["rosen"]), "abs": counted(fns["abs"])}


@pytest.mark.parametrize("k", range(len(W["meta"])))
def test_lbfgs_strong_wolfe_matches_reference(oracle_mod, k):
    """lbfgs(..., line_search_fn='strong_wolfe') on the device against iterates recorded
    from the reference (float64): same iterates (bound: 10x the oracle's own response to
    x0 changes of a few ulps) and, where that response leaves the branch decisions
    unchanged, the same sequence of objective calls (rows per call)."""
    from packing_amd.lbfgs_torch_parallel import lbfgs
    from test_oracle_lbfgs import WFNS
    name, opts = str(W["meta"][k]).split("|", 1)
    opts = ast.literal_eval(opts)
    x0 = W[f"x0_{k}"]
    calls = []
    x = torch.tensor(x0, device="cuda")
    lbfgs(torch_wolfe_fns(calls)[name], x, **opts)
    ref = W[f"x_{k}"]
    o = dict(opts)
    o.pop("line_search_fn")
    c0 = []
    xo = oracle_mod.lbfgs_wolfe(WFNS[name], x0, calls=c0, **o)
    spread, stable = 0.0, True
    for eps in (1e-15, -1e-15, 3e-15):
        cp = []
        xp = oracle_mod.lbfgs_wolfe(WFNS[name], x0 * (1 + eps), calls=cp, **o)
        spread = max(spread, np.abs(xp - xo).max())
        stable &= cp == c0
    tol = max(1e-9, 10 * spread)
    assert np.abs(x.cpu().numpy() - ref).max() <= tol, (np.abs(x.cpu().numpy() - ref).max(), tol)
    if stable:
        assert calls == list(W[f"calls_{k}"]), (calls, list(W[f"calls_{k}"]))


def test_strong_wolfe_float32_converges():
    """float32 with the line search: a well-conditioned batched quadratic reaches A^-1 b in
    fewer iterations than without it."""
    from packing_amd.lbfgs_torch_parallel import lbfgs
    rng = np.random.default_rng(4)
    M, n = 64, 30
    dd = torch.tensor(rng.uniform(1.0, 4.0, n), device="cuda", dtype=torch.float32)
    b = torch.tensor(rng.normal(size=n), device="cuda", dtype=torch.float32)

    def f(x, _m=True):
        return 0.5 * (dd * x * x).sum(1) - (b * x).sum(1), dd * x - b

    x = torch.zeros(M, n, device="cuda")
    lbfgs(f, x, max_iter=25, history_size=8, max_step=None, tolerance_grad=0, tolerance_change=0,
          line_search_fn="strong_wolfe")
    np.testing.assert_allclose(x.cpu().numpy(), np.broadcast_to((b / dd).cpu().numpy(), (M, n)), atol=1e-4)
