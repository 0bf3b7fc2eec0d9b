"""The oracle's batched L-BFGS (oracle.c or_lbfgs) against iterates recorded from the
reference implementation (tests/golden/lbfgs_golden.npz, written by
tests/golden/make_lbfgs_golden.py from core/lbfgs_torch_parallel.py)."""
import ast
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "lbfgs_golden.npz"))
A, B = G["quad_A"], G["quad_b"]


def quad(x):
    Ax = np.einsum("mij,mj->mi", A[: x.shape[0]], x)
    return 0.5 * (x * Ax).sum(1) - (B[: x.shape[0]] * x).sum(1), Ax - B[: x.shape[0]]


def rosen(x):
    a, b = x[:, :-1], x[:, 1:]
    f = (100 * (b - a ** 2) ** 2 + (1 - a) ** 2).sum(1)
    g = np.zeros_like(x)
    g[:, :-1] += -400 * a * (b - a ** 2) - 2 * (1 - a)
    g[:, 1:] += 200 * (b - a ** 2)
    return f, g


def absf(x):
    c = np.linspace(-1, 1, x.shape[1])
    return np.abs(x - c).sum(1) + 0.1 * (x ** 2).sum(1), np.sign(x - c) + 0.2 * x


FNS = {"quad": quad, "rosen": rosen, "abs": absf}


@pytest.mark.parametrize("k", range(len(G["meta"])))
def test_lbfgs_matches_reference(oracle_mod, k):
    name, opts = str(G["meta"][k]).split("|", 1)
    opts = ast.literal_eval(opts)
    x0 = G[f"x0_{k}"]
    x = oracle_mod.lbfgs(FNS[name], x0, **opts)
    ref = G[f"x_{k}"]
    # Same algorithm in float64; only summation order differs from torch.  Long runs on
    # Rosenbrock amplify rounding, so the bound is 10x the run's own response to a 1e-15
    # relative change of x0 (never looser than that, never tighter than 1e-9).
    xp = oracle_mod.lbfgs(FNS[name], x0 * (1 + 1e-15), **opts)
    tol = max(1e-9, 10 * np.abs(xp - x).max())
    assert np.abs(x - ref).max() <= tol, (np.abs(x - ref).max(), tol)


W = np.load(os.path.join(HERE, "golden", "lbfgs_wolfe_golden.npz"))


def quad_shared(x):
    Ax = x @ W["quad_A"].T
    return 0.5 * (x * Ax).sum(1) - (W["quad_b"] * x).sum(1), Ax - W["quad_b"]


WFNS = {"quad_shared": quad_shared, "rosen": rosen, "abs": absf}


@pytest.mark.parametrize("k", range(len(W["meta"])))
def test_lbfgs_wolfe_matches_reference(oracle_mod, k):
    """The oracle's strong-Wolfe L-BFGS against iterates recorded from the reference, and
    the same sequence of objective calls (rows per call: the batch-shared line-search
    loop evaluates only the rows still searching)."""
    name, opts = str(W["meta"][k]).split("|", 1)
    opts = ast.literal_eval(opts)
    opts.pop("line_search_fn")
    x0 = W[f"x0_{k}"]
    calls = []
    x = oracle_mod.lbfgs_wolfe(WFNS[name], x0, calls=calls, **opts)
    ref = W[f"x_{k}"]
    # the bound is 10x the run's own response to relative changes of x0 of a few ulps:
    # on the nonsmooth objective exact ties decide branches, and where they do the
    # response (and the bound) is large, i.e. that case pins nothing beyond it
    spread, stable = 0.0, True
    for eps in (1e-15, -1e-15, 3e-15):
        cp = []
        xp = oracle_mod.lbfgs_wolfe(WFNS[name], x0 * (1 + eps), calls=cp, **opts)
        spread = max(spread, np.abs(xp - x).max())
        stable &= cp == calls
    tol = max(1e-9, 10 * spread)
    assert np.abs(x - ref).max() <= tol, (np.abs(x - ref).max(), tol)
    if stable:
        assert calls == list(W[f"calls_{k}"]), (calls, list(W[f"calls_{k}"]))
