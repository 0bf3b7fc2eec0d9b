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
