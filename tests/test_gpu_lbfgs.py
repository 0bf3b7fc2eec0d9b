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
