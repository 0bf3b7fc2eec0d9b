"""autograd.Function over the overlap cost: backward against FP64 central differences
(h = 1e-6) and against the reference's gradient (the FP64 oracle's reverse-mode
Sutherland-Hodgman), on the differentiation configuration (256 instances, N = 50)."""
import numpy as np
import pytest
import torch

from gpu_util import GEOMS, assert_grad_close, dense_poses, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


@pytest.mark.parametrize("use_sep", [False, True])
def test_backward_vs_fd_and_reference(oracle_mod, use_sep):
    from packing_amd.autograd import overlap_cost
    from packing_amd.context import set_geometry
    from packing_amd import pack_cuda
    geom = GEOMS["random"]()
    set_geometry(geom)
    x = dense_poses(4, 256, 50)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    w = torch.linspace(0.5, 1.5, 256, device="cuda")
    (overlap_cost(xt, use_sep) * w).sum().backward()
    g = xt.grad.cpu().numpy() / w.cpu().numpy()[:, None, None]
    # reference gradient (FP64 oracle)
    _, g64 = oracle_mod.overlap(oracle_mod.OracleGeometry(geom), x.astype(np.float64), use_sep)
    if not use_sep:
        assert_grad_close(g, g64, 1e-5, "vs oracle")
    else:
        # the SAT depth is a minimum over axes: where two axes are within float32
        # rounding of each other the float32 and float64 evaluations may pick different
        # ones and the gradient jumps; those near-ties must stay rare
        scale = np.abs(g64).reshape(256, -1).max(1)[:, None, None]
        bad = np.abs(g - g64) > 1e-5 * (np.abs(g64) + scale)
        assert bad.mean() < 1e-3, bad.sum()
    # FP64 central differences with the FP64 kernel, every instance perturbed at once
    xd = torch.tensor(x, device="cuda", dtype=torch.float64)
    h = 1e-6
    rng = np.random.default_rng(0)
    coords = rng.choice(50 * 3, size=30, replace=False)
    c = torch.empty(256, dtype=torch.float64, device="cuda")
    for q in coords:
        t, k = divmod(int(q), 3)
        xp = xd.clone()
        xp[:, t, k] += h
        pack_cuda.overlap_multi_ensemble(xp, xp, use_sep, out_cost=c)
        cp_ = c.clone()
        xm = xd.clone()
        xm[:, t, k] -= h
        pack_cuda.overlap_multi_ensemble(xm, xm, use_sep, out_cost=c)
        num = ((cp_ - c) / (2 * h)).cpu().numpy()
        scale = np.abs(g64).reshape(256, -1).max(1)
        err = np.abs(num - g[:, t, k])
        # kinks (a vertex crossing an edge within h) make a few differences meaningless
        bad = err > 1e-4 * (np.abs(num) + scale)
        assert bad.mean() < 0.02, (q, bad.sum())


def test_lut_backward():
    from packing_amd.autograd import overlap_cost
    from packing_amd.context import set_geometry
    from packing_amd.pack_cuda_lut import LookupTable
    geom = GEOMS["tree"]()
    set_geometry(geom)
    lut = LookupTable.build_exact(64, 64, 32)
    x = torch.tensor(dense_poses(6, 8, 30), device="cuda", requires_grad=True)
    overlap_cost(x, lut=lut).sum().backward()
    assert torch.isfinite(x.grad).all() and x.grad.abs().sum() > 0
