"""Pose-space table: GPU builder against exact FP64 clipping at the grid nodes
(<= 1e-4 absolute), lookup kernel against the oracle's restatement of the reference
kernel on the same table, and the trilinear closed form."""
import numpy as np
import pytest
import torch

from gpu_util import GEOMS, assert_cost_close, assert_grad_close, dense_poses, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


@pytest.mark.parametrize("gname", ["tree", "random"])
def test_table_nodes_vs_exact(oracle_mod, gname):
    from packing_amd.pack_cuda_lut import LookupTable
    from packing_amd.context import set_geometry
    geom = GEOMS[gname]()
    set_geometry(geom)
    lut = LookupTable.build_exact(64, 64, 32)
    vals = lut.lut_d.cpu().numpy()
    og = oracle_mod.OracleGeometry(geom)
    rng = np.random.default_rng(0)
    idx = rng.integers(0, [64, 64, 32], size=(400, 3))
    xyt = np.zeros((400, 2, 3))
    xyt[:, 1, 0] = lut.X[idx[:, 0]]
    xyt[:, 1, 1] = lut.Y[idx[:, 1]]
    xyt[:, 1, 2] = lut.theta[idx[:, 2]]
    ref, _ = oracle_mod.overlap(og, xyt, False, grads=False)
    got = vals[idx[:, 0], idx[:, 1], idx[:, 2]]
    assert np.abs(got - ref).max() <= 1e-4
    assert (ref > 0).sum() > 50


@pytest.mark.parametrize("mirror", [False, True])
def test_lookup_kernel_vs_oracle(oracle_mod, mirror):
    from packing_amd.pack_cuda_lut import LookupTable, overlap_multi_ensemble
    from packing_amd.context import set_geometry
    geom = GEOMS["tree"]()
    set_geometry(geom)
    lut = LookupTable.build_exact(80, 80, 48, full_x=not mirror)
    vals = lut.lut_d.cpu().numpy()
    OL = oracle_mod.OracleLut(lut.X, lut.Y, lut.theta, vals, mirror=mirror)
    x = dense_poses(8, 16, 50)
    xt = torch.tensor(x, device="cuda")
    c = torch.empty(16, device="cuda")
    g = torch.empty_like(xt)
    overlap_multi_ensemble(xt, c, lut, out_grads=g)
    torch.cuda.synchronize()
    c64, g64 = OL.overlap(geom.radius, x.astype(np.float64))
    assert_cost_close(c.cpu().numpy(), c64, 1e-4, atol=1e-6)
    assert_grad_close(g.cpu().numpy(), g64, 1e-3)


@pytest.mark.parametrize("mirror", [False, True])
def test_lookup_table_edges(oracle_mod, mirror):
    """Relative poses at the table's edges, which random poses never hit: theta exactly at
    the first and last node (the clamped upper theta corner), theta beyond pi (wrapped),
    x on the last node (clamped upper x corner) and, for the half table, mirrored x."""
    from packing_amd.pack_cuda_lut import LookupTable, overlap_multi_ensemble
    from packing_amd.context import set_geometry
    geom = GEOMS["tree"]()
    set_geometry(geom)
    lut = LookupTable.build_exact(48, 40, 24, full_x=not mirror)
    vals = lut.lut_d.cpu().numpy()
    OL = oracle_mod.OracleLut(lut.X, lut.Y, lut.theta, vals, mirror=mirror)
    R = geom.radius
    T, X, Y = lut.theta, lut.X, lut.Y
    poses, inside = [], []
    for dx, dy, on_x in ((0.35 * R, 0.2 * R, False), (-0.4 * R, 0.1 * R, False), (X[-1], 0.0, True),
                         (0.0, Y[1], True), (0.6 * R, -0.5 * R, False)):
        for th, on_t in ((T[-1], True), (T[0], True), (T[-2] + 0.5 * (T[-1] - T[-2]), False),
                         (T[-1] + 0.05, False), (T[0] - 0.05, False), (0.0, False)):
            poses.append([[0.0, 0.0, 0.0], [dx, dy, th]])
            inside.append(not (on_x or on_t))
    x = np.array(poses, dtype=np.float32)
    xt = torch.tensor(x, device="cuda")
    c = torch.empty(len(poses), device="cuda")
    g = torch.empty_like(xt)
    overlap_multi_ensemble(xt, c, lut, out_grads=g)
    torch.cuda.synchronize()
    c64, g64 = OL.overlap(geom.radius, x.astype(np.float64))
    assert (c64 > 0).sum() >= 12
    # the interpolated value is continuous across cells: exact nodes included
    assert_cost_close(c.cpu().numpy(), c64, 1e-4, atol=1e-6)
    # on a node the gradient is one-sided, and float and double may pick neighbouring
    # cells; compare it strictly inside the cells only
    inside = np.array(inside)
    assert inside.sum() >= 10
    assert_grad_close(g.cpu().numpy()[inside], g64[inside], 1e-3)
