"""Pose-space table: GPU builder against exact FP64 clipping at the grid nodes
(<= 1e-4 absolute), lookup kernel against the oracle's restatement of the reference
kernel on the same table (cost to 1e-5 relative, gradient to 1e-5 away from cell faces),
and tables of the reference's default 900^3 size."""
import numpy as np
import pytest
import torch

from gpu_util import GEOMS, assert_cost_close, assert_grad_close, dense_poses, lut_face_guard, require_gpu

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
    # float32 relative poses and interpolation: measured <= 1.3e-7 relative on the cost,
    # <= 1.7e-6 on the gradient (relative to the instance's largest component)
    assert_cost_close(c.cpu().numpy(), c64, 1e-5, atol=1e-7)
    guard = lut_face_guard(lut, x, geom.radius)
    ok = ~guard
    assert ok.mean() > 0.9
    gg, gr = g.cpu().numpy(), g64
    assert_grad_close(np.where(ok[..., None], gg, 0.0), np.where(ok[..., None], gr, 0.0), 1e-5)
    # near a face the other cell's gradient may be taken: held to the jump between cells
    assert_grad_close(gg, gr, 5e-2)


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
    assert_cost_close(c.cpu().numpy(), c64, 1e-5, atol=1e-7)
    # on a node the gradient is one-sided, and float and double may pick neighbouring
    # cells; compare it strictly inside the cells only
    inside = np.array(inside) & ~lut_face_guard(lut, x, geom.radius).any(axis=1)
    assert inside.sum() >= 10
    assert_grad_close(g.cpu().numpy()[inside], g64[inside], 1e-5)


@pytest.mark.parametrize("gname", ["tree", "random"])
def test_reference_default_table_size(oracle_mod, gname):
    """CollisionCost(use_lookup_table=True) with the reference's defaults (900^3 grid,
    core/pack_cost.py:315-317; half table for the mirror-symmetric tree, full x range for
    the random polygon: 817 x 855 x 900 = 6.3e8 nodes after trimming, above 2^28 and 2^29).
    The table cost is within the grid's interpolation error of the exact area, and lookups
    equal the oracle's on 2x2x2 sub-blocks read back from the device table."""
    from packing_amd import pack_cost
    from packing_amd.context import set_geometry
    from packing_amd.pack_cuda_lut import overlap_multi_ensemble
    geom = GEOMS[gname]()
    set_geometry(geom)
    cc = pack_cost.CollisionCostOverlappingArea(use_lookup_table=True)
    cc._ensure_lut_initialized()
    L = cc._lut
    nodes = L.N_x * L.N_y * L.N_theta
    assert L.mirror_x == (gname == "tree")
    if gname == "random":
        assert nodes > 4e8
    x = dense_poses(8, 8, 30)
    xt = torch.tensor(x, device="cuda")
    c = torch.empty(8, device="cuda")
    overlap_multi_ensemble(xt, c, L)
    c64, _ = oracle_mod.overlap(oracle_mod.OracleGeometry(geom), x.astype(np.float64), False, grads=False)
    assert np.abs(c.cpu().numpy() - c64).max() <= 1e-4 * np.abs(c64).max()
    vals = L.lut_d
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(40):
        ix, iy, it = (int(v) for v in rng.integers(1, [L.N_x - 2, L.N_y - 2, L.N_theta - 2]))
        p = [float(A[i]) + f * (float(A[i + 1]) - float(A[i]))
             for A, i, f in ((L.X, ix, 0.37), (L.Y, iy, 0.61), (L.theta, it, 0.23))]
        if p[0] < 0 and L.mirror_x:
            continue
        pr = np.array([[[0.0, 0.0, 0.0], p]], np.float32)
        c1 = torch.empty(1, device="cuda")
        overlap_multi_ensemble(torch.tensor(pr, device="cuda"), c1, L)
        sub = vals[ix:ix + 2, iy:iy + 2, it:it + 2].cpu().numpy()
        OL = oracle_mod.OracleLut(L.X[ix:ix + 2], L.Y[iy:iy + 2], L.theta[it:it + 2], sub)
        v, _ = OL.lookup(*(float(a) for a in pr[0, 1]))
        assert abs(v - float(c1.item())) <= 1e-6 * max(1.0, abs(v)), (ix, iy, it, v, float(c1.item()))
        checked += 1
    assert checked >= 20
    del cc, L, vals
    torch.cuda.empty_cache()


def test_table_term_bound():
    """Lookup terms are summed in fixed point, each below 2000 in magnitude: a table whose
    values or node steps could exceed that (or that holds NaN / inf) is refused when its
    handle is made; the exact tables are far inside the bound."""
    from packing_amd.pack_cuda_lut import LookupTable
    X = np.linspace(-1.5, 1.5, 16)
    Y = np.linspace(-1.5, 1.5, 16)
    T = np.linspace(-np.pi, np.pi, 8)
    base = np.random.default_rng(0).uniform(0, 0.3, size=(16, 16, 8)).astype(np.float32)
    LookupTable(X, Y, T, base.copy(), mirror_x=False).handle("cuda")
    big = base.copy()
    big[3, 4, 5] = 2500.0                               # a value above the bound
    steep = base.copy()
    steep[7, 7, 3] = 150.0                              # |value| fine, slope x range is not
    nan = base.copy()
    nan[1, 2, 3] = np.nan
    for vals, what in ((big, "value"), (steep, "slope"), (nan, "nan")):
        with pytest.raises(ValueError, match="too large"):
            LookupTable(X, Y, T, vals, mirror_x=False).handle("cuda")
    # squared tables square the value bound: 40^2 = 1600 passes, 50^2 does not
    flat = np.full((16, 16, 8), 40.0, np.float32)
    LookupTable(X, Y, T, flat, apply_quadratic_transform=True, mirror_x=False).handle("cuda")
    with pytest.raises(ValueError, match="too large"):
        LookupTable(X, Y, T, flat + 10.0, apply_quadratic_transform=True, mirror_x=False).handle("cuda")
