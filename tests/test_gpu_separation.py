"""Exact-separation table (pk_sep_build) and CollisionCostExactSeparation on the GPU,
against the FP64 oracle's signed separation distance and its table cost."""
import numpy as np
import pytest
import torch

from gpu_util import GEOMS, assert_cost_close, assert_grad_close, dense_poses, lut_face_guard, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


@pytest.mark.parametrize("gname", ["tree", "random"])
def test_builder_vs_oracle(oracle_mod, gname):
    """Every node of a 41 x 37 x 24 grid equals the oracle's double value rounded to
    float32 (both compute in double; the roundings of the two may differ in the last
    float32 place)."""
    from packing_amd.context import set_geometry
    from packing_amd.pack_cuda_lut import LookupTable
    geom = GEOMS[gname]()
    set_geometry(geom)
    lut = LookupTable.build_separation(41, 37, 24, full_x=True, trim_zeros=False)
    got = lut.lut_d.cpu().numpy()
    ref = oracle_mod.separation_grid(oracle_mod.OracleGeometry(geom), lut.X, lut.Y, lut.theta)
    ulp = np.spacing(np.abs(ref).astype(np.float32))
    assert np.all(np.abs(got - ref) <= ulp + 1e-12), np.abs(got - ref).max()
    assert (ref > 0).sum() > 1000 and (ref < 0).sum() > 1000


def test_exact_separation_cost_and_cache(oracle_mod, tmp_path, monkeypatch):
    """CollisionCostExactSeparation: built once (small grid), saved to the cache, loaded
    by a second instance without building; its cost and gradient equal the oracle's
    quadratic table cost on the same table."""
    from packing_amd import pack_cost as PC
    from packing_amd.context import set_geometry
    from packing_amd.pack_cuda_lut import LookupTable
    from packing_amd.solutions import SolutionCollectionSquare
    monkeypatch.setenv("PACKING_AMD_CACHE", str(tmp_path))
    geom = GEOMS["tree"]()
    set_geometry(geom)
    c1 = PC.CollisionCostExactSeparation(use_lookup_table=True, lut_N_x=96, lut_N_y=96, lut_N_theta=64)
    c1._ensure_lut_initialized()
    assert c1._lut.apply_quadratic_transform and c1._lut.mirror_x
    files = list((tmp_path / "lut_cache").iterdir())
    assert len(files) == 1

    def no_build(*a, **k):
        raise AssertionError("the cached table should have been loaded")

    monkeypatch.setattr(LookupTable, "build_separation", classmethod(no_build))
    c2 = PC.CollisionCostExactSeparation(use_lookup_table=True, lut_N_x=96, lut_N_y=96, lut_N_theta=64)
    c2._ensure_lut_initialized()
    assert np.array_equal(c2._lut.lut_d.cpu().numpy(), c1._lut.lut_d.cpu().numpy())
    x = dense_poses(8, 16, 40)
    sol = SolutionCollectionSquare(torch.tensor(x, device="cuda"),
                                   torch.tensor([[20.0, 0.0, 0.0]] * 16, device="cuda"), use_fixed_h=True)
    cost, g, gb = c2.compute_cost_allocate(sol)
    L = c2._lut
    OL = oracle_mod.OracleLut(L.X, L.Y, L.theta, L.lut_d.cpu().numpy(), quad=True, mirror=True)
    c64, g64 = OL.overlap(geom.radius, x.astype(np.float64))
    assert (c64 > 0).all()
    assert_cost_close(cost.cpu().numpy(), c64, 1e-5, atol=1e-7)
    ok = ~lut_face_guard(L, x, geom.radius)
    gg = g.cpu().numpy()
    assert_grad_close(np.where(ok[..., None], gg, 0.0), np.where(ok[..., None], g64, 0.0), 1e-5)
    with pytest.raises(ValueError):
        PC.CollisionCostExactSeparation()._ensure_lut_initialized()


def test_exact_separation_relaxation_vs_oracle(oracle_mod, tmp_path, monkeypatch):
    """The GA's production objective (boundary + exact separation table) through the
    unchanged OptimizerBFGS: the fused engine in quadratic-table mode follows the FP64
    oracle's iterates over the same table."""
    from packing_amd import pack_cost as PC
    from packing_amd.context import set_geometry
    from packing_amd.pack_dynamics import OptimizerBFGS
    from packing_amd.solutions import SolutionCollectionSquare
    monkeypatch.setenv("PACKING_AMD_CACHE", str(tmp_path))
    geom = GEOMS["tree"]()
    set_geometry(geom)
    coll = PC.CollisionCostExactSeparation(use_lookup_table=True, lut_N_x=96, lut_N_y=96, lut_N_theta=64)
    cost = PC.CostCompound(costs=[PC.BoundaryDistanceCost(), coll])
    N = 20
    x = dense_poses(6, 4, N)
    h = np.array([[1.2 * np.sqrt(N), 0.6 * np.sqrt(N), 0.6 * np.sqrt(N)]] * 4)
    opt = OptimizerBFGS(cost=cost, n_iterations=4, history_size=3, max_step=0.1)
    sol = SolutionCollectionSquare(torch.tensor(x, device="cuda"), torch.tensor(h, device="cuda", dtype=torch.float32),
                                   use_fixed_h=True)
    o = opt.engine_options(sol)
    assert o is not None and o.mode == "lut" and o.lut.apply_quadratic_transform
    r = opt.run_simulation(sol)
    L = coll._lut
    OL = oracle_mod.OracleLut(L.X, L.Y, L.theta, L.lut_d.cpu().numpy(), quad=True, mirror=True)
    ref = oracle_mod.relax(oracle_mod.OracleGeometry(geom), x.astype(np.float64), h, max_iter=4, m=3,
                           max_step=0.1, lut=OL)
    d = np.abs(r.xyt.cpu().numpy() - ref).max()
    assert d < 3e-6, d
    assert np.abs(r.xyt.cpu().numpy() - x).max() > 1e-3


def test_reference_default_separation_table(oracle_mod, tmp_path, monkeypatch):
    """The reference's default 900^3 grid for the tree (half table, trimmed): built on the
    GPU in double; sampled nodes equal the oracle's."""
    import time
    from packing_amd import pack_cost as PC
    from packing_amd.context import set_geometry
    monkeypatch.setenv("PACKING_AMD_CACHE", str(tmp_path))
    geom = GEOMS["tree"]()
    set_geometry(geom)
    c = PC.CollisionCostExactSeparation(use_lookup_table=True)
    t0 = time.perf_counter()
    c._ensure_lut_initialized()
    L = c._lut
    print(f"separation table {L.N_x} x {L.N_y} x {L.N_theta} in {time.perf_counter() - t0:.2f} s")
    vals = L.lut_d
    og = oracle_mod.OracleGeometry(geom)
    rng = np.random.default_rng(0)
    for _ in range(200):
        i, j, k = (int(v) for v in rng.integers(0, [L.N_x, L.N_y, L.N_theta]))
        ref = np.float32(oracle_mod.separation(og, float(L.X[i]), float(L.Y[j]), float(L.theta[k])))
        got = float(vals[i, j, k].item())
        assert abs(got - ref) <= np.spacing(abs(ref)) + 1e-12, (i, j, k, got, ref)
    del c, L, vals
    torch.cuda.empty_cache()
