"""pack_cost drop-in on the reference's eight cost-test solutions (core/pack_test.py:137-213):
recorded known answers in float32 and float64, finite-difference gradients including the
symmetric phenotype backprop and the periodic finite-difference terms, and the
vectorised-equals-individual property (core/pack_test.py:372-447), all on the GPU."""
import json
import os

import numpy as np
import pytest
import torch

from gpu_util import require_gpu

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()
    from packing_amd.context import set_geometry
    from packing_amd.geometry import reference_tree
    set_geometry(reference_tree())


def make_solution(i, dtype):
    from packing_amd import solutions as S
    from test_oracle_known_answers import place_random
    sol = KA["solutions"][i]
    k = sol["kind"]
    d = dict(dtype=dtype, device="cuda")
    if k in ("sym180", "sym90", "square"):
        x = torch.tensor(place_random(*sol["place_random"])[None], **d)
        h = torch.tensor([sol["h"]], **d)
        if k == "sym180":
            s = S.SolutionCollectionSquareSymmetric180(x, h)
            s.canonicalize()
        elif k == "sym90":
            s = S.SolutionCollectionSquareSymmetric90(x, h)
            s.canonicalize()
        else:
            s = S.SolutionCollectionSquare(x, h, override_phenotype=True)
        return s
    x = torch.tensor(np.array(sol["xyt_over_4"])[None] / 4, **d)
    h = torch.tensor([sol["h"]], **d)
    if k == "lattice":
        return S.SolutionCollectionLattice(x, h)
    if k == "rectangle":
        return S.SolutionCollectionLatticeRectangle(x, h)
    return S.SolutionCollectionLatticeFixed(x, h, N_periodic=sol["n_periodic"],
                                            aspect_ratios=torch.tensor([sol["aspect"]], **d))


def cost_objects():
    from packing_amd import pack_cost as PC
    return {
        "AreaCost": PC.AreaCost(),
        "CollisionCostSeparation": PC.CollisionCostSeparation(scaling=5.0),
        "BoundaryDistanceCost_kernel": PC.BoundaryDistanceCost(scaling=5.0),
        "CollisionCostOverlappingArea": PC.CollisionCostOverlappingArea(scaling=3.0),
        "CostCompound_1.5xArea+Sep": PC.CostCompound(scaling=1.5, costs=[PC.AreaCost(), PC.CollisionCostSeparation()]),
    }


@pytest.mark.parametrize("name", list(cost_objects()))
def test_known_answers(name):
    c = cost_objects()[name]
    ref = KA["costs"][name]["f64"]
    for i in range(len(ref)):
        for dtype, rtol in ((torch.float32, 1e-5), (torch.float64, 1e-11)):
            sol = make_solution(i, dtype)
            cost, _, _ = c.compute_cost_allocate(sol)
            cost_ng, _, _ = c.compute_cost_allocate(sol, evaluate_gradient=False)
            assert torch.equal(cost, cost_ng)
            assert abs(cost.item() - ref[i]) <= rtol * abs(ref[i]) + 1e-12, (name, i, dtype, cost.item(), ref[i])


@pytest.mark.parametrize("name", ["CollisionCostSeparation", "CollisionCostOverlappingArea",
                                  "BoundaryDistanceCost_kernel", "CostCompound_1.5xArea+Sep"])
def test_gradients_fd(name):
    """float64: analytic (and, for lattices, the reference's finite-difference) gradients vs
    central differences with eps 1e-6, tolerance 1e-4 as core/pack_test.py:301-370."""
    c = cost_objects()[name]
    n_sol = 4 if name.startswith("Boundary") else 8
    for i in range(n_sol):
        sol = make_solution(i, torch.float64)
        _, g, gb = c.compute_cost_allocate(sol)
        eps = 1e-6
        flat = sol.xyt.reshape(-1)
        for q in range(flat.numel()):
            x0 = sol.xyt.clone()
            sol.xyt.view(-1)[q] += eps
            cp_ = c.compute_cost_allocate(sol, evaluate_gradient=False)[0].item()
            sol.xyt.copy_(x0)
            sol.xyt.view(-1)[q] -= eps
            cm_ = c.compute_cost_allocate(sol, evaluate_gradient=False)[0].item()
            sol.xyt.copy_(x0)
            num = (cp_ - cm_) / (2 * eps)
            assert abs(num - g.view(-1)[q].item()) <= 1e-4 + 1e-4 * abs(num), (name, i, q)
        for q in range(sol.h.numel()):
            h0 = sol.h.clone()
            sol.h.view(-1)[q] += eps
            cp_ = c.compute_cost_allocate(sol, evaluate_gradient=False)[0].item()
            sol.h.copy_(h0)
            sol.h.view(-1)[q] -= eps
            cm_ = c.compute_cost_allocate(sol, evaluate_gradient=False)[0].item()
            sol.h.copy_(h0)
            num = (cp_ - cm_) / (2 * eps)
            assert abs(num - gb.view(-1)[q].item()) <= 1e-4 + 1e-4 * abs(num), (name, i, "h", q)


@pytest.mark.parametrize("name", list(cost_objects()))
def test_vectorised_equals_individual(name):
    """Solutions 2+3 (and 4+5 for non-boundary costs) in one call give bitwise the same
    cost and gradients as separate calls (core/pack_test.py:372-447)."""
    c = cost_objects()[name]
    groups = [[2, 3]] if name.startswith("Boundary") else [[2, 3], [4, 5]]
    for grp in groups:
        sols = [make_solution(i, torch.float32) for i in grp]
        full = make_solution(grp[0], torch.float32)
        full.xyt = torch.cat([s.xyt for s in sols], 0)
        full.h = torch.cat([s.h for s in sols], 0)
        fc, fg, fb = c.compute_cost_allocate(full)
        for k, s in enumerate(sols):
            sc, sg, sb = c.compute_cost_allocate(s)
            assert torch.equal(fc[k], sc[0]) and torch.equal(fg[k], sg[0]) and torch.equal(fb[k], sb[0])


def test_pack_dynamics_paths_agree():
    """OptimizerBFGS on the reference loop (drop-in lbfgs over the cost objects) and on the
    fused engine (chosen automatically) agree to float32 rounding, fixed and free h."""
    from packing_amd import pack_cost as PC
    from packing_amd.pack_dynamics import OptimizerBFGS
    from packing_amd.solutions import SolutionCollectionSquare
    from gpu_util import dense_poses
    x = torch.tensor(dense_poses(2, 6, 12), device="cuda")
    h = torch.tensor([[3.0, 0.9, 0.9]] * 6, device="cuda")
    for fixed in (True, False):
        opt = OptimizerBFGS(n_iterations=3, history_size=4, max_step=0.1)   # default cost
        sol = SolutionCollectionSquare(x.clone(), h.clone(), use_fixed_h=fixed)
        assert opt.engine_options(sol) is not None
        a = opt.run_simulation(sol, engine="reference")
        b = opt.run_simulation(sol)
        assert torch.allclose(a.xyt, b.xyt, atol=5e-6), (a.xyt - b.xyt).abs().max()
        assert torch.allclose(a.h, b.h, atol=5e-6)
        if fixed:
            assert torch.equal(a.h, h) and torch.equal(b.h, h)
        else:
            assert not torch.equal(b.h, h)
    # settings the engine does not reproduce take the reference loop
    assert OptimizerBFGS(tolerance_rel_change=1e-3).engine_options(sol) is None
    assert OptimizerBFGS(stop_on_cost_increase=True).engine_options(sol) is None
    assert OptimizerBFGS().engine_options(SolutionCollectionSquare(x.double(), h.double())) is None


def test_pack_dynamics_table_cost_vs_oracle(oracle_mod):
    """The GA's table objective through the unchanged reference API: OptimizerBFGS with
    Area + Boundary + CollisionCostOverlappingArea(use_lookup_table=True) runs on the fused
    engine and its first iterates follow the FP64 oracle over the same table."""
    from packing_amd import pack_cost as PC
    from packing_amd.context import set_geometry
    from packing_amd.pack_cuda_lut import LookupTable
    from packing_amd.pack_dynamics import OptimizerBFGS
    from packing_amd.solutions import SolutionCollectionSquare
    from gpu_util import GEOMS, dense_poses
    geom = GEOMS["tree"]()
    set_geometry(geom)
    coll = PC.CollisionCostOverlappingArea(use_lookup_table=True)
    coll._lut = LookupTable.build_exact(96, 96, 64, full_x=False)
    cost = PC.CostCompound(costs=[PC.AreaCost(scaling=1e-2), PC.BoundaryDistanceCost(), coll])
    N = 20
    x = dense_poses(5, 4, N)
    h = np.array([[1.2 * np.sqrt(N), 0.6 * np.sqrt(N), 0.6 * np.sqrt(N)]] * 4)
    opt = OptimizerBFGS(cost=cost, n_iterations=4, history_size=3, max_step=0.1)
    sol = SolutionCollectionSquare(torch.tensor(x, device="cuda"), torch.tensor(h, device="cuda", dtype=torch.float32),
                                   use_fixed_h=False)
    o = opt.engine_options(sol)
    assert o is not None and o.mode == "lut"
    r = opt.run_simulation(sol)
    OL = oracle_mod.OracleLut(coll._lut.X, coll._lut.Y, coll._lut.theta, coll._lut.lut_d.cpu().numpy(), mirror=True)
    rx, rh, _ = oracle_mod.relax(oracle_mod.OracleGeometry(geom), x.astype(np.float64), h, max_iter=4, m=3,
                                 max_step=0.1, lut=OL, w_ar=1e-2, fixed_h=False, full=True)
    assert np.abs(r.xyt.cpu().numpy() - rx).max() < 3e-6
    assert np.abs(r.h.cpu().numpy() - rh).max() < 3e-6


def test_run_simulation_list_one_ragged_batch():
    """Solutions of different tree counts are relaxed in one ragged batch: each result is
    bitwise what relaxing that solution alone gives, and is written back in place."""
    from packing_amd.pack_dynamics import OptimizerBFGS, run_simulation_list
    from packing_amd.solutions import SolutionCollectionSquare
    from gpu_util import dense_poses
    opt = OptimizerBFGS(n_iterations=20, history_size=4, max_step=0.1)
    sols = []
    for k, n in enumerate((7, 12, 12, 30)):
        x = torch.tensor(dense_poses(30 + k, 2, n), device="cuda")
        h = torch.tensor([[1.2 * n ** 0.5, 0.6 * n ** 0.5, 0.6 * n ** 0.5]] * 2, device="cuda", dtype=torch.float32)
        sols.append(SolutionCollectionSquare(x, h, use_fixed_h=k % 2 == 0))
    alone = [opt.run_simulation(s) for s in sols]
    calls = []
    orig = opt._run_engine
    opt._run_engine = lambda ss, o: calls.append(len(ss)) or orig(ss, o)
    run_simulation_list(opt, sols)
    assert sorted(calls) == [2, 2]          # one batch for fixed h, one for free h
    for s, a in zip(sols, alone):
        assert torch.equal(s.xyt, a.xyt) and torch.equal(s.h, a.h)


def test_pack_dynamics_line_search_vs_oracle(oracle_mod):
    """OptimizerBFGS(use_line_search=True) (the reference loop with the device strong-Wolfe
    search) on FP64 packings follows the oracle's strong-Wolfe L-BFGS over the FP64
    overlap + boundary cost for its first iterations."""
    from packing_amd import pack_cost as PC
    from packing_amd.context import set_geometry
    from packing_amd.pack_dynamics import OptimizerBFGS
    from packing_amd.solutions import SolutionCollectionSquare
    from gpu_util import GEOMS, dense_poses
    geom = GEOMS["tree"]()
    set_geometry(geom)
    og = oracle_mod.OracleGeometry(geom)
    N, E = 12, 4
    x = dense_poses(9, E, N).astype(np.float64)
    h = np.array([[1.2 * np.sqrt(N), 0.6 * np.sqrt(N), 0.6 * np.sqrt(N)]] * E)
    cost = PC.CostCompound(costs=[PC.BoundaryDistanceCost(), PC.CollisionCostOverlappingArea()])
    opt = OptimizerBFGS(cost=cost, n_iterations=4, history_size=3, max_step=0.1, use_line_search=True)
    sol = SolutionCollectionSquare(torch.tensor(x, device="cuda"), torch.tensor(h, device="cuda"), use_fixed_h=True)
    assert opt.engine_options(sol) is None
    r = opt.run_simulation(sol)

    def f(rows):
        xyt = rows[:, :3 * N].reshape(-1, N, 3)
        hh = rows[:, 3 * N:]
        co, go = oracle_mod.overlap(og, xyt, False)
        cb, gb, _ = oracle_mod.boundary(og, xyt, hh)
        g = np.zeros_like(rows)
        g[:, :3 * N] = (go + gb).reshape(len(rows), -1)
        return co + cb, g

    x0 = np.concatenate([x.reshape(E, -1), h], 1)
    ref = oracle_mod.lbfgs_wolfe(f, x0, max_iter=4, history_size=3, max_step=0.1, tolerance_grad=0.0,
                                 tolerance_change=0.0)
    assert np.abs(r.xyt.cpu().numpy().reshape(E, -1) - ref[:, :3 * N]).max() < 1e-9
    assert not np.allclose(ref[:, :3 * N], x0[:, :3 * N])
