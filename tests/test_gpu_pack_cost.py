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
    """OptimizerBFGS drop-in path and fused engine give close results after a few iterations."""
    from packing_amd import pack_cost as PC
    from packing_amd.pack_dynamics import OptimizerBFGS, run_simulation_list
    from packing_amd.solutions import SolutionCollectionSquare
    from gpu_util import dense_poses
    x = torch.tensor(dense_poses(2, 6, 12), device="cuda")
    h = torch.tensor([[3.0, 0.9, 0.9]] * 6, device="cuda")
    cost = PC.CostCompound(costs=[PC.BoundaryDistanceCost(), PC.CollisionCostOverlappingArea()])
    sol = SolutionCollectionSquare(x.clone(), h.clone(), use_fixed_h=True)
    a = OptimizerBFGS(cost=cost, n_iterations=3, history_size=4, max_step=0.1).run_simulation(sol)
    b = OptimizerBFGS(cost=cost, n_iterations=3, history_size=4, max_step=0.1, fused=True).run_simulation(sol)
    assert torch.allclose(a.xyt, b.xyt, atol=2e-5)
    assert torch.equal(a.h, h) and torch.equal(b.h, h)
    # run_simulation_list groups by tree count and writes back in place
    sols = [SolutionCollectionSquare(x[:2].clone(), h[:2].clone(), use_fixed_h=True),
            SolutionCollectionSquare(x[2:3, :7].contiguous().clone(), h[2:3].clone(), use_fixed_h=True)]
    before = [s.xyt.clone() for s in sols]
    run_simulation_list(OptimizerBFGS(cost=cost, n_iterations=5, max_step=0.1), sols)
    assert not torch.equal(before[0], sols[0].xyt)
