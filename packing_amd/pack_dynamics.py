"""Drop-in for the reference module ``pack_dynamics`` (core/pack_dynamics.py) on MI355X.

``OptimizerBFGS.run_simulation(sol)`` flattens [xyt | h] per solution, runs the device
L-BFGS (lbfgs_torch_parallel drop-in) on the cost object, and unflattens, as
core/pack_dynamics.py:54-149 does (tolerance_grad = tolerance_change = 0).
``run_simulation_list`` groups solutions by tree count and relaxes each group as one
batch (:152-236).  ``OptimizerBFGS.fused`` selects the fused relaxation engine
(relax.RelaxBatch: cost and L-BFGS in one persistent kernel, any mix of tree counts in
one launch) when the cost is the standard overlap + boundary [+ area] composition on
square phenotype solutions.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field

import torch

from . import lbfgs_torch_parallel, pack_cost


@dataclass
class OptimizerBFGS:
    cost: object = None
    n_iterations: int = 100
    history_size: int = 3
    max_step: float = 0.01
    tolerance_rel_change: float = 0.0
    stop_on_cost_increase: bool = False
    use_line_search: bool = False
    fused: bool = False

    def __post_init__(self):
        if self.cost is None:
            self.cost = pack_cost.CostCompound(costs=[pack_cost.AreaCost(scaling=1e-2),
                                                      pack_cost.BoundaryDistanceCost(scaling=1.0),
                                                      pack_cost.CollisionCostOverlappingArea(scaling=1.0)])

    # -------------------------------------------------------------- fused path
    def _fused_terms(self, sol):
        """(w_overlap, w_boundary, w_area, mode) when the cost is overlap + boundary [+ area]
        on square phenotype solutions, else None."""
        from .solutions import SolutionCollectionSquare
        if not (isinstance(sol, SolutionCollectionSquare) and sol.is_phenotype()):
            return None
        terms = self.cost.costs if isinstance(self.cost, pack_cost.CostCompound) else [self.cost]
        s = self.cost.scaling if isinstance(self.cost, pack_cost.CostCompound) else 1.0
        w = {"ov": 0.0, "bd": 0.0, "ar": 0.0}
        mode = None
        for c in terms:
            if isinstance(c, pack_cost.AreaCost):
                w["ar"] += s * c.scaling
            elif isinstance(c, pack_cost.BoundaryDistanceCost):
                w["bd"] += s * c.scaling
            elif isinstance(c, pack_cost.CollisionCost) and not c.use_lookup_table:
                m = "sep" if c.use_separation else "area"
                if mode not in (None, m):
                    return None
                mode = m
                w["ov"] += s * c.scaling
            else:
                return None
        return w["ov"], w["bd"], w["ar"], mode or "area"

    def run_simulation(self, sol):
        sol.check_constraints()
        sol = copy.deepcopy(sol)
        terms = self._fused_terms(sol) if self.fused else None
        if terms is not None and not self.use_line_search and sol.xyt.dtype == torch.float32:
            from .relax import RelaxBatch, RelaxOptions
            w_ov, w_bd, w_ar, mode = terms
            opts = RelaxOptions(mode=mode, w_overlap=w_ov, w_boundary=w_bd, w_area=w_ar,
                                fixed_h=sol.use_fixed_h, history=self.history_size,
                                max_iter=self.n_iterations, max_step=self.max_step,
                                tol_rel=self.tolerance_rel_change,
                                stop_on_increase=self.stop_on_cost_increase)
            b = RelaxBatch(sol.xyt, sol.h, opts, device=sol.xyt.device)
            b.run()
            N = sol.xyt.shape[1]
            x = b.x.view(sol.N_solutions, 3 * N + 3)
            sol.xyt = x[:, :3 * N].reshape(sol.N_solutions, N, 3).contiguous()
            sol.h = x[:, 3 * N:].contiguous()
            return sol
        ns, nt = sol.N_solutions, sol.xyt.shape[1]
        split = nt * 3
        x0 = torch.cat([sol.xyt.reshape(ns, -1), sol.h.reshape(ns, -1)], dim=1).contiguous()
        tmp = copy.deepcopy(sol)
        cost = torch.zeros(ns, dtype=sol.xyt.dtype, device=sol.xyt.device)
        gx = torch.zeros_like(sol.xyt)
        gh = torch.zeros_like(sol.h)

        def f_torch(x, is_main_loop):
            tmp.xyt = x[:, :split].reshape(ns, nt, 3).contiguous()
            tmp.h = x[:, split:].reshape(ns, -1).contiguous()
            self.cost.compute_cost(tmp, cost, gx, gh)
            return cost.clone(), torch.cat([gx.reshape(ns, -1), gh.reshape(ns, -1)], dim=1)

        res = lbfgs_torch_parallel.lbfgs(f_torch, x0, tolerance_grad=0, tolerance_change=0,
                                         tolerance_rel_change=self.tolerance_rel_change,
                                         max_iter=self.n_iterations, history_size=self.history_size,
                                         max_step=self.max_step,
                                         line_search_fn="strong_wolfe" if self.use_line_search else None,
                                         stop_on_cost_increase=self.stop_on_cost_increase)
        sol.xyt = res[:, :split].reshape(ns, nt, 3).contiguous()
        sol.h = res[:, split:].reshape(ns, -1).contiguous()
        return sol


def run_simulation_list(simulator, solution_list):
    """Group by tree count, relax each group as one batch, scatter back in place
    (core/pack_dynamics.py:152-236)."""
    if len(solution_list) == 0:
        return
    groups: dict = {}
    for s in solution_list:
        groups.setdefault(s.xyt.shape[1], []).append(s)
    for _, grp in groups.items():
        if len(grp) == 1:
            r = simulator.run_simulation(grp[0])
            grp[0].xyt[:] = r.xyt
            grp[0].h[:] = r.h
            continue
        merged = copy.deepcopy(grp[0])
        merged.xyt = torch.cat([s.xyt for s in grp], 0)
        merged.h = torch.cat([s.h for s in grp], 0)
        r = simulator.run_simulation(merged)
        k = 0
        for s in grp:
            n = s.xyt.shape[0]
            s.xyt[:] = r.xyt[k:k + n]
            s.h[:] = r.h[k:k + n]
            k += n
