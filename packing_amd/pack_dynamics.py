"""Drop-in for the reference module ``pack_dynamics`` (core/pack_dynamics.py) on MI355X.

``OptimizerBFGS.run_simulation(sol)`` relaxes [xyt | h] per solution with L-BFGS at
tolerance_grad = tolerance_change = 0, as core/pack_dynamics.py:54-149 does, and
``run_simulation_list`` relaxes a list of solutions in place (:152-236).

Two engines, chosen per call with no caller-visible switch:

* the fused relaxation engine (relax.RelaxBatch over pk_relax_step: cost and L-BFGS in
  one persistent kernel per instance, any mix of tree counts in one launch) whenever it
  computes what the reference computes: float32 square phenotype solutions, no line
  search, a cost that is a weighted sum of AreaCost, BoundaryDistanceCost and one
  collision term (exact area, SAT separation, or a pose table, quadratic or not), and no
  relative-change or cost-increase stop;
* otherwise the reference loop: the lbfgs_torch_parallel drop-in (device L-BFGS kernels)
  over the cost object, as the reference runs it.

Why the two agree: the reference's batched L-BFGS keeps stepping systems that have
stopped while any system is active (core/lbfgs_torch_parallel.py:626), the engine stops
each instance alone.  With both tolerances 0 the only stops are an exactly zero
gradient (the direction is then exactly zero, so further steps leave the poses as they
are), the evaluation limit (reached by every system in the same iteration) and
max_iter, so the final poses are the same either way.  A relative-change or
cost-increase stop would freeze one system while the reference moves it on, so those
settings take the reference loop.  The engine computes in float32 with its own reduction
order, so its poses agree with the loop's to float32 rounding, not bit for bit.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass

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

    def __post_init__(self):
        if self.cost is None:
            self.cost = pack_cost.CostCompound(costs=[pack_cost.AreaCost(scaling=1e-2),
                                                      pack_cost.BoundaryDistanceCost(scaling=1.0),
                                                      pack_cost.CollisionCostOverlappingArea(scaling=1.0)])

    # -------------------------------------------------------------- engine choice
    def _cost_terms(self):
        """(w_overlap, w_boundary, w_area, mode, table) when the cost is a weighted sum of
        at most one collision term, boundary and area terms, else None."""
        c = self.cost
        if isinstance(c, pack_cost.CostCompound):
            terms, s = c.costs, c.scaling
        else:
            terms, s = [c], 1.0
        w = {"ov": 0.0, "bd": 0.0, "ar": 0.0}
        coll = None
        for t in terms:
            if isinstance(t, pack_cost.CostCompound):
                return None
            if isinstance(t, pack_cost.AreaCost):
                w["ar"] += s * t.scaling
            elif isinstance(t, pack_cost.BoundaryDistanceCost):
                w["bd"] += s * t.scaling
            elif isinstance(t, pack_cost.CollisionCost):
                if coll is not None:
                    return None
                coll = t
                w["ov"] += s * t.scaling
            elif isinstance(t, pack_cost.CostDummy):
                continue
            else:
                return None
        mode, lut = "area", None
        if coll is not None:
            if coll.use_lookup_table:
                coll._ensure_lut_initialized()
                mode, lut = "lut", coll._lut
            elif coll.use_separation:
                mode = "sep"
        return w["ov"], w["bd"], w["ar"], mode, lut

    def engine_options(self, sol):
        """RelaxOptions for the fused engine when it reproduces the reference on sol, else
        None (see the module docstring)."""
        from .relax import RelaxOptions
        from .solutions import SolutionCollectionSquare
        if self.use_line_search or self.tolerance_rel_change != 0 or self.stop_on_cost_increase:
            return None
        if not (1 <= int(self.history_size) <= 16):
            return None
        if not (isinstance(sol, SolutionCollectionSquare) and sol.is_phenotype() and not sol.periodic):
            return None
        if sol.xyt is None or sol.xyt.dtype != torch.float32 or not sol.xyt.is_cuda:
            return None
        terms = self._cost_terms()
        if terms is None:
            return None
        w_ov, w_bd, w_ar, mode, lut = terms
        return RelaxOptions(mode=mode, lut=lut, w_overlap=w_ov, w_boundary=w_bd, w_area=w_ar,
                            fixed_h=bool(sol.use_fixed_h), history=int(self.history_size),
                            max_iter=int(self.n_iterations), max_step=self.max_step,
                            tol_grad=0.0, tol_change=0.0)

    # -------------------------------------------------------------- paths
    def _run_engine(self, sols, opts):
        """Relax every solution of sols (one ragged batch, one launch per size bucket and
        chunk); returns [(xyt, h)] per solution."""
        from .relax import RelaxBatch
        xs, hs, where = [], [], []
        for k, s in enumerate(sols):
            for e in range(s.N_solutions):
                xs.append(s.xyt[e])
                hs.append(s.h[e])
                where.append(k)
        b = RelaxBatch(xs, torch.stack(hs), opts, device=sols[0].xyt.device)
        b.run()
        out = [([], []) for _ in sols]
        for i, k in enumerate(where):
            o, n = int(b.xoff_np[i]), int(b.sizes[i])
            out[k][0].append(b.x[o:o + 3 * n].view(n, 3))
            out[k][1].append(b.x[o + 3 * n:o + 3 * n + 3])
        return [(torch.stack(a).contiguous(), torch.stack(c).contiguous()) for a, c in out]

    def _run_reference_loop(self, sol):
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

    def run_simulation(self, sol, engine: str = "auto"):
        """Relaxed deep copy of sol.  engine: "auto" (the fused engine where it applies),
        "reference" (always the reference loop)."""
        sol.check_constraints()
        sol = copy.deepcopy(sol)
        opts = self.engine_options(sol) if engine == "auto" else None
        if opts is not None and sol.N_solutions > 0:
            ((xyt, h),) = self._run_engine([sol], opts)
            sol.xyt, sol.h = xyt, h
            return sol
        return self._run_reference_loop(sol)


def run_simulation_list(simulator, solution_list):
    """Relax a list of solutions in place (core/pack_dynamics.py:152-236).  Solutions the
    fused engine takes are relaxed together, every tree count in one ragged batch (one
    batch per distinct option set, e.g. fixed and free h); the rest are grouped by tree
    count and relaxed one group at a time, as the reference does."""
    if len(solution_list) == 0:
        return
    rest = list(solution_list)
    if isinstance(simulator, OptimizerBFGS):
        batches: dict = {}
        rest = []
        for s in solution_list:
            s.check_constraints()
            o = simulator.engine_options(s)
            if o is None or s.N_solutions == 0:
                rest.append(s)
                continue
            key = (o.fixed_h, s.xyt.device)
            batches.setdefault(key, (o, []))[1].append(s)
        for o, sols in batches.values():
            for s, (xyt, h) in zip(sols, simulator._run_engine(sols, o)):
                s.xyt[:] = xyt
                s.h[:] = h
    groups: dict = {}
    for s in rest:
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
