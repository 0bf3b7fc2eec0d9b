"""Drop-in for the GPU side of the reference module ``pack_cost`` (core/pack_cost.py).

Cost objects keep the reference's names, fields and call signatures:
``compute_cost(sol, cost, grad_xyt, grad_bound, evaluate_gradient=True)`` and
``compute_cost_allocate(sol, evaluate_gradient=True)``, with phenotype expansion and
gradient backprop, scaling, zeroed h-gradient for fixed h (core/pack_cost.py:140-188),
compound sums (:245-277) and the periodic path with its finite-difference
self-image theta gradient and h gradient (:630-722).  The shapely/finite-difference CPU
references (``compute_cost_ref``) are not part of this package: the FP64 oracle under
oracle/ plays that role in the tests.  Tensors are torch on the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import torch

from . import pack_cuda, pack_cuda_lut


@dataclass
class Cost:
    scaling: float = 1.0

    def compute_cost_allocate(self, sol, evaluate_gradient: bool = True):
        cost = torch.zeros(sol.N_solutions, dtype=sol.xyt.dtype, device=sol.xyt.device)
        grad_xyt = torch.zeros_like(sol.xyt)
        grad_bound = torch.zeros_like(sol.h)
        self.compute_cost(sol, cost, grad_xyt, grad_bound, evaluate_gradient=evaluate_gradient)
        return cost, grad_xyt, grad_bound

    def compute_cost(self, sol, cost, grad_xyt, grad_bound, evaluate_gradient: bool = True):
        if not sol.is_phenotype():
            ph = sol.convert_to_phenotype()
            gx = torch.zeros_like(ph.xyt)
            gh = torch.zeros_like(ph.h)
            self.compute_cost(ph, cost, gx, gh, evaluate_gradient=evaluate_gradient)
            if evaluate_gradient:
                sol.backprop_phenotype(gx, gh, grad_xyt, grad_bound)
            return
        self._compute_cost(sol, cost, grad_xyt, grad_bound, evaluate_gradient)
        if self.scaling != 1.0:
            cost *= self.scaling
            if evaluate_gradient:
                grad_xyt *= self.scaling
                if sol.use_fixed_h:
                    grad_bound.zero_()
                else:
                    grad_bound *= self.scaling
        elif sol.use_fixed_h:
            grad_bound.zero_()

    def _compute_cost(self, sol, cost, grad_xyt, grad_bound, evaluate_gradient):
        raise NotImplementedError


@dataclass
class CostCompound(Cost):
    costs: list = field(default_factory=list)

    def _compute_cost(self, sol, cost, grad_xyt, grad_bound, evaluate_gradient):
        cost.zero_()
        grad_xyt.zero_()
        grad_bound.zero_()
        tc = torch.zeros_like(cost)
        tg = torch.zeros_like(grad_xyt)
        tb = torch.zeros_like(grad_bound)
        for c in self.costs:
            tc.zero_()
            tg.zero_()
            tb.zero_()
            c.compute_cost(sol, tc, tg, tb, evaluate_gradient)
            cost += tc
            grad_xyt += tg
            grad_bound += tb


@dataclass
class CostDummy(Cost):
    def _compute_cost(self, sol, cost, grad_xyt, grad_bound, evaluate_gradient):
        cost.zero_()
        grad_xyt.zero_()
        grad_bound.zero_()


@dataclass
class AreaCost(Cost):
    """h0^2 for squares, the cell area for lattices, 0 for fixed h (core/pack_cost.py:1250-1284)."""

    def _compute_cost(self, sol, cost, grad_xyt, grad_bound, evaluate_gradient):
        grad_xyt.zero_()
        if sol.use_fixed_h:
            cost.zero_()
            grad_bound.zero_()
        else:
            sol.compute_area(cost, grad_bound)


@dataclass
class BoundaryDistanceCost(Cost):
    """core/pack_cost.py:1287-1549 (kernel path)."""
    use_kernel: bool = True

    def _compute_cost(self, sol, cost, grad_xyt, grad_bound, evaluate_gradient):
        assert not sol.periodic
        if evaluate_gradient:
            pack_cuda.boundary_distance_multi_ensemble(sol.xyt, sol.h, out_cost=cost, out_grads=grad_xyt,
                                                       out_grad_h=grad_bound)
        else:
            pack_cuda.boundary_distance_multi_ensemble(sol.xyt, sol.h, out_cost=cost)


@dataclass
class CollisionCost(Cost):
    """Pairwise collision cost, exact kernel or pose table (core/pack_cost.py:297-722)."""
    use_lookup_table: bool = False
    lut_N_x: int = 900
    lut_N_y: int = 900
    lut_N_theta: int = 900
    lut_trim_zeros: bool = True
    _lut: object = field(default=None, repr=False)
    use_separation = False

    def _ensure_lut_initialized(self):
        if self.use_lookup_table and self._lut is None:
            if self.use_separation:
                raise NotImplementedError("tables are built for the overlap area and the exact separation "
                                          "(CollisionCostExactSeparation) only")
            # the reference tabulates the x >= 0 half (core/pack_cuda_lut.py:128-131), which
            # holds for mirror-symmetric polygons only; others get the full x range
            from .context import get_geometry
            full_x = not get_geometry().mirror_symmetric
            self._lut = pack_cuda_lut.LookupTable.build_exact(self.lut_N_x, self.lut_N_y, self.lut_N_theta,
                                                             full_x=full_x, trim_zeros=self.lut_trim_zeros)

    def _internal(self, sol, cost, grad_xyt, evaluate_gradient, crystal_axes=None, only_self=False):
        pack_cuda.overlap_multi_ensemble(sol.xyt, sol.xyt, self.use_separation, out_cost=cost,
                                         out_grads=grad_xyt if evaluate_gradient else None,
                                         crystal_axes=crystal_axes, only_self_interactions=only_self,
                                         N_periodic=sol.N_periodic)

    def _compute_cost(self, sol, cost, grad_xyt, grad_bound, evaluate_gradient):
        if not sol.periodic:
            if self.use_lookup_table:
                self._ensure_lut_initialized()
                pack_cuda_lut.overlap_multi_ensemble(sol.xyt, cost, self._lut,
                                                     grad_xyt if evaluate_gradient else None)
            else:
                self._internal(sol, cost, grad_xyt, evaluate_gradient)
            if evaluate_gradient:
                grad_bound.zero_()
            return
        axes = sol.get_crystal_axes().contiguous()
        self._internal(sol, cost, grad_xyt, evaluate_gradient, crystal_axes=axes)
        if not evaluate_gradient:
            return
        # self images carry no kernel gradient: theta by central differences (:637-676)
        eps = 1e-6
        cp_ = torch.zeros_like(cost)
        cm_ = torch.zeros_like(cost)
        for t in range(sol.N_trees):
            orig = sol.xyt.clone()
            sol.xyt[:, t, 2] += eps
            self._internal(sol, cp_, None, False, crystal_axes=axes, only_self=True)
            sol.xyt.copy_(orig)
            sol.xyt[:, t, 2] -= eps
            self._internal(sol, cm_, None, False, crystal_axes=axes, only_self=True)
            sol.xyt.copy_(orig)
            grad_xyt[:, t, 2] += (cp_ - cm_) / (2.0 * eps) / 2
        # h gradient by central differences of the whole periodic cost (:678-722)
        grad_bound.zero_()
        for k in range(sol.h.shape[1]):
            h0 = sol.h.clone()
            sol.h[:, k] += eps
            self._internal(sol, cp_, None, False, crystal_axes=sol.get_crystal_axes().contiguous())
            sol.h.copy_(h0)
            sol.h[:, k] -= eps
            self._internal(sol, cm_, None, False, crystal_axes=sol.get_crystal_axes().contiguous())
            sol.h.copy_(h0)
            grad_bound[:, k] = (cp_ - cm_) / (2.0 * eps)


@dataclass
class CollisionCostOverlappingArea(CollisionCost):
    """Summed overlap area (core/pack_cost.py:724-818)."""
    use_separation = False


@dataclass
class CollisionCostSeparation(CollisionCost):
    """Summed squared SAT separation (core/pack_cost.py:821-1116)."""
    use_separation = True


@dataclass
class CollisionCostExactSeparation(CollisionCost):
    """Collision cost from the exact Minkowski separation distance, table only
    (core/pack_cost.py:1119-1247): per pair max(0, sep)^2, sep the signed separation of
    the two polygons (penetration depth positive).  The table is computed on the GPU in
    double (pk_sep_build; the reference uses shapely on the CPU) on the reference's grid
    and cached on disk in the binary format of lut_cache, keyed by the grid parameters
    and the polygon's fingerprint (the reference keys its pickle by an md5 of the tree
    vertices, :1131-1142)."""
    use_separation = False

    def _cache_key(self) -> str:
        from .context import get_geometry
        return (f"exact_sep_Nx{self.lut_N_x}_Ny{self.lut_N_y}_Nt{self.lut_N_theta}_trim{self.lut_trim_zeros}"
                f"_tree{get_geometry().fingerprint()[:8]}")

    def _ensure_lut_initialized(self):
        if not self.use_lookup_table:
            raise ValueError(f"{self.__class__.__name__} requires use_lookup_table=True")
        if self._lut is not None:
            return
        import os
        from . import lut_cache
        from .context import get_geometry
        key = self._cache_key()
        path = os.path.join(lut_cache.cache_dir(), key + ".pklut")
        hit = lut_cache.load(path, key)
        if hit is not None:
            X, Y, T, vals, quad, mirror = hit
            self._lut = pack_cuda_lut.LookupTable(X=X, Y=Y, theta=T, vals=vals, apply_quadratic_transform=quad,
                                                  mirror_x=mirror)
            return
        full_x = not get_geometry().mirror_symmetric
        lut = pack_cuda_lut.LookupTable.build_separation(self.lut_N_x, self.lut_N_y, self.lut_N_theta,
                                                         full_x=full_x, trim_zeros=self.lut_trim_zeros)
        try:
            lut_cache.save(path, key, lut)
        except OSError:
            pass   # an unwritable cache directory only costs a rebuild next time
        self._lut = lut

    def _internal(self, sol, cost, grad_xyt, evaluate_gradient, crystal_axes=None, only_self=False):
        raise NotImplementedError(f"{self.__class__.__name__} has no exact kernel; use_lookup_table=True")

    def _compute_cost(self, sol, cost, grad_xyt, grad_bound, evaluate_gradient):
        assert self.use_lookup_table, f"{self.__class__.__name__} requires use_lookup_table=True"
        assert not sol.periodic, f"{self.__class__.__name__} does not support periodic boundaries"
        super()._compute_cost(sol, cost, grad_xyt, grad_bound, evaluate_gradient)
