"""Drop-in for the reference module ``lap_batch`` (core/lap_batch.py) on MI355X.

``solve_lap_batch`` runs the wave-per-problem Hungarian kernel (pk_lap_hungarian);
``compute_diversity_shortcut_kernel`` the fused transform + min-cost-row kernel
(pk_diversity).  Inputs and outputs are torch tensors on the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Literal

import numpy as np
import torch

from . import _native as N


@dataclass
class LAPConfig:
    """core/lap_batch.py:28-38"""
    algorithm: Literal["hungarian", "min_cost_row", "min_cost_col"] = "hungarian"
    use_diversity_kernel: bool = True


def solve_lap_batch(cost_matrices_gpu: torch.Tensor, config: LAPConfig | None = None, stream=None,
                    exact_int_total: bool = False):
    """Batched assignment (core/lap_batch.py:457-534).

    cost (B, N, N) float32/float64/int32 on the GPU, N <= 200 as in the reference.
    Returns (row_match int32 (B, N), total_costs float32 (B,)); for int32 input with
    ``exact_int_total`` the totals are the exact int64 sums instead.  The min-cost
    heuristics leave row_match at -1, as the reference does.
    """
    if config is None:
        config = LAPConfig()
    if cost_matrices_gpu.ndim != 3 or cost_matrices_gpu.shape[1] != cost_matrices_gpu.shape[2]:
        raise ValueError("cost matrices must be (B, N, N)")
    B, n, _ = cost_matrices_gpu.shape
    assert n <= 200, f"N={n} exceeds maximum of 200"
    dev = cost_matrices_gpu.device
    row_match = torch.full((B, n), -1, dtype=torch.int32, device=dev)
    if config.algorithm == "hungarian":
        if cost_matrices_gpu.dtype == torch.int32:
            c = cost_matrices_gpu.contiguous()
            code = N.PK_I32
        else:
            c = cost_matrices_gpu.to(torch.float32).contiguous()
            code = N.PK_F32
        tot32 = torch.empty(B, dtype=torch.float32, device=dev)
        tot = torch.empty(B, dtype=torch.int64, device=dev) if code == N.PK_I32 else tot32
        if B and n:
            N.check(N.lib().pk_lap_hungarian(code, c.data_ptr(), B, n, row_match.data_ptr(), tot.data_ptr(),
                                             tot32.data_ptr(), N.stream_handle(stream)), "pk_lap_hungarian")
        else:
            tot32.zero_()
        return row_match, (tot if (exact_int_total and code == N.PK_I32) else tot32)
    c = cost_matrices_gpu.to(torch.float32)
    if config.algorithm == "min_cost_row":
        return row_match, c.amin(dim=2).sum(dim=1).to(torch.float32)
    if config.algorithm == "min_cost_col":
        return row_match, c.amin(dim=1).sum(dim=1).to(torch.float32)
    raise ValueError(f"Unknown LAPConfig.algorithm={config.algorithm!r}")


def compute_diversity_shortcut_kernel(pop_xyt: torch.Tensor, ref_xyt: torch.Tensor, cos_a: float,
                                      sin_a: float, do_mirror: bool, stream=None) -> torch.Tensor:
    """(N_pop, N_ref) sums of nearest-tree distances (core/lap_batch.py:537-608)."""
    orig = pop_xyt.dtype
    pop = pop_xyt.to(torch.float32).contiguous()
    ref = ref_xyt.to(torch.float32).contiguous()
    n_pop, n_trees = pop.shape[0], pop.shape[1]
    n_ref = ref.shape[0]
    out = torch.empty(n_pop * n_ref, dtype=torch.float32, device=pop.device)
    N.check(N.lib().pk_diversity(pop.data_ptr(), ref.data_ptr(), n_pop, n_ref, n_trees,
                                 float(np.float32(cos_a)), float(np.float32(sin_a)), int(bool(do_mirror)),
                                 out.data_ptr(), N.stream_handle(stream)), "pk_diversity")
    out = out.view(n_pop, n_ref)
    return out.to(orig) if orig == torch.float64 else out
