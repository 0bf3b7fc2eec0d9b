"""Synthetic inputs of the benchmark configurations (seeds fixed here; no dataset).

* ragged batch: n_inst instances, N ~ U{n_lo..n_hi}, poses x, y ~ U(0, 1.2 sqrt N),
  theta ~ U(0, 2 pi), square box h = [1.2 sqrt N, 0.6 sqrt N, 0.6 sqrt N];
* uniform batch: n_ens instances of N trees, same pose and box rule.
"""
from __future__ import annotations

import numpy as np

from .geometry import box_for, make_poses


def ragged_workload(n_inst: int, n_lo: int = 10, n_hi: int = 200, seed: int = 3):
    """(sizes, list of (N_i, 3) float32 poses, (n_inst, 3) float32 boxes)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_lo, n_hi + 1, size=n_inst)
    xs = [make_poses(rng, 1, int(n), np.float32)[0] for n in sizes]
    hs = np.stack([box_for(int(n)) for n in sizes]).astype(np.float32)
    return sizes, xs, hs


def uniform_workload(n_ens: int, n_trees: int, seed: int, dtype=np.float32):
    """((n_ens, N, 3) poses, (n_ens, 3) boxes) in dtype."""
    x = make_poses(np.random.default_rng(seed), n_ens, n_trees, dtype)
    h = np.tile(box_for(n_trees), (n_ens, 1)).astype(dtype)
    return x, h
