"""Tile-shape choice for the relaxation kernels.

Candidates: 1 = one wave per instance with the state and L-BFGS history in registers
(up to 212 trees, history <= 8), or 64 / 128 / 256 threads per instance with the vectors
in LDS.  0 lets the library choose (1 where it applies, else 256).  ``choose_block`` reads
a table keyed by (mode, instance-size bucket); ``tune`` measures the candidates on the
device for a given batch and stores the winner in that table.  Every candidate gives
bitwise the same results (all reductions are fixed-order or fixed-point), so the choice
only affects speed.
"""
from __future__ import annotations

import json
import os
import time

BLOCKS = (1, 64, 128, 256)
WAVE_MAX_TREES = 212
BUCKETS = (32, 64, 128, 256, 512, 4096)

_HERE = os.path.dirname(os.path.abspath(__file__))
_TABLE_PATH = os.path.join(_HERE, "tiles.json")
_table: dict | None = None


def _bucket(n: int) -> int:
    for b in BUCKETS:
        if n <= b:
            return b
    return BUCKETS[-1]


def _default(mode: str, n: int, n_pieces: int) -> int:
    return 0


def table() -> dict:
    global _table
    if _table is None:
        _table = {}
        if os.path.exists(_TABLE_PATH):
            try:
                with open(_TABLE_PATH) as f:
                    _table = json.load(f)
            except (OSError, ValueError):
                _table = {}
    return _table


def choose_block(mode: str, n_max: int, n_pieces: int) -> int:
    key = f"{mode}:{_bucket(n_max)}:{n_pieces}"
    v = table().get(key)
    return int(v) if v in BLOCKS else _default(mode, n_max, n_pieces)


def tune(make_batch, mode: str, n_max: int, n_pieces: int, iters: int = 20, save: bool = False) -> dict:
    """Time ``iters`` fused iterations for each block size on fresh batches from
    ``make_batch()`` (a RelaxBatch factory); record and return {block: seconds}."""
    import torch
    res = {}
    for blk in BLOCKS:
        if blk == 1 and n_max > WAVE_MAX_TREES:
            continue
        b = make_batch()
        b.opts.block = blk
        b.launch(1)                      # first evaluation + one iteration (warm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.launch(iters)
        torch.cuda.synchronize()
        res[blk] = time.perf_counter() - t0
    best = min(res, key=res.get)
    table()[f"{mode}:{_bucket(n_max)}:{n_pieces}"] = best
    if save:
        with open(_TABLE_PATH, "w") as f:
            json.dump(table(), f, indent=1, sort_keys=True)
    return res
