"""Tile choice for the relaxation kernels, measured per size bucket.

Candidates: 1 = one wave per instance with the state and L-BFGS history in registers
(up to 212 trees, history <= 8), 2 = the same registers split over two waves (rows 0-4
and 5-9), or 64 / 128 / 256 threads per instance with the vectors in LDS.  Every candidate gives bitwise the same results (all reductions are fixed-order
or fixed-point), so the choice only affects speed.

``tune_batch`` times every candidate on each size bucket of a RelaxBatch (the bucket's
own instances, from the batch's initial state, HIP events on the launch stream) and
records the fastest in a table keyed by (mode, size class, pieces, history); the table
is saved as ``tiles.json`` next to this file and read by ``choose_block``.  A size
class is the register row count R = ceil((3N + 3) / 64) of the one-wave kernel up to
its limit, then powers of two.  Without an entry the library's own rule applies (block
0: the one-wave kernel where it fits, else 256 threads).
"""
from __future__ import annotations

import json
import os

BLOCKS = (1, 2, 64, 128, 256)
WAVE_MAX_TREES = 212
WAVE_MAX_HISTORY = 8

_HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(_HERE, "tiles.json")
_table: dict | None = None


def size_class(n: int) -> str:
    if n <= WAVE_MAX_TREES:
        return f"R{-(-(3 * n + 3) // 64)}"
    c = 256
    while c < n:
        c *= 2
    return f"N{c}"


def key(mode: str, n_max: int, n_pieces: int, history: int) -> str:
    return f"{mode}:{size_class(n_max)}:P{n_pieces}:m{history}"


def table() -> dict:
    global _table
    if _table is None:
        _table = {}
        if os.path.exists(TABLE_PATH):
            try:
                with open(TABLE_PATH) as f:
                    _table = json.load(f)
            except (OSError, ValueError):
                _table = {}
    return _table


def choose_block(mode: str, n_max: int, n_pieces: int, history: int = 8) -> int:
    """Measured tile for this bucket, or 0 (the library's rule) without a table entry."""
    v = table().get(key(mode, n_max, n_pieces, history))
    if isinstance(v, dict):
        v = v.get("block")
    return int(v) if v in BLOCKS else 0


def candidates(n_max: int, history: int):
    for blk in BLOCKS:
        if blk in (1, 2) and (n_max > WAVE_MAX_TREES or history > WAVE_MAX_HISTORY):
            continue
        yield blk


def tune_batch(batch, iters: int = 20, reps: int = 2, save: bool = False, stream=None, path=None) -> list:
    """Time every candidate tile on each size bucket of ``batch`` (a RelaxBatch): ``iters``
    iterations from the batch's initial state, best of ``reps``.  Records the winners in
    the table (and with ``save`` writes the table to ``path``, default tiles.json next to
    this file); leaves the batch reset.  Returns one {key, n_max, instances, ms: {block:
    ms}, block} per bucket."""
    import torch
    st = stream if stream is not None else torch.cuda.Stream(device=batch.device)
    o = batch.opts
    n_pieces = batch.geometry.n_pieces
    out = []
    with torch.cuda.stream(st):
        for k, (_, cnt, nmax) in enumerate(batch.buckets()):
            res = {}
            for blk in candidates(nmax, o.history):
                best = None
                for _ in range(reps + 1):       # the first launch loads the code object
                    batch.reset()
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(st)
                    batch.launch_bucket(k, iters, stream=st, block=blk)
                    e1.record(st)
                    st.synchronize()
                    ms = e0.elapsed_time(e1)
                    best = ms if best is None else min(best, ms)
                res[blk] = round(best, 4)
            win = min(res, key=res.get)
            kk = key(o.mode, nmax, n_pieces, o.history)
            table()[kk] = {"block": win, "ms": {str(b): v for b, v in res.items()}, "iters": iters,
                           "instances": int(cnt)}
            out.append({"key": kk, "n_max": int(nmax), "instances": int(cnt), "ms": res, "block": win})
    batch.reset()
    torch.cuda.synchronize()
    if save:
        with open(path or TABLE_PATH, "w") as f:
            json.dump(table(), f, indent=1, sort_keys=True)
            f.write("\n")
    return out
