"""Tile choice for the relaxation kernels, measured per size bucket.

Candidates: 1 = one wave per instance with the state and L-BFGS history in registers
(up to 212 trees, history <= 8), 2 = the same registers split over two waves (rows 0-4
and 5-9), or 64 / 128 / 256 threads per instance with the vectors in LDS.  Every candidate gives bitwise the same results (all reductions are fixed-order
or fixed-point), so the choice only affects speed.

``tune_batch`` times every candidate on each size bucket of a RelaxBatch (the bucket's
own instances, from the batch's initial state, HIP events on the launch stream), then,
since the buckets of a batch run concurrently and share the GPU, picks each size class's
tile by the time of the batch's whole relaxation (all buckets at once, one class at a
time with the others at their current choice): a tile that holds an instance in half the registers at
a slightly longer launch of its own bucket leaves room for the other buckets (on the
4096-instance table workload the one-wave kernel for the N <= 103 bucket, 0.98 ms per
launch against 0.95 with two waves, took the batch from 12.4 to 11.4 ms).  The winners
go into a table keyed by (mode, size class, pieces, history), saved as ``tiles.json``
next to this file and read by ``choose_block``; each entry also records the bucket's own
times and the workgroups per CU of each tile (``pk_relax_occupancy``).  A size class is
the register row count R = ceil((3N + 3) / 64) of the one-wave kernel up to its limit,
then powers of two.  Without an entry the library's own rule applies (block
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


def occupancy(batch, n_max: int, blk: int) -> int:
    """Instances of this bucket's kernel for tile ``blk`` that one CU holds at once."""
    import ctypes as C
    import torch
    from . import _native as N
    from .relax import MODES
    o = batch.opts
    nb = C.c_int(0)
    with torch.cuda.device(batch.device):
        N.check(N.lib().pk_relax_occupancy(batch._dg.handle, MODES[o.mode], int(n_max), int(o.history),
                                           int(o.hist_lds), int(blk), C.byref(nb)), "pk_relax_occupancy")
    return max(1, nb.value)


def candidates(n_max: int, history: int):
    for blk in BLOCKS:
        if blk in (1, 2) and (n_max > WAVE_MAX_TREES or history > WAVE_MAX_HISTORY):
            continue
        yield blk


def tune_batch(batch, iters: int = 20, reps: int = 2, save: bool = False, stream=None, path=None,
               passes: int = 2) -> list:
    """Time every candidate tile on each size bucket of ``batch`` (a RelaxBatch, opts.block
    0): ``iters`` iterations from the batch's initial state, best of ``reps``, for the
    bucket alone, then (several buckets) the whole relaxation of the batch (opts.max_iter
    iterations in launches of ``iters``) with each size class's candidates in turn
    (``passes`` rounds over the classes).  Records the winners in the table (and with ``save`` writes the
    table to ``path``, default tiles.json next to this file); leaves the batch reset.
    Returns one {key, n_max, instances, ms: {block: ms}, batch_ms: {block: ms},
    blocks_per_cu: {block: n}, block} per bucket."""

    def timed(launch):
        best = None
        for _ in range(reps + 1):       # the first launch loads the code object
            batch.reset()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            launch()
            e1.record(st)
            st.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
        return round(best, 4)

    import torch
    st = stream if stream is not None else torch.cuda.Stream(device=batch.device)
    o = batch.opts
    n_pieces = batch.geometry.n_pieces
    out = []
    with torch.cuda.stream(st):
        for k, (_, cnt, nmax) in enumerate(batch.buckets()):
            res, occ = {}, {}
            for blk in candidates(nmax, o.history):
                occ[blk] = occupancy(batch, nmax, blk)
                res[blk] = timed(lambda: batch.launch_bucket(k, iters, stream=st, block=blk))
            win = min(res, key=res.get)
            kk = key(o.mode, nmax, n_pieces, o.history)
            table()[kk] = {"block": win, "ms": {str(b): v for b, v in res.items()},
                           "blocks_per_cu": {str(b): v for b, v in occ.items()}, "iters": iters,
                           "instances": int(cnt)}
            out.append({"key": kk, "n_max": int(nmax), "instances": int(cnt), "ms": res, "blocks_per_cu": occ,
                        "block": win})
        chunks = -(-o.max_iter // iters)        # the whole relaxation, as the batch runs it
        for _ in range(passes if len(out) > 1 else 0):
            done = set()
            for r in out:
                kk = r["key"]
                if kk in done:
                    continue
                done.add(kk)
                bres = {}
                for blk in r["ms"]:
                    table()[kk]["block"] = blk
                    bres[blk] = timed(lambda: batch.launch(iters, stream=st, chunks=chunks))
                win = min(bres, key=bres.get)
                table()[kk]["block"] = win
                table()[kk]["batch_ms"] = {str(b): v for b, v in bres.items()}
                for q in out:
                    if q["key"] == kk:
                        q["batch_ms"] = bres
                        q["block"] = win
    batch.reset()
    torch.cuda.synchronize()
    if save:
        with open(path or TABLE_PATH, "w") as f:
            json.dump(table(), f, indent=1, sort_keys=True)
            f.write("\n")
    return out
