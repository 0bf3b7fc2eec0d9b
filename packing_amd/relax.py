"""Batched relaxation engine over the fused HIP kernel (pk_relax_step).

One workgroup relaxes one instance: cost evaluation and the L-BFGS update run inside
the kernel, so a ragged batch of thousands of instances advances by ``iters``
iterations per launch (``iters = max_iter`` relaxes the whole batch in one launch).
Each instance follows core/lbfgs_torch_parallel.py:411-729 (no line search) on its
own, with the objective of core/pack_dynamics.py:93-123 (overlap + boundary [+ area],
h fixed by default), so its result is the same alone, in any batch and with any
workgroup size.

Paths:
  * ``RelaxBatch.run()``            batched: all instances per launch, chunked launches;
  * ``RelaxBatch.run_plain()``      plain: one launch per instance per iteration (the
                                    baseline; bitwise the same results);
  * ``RelaxBatch.capture_graph()``  the chunk launch captured as a HIP graph.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native as N
from .context import device_geometry, get_geometry

MODES = {"area": N.PK_MODE_AREA, "sep": N.PK_MODE_SEP, "lut": N.PK_MODE_LUT}


# Size buckets launched concurrently, one stream each: at most the 4 hardware queues a
# process gets (more were measured slower and are not exercised).
MAX_BUCKETS = 4


@dataclass
class RelaxOptions:
    mode: str = "area"
    w_overlap: float = 1.0
    w_boundary: float = 1.0
    w_area: float = 0.0
    fixed_h: bool = True
    history: int = 8
    max_iter: int = 500
    max_eval: int | None = None
    lr: float = 1.0
    max_step: float | None = 0.1
    tol_grad: float = 0.0
    tol_change: float = 0.0
    tol_rel: float = 0.0
    stop_on_increase: bool = False
    block: int = 0          # 0: tile table (autotune.py), else the library's choice
    buckets: int = 4        # size buckets launched concurrently (1: one launch; at most 4,
                            # the hardware queues of one process)
    hist_lds: int = -1      # L-BFGS history in LDS: -1 when it fits, 0 never, 1 always

    def __post_init__(self):
        if not 1 <= int(self.buckets or 1) <= MAX_BUCKETS:
            raise ValueError(f"buckets must be 1..{MAX_BUCKETS} (one stream per hardware queue), "
                             f"not {self.buckets}")
    lut: object = None      # pack_cuda_lut.LookupTable for mode "lut"


def state_size(n_trees) -> np.ndarray:
    return 3 * np.asarray(n_trees, dtype=np.int64) + 3


class RelaxBatch:
    """Device-resident state of a ragged batch of relaxations.

    xyt: list of (N_i, 3) arrays/tensors, or a (M, N, 3) array; h: (M, 3) boxes.
    """

    def __init__(self, xyt, h, opts: RelaxOptions | None = None, geometry=None, device="cuda",
                 capacity_trees: int | None = None):
        self.opts = opts or RelaxOptions()
        self.device = torch.device(device)
        self.geometry = geometry or get_geometry()
        if isinstance(xyt, (np.ndarray, torch.Tensor)) and np.ndim(xyt) == 3:
            xyt = [xyt[i] for i in range(xyt.shape[0])]
        sizes = np.array([int(a.shape[0]) for a in xyt], dtype=np.int64)
        self.M = len(sizes)
        self.sizes = sizes
        n = state_size(sizes)
        xoff = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
        if n.sum() >= 2 ** 31:
            raise ValueError("batch too large for int32 state offsets")
        self.n_state = int(n.sum())
        self.xoff_np = xoff
        h = np.asarray(h.cpu() if isinstance(h, torch.Tensor) else h, dtype=np.float32).reshape(self.M, 3)
        xs = np.empty(self.n_state, dtype=np.float32)
        for i, a in enumerate(xyt):
            a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
            o = xoff[i]
            xs[o:o + 3 * sizes[i]] = a.astype(np.float32).reshape(-1)
            xs[o + 3 * sizes[i]:o + 3 * sizes[i] + 3] = h[i]
        dev = self.device
        m = self.opts.history
        self.x = torch.tensor(xs, device=dev)
        self.g = torch.zeros_like(self.x)
        self.s_hist = torch.zeros(m * self.n_state, dtype=torch.float32, device=dev)
        self.y_hist = torch.zeros_like(self.s_hist)
        self.state = torch.zeros(self.M * N.PK_LBFGS_SLOT_BYTES, dtype=torch.uint8, device=dev)
        self.xoff = torch.tensor(xoff.astype(np.int32), device=dev)
        self.ntree = torch.tensor(sizes.astype(np.int32), device=dev)
        # longest-processing-time first: big instances start first
        order = np.argsort(-sizes, kind="stable").astype(np.int32)
        self.order = torch.tensor(order, device=dev)
        self.status = torch.zeros(self.M, dtype=torch.int32, device=dev)
        self.loss = torch.zeros(self.M, dtype=torch.float32, device=dev)
        self.stats = torch.zeros(2, dtype=torch.int64, device=dev)
        self.n_tree_max = int(sizes.max()) if self.M else 1
        self._dg = device_geometry(dev, self.geometry)
        self._graph = None
        self.x0 = self.x.clone()
        self._buckets = self._make_buckets(order)
        self._side = [torch.cuda.Stream(device=dev) for _ in self._buckets[1:]] if self.M else []

    def _make_buckets(self, order):
        """Split the size-sorted order into up to opts.buckets equal-count slices:
        (offset, count, n_tree_max).  Dynamic LDS is sized per launch by its largest
        instance, so a slice of small instances gets small workgroups and high occupancy."""
        if not 1 <= int(self.opts.buckets or 1) <= MAX_BUCKETS:
            raise ValueError(f"buckets must be 1..{MAX_BUCKETS} (one stream per hardware queue), "
                             f"not {self.opts.buckets}")
        k = max(1, min(int(self.opts.buckets or 1), self.M))
        if self.M == 0:
            return []
        per = -(-self.M // k)
        out = []
        for b in range(0, self.M, per):
            cnt = min(per, self.M - b)
            out.append((b, cnt, int(self.sizes[order[b]])))
        return out

    def reset(self) -> None:
        """Back to the initial poses with fresh optimizer state (device-side copies, no
        allocation; history needs no clearing since a fresh slot never reads it)."""
        torch.mul(self.x0, 1.0, out=self.x)   # a kernel, so it can sit in a captured graph
        self.state.zero_()
        self.status.zero_()
        self.stats.zero_()

    # ------------------------------------------------------------------ params
    def _params(self, iters, order=None, n_launch=None, n_tree_max=None, block=None) -> N.RelaxParams:
        o = self.opts
        p = N.RelaxParams()
        p.n_slots = self.M
        p.xoff = self.xoff.data_ptr()
        p.ntree = self.ntree.data_ptr()
        p.x = self.x.data_ptr()
        p.g = self.g.data_ptr()
        p.s_hist = self.s_hist.data_ptr()
        p.y_hist = self.y_hist.data_ptr()
        p.state = self.state.data_ptr()
        p.order = (self.order if order is None else order).data_ptr()
        p.n_launch = self.M if n_launch is None else n_launch
        p.n_tree_max = self.n_tree_max if n_tree_max is None else n_tree_max
        p.mode = MODES[o.mode]
        p.w_overlap, p.w_boundary, p.w_area = o.w_overlap, o.w_boundary, o.w_area
        p.fixed_h = int(o.fixed_h)
        if o.mode == "lut":
            if o.lut is None:
                raise ValueError("mode 'lut' needs opts.lut")
            p.lut = o.lut.handle(self.device)
        p.history = o.history
        p.max_iter = o.max_iter
        p.max_eval = 0 if o.max_eval is None else o.max_eval
        p.lr = o.lr
        p.max_step = 0.0 if o.max_step is None else o.max_step
        p.tol_grad, p.tol_change, p.tol_rel = o.tol_grad, o.tol_change, o.tol_rel
        p.stop_on_increase = int(o.stop_on_increase)
        p.iters = int(iters)
        p.block = int(self.opts.block if block is None else block) or self._auto_block()
        p.status_out = self.status.data_ptr()
        p.loss_out = self.loss.data_ptr()
        p.stats = self.stats.data_ptr()
        p.hist_lds = int(o.hist_lds)
        return p

    def _auto_block(self) -> int:
        from .autotune import choose_block
        return choose_block(self.opts.mode, self.n_tree_max, self.geometry.n_pieces, self.opts.history)

    def _launch_one(self, iters, stream, **kw) -> None:
        p = self._params(iters, **kw)
        N.check(N.lib().pk_relax_step(self._dg.handle, C.byref(p), N.stream_handle(stream)),
                "pk_relax_step")

    def launch(self, iters: int, stream=None, chunks: int = 1, **kw) -> None:
        """Advance by up to ``chunks`` x ``iters`` iterations.  Without explicit
        order/n_launch, each size bucket runs its ``chunks`` launches back to back on its
        own stream (the largest on ``stream``, the others on side streams forked from it),
        joined back once at the end: a bucket that finishes a chunk early starts its next
        one instead of waiting for the slowest bucket (a graph captures the fork and join).
        Instances are independent, so the results are those of ``chunks`` calls."""
        if kw or len(self._buckets) <= 1:
            for _ in range(chunks):
                self._launch_one(iters, stream, **kw)
            return
        main = stream if stream is not None else torch.cuda.current_stream(self.device)
        fork = torch.cuda.Event()
        fork.record(main)
        joins = []
        for k, (off, cnt, nmax) in enumerate(self._buckets):
            s = main if k == 0 else self._side[k - 1]
            if k:
                s.wait_event(fork)
            blk = int(self.opts.block) or self._block_for(nmax)
            for _ in range(chunks):
                self._launch_one(iters, s, order=self.order[off:off + cnt], n_launch=cnt,
                                 n_tree_max=nmax, block=blk)
            if k:
                e = torch.cuda.Event()
                e.record(s)
                joins.append(e)
        for e in joins:
            main.wait_event(e)

    def buckets(self):
        """Size buckets as (first position in the size order, count, largest tree count)."""
        return list(self._buckets)

    def launch_bucket(self, k: int, iters: int, stream=None, block: int | None = None) -> None:
        """Advance only the instances of size bucket k (one kernel launch); block: the
        tile (default: opts.block, else the measured tile table, else the library's rule)."""
        off, cnt, nmax = self._buckets[k]
        blk = int(block) if block is not None else (int(self.opts.block) or self._block_for(nmax))
        self._launch_one(iters, stream, order=self.order[off:off + cnt], n_launch=cnt, n_tree_max=nmax,
                         block=blk)

    def _block_for(self, n_max: int) -> int:
        from .autotune import choose_block
        return choose_block(self.opts.mode, n_max, self.geometry.n_pieces, self.opts.history)

    # ------------------------------------------------------------------ paths
    def run(self, iters_per_launch: int | None = None, stream=None) -> None:
        """Batched path: every instance in each launch, until max_iter."""
        k = iters_per_launch or self.opts.max_iter
        self.launch(k, stream=stream, chunks=-(-self.opts.max_iter // k))

    def run_noalloc(self, stream=None) -> None:
        """Relax the batch from its initial poses to the end in one call that allocates
        nothing: gradient, history and optimizer records in the library's static scratch
        (pk_relax_run_noalloc); poses, h, loss and status land in this batch's tensors.
        The batch's own g, history and records are not touched."""
        p = self._params(self.opts.max_iter)
        p.g = p.s_hist = p.y_hist = p.state = None
        p.stats = None
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        with torch.cuda.stream(st):
            torch.mul(self.x0, 1.0, out=self.x)   # a kernel: nothing is allocated
        N.check(N.lib().pk_relax_run_noalloc(self._dg.handle, C.byref(p), self.n_state,
                                             N.stream_handle(stream)), "pk_relax_run_noalloc")

    def run_plain(self, instances=None, stream=None) -> None:
        """Baseline path: one launch per instance per iteration."""
        idx = range(self.M) if instances is None else instances
        one = torch.empty(1, dtype=torch.int32, device=self.device)
        for i in idx:
            one.fill_(int(i))
            for _ in range(self.opts.max_iter):
                self.launch(1, order=one, n_launch=1, n_tree_max=int(self.sizes[i]), stream=stream)

    def capture_graph(self, iters_per_launch: int, stream=None, chunks: int = 1):
        """Capture ``chunks`` chunk launches (all instances, iters_per_launch iterations
        each; each size bucket's chunks back to back) as a graph; one replay advances the
        batch like launch(iters_per_launch, chunks=chunks)."""
        s = stream or torch.cuda.Stream(device=self.device)
        g = torch.cuda.CUDAGraph()
        # warm the kernel attribute / lazy state outside the capture on the same stream
        with torch.cuda.stream(s):
            p = self._params(iters_per_launch)
            p.iters = 0
            N.check(N.lib().pk_relax_step(self._dg.handle, C.byref(p), N.stream_handle(s)), "warmup")
            torch.cuda.current_stream().synchronize()
            with torch.cuda.graph(g, stream=s, capture_error_mode="relaxed"):
                self.launch(iters_per_launch, stream=s, chunks=chunks)
        self._graph = g
        return g

    # ------------------------------------------------------------------ results
    def done(self) -> torch.Tensor:
        return self.status.bool()

    def xyt(self, i: int) -> torch.Tensor:
        o = int(self.xoff_np[i])
        n = int(self.sizes[i])
        return self.x[o:o + 3 * n].view(n, 3)

    def xyt_list(self):
        return [self.xyt(i) for i in range(self.M)]

    def xyt_flat(self) -> torch.Tensor:
        """(sum N, 3) poses of all instances back to back."""
        idx = np.concatenate([np.arange(o, o + 3 * n) for o, n in zip(self.xoff_np, self.sizes)])
        return self.x[torch.tensor(idx, device=self.device)].view(-1, 3)

    def slot_losses(self) -> np.ndarray:
        """Per-instance (loss, initial loss) in double, as the stopping tests compare them."""
        raw = self.state.view(self.M, N.PK_LBFGS_SLOT_BYTES).cpu().numpy()
        return raw[:, 0:16].copy().view(np.float64)

    def slot_records(self) -> np.ndarray:
        """Per-instance (n_iter, n_eval, hist_len, status, reason) read back from the device."""
        raw = self.state.view(self.M, N.PK_LBFGS_SLOT_BYTES).cpu().numpy()
        ints = raw[:, 152:184].view(np.int32)
        return ints[:, [0, 1, 2, 4, 5]]
