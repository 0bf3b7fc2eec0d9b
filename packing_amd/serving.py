"""Continuous-batching relaxation server on one GPU.

Requests (initial poses + box, in host memory) arrive over time.  The server keeps a
device slot table (fixed-capacity slots of up to ``n_max`` trees: state, gradient, L-BFGS
history and optimizer record in HBM) and admits a request into a free slot when its
memory fits the HBM budget: its poses are staged in pinned host memory then and copied
to the device with the next step (one copy per run of consecutive slots), inside its
latency.  A step advances every occupied slot by ``chunk`` iterations of the fused
engine (pk_relax_step), one launch per size class (the register rows R = ceil((3N + 3)
/ 64) of the one-wave kernel: R <= 3, 4-5, 6-8, 9-10, and larger), each with the kernel
for its own largest request, on forked streams that run concurrently.  Two steps are
kept in flight, so the GPU never waits for the host to retire and admit: a request
admitted while step k runs joins step k + 1 or k + 2.  Finished slots are retired from
the status copy of the step that finished them; results and the FP64 feasibility check
(pk_feasible) come back to the host on a separate stream while the next step runs, the
host polling for them instead of waiting, and a request's latency ends when its result
is in host memory.

Each slot runs exactly the computation it would run alone (per-instance L-BFGS,
fixed-point reductions), so a served result is bitwise the same as relaxing the request
by itself, whatever it was batched with.
"""
from __future__ import annotations

import ctypes as C
import time
from collections import deque
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as N
from .context import device_geometry, get_geometry
from .relax import MODES, RelaxOptions


@dataclass
class Request:
    rid: int
    xyt: np.ndarray          # (N, 3) float32
    h: np.ndarray            # (3,)
    arrival: float = 0.0     # seconds from the start of the session
    t_admit: float = -1.0
    t_done: float = -1.0
    slot: int = -1
    result: torch.Tensor | None = None    # (N, 3) float32, host memory
    loss: float = float("nan")
    overlapping_pairs: int = -1


def check_request(req: Request, n_max: int) -> None:
    """Raise ValueError unless the request's poses are (N, 3) with 1 <= N <= n_max and its
    box is 3 numbers."""
    shp = tuple(np.shape(req.xyt))
    if len(shp) != 2 or shp[1] != 3 or not (1 <= shp[0] <= n_max):
        raise ValueError(f"request {req.rid}: poses of shape {shp}; a slot takes (N, 3) with "
                         f"1 <= N <= n_max = {n_max}")
    if np.size(req.h) != 3:
        raise ValueError(f"request {req.rid}: box h must hold 3 numbers")


def slot_bytes(n_max: int, history: int) -> int:
    """HBM bytes one slot holds: x, g (n), s and y history (history * n), record."""
    n = 3 * n_max + 3
    return 4 * n * (2 + 2 * history) + N.PK_LBFGS_SLOT_BYTES + 16


# Size classes by register rows R of the one-wave kernel (upper bounds; beyond the last,
# one class for the workgroup kernel).
R_CLASS_MAX = (3, 5, 8, 10)


def size_class(n_trees: int) -> int:
    r = -(-(3 * n_trees + 3) // 64)
    for k, hi in enumerate(R_CLASS_MAX):
        if r <= hi:
            return k
    return len(R_CLASS_MAX)


class _Step:
    """One step in flight: its completion event, the status copy it ends with, and the
    request of every slot it advanced."""
    __slots__ = ("event", "status_h", "slots")

    def __init__(self, event, status_h, slots):
        self.event, self.status_h, self.slots = event, status_h, slots


class RelaxServer:
    """Continuous-batching server (see the module docstring).  chunk: iterations per step
    (100 measured at least as good as 50 for p50/p99 at 2000 requests/s and better for
    capacity; 25 and 200 worse)."""

    def __init__(self, opts: RelaxOptions, n_slots: int = 4096, n_max: int = 200,
                 hbm_budget_bytes: float = 64e9, chunk: int = 100, device="cuda", geometry=None,
                 in_flight: int = 2):
        self.opts = opts
        self.dev = torch.device(device)
        self.geometry = geometry or get_geometry()
        self.n_max = n_max
        per = slot_bytes(n_max, opts.history)
        if per * n_slots > hbm_budget_bytes:
            n_slots = int(hbm_budget_bytes // per)
        if n_slots < 1:
            raise MemoryError("HBM budget below one slot")
        self.S = n_slots
        self.budget = hbm_budget_bytes
        self.chunk = chunk
        self.in_flight = max(1, int(in_flight))
        n = 3 * n_max + 3
        self.stride = n
        dev = self.dev
        m = opts.history
        self.x = torch.zeros(n_slots * n, dtype=torch.float32, device=dev)
        self.g = torch.zeros_like(self.x)
        self.sh = torch.zeros(m * n_slots * n, dtype=torch.float32, device=dev)
        self.yh = torch.zeros_like(self.sh)
        self.state = torch.zeros(n_slots * N.PK_LBFGS_SLOT_BYTES, dtype=torch.uint8, device=dev)
        self.xoff = torch.arange(n_slots, dtype=torch.int32, device=dev) * n
        self.ntree = torch.zeros(n_slots, dtype=torch.int32, device=dev)
        self.status = torch.zeros(n_slots, dtype=torch.int32, device=dev)
        self.loss = torch.zeros(n_slots, dtype=torch.float32, device=dev)
        self.order = torch.zeros((self.in_flight + 1) * n_slots, dtype=torch.int32, device=dev)
        self.stats = torch.zeros(2, dtype=torch.int64, device=dev)
        # host staging of admitted poses (slot s at s * stride) and of the per-step lists
        self.x_host = torch.zeros(n_slots * n, dtype=torch.float32).pin_memory()
        self.order_h = torch.zeros((self.in_flight + 1) * n_slots, dtype=torch.int32).pin_memory()
        # one admission buffer per step that can be queued: a step's non-blocking copy reads
        # its buffer when the stream reaches it, so the next step must not overwrite it
        self.adm_h = torch.zeros((self.in_flight + 1, 2 * n_slots), dtype=torch.int64).pin_memory()
        self.status_h = [torch.zeros(n_slots, dtype=torch.int32).pin_memory() for _ in range(self.in_flight + 1)]
        self.stream = torch.cuda.Stream(device=dev)
        self.side = [torch.cuda.Stream(device=dev) for _ in range(len(R_CLASS_MAX))]
        self.retire_stream = torch.cuda.Stream(device=dev)
        self._dg = device_geometry(dev, self.geometry)
        self.free = list(range(n_slots))[::-1]
        self.busy: dict[int, Request] = {}
        self.pending: list[int] = []       # admitted slots whose device reset is not yet enqueued
        self.bytes_in_use = 0
        self.per_slot = per
        self._steps: deque = deque()
        self._retiring: deque = deque()    # retirements whose results are still on their way
        self._k = 0                        # steps launched (selects the staging buffers)

    # ---------------------------------------------------------------- admission
    def check_request(self, req: Request) -> None:
        """A slot holds at most n_max trees: a larger request would spill its state and
        history into the next slot, so it is refused."""
        check_request(req, self.n_max)

    def _admit(self, req: Request, now: float) -> None:
        """Take a free slot and stage the request's poses and box in the slot's pinned host
        region; the host -> device copy is enqueued with the next step (_reset_pending)."""
        self.check_request(req)
        s = self.free.pop()
        n = req.xyt.shape[0]
        o = s * self.stride
        xh = self.x_host[o:o + 3 * n + 3].numpy()
        xh[:3 * n] = np.asarray(req.xyt, dtype=np.float32).reshape(-1)
        xh[3 * n:] = np.asarray(req.h, dtype=np.float32).reshape(-1)
        req.slot = s
        req.t_admit = now
        self.busy[s] = req
        self.pending.append(s)
        self.bytes_in_use += self.per_slot

    def _reset_pending(self) -> None:
        """The slots admitted since the last step: their staged poses copied to the device
        (one copy per run of consecutive slots; a slot's whole region is copied, the part
        beyond its 3N + 3 values is never read), fresh optimizer records, status and tree
        counts (one index copy for all of them)."""
        if not self.pending:
            return
        k = len(self.pending)
        buf = self.adm_h[self._k % (self.in_flight + 1)]
        a = buf.numpy()
        a[:k] = self.pending
        a[k:2 * k] = [self.busy[s].xyt.shape[0] for s in self.pending]
        runs = np.sort(np.asarray(self.pending))
        cut = np.flatnonzero(np.diff(runs) != 1) + 1
        self.pending = []
        with torch.cuda.stream(self.stream):
            w = self.stride
            for r in np.split(runs, cut):
                lo, hi = int(r[0]) * w, (int(r[-1]) + 1) * w
                self.x[lo:hi].copy_(self.x_host[lo:hi], non_blocking=True)
            d = buf[:2 * k].to(self.dev, non_blocking=True)
            idx, nt = d[:k], d[k:].to(torch.int32)
            self.state.view(self.S, -1).index_fill_(0, idx, 0)
            self.status.index_fill_(0, idx, 0)
            self.ntree.index_copy_(0, idx, nt)

    def _params(self, order_ptr: int, n_launch: int, n_tree_max: int) -> N.RelaxParams:
        o = self.opts
        p = N.RelaxParams()
        p.n_slots = self.S
        p.xoff = self.xoff.data_ptr()
        p.ntree = self.ntree.data_ptr()
        p.x, p.g = self.x.data_ptr(), self.g.data_ptr()
        p.s_hist, p.y_hist = self.sh.data_ptr(), self.yh.data_ptr()
        p.state = self.state.data_ptr()
        p.order = order_ptr
        p.n_launch = n_launch
        p.n_tree_max = n_tree_max
        p.mode = MODES[o.mode]
        p.w_overlap, p.w_boundary, p.w_area = o.w_overlap, o.w_boundary, o.w_area
        p.fixed_h = int(o.fixed_h)
        if o.mode == "lut":
            p.lut = o.lut.handle(self.dev)
        p.history, p.max_iter = o.history, o.max_iter
        p.max_eval = 0 if o.max_eval is None else o.max_eval
        p.lr = o.lr
        p.max_step = 0.0 if o.max_step is None else o.max_step
        p.tol_grad, p.tol_change, p.tol_rel = o.tol_grad, o.tol_change, o.tol_rel
        p.stop_on_increase = int(o.stop_on_increase)
        p.iters = self.chunk
        from .autotune import choose_block
        p.block = o.block or choose_block(o.mode, n_tree_max, self.geometry.n_pieces, o.history)
        p.status_out, p.loss_out, p.stats = self.status.data_ptr(), self.loss.data_ptr(), self.stats.data_ptr()
        return p

    def _launch(self) -> None:
        """One step: every busy slot advanced by chunk iterations, one launch per size
        class on forked streams, then the status copy this step is retired from."""
        self._reset_pending()
        slots = list(self.busy)
        sizes = np.array([self.busy[s].xyt.shape[0] for s in slots])
        cls = np.array([size_class(int(n)) for n in sizes])
        buf = self._k % (self.in_flight + 1)
        base = buf * self.S
        oh = self.order_h.numpy()
        groups, pos = [], base
        for c in sorted(set(cls.tolist()), reverse=True):
            sel = np.where(cls == c)[0]
            sel = sel[np.argsort(-sizes[sel], kind="stable")]
            oh[pos:pos + len(sel)] = [slots[i] for i in sel]
            groups.append((pos, len(sel), int(sizes[sel].max())))
            pos += len(sel)
        main = self.stream
        lib = N.lib()
        with torch.cuda.stream(main):
            self.order[base:pos].copy_(self.order_h[base:pos], non_blocking=True)
            fork = torch.cuda.Event()
            fork.record(main)
        joins = []
        for k, (off, cnt, nmax) in enumerate(groups):
            st = main if k == 0 else self.side[k - 1]
            if k:
                st.wait_event(fork)
            p = self._params(self.order.data_ptr() + 4 * off, cnt, nmax)
            N.check(lib.pk_relax_step(self._dg.handle, C.byref(p), st.cuda_stream), "pk_relax_step")
            if k:
                e = torch.cuda.Event()
                e.record(st)
                joins.append(e)
        for e in joins:
            main.wait_event(e)
        sh = self.status_h[buf]
        with torch.cuda.stream(main):
            sh.copy_(self.status, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(main)
        self._steps.append(_Step(ev, sh, {s: self.busy[s] for s in slots}))
        self._k += 1

    def _retire_begin(self, step: _Step) -> None:
        """Retire the slots that the finished step left done: their results and the FP64
        feasibility check are enqueued on the retire stream (the next step keeps running)
        into pinned host buffers, with an event that _retire_end waits for; the host is not
        blocked.  The slots stay allocated until then (a new request could otherwise
        overwrite a result before it is copied out)."""
        st = step.status_h.numpy()
        bad = [s for s in step.slots if st[s] < 0]
        if bad:
            raise RuntimeError(f"relaxation refused slots {bad[:8]}: tree count above the launch's bound")
        fin = [s for s, r in step.slots.items() if st[s] == 1 and self.busy.get(s) is r]
        if not fin:
            return
        reqs = [self.busy.pop(s) for s in fin]
        from .feasibility import overlapping_pairs
        rs = self.retire_stream
        rs.wait_event(step.event)
        sizes = [r.xyt.shape[0] for r in reqs]
        k = len(fin)
        # slot ids and instance offsets through pinned memory: the copies stay asynchronous
        meta = torch.empty(2 * k + 1, dtype=torch.int64, pin_memory=True)
        mv = meta.numpy()
        mv[:k] = fin
        mv[k:] = np.concatenate([[0], np.cumsum(sizes)])
        host = torch.empty(3 * int(mv[-1]), dtype=torch.float32, pin_memory=True)
        pairs_h = torch.empty(k, dtype=torch.int32, pin_memory=True)
        loss_h = torch.empty(k, dtype=torch.float32, pin_memory=True)
        with torch.cuda.stream(rs):
            # done slots are never written again by a step, so reading them races with nothing
            flat = torch.cat([self.x[s * self.stride:s * self.stride + 3 * n] for s, n in zip(fin, sizes)])
            md = meta.to(self.dev, non_blocking=True)
            off = md[k:].to(torch.int32)
            pairs = overlapping_pairs(flat.view(-1, 3), off, max(sizes), stream=rs, geometry=self.geometry)
            host.copy_(flat, non_blocking=True)
            pairs_h.copy_(pairs, non_blocking=True)
            loss_h.copy_(self.loss[md[:k]], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(rs)
        self._retiring.append((ev, fin, reqs, sizes, host, pairs_h, loss_h, meta))

    def _retire_end(self, t0: float, done_list: list) -> None:
        """Hand the oldest retirement's results to its requests (its event has completed)
        and free their slots."""
        ev, fin, reqs, sizes, host, pairs_h, loss_h, _ = self._retiring.popleft()
        now = time.perf_counter() - t0
        k = 0
        for j, (s, r) in enumerate(zip(fin, reqs)):
            n = sizes[j]
            r.result = host[k:k + 3 * n].view(n, 3)
            k += 3 * n
            r.t_done = now
            r.overlapping_pairs = int(pairs_h[j])
            r.loss = float(loss_h[j])
            self.free.append(s)
            self.bytes_in_use -= self.per_slot
        done_list.extend(reqs)

    # ---------------------------------------------------------------- session
    def serve(self, requests: list, realtime: bool = True, poll_s: float = 2e-5) -> dict:
        """Run a session: requests sorted by arrival (seconds from start).  With realtime a
        request is admitted only once its arrival time has passed (latency = arrival to
        result in host memory); without, all are available at the start (saturation:
        throughput = capacity; latency = admission to result)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            self.check_request(r)
        torch.cuda.synchronize()
        done: list = []
        k = 0
        n_launch = 0
        t0 = time.perf_counter()
        while len(done) < len(reqs):
            now = time.perf_counter() - t0
            while k < len(reqs) and self.free and (not realtime or reqs[k].arrival <= now) \
                    and self.bytes_in_use + self.per_slot <= self.budget:
                self._admit(reqs[k], now)
                k += 1
            if len(self._steps) < self.in_flight and self.busy:
                self._launch()
                n_launch += 1
                continue
            if self._steps and self._steps[0].event.query():
                self._retire_begin(self._steps.popleft())
                continue
            if self._retiring and self._retiring[0][0].query():
                self._retire_end(t0, done)
                continue
            if self._steps or self._retiring:
                time.sleep(poll_s)
            elif k < len(reqs):
                time.sleep(max(0.0, min(reqs[k].arrival - (time.perf_counter() - t0), 0.01)))
        while self._steps:                       # steps launched after the last retirement
            self._steps.popleft().event.synchronize()
        wall = time.perf_counter() - t0
        arr = np.array([r.arrival if realtime else r.t_admit for r in done])
        lat = np.array([r.t_done for r in done]) - arr
        wait = np.array([r.t_admit for r in done]) - arr
        svc = np.array([r.t_done - r.t_admit for r in done])
        return {"requests": len(done), "wall_s": wall, "throughput": len(done) / wall,
                "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "queue_p50_s": float(np.percentile(wait, 50)), "queue_p99_s": float(np.percentile(wait, 99)),
                "service_p50_s": float(np.percentile(svc, 50)), "service_p99_s": float(np.percentile(svc, 99)),
                "launches": n_launch, "feasible_fraction": float(np.mean([r.overlapping_pairs == 0 for r in done])),
                "slots": self.S, "slot_bytes": self.per_slot, "slot_table_bytes": self.S * self.per_slot}


def serve_plain(requests: list, opts: RelaxOptions, realtime: bool = True, geometry=None, device="cuda") -> dict:
    """Baseline: requests one at a time in arrival order, each relaxed alone with one
    launch per iteration (the plain path)."""
    from .feasibility import overlapping_pairs
    from .relax import RelaxBatch
    reqs = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    for r in reqs:
        now = time.perf_counter() - t0
        if realtime and r.arrival > now:
            time.sleep(r.arrival - now)
        b = RelaxBatch([r.xyt], r.h[None], opts, geometry=geometry, device=device)
        r.t_admit = time.perf_counter() - t0
        b.run_plain()
        r.overlapping_pairs = int(overlapping_pairs(b.xyt(0)[None].contiguous(), geometry=geometry)[0])
        r.result = b.xyt(0).cpu()
        r.t_done = time.perf_counter() - t0
    wall = time.perf_counter() - t0
    lat = np.array([r.t_done - r.arrival for r in reqs])
    return {"requests": len(reqs), "wall_s": wall, "throughput": len(reqs) / wall,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "feasible_fraction": float(np.mean([r.overlapping_pairs == 0 for r in reqs]))}


def poisson_requests(n: int, rate: float, n_lo: int = 10, n_hi: int = 200, seed_arrivals: int = 6,
                     seed_requests: int = 7) -> list:
    """Poisson arrivals at `rate` per second; N log-uniform in [n_lo, n_hi]; poses as the
    benchmark starts (x, y ~ U(0, 1.2 sqrt N), theta ~ U(0, 2 pi))."""
    from .geometry import box_for, make_poses
    ra = np.random.default_rng(seed_arrivals)
    rr = np.random.default_rng(seed_requests)
    t = np.cumsum(ra.exponential(1.0 / rate, size=n))
    out = []
    for i in range(n):
        N_ = int(round(np.exp(rr.uniform(np.log(n_lo), np.log(n_hi)))))
        out.append(Request(i, make_poses(rr, 1, N_, np.float32)[0], box_for(N_).astype(np.float32), float(t[i])))
    return out
