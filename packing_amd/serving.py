"""Continuous-batching relaxation server on one GPU.

Requests (initial poses + box) arrive over time.  The server keeps a device slot table
(fixed-capacity slots of up to ``n_max`` trees: state, gradient, L-BFGS history and
optimizer record in HBM) and admits a request into a free slot when its memory fits
the HBM budget.  One persistent launch advances every occupied slot by ``chunk``
iterations (the fused engine, pk_relax_step); slots finish independently and are
freed as soon as the host sees their status, so requests that arrive while others run
join the next launch.  Each slot runs exactly the computation it would run alone
(per-instance L-BFGS, fixed-point reductions), so a served result is bitwise the same
as relaxing the request by itself.

Final packings are checked in FP64 on the GPU (pk_feasible) as they complete.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

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
    result: np.ndarray | None = None
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


class RelaxServer:
    def __init__(self, opts: RelaxOptions, n_slots: int = 4096, n_max: int = 200,
                 hbm_budget_bytes: float = 64e9, chunk: int = 100, device="cuda", geometry=None):
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
        self.order = torch.zeros(n_slots, dtype=torch.int32, device=dev)
        self.stats = torch.zeros(2, dtype=torch.int64, device=dev)
        self.status_h = torch.zeros(n_slots, dtype=torch.int32).pin_memory()
        self.stream = torch.cuda.Stream(device=dev)
        self._dg = device_geometry(dev, self.geometry)
        self.free = list(range(n_slots))[::-1]
        self.busy: dict[int, Request] = {}
        self.bytes_in_use = 0
        self.per_slot = per

    # ---------------------------------------------------------------- admission
    def check_request(self, req: Request) -> None:
        """A slot holds at most n_max trees: a larger request would spill its state and
        history into the next slot, so it is refused."""
        check_request(req, self.n_max)

    def _admit(self, req: Request, staged: torch.Tensor, now: float) -> None:
        self.check_request(req)
        s = self.free.pop()
        n = req.xyt.shape[0]
        o = s * self.stride
        with torch.cuda.stream(self.stream):
            self.x[o:o + 3 * n].copy_(staged[:3 * n], non_blocking=True)
            self.x[o + 3 * n:o + 3 * n + 3].copy_(staged[3 * n:3 * n + 3], non_blocking=True)
            self.state[s * N.PK_LBFGS_SLOT_BYTES:(s + 1) * N.PK_LBFGS_SLOT_BYTES].zero_()
            self.status[s:s + 1].zero_()
            self.ntree[s:s + 1].fill_(n)
        req.slot = s
        req.t_admit = now
        self.busy[s] = req
        self.bytes_in_use += self.per_slot

    def _launch(self) -> None:
        slots = sorted(self.busy, key=lambda s: -self.busy[s].xyt.shape[0])
        ordh = torch.tensor(slots, dtype=torch.int32).pin_memory()
        o = self.opts
        with torch.cuda.stream(self.stream):
            self.order[:len(slots)].copy_(ordh, non_blocking=True)
        p = N.RelaxParams()
        p.n_slots = self.S
        p.xoff = self.xoff.data_ptr()
        p.ntree = self.ntree.data_ptr()
        p.x, p.g = self.x.data_ptr(), self.g.data_ptr()
        p.s_hist, p.y_hist = self.sh.data_ptr(), self.yh.data_ptr()
        p.state = self.state.data_ptr()
        p.order = self.order.data_ptr()
        p.n_launch = len(slots)
        p.n_tree_max = max(self.busy[s].xyt.shape[0] for s in slots)
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
        p.block = o.block or choose_block(o.mode, p.n_tree_max, self.geometry.n_pieces)
        p.status_out, p.loss_out, p.stats = self.status.data_ptr(), self.loss.data_ptr(), self.stats.data_ptr()
        N.check(N.lib().pk_relax_step(self._dg.handle, C.byref(p), self.stream.cuda_stream), "pk_relax_step")
        with torch.cuda.stream(self.stream):
            self.status_h.copy_(self.status, non_blocking=True)
        self._ev = torch.cuda.Event()
        self._ev.record(self.stream)
        self._keep = ordh

    def _retire(self, now: float, done_list: list) -> None:
        """Free every slot of the last launch whose status reads finished."""
        sh = self.status_h.numpy()
        fin = [s for s in self._launched if sh[s] == 1]
        if not fin:
            return
        reqs = [self.busy.pop(s) for s in fin]
        from .feasibility import overlapping_pairs
        with torch.cuda.stream(self.stream):
            res = [self.x[s * self.stride:s * self.stride + 3 * r.xyt.shape[0]].view(-1, 3).clone()
                   for s, r in zip(fin, reqs)]
            sizes = [r.xyt.shape[0] for r in reqs]
            off = torch.tensor(np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)).to(self.dev, non_blocking=True)
            pairs = overlapping_pairs(torch.cat(res, 0), off, max(sizes), stream=self.stream,
                                      geometry=self.geometry)
            idx = torch.tensor(fin, dtype=torch.int64).to(self.dev, non_blocking=True)
            losses = self.loss[idx]
            self.ntree[idx] = 0
        self.stream.synchronize()
        pairs = pairs.cpu().numpy()
        losses = losses.cpu().numpy()
        for k, (s, r) in enumerate(zip(fin, reqs)):
            r.t_done = now
            r.result = res[k]
            r.overlapping_pairs = int(pairs[k])
            r.loss = float(losses[k])
            self.free.append(s)
            self.bytes_in_use -= self.per_slot
        done_list.extend(reqs)

    # ---------------------------------------------------------------- session
    def serve(self, requests: list, realtime: bool = True, poll_s: float = 5e-5) -> dict:
        """Run a session: requests sorted by arrival (seconds from start).  With realtime
        the host admits a request only once its arrival time has passed; admissions made
        while a launch runs are queued on the stream and join the next launch."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            self.check_request(r)
        # inputs resident in HBM before the session starts
        staged = [torch.tensor(np.concatenate([r.xyt.reshape(-1), r.h]).astype(np.float32), device=self.dev)
                  for r in reqs]
        torch.cuda.synchronize()
        done: list = []
        k = 0
        t0 = time.perf_counter()
        inflight = False
        n_launch = 0
        while len(done) < len(reqs):
            now = time.perf_counter() - t0
            while k < len(reqs) and self.free and (not realtime or reqs[k].arrival <= now) \
                    and self.bytes_in_use + self.per_slot <= self.budget:
                self._admit(reqs[k], staged[k], now)
                k += 1
            if not inflight and self.busy:
                self._launched = list(self.busy)
                self._launch()
                n_launch += 1
                inflight = True
            if inflight:
                if self._ev.query():
                    inflight = False
                    self._retire(time.perf_counter() - t0, done)
                else:
                    time.sleep(poll_s)
            elif k < len(reqs):
                time.sleep(max(0.0, min(reqs[k].arrival - (time.perf_counter() - t0), 0.01)))
        wall = time.perf_counter() - t0
        lat = np.array([r.t_done - r.arrival for r in done]) if realtime else \
            np.array([r.t_done - r.t_admit for r in done])
        return {"requests": len(done), "wall_s": wall, "throughput": len(done) / wall,
                "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "launches": n_launch, "feasible_fraction": float(np.mean([r.overlapping_pairs == 0 for r in done])),
                "slots": self.S, "slot_bytes": self.per_slot}


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
        r.result = b.xyt(0).clone()
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
