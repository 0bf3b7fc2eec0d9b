"""Benchmark: ragged batched relaxation throughput on one MI355X.

Workload (the ragged batching configuration): 4096 instances of the seed-0 random
15-vertex polygon, N ~ U{10..200} trees each, poses x, y ~ U(0, 1.2 sqrt N),
theta ~ U(0, 2 pi), square h = [1.2 sqrt N, 0.6 sqrt N, 0.6 sqrt N] (fixed), cost =
pose-table overlap (256 x 256 x 128 float32 table built on the GPU from exact
clipping) + boundary term, 500 L-BFGS iterations (history 8, max_step 0.1, no line
search).  One step = relaxing the whole batch from the initial poses (reset + all
iterations, replayed from a HIP graph: each size bucket's chunked persistent launches back
to back on a stream of its own, forked once and joined once).

value = relaxations per second (instances x steps / time).  Printed as one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(n_inst, n_lo, n_hi, seed):
    from packing_amd.workloads import ragged_workload
    return ragged_workload(n_inst, n_lo, n_hi, seed)


# Arithmetic of the exact area kernel (add, sub, mul, reciprocal; a fused multiply-add
# counts 2), counted from the fixed, fully unrolled code of one evaluation:
#   per tree:            sincos (counted 20) + rotation of P*4 piece vertices (6 each)
#   per tree-pair test:  centre distance (5)
#   per piece-pair test: both piece bounding-circle centres and the distance (21)
#   per clipped piece pair: 623 (clip of A's edges against B (200, 88 of them FP64),
#       Green's sums and shape derivative over A's edges (124), clip of B against A
#       (152, 88 FP64), sums over B's edges (124), closing terms (8), relative vertices (8),
#       fixed-point conversion of cost and 6 gradient terms (7))
FLOP_TREE_BASE, FLOP_VERT, FLOP_TP, FLOP_PP_TEST, FLOP_PP = 20, 24, 5, 21, 623


def area_eval_flops(n_trees, n_pieces, tree_pairs, piece_pairs):
    """Algorithmic FLOP of one exact area + gradient evaluation over a batch."""
    return (n_trees * (FLOP_TREE_BASE + FLOP_VERT * n_pieces) + tree_pairs * FLOP_TP
            + tree_pairs * n_pieces * n_pieces * FLOP_PP_TEST + piece_pairs * FLOP_PP)


def overlap_kernel_roofline(geom, dev, n_ens=1024, n_trees=100, reps=10, seed=1):
    """Exact overlap area + gradient (pk_overlap) on the plain-method shape: n_ens instances
    of n_trees trees at the benchmark's initial pose distribution.  Kernel time from HIP
    events on the launch stream; FLOP from the kernel's own pair counters."""
    import torch
    from packing_amd import pack_cuda
    from packing_amd.geometry import make_poses
    x = make_poses(np.random.default_rng(seed), n_ens, n_trees, np.float32)
    xt = torch.tensor(x, device=dev)
    c = torch.empty(n_ens, device=dev)
    g = torch.empty_like(xt)
    st = torch.zeros(2, dtype=torch.int64, device=dev)
    s = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(s):
        pack_cuda.overlap_multi_ensemble(xt, xt, False, out_cost=c, out_grads=g, stream=s)
        s.synchronize()
        pack_cuda.overlap_multi_ensemble(xt, xt, False, out_cost=c, out_grads=g, stream=s, stats=st)
        s.synchronize()
        tp, pp = (int(v) for v in st.cpu().numpy())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            pack_cuda.overlap_multi_ensemble(xt, xt, False, out_cost=c, out_grads=g, stream=s)
        e1.record(s)
        s.synchronize()
    ms = e0.elapsed_time(e1) / reps
    fl = area_eval_flops(n_ens * n_trees, geom.n_pieces, tp, pp)
    tf = fl / (ms * 1e-3) / 1e12
    traffic, src = profiled_traffic("pk::k_overlap<float, 0, true, false, 256>")
    return {"kernel": "k_overlap<float, area, grad>", "bound": "valu", "traffic": traffic,
            "traffic_source": src,
            "workload": f"{n_ens} instances x {n_trees} trees, initial poses, random 15-vertex "
                        f"polygon ({geom.n_pieces} pieces)",
            "avg_launch_ms": round(ms, 4), "tree_pairs": tp, "piece_pairs": pp,
            "flop_per_launch": fl, "achieved": round(tf, 3), "unit": "TFLOP/s",
            "peak": 157.3, "frac": round(tf / 157.3, 4),
            "frac_of_nonpacked_fp32": round(tf / 78.6, 4)}


def serving_section(geom, lut, dev, n_req, n_plain, iters, rate=2000.0, n_sat=4096):
    """Continuous-batching server under Poisson arrivals (rate per second, N log-uniform
    in [10, 200], iters L-BFGS iterations per request), the same server at saturation
    (n_sat requests all available at the start), and the plain per-request path on the
    first n_plain requests of the Poisson stream."""
    from packing_amd.relax import RelaxOptions
    from packing_amd.serving import RelaxServer, poisson_requests, serve_plain
    opts = RelaxOptions(mode="lut", lut=lut, max_iter=iters, history=8, max_step=0.1)
    srv = RelaxServer(opts, n_slots=4096, n_max=200, hbm_budget_bytes=64e9, chunk=100, device=dev,
                      geometry=geom)
    warm = poisson_requests(64, rate=rate, seed_arrivals=1, seed_requests=2)
    srv.serve(warm, realtime=False)
    log("serving: warm-up session done")
    reqs = poisson_requests(n_req, rate=rate)
    rep = srv.serve(reqs)
    log(f"serving: Poisson session done ({rep['requests']} requests, {rep['wall_s']:.2f} s)")
    ms = lambda v: round(v * 1e3, 3)
    out = {"requests": rep["requests"], "rate_per_s": rate, "iters": iters,
           "throughput": round(rep["throughput"], 2), "p50_ms": ms(rep["p50_s"]), "p99_ms": ms(rep["p99_s"]),
           # latency = arrival -> result in host memory: queue (arrival -> admission with
           # the host-to-device copy) + service (admission -> result back on the host)
           "queue_p50_ms": ms(rep["queue_p50_s"]), "queue_p99_ms": ms(rep["queue_p99_s"]),
           "service_p50_ms": ms(rep["service_p50_s"]), "service_p99_ms": ms(rep["service_p99_s"]),
           "launch_steps": rep["launches"], "steps_in_flight": srv.in_flight,
           "feasible_fraction_fp64": rep["feasible_fraction"], "slots": rep["slots"],
           "slot_table_gb": round(rep["slot_table_bytes"] / 1e9, 3), "hbm_budget_gb": 64}
    # saturation: every request available at the start, so throughput is the server's capacity
    sat = poisson_requests(n_sat, rate=rate, seed_arrivals=8, seed_requests=9)
    rs = srv.serve(sat, realtime=False)
    log(f"serving: saturation session done ({rs['requests']} requests, {rs['wall_s']:.2f} s)")
    out["saturation"] = {"requests": rs["requests"], "throughput": round(rs["throughput"], 2),
                         "service_p50_ms": ms(rs["service_p50_s"]), "service_p99_ms": ms(rs["service_p99_s"]),
                         "feasible_fraction_fp64": rs["feasible_fraction"]}
    if n_plain > 0:
        pr = poisson_requests(n_req, rate=rate)[:n_plain]
        pl = serve_plain(pr, opts, geometry=geom, device=dev)
        svc = np.array([r.t_done - r.t_admit for r in pr])
        out["plain"] = {"requests": n_plain, "throughput": round(pl["throughput"], 3),
                        "p50_ms": round(pl["p50_s"] * 1e3, 3), "p99_ms": round(pl["p99_s"] * 1e3, 3),
                        "service_p50_ms": round(float(np.percentile(svc, 50)) * 1e3, 3),
                        "feasible_fraction_fp64": pl["feasible_fraction"]}
        same = all(np.array_equal(a.result.numpy(), b.result.numpy())
                   for a, b in zip(pr, reqs[:n_plain]))
        out["plain"]["bitwise_equal_to_served"] = bool(same)
        out["throughput_vs_plain"] = round(out["throughput"] / max(pl["throughput"], 1e-12), 2)
    return out


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` recorded from rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    (calibrated by a plain copy) in the committed profile summaries, or None."""
    return profiled_value("traffic_bytes_per_launch", kernel)


def profiled_value(name, kernel):
    """A figure `name[kernel] = value` recorded in the committed profile summaries, with the
    file it came from, or (None, None)."""
    import glob
    key = f"{name}[{kernel}]"
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(f) as fh:
            for line in fh:
                line = line.strip().strip("`")
                if line.startswith(key):
                    try:
                        return float(line.split("=", 1)[1]), os.path.relpath(f, ROOT)
                    except ValueError:
                        pass
    return None, None


def bucket_roofline(b, sizes, opts, chunk, n_chunks, stream):
    """Each size bucket's kernel launched on its own (same chunks as the batched run, one
    bucket at a time, HIP events on the launch stream).  Algorithmic bytes per launch:
    table gathers (8 corners x 4 B per pair looked up, counted by the kernel) plus, for
    every instance still running at the start of the launch, its state x and gradient
    read and written once (16 B per state element) and its L-BFGS history rows read at
    the start and written at the end (2 x 4 B per element and row); within a launch the
    state and history stay on chip.  Returns the dominant (longest) bucket's figures and
    a per-bucket list."""
    import torch
    # load every bucket's kernel before timing (the first launch of a kernel pays for
    # loading its code object on the host), then start over from the initial poses
    for k in range(len(b.buckets())):
        b.launch_bucket(k, 1, stream=stream)
    stream.synchronize()
    b.reset()
    torch.cuda.synchronize()
    n_state = 3 * sizes.astype(np.int64) + 3
    order = b.order.cpu().numpy()
    bk = b.buckets()
    per = [{"ms": [], "lookups": 0, "bytes": 0.0, "n_max": int(nmax), "instances": int(cnt)}
           for (_, cnt, nmax) in bk]
    rec0 = b.slot_records()
    with torch.cuda.stream(stream):
        for c in range(n_chunks):
            for k, (off, cnt, nmax) in enumerate(bk):
                idx = order[off:off + cnt]
                b.stats.zero_()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                b.launch_bucket(k, chunk, stream=stream)
                e1.record(stream)
                stream.synchronize()
                rec1 = b.slot_records()
                run = rec0[idx, 3] != 2
                rows = rec0[idx, 2].astype(np.int64) + rec1[idx, 2].astype(np.int64)
                per[k]["ms"].append(e0.elapsed_time(e1))
                per[k]["lookups"] += int(b.stats[1].item())
                per[k]["bytes"] += float(np.sum((n_state[idx] * 4 * (4 + 2 * rows))[run]))
                rec0[idx] = rec1[idx]
    from packing_amd.autotune import choose_block
    for k, q in enumerate(per):
        q["bytes"] += q["lookups"] * 32.0
        q["block"] = b._block_for(q["n_max"])
        q["tile_source"] = "tiles.json" if choose_block(opts.mode, q["n_max"], b.geometry.n_pieces,
                                                        opts.history) else "library rule"
        rows = -(-(3 * q["n_max"] + 3) // 64)
        blk = q["block"] or (1 if rows <= 3 else 2)      # 0: the library's rule
        q["kernel"] = (f"pk::k_relax_w<2, {rows}, 8, 1>" if blk == 1 else
                       "pk::k_relax_w<2, 10, 8, 2>" if blk == 2 else f"pk::k_relax<2, {blk}, ...>")
    dom = max(per, key=lambda q: sum(q["ms"]))
    avg_s = float(np.mean(dom["ms"])) / 1e3
    alg = dom["bytes"] / n_chunks
    achieved = alg / avg_s / 1e9
    traffic, src = profiled_traffic(dom["kernel"])
    issue, _ = profiled_value("issue_fraction", dom["kernel"])
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": 8000.0, "unit": "GB/s",
            "frac": round(achieved / 8000.0, 5), "traffic": traffic,
            "traffic_source": src,
            # share of the wave's cycles spent issuing instructions (two waves per SIMD):
            # what actually bounds this kernel, from the same profile
            "issue_frac": issue,
            "kernel": dom["kernel"] + f" (table mode, size bucket of {dom['instances']} instances, "
                                      f"N <= {dom['n_max']})",
            "avg_launch_ms": round(avg_s * 1e3, 3), "alg_bytes_per_launch": alg,
            "lookups_per_launch": dom["lookups"] / n_chunks,
            "buckets": [{"kernel": q["kernel"], "instances": q["instances"], "n_max": q["n_max"],
                         "block": q["block"], "tile_source": q["tile_source"],
                         "avg_launch_ms": round(float(np.mean(q["ms"])), 3),
                         "alg_bytes_per_launch": q["bytes"] / n_chunks,
                         "achieved_GBps": round(q["bytes"] / n_chunks / (float(np.mean(q["ms"])) / 1e3) / 1e9, 2)}
                        for q in per]}


def c1_plain_section(dev, iters=500):
    """Plain method on the GPU configuration: 1024 instances x 100 trees, exact overlap
    area + boundary, fixed square boxes, 500 L-BFGS iterations (history 8, max_step 0.1),
    from the same initial poses through the reference API (pack_dynamics.OptimizerBFGS):
    the reference loop (one pack_cost evaluation + the drop-in batched L-BFGS kernels per
    iteration, engine="reference") and the fused engine (engine="auto")."""
    import torch
    from packing_amd import pack_cost as PC
    from packing_amd.feasibility import overlapping_pairs
    from packing_amd.pack_dynamics import OptimizerBFGS
    from packing_amd.solutions import SolutionCollectionSquare
    from packing_amd.workloads import uniform_workload
    x, h = uniform_workload(1024, 100, seed=2)
    cost = PC.CostCompound(costs=[PC.BoundaryDistanceCost(), PC.CollisionCostOverlappingArea()])

    def sol():
        return SolutionCollectionSquare(torch.tensor(x, device=dev), torch.tensor(h, device=dev), use_fixed_h=True)

    out = {"workload": "1024 x 100 trees, exact area + boundary, 500 iterations, history 8, FP32"}
    for engine in ("reference", "auto"):
        OptimizerBFGS(cost=cost, n_iterations=3, history_size=8, max_step=0.1).run_simulation(sol(), engine=engine)
        torch.cuda.synchronize()
        opt = OptimizerBFGS(cost=cost, n_iterations=iters, history_size=8, max_step=0.1)
        s0 = sol()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = opt.run_simulation(s0, engine=engine)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        pairs = overlapping_pairs(r.xyt).cpu().numpy()
        out[engine] = {"value": round(1024 / t, 2), "unit": "relaxations/s", "seconds": round(t, 4),
                       "feasible_fraction_fp64": float((pairs == 0).mean())}
    out["fused_vs_reference_loop"] = round(out["auto"]["value"] / out["reference"]["value"], 2)
    return out


def c3_section(dev, cpu_sample=8):
    """Differentiation + assignment configuration: the autograd overlap cost (forward with
    the analytic gradient saved, backward) on 256 instances x 50 trees in FP32, and the
    batched Hungarian on 512 int32 200 x 200 cost matrices with entries U[0, 10^6)
    (assignments bit-exact: tested in tests/test_gpu_lap.py).  The CPU baseline of the
    assignment is scipy's linear_sum_assignment on one core, on the first cpu_sample
    matrices."""
    import torch
    from scipy.optimize import linear_sum_assignment
    from packing_amd.autograd import overlap_cost
    from packing_amd.geometry import make_poses
    from packing_amd.lap_batch import solve_lap_batch
    out = {}
    x = torch.tensor(make_poses(np.random.default_rng(4), 256, 50, np.float32), device=dev)

    def fwd_bwd():
        xr = x.clone().requires_grad_(True)
        overlap_cost(xr).sum().backward()
        return xr.grad

    for _ in range(3):
        fwd_bwd()
    torch.cuda.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        fwd_bwd()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / reps
    out["autograd"] = {"workload": "256 instances x 50 trees, exact area, forward + backward, FP32",
                       "ms_per_call": round(t * 1e3, 4), "instances_per_s": round(256 / t, 1)}
    rng = np.random.default_rng(6)
    C = rng.integers(0, 10 ** 6, size=(512, 200, 200)).astype(np.int32)
    Cd = torch.tensor(C, device=dev)
    s = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(s):
        solve_lap_batch(Cd, stream=s)
        s.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(3):
            solve_lap_batch(Cd, stream=s)
        e1.record(s)
        s.synchronize()
    ms = e0.elapsed_time(e1) / 3
    t0 = time.perf_counter()
    for b in range(cpu_sample):
        linear_sum_assignment(C[b])
    t_cpu = (time.perf_counter() - t0) / cpu_sample
    out["assignment"] = {"workload": "512 int32 cost matrices 200 x 200, entries U[0, 10^6), Hungarian",
                         "ms_per_batch": round(ms, 3), "problems_per_s": round(512 / (ms / 1e3), 1),
                         "cpu_scipy": {"problems_per_s": round(1.0 / t_cpu, 1), "cores": 1,
                                       "sample": f"{cpu_sample} matrices"}}
    return out


def c0_section(geom, dev, budget_s, iters=200):
    """The CPU configuration: 16 instances x 20 trees in FP64, exact overlap area +
    boundary, 200 L-BFGS iterations (history 8): the FP64 oracle on the host cores (OpenMP
    over instances) and the drop-in path on the GPU in FP64 (pack_dynamics reference loop)."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from packing_amd import pack_cost as PC
    from packing_amd.pack_dynamics import OptimizerBFGS
    from packing_amd.solutions import SolutionCollectionSquare
    from packing_amd.workloads import uniform_workload
    x, h = uniform_workload(16, 20, seed=1, dtype=np.float64)
    cores = len(os.sched_getaffinity(0))
    og = O.OracleGeometry(geom)
    reps, t_cpu = 0, 0.0
    while t_cpu < budget_s and reps < 50:
        t0 = time.perf_counter()
        O.relax(og, x, h, max_iter=iters, m=8, max_step=0.1, n_threads=cores)
        t_cpu += time.perf_counter() - t0
        reps += 1
    cost = PC.CostCompound(costs=[PC.BoundaryDistanceCost(), PC.CollisionCostOverlappingArea()])
    opt = OptimizerBFGS(cost=cost, n_iterations=iters, history_size=8, max_step=0.1)

    def sol():
        return SolutionCollectionSquare(torch.tensor(x, device=dev), torch.tensor(h, device=dev), use_fixed_h=True)

    opt.run_simulation(sol())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    opt.run_simulation(sol())
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    return {"workload": "16 x 20 trees, FP64, exact area + boundary, 200 iterations, history 8",
            "cpu": {"value": round(16 * reps / t_cpu, 2), "unit": "relaxations/s", "cores": cores,
                    "kind": "port", "sample": f"{reps} runs of the whole configuration, {t_cpu:.1f} s"},
            "gpu_reference_loop_fp64": {"value": round(16 / t_gpu, 2), "unit": "relaxations/s",
                                        "seconds": round(t_gpu, 4)}}


def infeasible_breakdown(b, sizes, pairs, lut):
    """The final packings that still overlap in FP64: how many the table sees as free (zero
    table cost: the interpolated table, not the exact polygon) and their stop reasons."""
    import torch
    from packing_amd.pack_cuda_lut import overlap_multi_ensemble
    bad = np.where(pairs > 0)[0]
    rec = b.slot_records()
    zero = 0
    for i in bad:
        c = torch.empty(1, device=b.device)
        overlap_multi_ensemble(b.xyt(int(i))[None].contiguous(), c, lut)
        zero += float(c.item()) == 0.0
    names = {1: "grad", 2: "dir", 4: "param", 8: "loss", 16: "rel", 32: "increase", 64: "eval", 128: "iter"}
    reasons = {}
    for i in bad:
        r = int(rec[i, 4])
        key = "+".join(v for k, v in names.items() if r & k) or "none"
        reasons[key] = reasons.get(key, 0) + 1
    return {"overlapping": int(len(bad)), "zero_table_cost": int(zero),
            "overlapping_pairs_total": int(pairs[bad].sum()) if len(bad) else 0,
            "stop_reasons": reasons,
            "mean_trees": round(float(sizes[bad].mean()), 1) if len(bad) else None}


def cpu_baseline(geom, lut, sizes, xs, hs, opts, budget_s):
    """Same workload on the host: the FP64 oracle (per-instance L-BFGS, table overlap +
    boundary), OpenMP over instances on every host core, on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    cores = len(os.sched_getaffinity(0))
    og = O.OracleGeometry(geom)
    OL = O.OracleLut(lut.X, lut.Y, lut.theta, lut.lut_d.cpu().numpy(), mirror=lut.mirror_x)
    done, t_used, k = 0, 0.0, 0
    # probe the cost of one relaxation of the mean size, then size the sample
    order = np.argsort(np.abs(sizes - np.mean(sizes)))
    while t_used < budget_s and k < len(order):
        batch = order[k:k + cores]
        k += len(batch)
        t0 = time.perf_counter()
        for n in np.unique(sizes[batch]):
            idx = [i for i in batch if sizes[i] == n]
            X = np.stack([xs[i] for i in idx]).astype(np.float64)
            H = np.stack([hs[i] for i in idx]).astype(np.float64)
            O.relax(og, X, H, max_iter=opts.max_iter, m=opts.history, max_step=opts.max_step,
                    n_threads=cores, lut=OL)
        t_used += time.perf_counter() - t0
        done += len(batch)
    part = "all instances" if done >= len(sizes) else \
        f"the {done} instances closest to the mean size N={np.mean(sizes):.0f}"
    return {"value": done / t_used, "unit": "relaxations/s", "cores": cores, "kind": "port",
            "sample": f"{part} of the workload, {opts.max_iter} iterations each, FP64 oracle "
                      f"(all-pairs loop within the cutoff, as the reference's table kernel), {t_used:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--iters", type=int, default=500)
    ap.add_argument("--chunk", type=int, default=100, help="iterations per persistent launch")
    ap.add_argument("--plain-sample", type=int, default=6)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests of the served session (0: skip serving)")
    ap.add_argument("--serve-plain", type=int, default=10, help="requests timed on the plain path")
    ap.add_argument("--serve-iters", type=int, default=2000)
    ap.add_argument("--no-overlap", action="store_true", help="skip the exact-kernel roofline")
    ap.add_argument("--overlap-only", action="store_true",
                    help="only the exact-kernel roofline (the launches profiled with rocprofv3)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the per-bucket kernel measurement (the launches profiled with rocprofv3)")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--tune", action="store_true",
                    help="only measure every tile on each size bucket and save packing_amd/tiles.json")
    ap.add_argument("--tiles-out", default=None, help="where --tune writes the table")
    ap.add_argument("--no-configs", action="store_true", help="skip the plain-method configuration lines")
    ap.add_argument("--serve-saturation", type=int, default=4096,
                    help="requests of the saturation session")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every instance is independent and this build runs one "
              "process on one MI355X; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from packing_amd import _native
    from packing_amd.context import set_geometry
    from packing_amd.feasibility import overlapping_pairs
    from packing_amd.geometry import random_polygon
    from packing_amd.pack_cuda_lut import LookupTable
    from packing_amd.relax import RelaxBatch, RelaxOptions

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    info = _native.device_info()
    geom = random_polygon(0)
    set_geometry(geom)
    if args.overlap_only:
        print(json.dumps({"overlap_kernel": overlap_kernel_roofline(geom, dev)}))
        return
    t0 = time.perf_counter()
    lut = LookupTable.build_exact(256, 256, 128, device=dev)
    torch.cuda.synchronize()
    t_lut = time.perf_counter() - t0
    log(f"table 256x256x128 built in {t_lut:.3f} s; pieces={geom.n_pieces}")

    sizes, xs, hs = make_workload(args.instances, 10, 200, args.seed)
    opts = RelaxOptions(mode="lut", lut=lut, max_iter=args.iters, history=8, max_step=0.1)
    b = RelaxBatch(xs, hs, opts, geometry=geom, device=dev)
    stream = torch.cuda.Stream(device=dev)
    chunk = min(args.chunk, args.iters)
    n_chunks = -(-args.iters // chunk)
    if args.roofline_only:
        print(json.dumps({"roofline": bucket_roofline(b, sizes, opts, chunk, n_chunks, stream)}))
        return
    if args.tune:
        # the headline workload (table mode, every size bucket) and the plain-method GPU
        # configuration's shape in exact-area mode (1024 x 100)
        from packing_amd.autotune import tune_batch
        from packing_amd.workloads import uniform_workload
        res = tune_batch(b, iters=chunk, stream=stream)
        x1, h1 = uniform_workload(1024, 100, seed=2)
        b1 = RelaxBatch(list(x1), h1, RelaxOptions(mode="area", max_iter=args.iters, history=8, max_step=0.1),
                        geometry=geom, device=dev)
        res += tune_batch(b1, iters=chunk, save=True, stream=stream, path=args.tiles_out)
        print(json.dumps({"tiles": res}))
        return

    # one graph = reset + all chunked launches of one relaxation of the batch
    with torch.cuda.stream(stream):
        b.reset()
        b.launch(1, stream=stream)          # sets kernel attributes outside capture
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream, capture_error_mode="relaxed"):
            b.reset()
            b.launch(chunk, stream=stream, chunks=n_chunks)
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        graph.replay()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = ev0.elapsed_time(ev1) / 1e3
    elapsed = max(wall, gpu_s)
    value = args.instances * args.steps / elapsed
    log(f"batched: {args.steps} steps in {elapsed:.3f} s -> {value:.1f} relaxations/s")

    # FP64 feasibility of every final packing on the GPU
    flat = b.xyt_flat().contiguous()
    off = torch.tensor(np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32), device=dev)
    pairs = overlapping_pairs(flat, off, int(sizes.max())).cpu().numpy()
    feasible_frac = float((pairs == 0).mean())
    rec = b.slot_records()
    loss = b.loss.cpu().numpy()
    infeasible = infeasible_breakdown(b, sizes, pairs, lut)
    log(f"infeasible: {infeasible}")

    roofline = bucket_roofline(b, sizes, opts, chunk, n_chunks, stream)
    log(f"roofline: {roofline}")

    # plain path: one launch per instance per iteration, same kernel and numerics
    plain = None
    if args.plain_sample > 0:
        pick = np.argsort(np.abs(sizes - np.mean(sizes)))[: args.plain_sample]
        pb = RelaxBatch([xs[i] for i in pick], hs[pick], opts, geometry=geom, device=dev)
        pb.run_plain(instances=[0])            # warm
        pb.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pb.run_plain()
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        plain = {"value": round(len(pick) / tp, 3), "unit": "relaxations/s",
                 "sample": f"{len(pick)} instances near mean N, {args.iters} launches each"}
        # bitwise agreement with the batched run for the sampled instances
        same = all(np.array_equal(pb.xyt(k).cpu().numpy(), b.xyt(int(i)).cpu().numpy())
                   for k, i in enumerate(pick))
        plain["bitwise_equal_to_batched"] = bool(same)
        log(f"plain: {plain}")

    ovl = None
    if not args.no_overlap:
        ovl = overlap_kernel_roofline(geom, dev)
        log(f"overlap kernel: {ovl}")

    serving = None
    if args.serve_requests > 0:
        serving = serving_section(geom, lut, dev, args.serve_requests, args.serve_plain, args.serve_iters,
                                  n_sat=args.serve_saturation)
        log(f"serving: {serving}")

    c1 = c0 = c3 = None
    if not args.no_configs:
        c1 = c1_plain_section(dev)
        log(f"plain-method GPU configuration: {c1}")
        c3 = c3_section(dev)
        log(f"differentiation + assignment configuration: {c3}")
        if not args.no_cpu:
            c0 = c0_section(geom, dev, budget_s=5.0)
            log(f"plain-method CPU configuration: {c0}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(geom, lut, sizes, xs, hs, opts, args.cpu_seconds)
        log(f"cpu: {cpu}")

    out = {
        "metric": "ragged relaxations/sec (FP64-verified zero-overlap feasibility)",
        "value": round(value, 2),
        "unit": "relaxations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "ragged batching: 4096 instances, N~U{10..200}, random 15-vertex polygon "
                               "(seed 0), pose table 256x256x128, 500 L-BFGS iterations, history 8",
                   "instances": args.instances, "iters": args.iters, "chunk": chunk,
                   "total_trees": int(sizes.sum()), "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain": plain,
        "speedup_vs_plain": round(value / plain["value"], 2) if plain else None,
        "feasible_fraction_fp64": feasible_frac,
        "infeasible": infeasible,
        "iterations_mean": round(float(rec[:, 0].mean()), 2),
        "stopped_before_max_iter": round(float((rec[:, 0] < args.iters).mean()), 4),
        "serving": serving,
        "overlap_kernel": ovl,
        "plain_method_gpu_config": c1,
        "plain_method_cpu_config": c0,
        "differentiation_assignment_config": c3,
        "final_loss_median": float(np.median(loss)),
        "table_build_s": round(t_lut, 4),
        "device": info,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
