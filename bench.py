#!/usr/bin/env python3
"""Benchmark: the curvelet-construction path on one MI355X.

Workload: one synthetic 1920x1080 third-order edge map (~207k edges; 1500
curves, 40 T-junction patches, 30% clutter; seed 1003), group_max_sz = 7,
other parameters at the reference defaults.  A step is one pass of the whole
path over that map on the reference's timed span (cpu_curvelet/main.cpp:43-66):
edges in host memory -> spatial hash + CSR neighbour graph + pairwise bundles +
chain growth + dedup + record -> compacted chains and info in host memory.

Prints one JSON line.  `value` is edges / the median step time over the timed
steps, each step timed with HIP events on the context's stream from before
the edges' upload to after the results' download (page-locked host buffers on
both sides).  Extra keys: the rate with edges resident in HBM and results left
there, the rate when consecutive maps overlap their transfers, and the text
end-to-end time (reader + path + writers).  `roofline` is for the dominant
kernel (k_curvelets, the fused bundle + growth kernel), timed with HIP events
on the context's stream; its algorithmic FP64 work is what the kernel
executes: the transport cells (1,621 FLOP per neighbour pair; the 14 sines per
pair come precomputed) and the growth intersections (189 FLOP each).
`cpu_baseline` is the reference's own cpu_curvelet (built in place under
oracle/_ref) on this host's cores, timed on the same map with span (ii):
CSR build + build_curvelets_greedy; without oracle/_ref, the repository's
restatement of it (kind "port").
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

WORKLOAD = {"preset": "C3", "seed": 1003, "group_max_sz": 7}
F_CELLS = 1621.0          # FP64 FLOP per bundle pair in the fused kernel: 63 cells x ~25 + per-row/pair terms
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector peak
PEAK_HBM_GBS = 8000.0


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    lim = os.environ.get("OMP_NUM_THREADS")
    if lim and lim.isdigit():
        n = min(n, int(lim))
    return max(1, n)


def fused_traffic():
    """HBM bytes per step of the fused kernel's launches, from the committed
    rocprofv3 PMC summary of this workload (FETCH_SIZE + WRITE_SIZE, each scaled
    by its copy-kernel calibration factor); None when it is absent."""
    f = ROOT / "profiles" / "k_curvelets_1080p_traffic.json"
    try:
        return json.loads(f.read_text())["hbm_bytes_per_step"]
    except (OSError, KeyError, ValueError):
        return None


def host_info() -> dict:
    """CPU model and glibc of this host (recorded beside every CPU timing)."""
    import platform
    model = platform.processor() or ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "glibc": "-".join(platform.libc_ver())}


def cpu_time(edges, params, threads, reps=3):
    """Median seconds of span (ii) (CSR build + build_curvelets_greedy) on
    `threads` host threads: the reference's own code built in place
    (oracle/_ref) when it is there, else the repository's restatement of it
    (oracle/liboracle.so, bit-exact with it on every fixture)."""
    from oracle import Oracle, Ref
    times, rows = [], 0
    if Ref.available():
        ref, kind = Ref(), "reference"
        for _ in range(reps):
            r = ref.run(edges, params, nthreads=threads, valid_pairs=False)
            times.append(r["times"]["csr"] + r["times"]["greedy"])
            rows = len(r["chains"])
    else:
        orc, kind = Oracle(), "port"
        for _ in range(reps):
            t0 = time.perf_counter()
            orc.build_csr(edges, params, nthreads=threads)
            chains, _ = orc.greedy(params, nthreads=threads)
            times.append(time.perf_counter() - t0)
            rows = len(chains)
    return {"seconds": statistics.median(times), "kind": kind, "rows": int(rows)}


def cpu_baseline(edges, params):
    """The CPU baseline on this host's cores, span (ii), on the whole map."""
    try:
        cores = host_cores()
        t = cpu_time(edges, params, cores)
        what = ("the reference's code (oracle/_ref, built in place from cpu_curvelet)" if t["kind"] == "reference"
                else "the repository's restatement of cpu_curvelet (oracle/liboracle.so; oracle/_ref absent)")
        return {"value": len(edges) / t["seconds"], "unit": "edges/s", "cores": cores, "kind": t["kind"],
                "sample": f"the full {len(edges)}-edge map, CSR + build_curvelets_greedy, median of 3 runs "
                          f"({t['seconds']:.3f} s each on {cores} threads) of {what}",
                "seconds": t["seconds"], "num_curvelets": t["rows"], **host_info()}
    except Exception as exc:  # the baseline is reported, never required
        return {"error": repr(exc)}


BATCH_SIZES = [(320, 240), (640, 480), (1280, 720), (1920, 1080), (3840, 2160)]


def batch_maps(count=64, seed=4001):
    """64 synthetic maps: sizes uniform over BATCH_SIZES, clutter U[0, 0.4],
    curves and junction patches in proportion to area (the 1080p map's
    generator: 1500 curves and 40 patches per 1920x1080)."""
    import numpy as np
    from curvelet_construction_on_gpu_amd import synth_edges
    rng = np.random.default_rng(seed)
    maps, sizes = [], []
    for m in range(count):
        w, h = BATCH_SIZES[rng.integers(len(BATCH_SIZES))]
        area = w * h / (1920 * 1080)
        maps.append(synth_edges("C3", seed * 1000 + m, width=w, height=h, num_curves=max(1, round(1500 * area)),
                                junction_patches=round(40 * area), clutter_fraction=float(rng.uniform(0, 0.4))))
        sizes.append(f"{w}x{h}")
    return maps, sizes


def bench_batch(args) -> int:
    """maps/s of the batched form (one pass over all 64 maps) against the plain
    loop (one cvg_run_device per map), inputs resident in HBM for both, every
    run queued without stats (no wait between maps or batches)."""
    import numpy as np
    from curvelet_construction_on_gpu_amd import Context, CurveletParams
    maps, sizes = batch_maps()
    params = CurveletParams()
    counts = [len(m) for m in maps]
    ctx = Context(0)
    allm = np.concatenate(maps)
    dall = ctx.upload(allm)
    eoff = np.concatenate([[0], np.cumsum(counts)])
    step = 8 * allm.shape[1]
    for _ in range(max(1, args.warmup // 2)):
        ctx.run_batch_device(dall, counts, 4, params)
        for m in range(len(maps)):
            ctx.run_device(dall + int(eoff[m]) * step, counts[m], 4, params)
    ctx.synchronize()
    # launch shape for this mix (untimed; results do not depend on it)
    default_tuning = ctx.tuning
    tuning, tuned_ms, trials = ctx.autotune_batch(dall, counts, 4, params, reps=1)
    reps = max(1, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.run_batch_device(dall, counts, 4, params, stats=False)
    ctx.synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    ctx.tuning = default_tuning  # the plain loop runs one map at a time: its own (automatic) launch shape
    t0 = time.perf_counter()
    for _ in range(reps):
        for m in range(len(maps)):
            ctx.run_device(dall + int(eoff[m]) * step, counts[m], 4, params, stats=False)
    ctx.synchronize()
    t_plain = (time.perf_counter() - t0) / reps
    ctx.tuning = tuning
    st = ctx.run_batch_device(dall, counts, 4, params)  # counts for the report (untimed)
    # the batch's rows of a few maps against their plain-path rows
    same = True
    ctx.run_batch_device(dall, counts, 4, params)
    got = {m: ctx.batch_curvelets(m) for m in (0, 1, len(maps) - 1)}
    for m, (ch, inf) in got.items():
        ctx.run_device(dall + int(eoff[m]) * step, counts[m], 4, params)
        c2, i2 = ctx.curvelets()
        same = same and np.array_equal(ch, c2) and np.array_equal(inf.view(np.uint64), i2.view(np.uint64))
    ctx.close()
    n = int(eoff[-1])
    line = {"metric": "maps/s, 64-map batch (sizes 320x240..3840x2160)", "value": len(maps) / t_batch,
            "unit": "maps/s", "n_gpus": 1, "steps": reps, "warmup": max(1, args.warmup // 2),
            "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "64 synthetic maps, sizes uniform over 320x240/640x480/1280x720/1920x1080/"
                                   "3840x2160, clutter U[0,0.4], curves and junctions per area, defaults (G=4)",
                       "edges": n, "maps": len(maps), "sizes": {k: sizes.count(k) for k in sorted(set(sizes))},
                       "neighbor_pairs": st["total_neighbor_pairs"], "curvelets": st["num_curvelets"]},
            "plain_loop_maps_per_s": len(maps) / t_plain, "batch_speedup_vs_plain_loop": t_plain / t_batch,
            "batch_edges_per_s": n / t_batch, "plain_loop_edges_per_s": n / t_plain,
            "sampled_maps_identical": bool(same),
            "autotune": {"default": default_tuning, "default_ms": trials[0][1], "chosen": tuning,
                         "chosen_ms": tuned_ms, "trials": len(trials)}}
    print(json.dumps(line))
    return 0


def serve_pool(count=96, seed=5001):
    """Request maps: area log-uniform between 320x240 and 3840x2160 (16:9),
    curves and junction patches per area, clutter U[0, 0.4]."""
    import math
    import numpy as np
    from curvelet_construction_on_gpu_amd import synth_edges
    rng = np.random.default_rng(seed)
    pool, areas = [], []
    lo, hi = math.log(320 * 240), math.log(3840 * 2160)
    for m in range(count):
        area = math.exp(rng.uniform(lo, hi))
        w = max(32, int(round(math.sqrt(area * 16 / 9))))
        h = max(32, int(round(area / w)))
        a = w * h / (1920 * 1080)
        pool.append(synth_edges("C3", seed * 1000 + m, width=w, height=h, num_curves=max(1, round(1500 * a)),
                                junction_patches=round(40 * a), clutter_fraction=float(rng.uniform(0, 0.4))))
        areas.append(w * h)
    return pool, areas


def bench_serve(args) -> int:
    """The served form: (1) saturation maps/s of continuous batching against
    the plain one-map-at-a-time loop (the same server with one worker and
    batches of one) on 400 requests submitted at once: the median of three
    runs after four warm-up runs of the same size (the page-locked pools grown
    to their steady-state size: pinning new host memory costs more than
    serving); (2) a Poisson stream at
    200 maps/s (2,000 requests after 100 warm-up ones): p50/p99 latency from
    submit to results in host memory, overall and for 1080p-class maps; every
    20th served result of the stream (5%) is checked afterwards against the
    oracle with the device's correctly rounded math (bit for bit, by hash)."""
    import threading
    import numpy as np
    import xxhash
    from curvelet_construction_on_gpu_amd import CurveletParams, Server
    pool, areas = serve_pool()
    params = CurveletParams()
    rng = np.random.default_rng(6001)

    def saturate(srv, nreq):
        picks = rng.integers(len(pool), size=nreq)
        t0 = time.perf_counter()
        ids = [srv.submit(pool[k]) for k in picks]
        for i in ids:
            srv.view(i)
            srv.release(i)
        return nreq / (time.perf_counter() - t0), int(sum(len(pool[k]) for k in picks))

    out = {}
    nsat, nwarm_sat = 400, 4

    def steady(srv):
        for _ in range(nwarm_sat):
            saturate(srv, nsat)
        runs = [saturate(srv, nsat) for _ in range(3)]
        return statistics.median(r for r, _ in runs), runs[0][1]

    with Server(params, workers=1, max_batch_maps=1) as plain:
        out["plain_loop_maps_per_s"], _ = steady(plain)
    with Server(params) as srv:
        sat, sat_edges = steady(srv)
        out["saturation_maps_per_s"] = sat
        cfg = srv.config
        # Poisson stream
        rate, nwarm, nreq = 200.0, 100, 2000
        gaps = rng.exponential(1.0 / rate, size=nwarm + nreq)
        picks = rng.integers(len(pool), size=nwarm + nreq)
        ids = [None] * (nwarm + nreq)
        t_start = time.perf_counter()

        def client():
            t = t_start
            for k in range(nwarm + nreq):
                t += gaps[k]
                d = t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[k] = srv.submit(pool[picks[k]])

        th = threading.Thread(target=client)
        th.start()
        lat, lat_1080, batch_sizes, sample = [], [], [], []
        for k in range(nwarm + nreq):
            while ids[k] is None:
                time.sleep(0.0002)
            ch, inf, r = srv.view(ids[k])
            if k >= nwarm and (k - nwarm) % 20 == 0:  # the 5% sample checked against the oracle afterwards
                sample.append((int(picks[k]), xxhash.xxh3_128_hexdigest(np.ascontiguousarray(ch)),
                               xxhash.xxh3_128_hexdigest(np.ascontiguousarray(inf))))
            srv.release(ids[k])
            if k >= nwarm:
                lat.append(r["latency_ms"])
                batch_sizes.append(r["batch_maps"])
                if 0.65 <= areas[picks[k]] / (1920 * 1080) <= 1.35:
                    lat_1080.append(r["latency_ms"])
            if k == nwarm:
                t_first = time.perf_counter()
        th.join()
        t_end = time.perf_counter()
    stream_rate = nreq / (t_end - t_first)
    # the sample against the restatement with the device's correctly rounded
    # math: every chain and info bit (one oracle run per distinct map)
    oracle_check = {"sampled": len(sample), "maps": 0, "mismatches": 0}
    try:
        from oracle import Oracle
        orc, want = Oracle(cr=True), {}
        for k, hc, hi in sample:
            if k not in want:
                w = orc.run(pool[k], params, nthreads=0, valid_pairs=False)
                want[k] = (xxhash.xxh3_128_hexdigest(np.ascontiguousarray(w["chains"].astype(np.uint32))),
                           xxhash.xxh3_128_hexdigest(np.ascontiguousarray(w["info"])))
            oracle_check["mismatches"] += int((hc, hi) != want[k])
        oracle_check["maps"] = len(want)
    except Exception as exc:  # reported, never required
        oracle_check["error"] = repr(exc)
    line = {"metric": "maps/s (served, continuous batching, saturation)", "value": out["saturation_maps_per_s"],
            "unit": "maps/s", "n_gpus": 1, "steps": nreq, "warmup": nwarm, "ms_per_step": 1e3 / out["saturation_maps_per_s"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "request maps with area log-uniform 320x240..3840x2160, curves/junctions per "
                                   "area, clutter U[0,0.4], defaults (G=4); pool of 96 seeded maps",
                       "server": cfg, "saturation_requests": nsat, "saturation_warmup_runs": nwarm_sat,
                       "saturation_statistic": "median of 3 runs",
                       "saturation_edges": sat_edges},
            "plain_loop_maps_per_s": out["plain_loop_maps_per_s"],
            "speedup_vs_plain_loop": out["saturation_maps_per_s"] / out["plain_loop_maps_per_s"],
            "poisson": {"rate_maps_per_s": rate, "requests": nreq, "sustained_maps_per_s": stream_rate,
                        "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
                        "latency_ms_p50_1080p": float(np.percentile(lat_1080, 50)) if lat_1080 else None,
                        "latency_ms_p99_1080p": float(np.percentile(lat_1080, 99)) if lat_1080 else None,
                        "requests_1080p": len(lat_1080), "mean_batch_maps": float(np.mean(batch_sizes)),
                        "oracle_sample": oracle_check}}
    print(json.dumps(line))
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["map", "batch", "serve"], default="map",
                    help="map: the 1080p map (the headline line); batch: 64 mixed-size maps, batched "
                         "form against the plain per-map loop; serve: a Poisson request stream through "
                         "the served form (continuous batching), latency and maps/s")
    ap.add_argument("--span-only", action="store_true",
                    help="map workload: only the protocol-span and resident steps of the 1080p map on one "
                         "context (no CPU baseline, two-context pipelined run, C1 map or text end-to-end), "
                         "so that a profiler's trace splits into whole steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        import torch
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)                   # the runtime's blit kernel
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)     # an elementwise kernel: reads nbytes, writes nbytes
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "repeats": 3,
                          "kernels": ["__amd_rocclr_copyBuffer", "vectorized_elementwise_kernel (mul)"]}))
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); maps shard across GPUs "
              f"without communication, so per-GPU throughput is the figure that matters.", file=sys.stderr)
        return 2

    from curvelet_construction_on_gpu_amd import Context, CurveletParams, synth_edges
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)

    edges = synth_edges(WORKLOAD["preset"], WORKLOAD["seed"])
    params = CurveletParams(group_max_sz=WORKLOAD["group_max_sz"])
    n, sz = edges.shape
    ctx = Context(0)

    # --- the protocol span: host edges in -> host chains and info out ---------
    # (the reference's timed span, main.cpp:43-66): H2D of the edges, every
    # kernel, D2H of the compacted chains and info, timed with HIP events on
    # the context's stream; page-locked host buffers on both sides
    proto_ms, proto_wall_ms, counts = protocol_span(ctx, edges, params, args.warmup, args.steps)
    ms_step = statistics.median(proto_ms)

    # --- resident in HBM: edges already on the device, results left there ----
    dptr = ctx.upload(edges)
    for _ in range(args.warmup):
        ctx.run_device(dptr, n, sz, params)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.run_device(dptr, n, sz, params, stats=False)
    ctx.synchronize()
    resident_s = (time.perf_counter() - t0) / args.steps
    # the fused kernel's span per step: the same steps again with the
    # library's HIP timing events on (they bracket its launches)
    ctx.set_event_timing(True)
    for _ in range(args.steps):
        ctx.run_device(dptr, n, sz, params, stats=False)
    fused = ctx.recent_fused_ms(args.steps)
    ctx.set_event_timing(False)
    st = ctx.run_device(dptr, n, sz, params)  # counts and stage times (untimed)
    pipelined = None if args.span_only else pipelined_pcie(edges, params, max(args.steps, 20))

    pairs = st["total_neighbor_pairs"]
    fused_ms = statistics.mean(fused)
    # FP64 work executed inside k_curvelets: the transport cells (F_CELLS per
    # pair; its sines come precomputed from k_edge_trig / k_pair_trig) and the
    # growth intersections (3 FLOP per cell per intersection)
    flops = pairs * F_CELLS + st["intersections"] * 63 * 3
    achieved = flops / (fused_ms * 1e-3) / 1e12
    line = {
        "metric": "edges/s per map at 1080p (~200k edges)",
        "value": n / (ms_step * 1e-3),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "1920x1080 synthetic third-order edge map, ~207k edges, group_max_sz=7 "
                               "(preset C3, seed 1003)",
                   "edges": n, "neighbor_pairs": pairs, "valid_pairs": st["valid_pairs"],
                   "curvelets": st["num_curvelets"], "group_max_sz": WORKLOAD["group_max_sz"],
                   "span": "host edges in -> host chains and info out (H2D + all kernels + D2H), page-locked "
                           "host buffers, HIP events on the context stream, median of the timed runs after "
                           "the warm-up runs",
                   "d2h_bytes": counts["d2h_bytes"], "h2d_bytes": counts["h2d_bytes"]},
        "ms_per_step_spread": {"min": min(proto_ms), "max": max(proto_ms), "host_wall_median": statistics.median(proto_wall_ms)},
        "bundle_pairs_per_s": pairs / (ms_step * 1e-3),
        "resident_edges_per_s": n / resident_s,
        "resident_ms_per_step": resident_s * 1e3,
        "pipelined_pcie_edges_per_s": pipelined,
        "fused_kernel_ms_median": statistics.median(fused),
        "stage_ms": {"graph": st["ms_graph"], "curvelets": st["ms_curvelets"], "fused_kernel": fused_ms},
        "lds_bytes_per_workgroup": st["lds_bytes_per_workgroup"],
        "roofline": {"bound": "fp64_valu", "pipe": "fp64 vector ALU (no MFMA in this path)",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": fused_traffic(),
                     "kernel": "k_curvelets", "kernel_ms": fused_ms,
                     "flop_per_launch": flops,
                     "flop_model": f"{F_CELLS:.0f} per neighbour pair + 189 per growth intersection, "
                                   "executed inside k_curvelets"},
    }
    traffic = fused_traffic()
    if traffic:
        # the same kernel against the HBM roofline: rocprof bytes per step over its time
        gbps = traffic / (fused_ms * 1e-3) / 1e9
        line["roofline_hbm"] = {"achieved": gbps, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                "frac": gbps / PEAK_HBM_GBS, "traffic": traffic}
    if args.span_only:
        line["passes"] = {"protocol_span": 1 + args.warmup + args.steps,
                          "resident": args.warmup + 2 * args.steps + 1}
        ctx.close()
        print(json.dumps(line))
        return 0
    if not args.no_cpu_baseline:
        cb = cpu_baseline(edges, params)
        if cb is not None:
            line["cpu_baseline"] = cb
            if "value" in cb:
                line["speedup_vs_cpu_baseline"] = line["value"] / cb["value"]
        line["c1_single_thread"] = c1_single_thread(ctx, args.warmup, args.steps)
    line["end_to_end_with_text_io"] = end_to_end_text(ctx, edges, params)
    ctx.close()
    print(json.dumps(line))
    return 0


def protocol_span(ctx, edges, params, warmup, runs):
    """H2D + kernels + D2H per run (HIP events on the context stream, ms),
    host wall time of the same runs, and the transfer sizes.  Each run is one
    cvg_run_to_host into page-locked outputs; the call returns with every row
    in host memory."""
    import numpy as np
    from curvelet_construction_on_gpu_amd import pinned_empty
    from curvelet_construction_on_gpu_amd.hip import Event
    n, sz = edges.shape
    h_in = pinned_empty((n, sz), np.float64)
    h_in[:] = edges
    ctx.run(h_in, params)
    nc, w = ctx.shape()
    h_chains = pinned_empty((nc, w), np.uint32)
    h_info = pinned_empty((nc, 10), np.float64)
    stream = ctx.stream
    e0, e1 = Event(), Event()
    dev_ms, wall_ms = [], []
    for k in range(warmup + runs):
        t0 = time.perf_counter()
        e0.record(stream)
        got, _ = ctx.run_to_host(h_in, params, h_chains, h_info)
        e1.record(stream)
        e1.synchronize()
        t1 = time.perf_counter()
        if got != nc:
            raise RuntimeError(f"row count changed between runs: {got} != {nc}")
        if k >= warmup:
            dev_ms.append(e0.elapsed_ms(e1))
            wall_ms.append((t1 - t0) * 1e3)
    return dev_ms, wall_ms, {"h2d_bytes": int(h_in.nbytes), "d2h_bytes": int(h_chains.nbytes + h_info.nbytes)}


def pipelined_pcie(edges, params, maps):
    """Host-to-host edges/s when consecutive maps overlap: two contexts (own
    streams) on two host threads, each running map after map from page-locked
    input to page-locked outputs, so one map's transfers run beside another's
    kernels (the served form's pattern)."""
    import threading
    import numpy as np
    from curvelet_construction_on_gpu_amd import Context, pinned_empty
    n, sz = edges.shape
    ctxs = [Context(0) for _ in range(2)]
    bufs = []
    for c in ctxs:
        h_in = pinned_empty((n, sz), np.float64)
        h_in[:] = edges
        c.run(h_in, params)
        nc, w = c.shape()
        bufs.append((h_in, pinned_empty((nc, w), np.uint32), pinned_empty((nc, 10), np.float64)))
    per = maps // 2

    def work(i):
        c, (h_in, hc, hi) = ctxs[i], bufs[i]
        for _ in range(per):
            c.run(h_in, params, stats=False)
            c.curvelets_into(hc, hi)

    for i in range(2):  # warm-up
        work(i)
    ths = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    return 2 * per * n / dt


def c1_single_thread(ctx, warmup, runs):
    """The CPU-runnable config (320x240, ~6k edges, defaults): the reference's
    code on one host thread against the GPU's protocol span on the same map."""
    from curvelet_construction_on_gpu_amd import CurveletParams, synth_edges
    try:
        e = synth_edges("C1", 1001)
        p = CurveletParams()
        gpu_ms, _, _ = protocol_span(ctx, e, p, warmup, runs)
        gpu = len(e) / (statistics.median(gpu_ms) * 1e-3)
        cpu = cpu_time(e, p, 1, reps=3)
        return {"edges": int(len(e)), "gpu_edges_per_s": gpu, "cpu_1_thread_edges_per_s": len(e) / cpu["seconds"],
                "cpu_kind": cpu["kind"], "speedup": gpu / (len(e) / cpu["seconds"])}
    except Exception as exc:  # reported, never required
        return {"error": repr(exc)}


def end_to_end_text(ctx, edges, params, reps=3):
    """Edge file text in -> chain and info files out: the library's reader,
    the host-to-host path and its writers (the reference parses and writes
    outside its timer, main.cpp:37-44 and :68-91); median of reps."""
    import tempfile
    import numpy as np
    from curvelet_construction_on_gpu_amd import read_TO_edges_from_file, write_chain_file, write_info_file
    with tempfile.TemporaryDirectory() as d:
        src = Path(d) / "edges.txt"
        np.savetxt(src, edges, fmt="%.17g", delimiter=" ")
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            e = read_TO_edges_from_file(src)
            ctx.run(e, params, stats=False)
            chains, info = ctx.curvelets()
            write_chain_file(Path(d) / "chain.txt", chains)
            write_info_file(Path(d) / "info.txt", info)
            times.append(time.perf_counter() - t0)
        t = statistics.median(times)
        return {"edges_per_s": len(edges) / t, "ms": t * 1e3, "input_bytes": src.stat().st_size,
                "output_bytes": (Path(d) / "chain.txt").stat().st_size + (Path(d) / "info.txt").stat().st_size,
                "runs": reps, "statistic": "median"}


if __name__ == "__main__":
    sys.exit(main())
