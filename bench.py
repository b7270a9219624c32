"""RCB throughput on one MI355X.

Default workload: 512^3 = 2^27 jittered hex-cell centroids (f64), lognormal
integer-valued f64 weights, iter_count 12 (4096 parts), tolerance 0.01.
A step is one full RCB of the whole point set with inputs already resident in
HBM (coupe_hip_rcb_device): prologue, 12 levels, part ids written to HBM.

Prints one JSON line: value = points / second (Mpts/s) for the step, the
dominant kernel's achieved bandwidth against the 8 TB/s HBM peak (HIP events
on the launch stream), and the CPU oracle (fork-join C restatement of coupe's
Rcb) timed on the same input on this host, with an id-for-id parity check.

Two more measurements ride in the same line (skip them with --plain-only):
  "batched": 4096 independent instances (sizes log-uniform in [10^3, 10^6],
             D in {2, 3}, integer weights in [1, 16], iter_count 4..8,
             tolerance 0.02, seed 3) partitioned by coupe_hip_rcb_batch (one
             call per dimension) versus one coupe_hip_rcb_device call per
             instance on the same inputs, ids compared instance by instance.
  "served":  a Poisson request stream (mean 200 req/s, sizes log-uniform in
             [10^4, 10^7] 3D points, weights 1..100, iter_count 6..12, seed 4)
             through the request server; 30% of requests re-partition a mesh
             resident in HBM after every point moved by N(0, 0.01) and 5% of
             its weights were resampled.  p50/p99 latency (submit -> ids in
             the caller's buffer) and aggregate Mpts/s.

Also in the line: "stream" (a STREAM-style probe of the device with the
partition's shape: 4 f32 arrays read and 4 written; the practical HBM
ceiling next to the 8 TB/s data-sheet peak), "end_to_end_host" (coupe_rcb
from host memory, PCIe copies included), "configs" (the 2^20-point 2D and
2^24-point Gaussian 3D workloads, device-resident), and the CPU baseline as
the median of 3 runs on all host threads plus a 1-thread run on a 2^24-point
sample of the same generator.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                       [--no-cpu] [--cpu-threads T] [--plain-only]
                       [--serve-requests R] [--grid-only]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RCB Mpts/s (3D, 2^27 pts, 4096 parts) batched vs plain on 1 GPU; p99 latency"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "jgrid512": "512^3 = 2^27 jittered hex-cell centroids, lognormal integer weights, iter_count 12 (4096 parts), "
                "tolerance 0.01, seed 2",
    "gauss3d": "2^24 points from 8 Gaussians (sigma 0.05), weights 1..100, iter_count 10 (1024 parts), "
               "tolerance 0.01, seed 1",
    "uniform2d": "2^20 points uniform in [0,1)^2, unit weights, iter_count 3 (8 parts), tolerance 0.05, seed 0",
    "grid512_unit": "512^3 unjittered cell centres, unit weights, iter_count 12, tolerance 0.01",
}


def algo_bytes_per_point(dim: int, levels: int) -> int:
    """Bytes per point of RCB done the straightforward way with coupe's
    working types (f32 coordinates, f64 weights, u32 index, u64 ids): prologue
    (12D+20) + one split sweep per level 12 L + a move after every level but
    the last (L-1)(8D+24) + ids 12.  The engine moves less than this (two
    levels per move, 4-byte integer weights); ``whole_job.engine_*`` counts
    what it does read and write."""
    return (12 * dim + 20) + 12 * levels + (levels - 1) * (8 * dim + 24) + 12


TRAFFIC_PROFILES = {("jgrid512", "partition"): "profiles/rcb_jgrid512_hbm_traffic.md",
                    ("gauss3d", "hist"): "profiles/rcb_gauss3d_hbm_traffic.md",
                    ("gauss3d", "partition"): "profiles/rcb_gauss3d_hbm_traffic.md"}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a 1 GiB copy), or
    None when no such profile exists for this workload and kernel."""
    path = TRAFFIC_PROFILES.get((workload, kernel))
    if not path or not os.path.exists(os.path.join(ROOT, path)):
        return None
    text = open(os.path.join(ROOT, path)).read()
    block = text.split("```json", 1)[-1].split("```", 1)[0] if "```json" in text else ""
    try:
        data = json.loads(block)
        if "classes" in data:   # per kernel class, averaged over the class's dispatches
            return int(data["classes"][kernel]["traffic_bytes"])
        return int(data["traffic_bytes"])
    except (ValueError, KeyError):
        return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="jgrid512", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a plain device-to-device copy of this many MiB (counter calibration)")
    ap.add_argument("--plain-only", action="store_true", help="skip the batched and served measurements")
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--grid-only", action="store_true",
                    help="only coupe's Grid benchmark (10000 x 10000, L = 12); prints its JSON object")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the partitioner runs on a "
              f"single device, multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import torch

    import coupe_amd
    from coupe_amd import workloads

    if args.grid_only:
        print(json.dumps(bench_grid(torch, coupe_amd, with_cpu=not args.no_cpu, steps=args.steps)), flush=True)
        return

    stream_probe = probe_stream()
    t0 = time.time()
    pts, w, iters, tol = workloads.standard(args.workload)
    n, dim = pts.shape
    log(f"workload {args.workload}: n={n} dim={dim} L={iters} tol={tol} generated in {time.time() - t0:.1f}s")

    dev = torch.device("cuda:0")
    d_pts = torch.from_numpy(pts).to(dev)
    d_w = torch.from_numpy(w).to(dev)
    d_out = torch.zeros(n, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    ctx = coupe_amd.DeviceContext(0)
    rcb = coupe_amd.Rcb(iter_count=iters, tolerance=tol)
    stream = torch.cuda.current_stream(dev)

    def step():
        rcb.partition(d_out, (d_pts, d_w), context=ctx, stream=stream.cuda_stream)

    copy_probe = None
    if args.copy_probe_mib:
        nbytes = args.copy_probe_mib << 20
        src = torch.empty(nbytes, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        c0.record(stream)
        for _ in range(reps):
            dst.copy_(src)
        c1.record(stream)
        torch.cuda.synchronize()
        ms = c0.elapsed_time(c1) / reps
        copy_probe = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": round(ms, 4),
                      "GBps_read_plus_write": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}
        del src, dst
        log(f"copy probe {args.copy_probe_mib} MiB: {ms:.3f} ms, {copy_probe['GBps_read_plus_write']} GB/s")

    # A first call on a fresh context (buffers allocated, no hints).
    cold = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_cold = time.perf_counter()
    cold[0].record(stream)
    step()
    cold[1].record(stream)
    torch.cuda.synchronize()
    cold_wall_ms = (time.perf_counter() - t_cold) * 1e3
    cold_ms = cold[0].elapsed_time(cold[1])
    log(f"cold call (fresh context): {cold_ms:.3f} ms on the stream, {cold_wall_ms:.1f} ms wall")
    for i in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} done")

    # Timed region: K steps between two events on the launch stream.
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    per_step = []
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(args.steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        step()
        e1.record(stream)
        per_step.append((e0, e1))
    stop.record(stream)
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    ms_per_step = total_ms / args.steps
    step_ms = sorted(a.elapsed_time(b) for a, b in per_step)
    stats = ctx.stats()
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, median {step_ms[len(step_ms) // 2]:.3f} ms")

    # The same shape with other data (another seed) on the warm context: its
    # hints come from the benchmark input, not from this input.
    fresh = None
    if args.workload in ("jgrid512", "gauss3d"):
        fp, fw = (workloads.jgrid(512, 7, True) if args.workload == "jgrid512"
                  else workloads.gauss_mix(pts.shape[0], 7, 0.05))
        d_fp, d_fw = torch.from_numpy(fp).to(dev), torch.from_numpy(fw).to(dev)
        f_out = torch.zeros(n, dtype=torch.int64, device=dev)
        torch.cuda.synchronize()
        fe = []
        for _ in range(2):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            rcb.partition(f_out, (d_fp, d_fw), context=ctx, stream=stream.cuda_stream)
            e1.record(stream)
            fe.append((e0, e1))
            torch.cuda.synchronize()
        fresh = {"seed": 7, "first_ms": round(fe[0][0].elapsed_time(fe[0][1]), 4),
                 "second_ms": round(fe[1][0].elapsed_time(fe[1][1]), 4),
                 "reruns": ctx.stats()["reruns"]}
        del d_fp, d_fw, f_out, fp, fw
        step()   # back to the benchmark input's hints
        torch.cuda.synchronize()
        log(f"fresh data (seed 7), warm context: {fresh['first_ms']} ms, then {fresh['second_ms']} ms")

    # Per-kernel timing: the same steps with events around every launch.
    ctx.set_profiling(True)
    kt_sum = {}
    prof_ms = []
    for _ in range(args.steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        step()
        e1.record(stream)
        torch.cuda.synchronize()
        prof_ms.append(e0.elapsed_time(e1))
        for k, v in ctx.kernel_times().items():
            acc = kt_sum.setdefault(k, {"launches": 0, "total_ms": 0.0, "algo_bytes": 0.0})
            for f in acc:
                acc[f] += v[f]
    ctx.set_profiling(False)
    kernels = {}
    for k, v in kt_sum.items():
        if v["launches"] == 0:
            continue
        avg_ms = v["total_ms"] / v["launches"]
        bytes_per_launch = v["algo_bytes"] / v["launches"]
        kernels[k] = {
            "launches_per_step": v["launches"] / args.steps,
            "ms_per_step": v["total_ms"] / args.steps,
            "avg_launch_ms": avg_ms,
            "algo_bytes_per_launch": bytes_per_launch,
            "GBps": (bytes_per_launch / (avg_ms * 1e-3) / 1e9) if avg_ms > 0 and bytes_per_launch > 0 else None,
        }
    dominant = max((k for k in kernels if kernels[k]["algo_bytes_per_launch"] > 0),
                   key=lambda k: kernels[k]["ms_per_step"])
    dk = kernels[dominant]
    stream_copy = stream_probe["copy_GBps"] if stream_probe else None
    # The same launches priced on the survey's basis: the partition moves two
    # levels per launch, each (8D + 24) B/pt with coupe's working types (f32
    # coordinates, f64 weight, u64 index read and written).
    s8d_bytes = None
    if dominant == "partition" and stats["weight_mode"] == 0:
        s8d_bytes = 2 * (8 * dim + 24) * n
    s8d_gbs = s8d_bytes / (dk["avg_launch_ms"] * 1e-3) / 1e9 if s8d_bytes else None
    roofline = {
        "kernel": dominant,
        "bound": "hbm",
        "achieved": round(dk["GBps"], 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(dk["GBps"] / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(args.workload, dominant),
        "basis": "engine: the bytes the move must read and write in this layout (f32 coordinates, 16-bit "
                 "weights, no index stream; two levels per launch), checked against FETCH_SIZE/WRITE_SIZE "
                 "(traffic); frac_s8d prices the same launches at the survey's (8D+24) B/pt per level",
        "frac_engine": round(dk["GBps"] / HBM_PEAK_GBS, 4),
        "achieved_s8d": round(s8d_gbs, 1) if s8d_gbs else None,
        "frac_s8d": round(s8d_gbs / HBM_PEAK_GBS, 4) if s8d_gbs else None,
        "avg_launch_ms": round(dk["avg_launch_ms"], 4),
        "algo_bytes_per_launch": dk["algo_bytes_per_launch"],
        "stream_copy_GBps": stream_copy,
        "frac_of_stream_copy": round(dk["GBps"] / stream_copy, 4) if stream_copy else None,
    }
    value = n / (ms_per_step * 1e-3) / 1e6
    bpp = algo_bytes_per_point(dim, iters)
    # The engine's own algorithmic bytes per step: the sum over kernel classes
    # of (launches x bytes per launch), over the profiled steps.
    engine_bytes = sum(v["algo_bytes_per_launch"] * v["launches_per_step"] for v in kernels.values())
    engine_gbs = engine_bytes / (ms_per_step * 1e-3) / 1e9
    whole_job = {
        "engine_bytes_per_point": round(engine_bytes / n, 2),
        "engine_GBps": round(engine_gbs, 1),
        "frac_of_hbm_peak": round(engine_gbs / HBM_PEAK_GBS, 4),
        "frac_of_stream_copy": round(engine_gbs / stream_copy, 4) if stream_copy else None,
        "one_move_per_level_bytes_per_point": bpp,
    }

    # CPU baseline and parity on the same input.
    cpu = None
    parity = None
    gpu_ids = d_out.cpu().numpy().astype(np.uint64)
    if not args.no_cpu:
        from oracle import oracle

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        runs = []
        for r in range(3):
            log(f"cpu oracle on {threads} threads, run {r + 1}/3 ...")
            t1 = time.perf_counter()
            err, cpu_ids, cst = oracle.rcb(pts, w, iters, tol, parallel=True, nthreads=threads)
            runs.append(time.perf_counter() - t1)
            assert err == 0
        cpu_s = sorted(runs)[1]
        one = cpu_one_thread(oracle, workloads, args.workload)
        cpu = {"value": round(n / cpu_s / 1e6, 2), "unit": "Mpts/s", "cores": threads, "kind": "port",
               "sample": f"the full workload ({n} points), median of 3 runs "
                         f"({', '.join(f'{x:.2f}' for x in runs)} s), fork-join C restatement of coupe::Rcb "
                         f"(OpenMP tasks for rayon::join, chunked parallel split passes)",
               "one_thread": one, "host": host_info()}
        same = bool(np.array_equal(gpu_ids, cpu_ids))
        parity = {"ids_equal_to_cpu_oracle": same,
                  "differing_ids": int(np.count_nonzero(gpu_ids != cpu_ids)),
                  "ambiguous_pivots": int(cst["ambiguous_pivots"])}
        log(f"cpu {cpu_s:.2f}s, parity {same}")

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/i64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload], "name": args.workload, "points": n, "dim": dim,
                   "iter_count": iters, "tolerance": tol, "weights": "f64, integer-valued",
                   "weight_mode": "exact int64 sums" if stats["weight_mode"] == 0 else "f64 sums",
                   "passes": stats["passes"], "max_passes_per_level": stats["max_passes"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "whole_job": whole_job,
        "copy_probe": copy_probe,
        "stream": stream_probe,
        "step_ms": {"median": round(step_ms[len(step_ms) // 2], 4), "min": round(step_ms[0], 4),
                    "max": round(step_ms[-1], 4), "profiled_median": round(sorted(prof_ms)[len(prof_ms) // 2], 4)},
        "cold_ms": round(cold_ms, 4),
        "cold_wall_ms": round(cold_wall_ms, 2),
        "fresh_data_ms": fresh["first_ms"] if fresh else None,
        "fresh_data": fresh,
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
    }
    out["end_to_end_host"] = end_to_end_host(coupe_amd, pts, w, iters, tol, gpu_ids)
    if not args.plain_only:
        del d_pts, d_w, d_out
        torch.cuda.empty_cache()
        out["configs"] = bench_configs(torch, coupe_amd, workloads)
        out["grid"] = bench_grid(torch, coupe_amd, with_cpu=not args.no_cpu)
        torch.cuda.empty_cache()
        out["batched"] = bench_batched(torch, coupe_amd, workloads)
        torch.cuda.empty_cache()
        out["served"] = bench_served(torch, coupe_amd, workloads, args.serve_requests)
    print(json.dumps(out), flush=True)


def probe_stream():
    """STREAM-style bandwidth of the device with the partition's shape
    (libcoupe_probe.so): 4 x 2^27 f32 copied, read, written."""
    import ctypes as C
    path = os.path.join(ROOT, "coupe_amd", "libcoupe_probe.so")
    if not os.path.exists(path):
        log("stream probe library missing; build with `python -m coupe_amd.build probe`")
        return None
    import torch  # noqa: F401  (one HIP runtime in the process)
    L = C.CDLL(path)
    L.coupe_probe_stream.argtypes = [C.c_int, C.c_uint64, C.c_int] + [C.POINTER(C.c_double)] * 3
    L.coupe_probe_stream.restype = C.c_int
    cp, rd, wr = C.c_double(), C.c_double(), C.c_double()
    err = L.coupe_probe_stream(0, 1 << 27, 10, C.byref(cp), C.byref(rd), C.byref(wr))
    if err != 0:
        log(f"stream probe failed: HIP error {err}")
        return None
    res = {"copy_GBps": round(cp.value, 1), "read_GBps": round(rd.value, 1), "write_GBps": round(wr.value, 1),
           "shape": "4 arrays of 2^27 f32 read and 4 written (copy), read only, written only; "
                    "best of 4 and 8 workgroups per CU, 10 launches each"}
    log(f"stream probe: copy {res['copy_GBps']} GB/s, read {res['read_GBps']}, write {res['write_GBps']}")
    return res


def host_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_one_thread(oracle, workloads, workload):
    """The sequential oracle on a 2^24-point sample of the same generator
    (jgrid: the 256^3 grid; other workloads: the 2^24 first points)."""
    if workload == "jgrid512":
        p, w = workloads.jgrid(256, 2, True)
        iters, tol = 12, 0.01
    else:
        p, w, iters, tol = workloads.standard(workload)
        p, w = p[: 1 << 24], w[: 1 << 24]
    t1 = time.perf_counter()
    err, _, _ = oracle.rcb(p, w, iters, tol)
    s = time.perf_counter() - t1
    assert err == 0
    log(f"cpu oracle, 1 thread, {p.shape[0]} points: {s:.2f} s")
    return {"value": round(p.shape[0] / s / 1e6, 2), "unit": "Mpts/s", "cores": 1,
            "sample": f"{p.shape[0]} points of the same generator, iter_count {iters}, one run, {s:.2f} s"}


def end_to_end_host(coupe_amd, pts, w, iters, tol, gpu_ids):
    """coupe_rcb from host memory: H2D of points and weights, the kernels, D2H of the ids."""
    part = np.zeros(pts.shape[0], dtype=np.uint64)
    rcb = coupe_amd.Rcb(iter_count=iters, tolerance=tol)
    rcb.partition(part, (pts, w))   # warm-up (staging buffers)
    times = []
    for _ in range(3):
        t1 = time.perf_counter()
        rcb.partition(part, (pts, w))
        times.append(time.perf_counter() - t1)
    s = sorted(times)[1]
    same = bool(np.array_equal(part, gpu_ids))
    log(f"end to end from host memory: {s * 1e3:.1f} ms, ids equal {same}")
    return {"ms": round(s * 1e3, 2), "Mpts_s": round(pts.shape[0] / s / 1e6, 1),
            "note": "coupe_rcb with host (pageable) buffers: PCIe copies of points, weights and ids included; "
                    "median of 3", "ids_equal_to_device_entry": same}


def bench_configs(torch, coupe_amd, workloads, steps=10, warmup=3):
    """Device-resident throughput of the 2^20-point 2D and 2^24-point Gaussian 3D workloads."""
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    res = {}
    for name in ("uniform2d", "gauss3d"):
        pts, w, iters, tol = workloads.standard(name)
        d_pts, d_w = torch.from_numpy(pts).to(dev), torch.from_numpy(w).to(dev)
        d_out = torch.zeros(pts.shape[0], dtype=torch.int64, device=dev)
        ctx = coupe_amd.DeviceContext(0)
        rcb = coupe_amd.Rcb(iter_count=iters, tolerance=tol)
        for _ in range(warmup):
            rcb.partition(d_out, (d_pts, d_w), context=ctx, stream=stream.cuda_stream)
        torch.cuda.synchronize()
        ev = []
        for _ in range(steps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            rcb.partition(d_out, (d_pts, d_w), context=ctx, stream=stream.cuda_stream)
            e1.record(stream)
            ev.append((e0, e1))
        torch.cuda.synchronize()
        ms = sorted(a.elapsed_time(b) for a, b in ev)[steps // 2]
        st = ctx.stats()
        ctx.close()
        res[name] = {"workload": WORKLOADS[name], "points": int(pts.shape[0]), "median_ms": round(ms, 4),
                     "Mpts_s": round(pts.shape[0] / (ms * 1e-3) / 1e6, 1), "passes": st["passes"]}
        log(f"config {name}: {ms:.3f} ms, {res[name]['Mpts_s']} Mpts/s")
        del d_pts, d_w, d_out
    return res


def bench_grid(torch, coupe_amd, with_cpu=True, side=10000, iters=12, threads=16, steps=10):
    """coupe's own Grid benchmark (coupe/benches/rcb_cartesian.rs): a
    10000 x 10000 grid, weight of cell i = i as f64, iter_count 12; the CPU
    oracle (sequential restatement) on the 2000 x 2000 grid of the same shape
    as the baseline and parity sample."""
    dev = torch.device("cuda:0")
    n = side * side
    w = torch.arange(n, dtype=torch.float64, device=dev)
    out = torch.zeros(n, dtype=torch.int64, device=dev)
    ctx = coupe_amd.DeviceContext(0)
    grid = coupe_amd.Grid.new_2d(side, side)
    stream = torch.cuda.current_stream(dev)
    for _ in range(3):
        grid.rcb(out, w, iters, threads=threads, context=ctx, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    ev = []
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        grid.rcb(out, w, iters, threads=threads, context=ctx, stream=stream.cuda_stream)
        e1.record(stream)
        ev.append((e0, e1))
    torch.cuda.synchronize()
    ms = sorted(a.elapsed_time(b) for a, b in ev)[steps // 2]
    # bytes of the per-level formulation (one read of the weights per level,
    # one write of the ids), and of the summed-area path that runs here (two
    # reads of the weights, one write of the table, one of the ids)
    gbps = n * (8.0 * iters + 8.0) / (ms * 1e-3) / 1e9
    res = {"workload": f"{side}x{side} grid, weight of cell i = i (f64), iter_count {iters}, rayon pool size "
                       f"{threads} for the median's chunking (coupe/benches/rcb_cartesian.rs)",
           "cells": n, "median_ms": round(ms, 4), "Mcells_s": round(n / (ms * 1e-3) / 1e6, 1),
           "algo_GBps": round(gbps, 1),
           "path_GBps": round(n * 32.0 / (ms * 1e-3) / 1e9, 1)}
    log(f"grid {side}^2: {ms:.3f} ms, {res['Mcells_s']} Mcells/s")
    del w, out
    if with_cpu:
        from oracle import oracle
        s = 2000
        ws = np.arange(s * s, dtype=np.float64)
        part = np.zeros(s * s, dtype=np.uint64)
        grid2 = coupe_amd.Grid.new_2d(s, s)
        grid2.rcb(part, ws, iters, threads=threads, context=ctx)
        t1 = time.perf_counter()
        exp = oracle.grid_rcb([s, s], ws, iters, threads=threads)
        cpu_s = time.perf_counter() - t1
        res["cpu_baseline"] = {"value": round(s * s / cpu_s / 1e6, 2), "unit": "Mcells/s", "cores": 1,
                               "kind": "port", "sample": f"{s}x{s} grid of the same shape, sequential C "
                                                         f"restatement of Grid::rcb, {cpu_s:.2f} s"}
        res["ids_equal_to_cpu_oracle_on_sample"] = bool(np.array_equal(part, exp))
    ctx.close()
    return res


def bench_batched(torch, coupe_amd, workloads, count=4096, seed=3):
    """Batched RCB versus per-instance launches on the same inputs: the 4096
    instances of both dimensions as one coupe_hip_rcb_batch_mixed call (after
    the autotuner chose the switches for the batch's class; calls of a known
    shape run as one captured graph) against one coupe_hip_rcb_device call
    per instance."""
    t0 = time.time()
    n, dim, lo, ext, iters = workloads.batch_header(count, seed=seed, lg_lo=3.0, lg_hi=6.0, l_lo=4, l_hi=8)
    tol = 0.02
    flat, ws = [], []
    for q in range(count):
        d = int(dim[q])
        p, w = workloads.box(int(n[q]), d, lo[q, :d], ext[q, :d], workloads.instance_seed(seed, q), 16)
        flat.append(p.ravel())
        ws.append(w)
    flat, ws = np.concatenate(flat), np.concatenate(ws)
    specs = [(int(n[q]), int(dim[q]), int(iters[q]), tol) for q in range(count)]
    total = int(n.sum())
    log(f"batched: {count} instances, {total} points generated in {time.time() - t0:.1f}s")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    ctx = coupe_amd.DeviceContext(0)
    d_flat, d_w = torch.from_numpy(flat).to(dev), torch.from_numpy(ws).to(dev)
    out_b = torch.zeros(total, dtype=torch.int64, device=dev)
    out_p = torch.zeros(total, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    tuning = coupe_amd.autotune(out_b, d_flat, d_w, specs, trials=2, context=ctx, stream=stream.cuda_stream)
    log(f"autotune: {tuning} in {time.perf_counter() - t1:.1f}s")

    def run_batched():
        coupe_amd.partition_batch_mixed(out_b, d_flat, d_w, specs, context=ctx, stream=stream.cuda_stream)

    offs = np.concatenate([[0], np.cumsum([k * d for k, d, _, _ in specs])])
    poffs = np.concatenate([[0], np.cumsum([k for k, _, _, _ in specs])])

    def run_plain():
        for q, (k, d, it, _) in enumerate(specs):
            o, po = int(offs[q]), int(poffs[q])
            coupe_amd.Rcb(it, tol).partition(out_p[po:po + k], (d_flat[o:o + k * d].view(k, d), d_w[po:po + k]),
                                             context=ctx, stream=stream.cuda_stream)

    res = {}
    for fn, name, reps in ((run_batched, "batched", 5), (run_plain, "plain", 1)):
        for _ in range(2):
            fn()   # warm-up: buffers, hints, the graph of the batch
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t1 = time.perf_counter()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        wall_ms = (time.perf_counter() - t1) * 1e3 / reps
        res[name] = {"ms": round(max(e0.elapsed_time(e1) / reps, wall_ms), 3)}
        res[name]["Mpts_s"] = round(total / (res[name]["ms"] * 1e-3) / 1e6, 1)
        if name == "batched":
            res[name]["stats"] = ctx.stats()
        log(f"batched bench: {name} {res[name]['ms']:.1f} ms, {res[name]['Mpts_s']} Mpts/s")
    same = bool(torch.equal(out_b, out_p))
    st = res["batched"]["stats"]
    ctx.close()
    return {"workload": f"{count} instances, sizes log-uniform in [1e3, 1e6], D in {{2,3}}, weights 1..16, "
                        f"iter_count 4..8, tolerance {tol}, seed {seed}; one call for both dimensions",
            "points": total, "value": res["batched"]["Mpts_s"], "unit": "Mpts/s",
            "batched_ms": res["batched"]["ms"], "per_instance_ms": res["plain"]["ms"],
            "per_instance_Mpts_s": res["plain"]["Mpts_s"],
            "speedup_vs_per_instance": round(res["plain"]["ms"] / res["batched"]["ms"], 2),
            "ids_equal_to_per_instance": same,
            "graph": {0: "kernel by kernel", 1: "one graph launch (replay)", 2: "captured"}[st["graph"]],
            "tuning": {k: v for k, v in tuning.items()}, "tuned": bool(st["tuned"])}


def bench_served(torch, coupe_amd, workloads, n_requests, seed=4, rate=200.0, resident=8, pool=24,
                 sat_rates=(400.0, 800.0, 1600.0, 3200.0), sat_requests=1500, plain_requests=120):
    """Poisson request stream through the request server: new requests from
    pinned host memory, and repartitions of meshes resident in HBM (mesh
    handles) after every point moved by N(0, 0.01) and 5% of the weights were
    resampled, on the device.  Then the same stream at rising rates until
    p99 breaks (served capacity), and the same kind of requests as one plain
    coupe_rcb call at a time (the per-instance baseline)."""
    import threading

    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    t0 = time.time()
    # Host meshes for new partition requests (pinned, reused across requests).
    sizes = np.round(10 ** rng.uniform(4, 7, pool)).astype(np.int64)
    host = []
    for k, m in enumerate(sizes):
        p, w = workloads.uniform(int(m), 3, seed * 1000 + k, 100)
        host.append((torch.from_numpy(p).pin_memory(), torch.from_numpy(w).pin_memory(),
                     torch.zeros(int(m), dtype=torch.int64).pin_memory()))
    srv = coupe_amd.Server(device=0, slots=4, memory_budget=64 << 30, max_batch_points=1 << 26,
                           max_batch_requests=64)
    # Meshes resident in HBM: the server's working copies, moved on the device.
    rsizes = np.round(10 ** rng.uniform(4, 7, resident)).astype(np.int64)
    res = []
    for k, m in enumerate(rsizes):
        p, w = workloads.uniform(int(m), 3, seed * 2000 + k, 100)
        dp, dw = torch.from_numpy(p).to(dev), torch.from_numpy(w).to(dev)
        res.append({"mesh": srv.create_mesh(dp, dw), "p": dp, "w": dw,
                    "out": torch.zeros(int(m), dtype=torch.int64).pin_memory()})
    torch.cuda.synchronize()
    log(f"served: pools ready in {time.time() - t0:.1f}s")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)

    def stream(rate, count, srng):
        lat, tickets = [], []
        lock = threading.Lock()
        done_evt = threading.Event()

        def waiter():
            i = 0
            while True:
                with lock:
                    item = tickets[i] if i < len(tickets) else None
                if item is None:
                    if done_evt.is_set():
                        with lock:
                            if i >= len(tickets):
                                return
                        continue
                    time.sleep(0.0002)
                    continue
                st = srv.wait(item[0])
                st["points"] = item[1]
                lat.append(st)
                i += 1

        th = threading.Thread(target=waiter)
        th.start()
        start = time.perf_counter()
        next_t = start
        n_repart = 0
        for r in range(count):
            next_t += srng.exponential(1.0 / rate)
            iters = int(6 + np.floor(7 * srng.random()))
            repart = srng.random() < 0.3
            k = int(srng.integers(0, resident if repart else pool))
            now = time.perf_counter()
            if next_t > now:
                time.sleep(next_t - now)
            if repart:
                m = res[k]
                # the mesh moves (on the device), then its working copy is updated
                # (after its queued requests) and it is repartitioned
                m["p"].add_(torch.randn(m["p"].shape, generator=gen, device=dev, dtype=m["p"].dtype), alpha=0.01)
                sel = torch.rand(m["w"].shape[0], generator=gen, device=dev) < 0.05
                m["w"][sel] = torch.floor(torch.rand(int(sel.sum()), generator=gen, device=dev,
                                                     dtype=m["w"].dtype) * 100) + 1
                m["mesh"].update(points=m["p"], weights=m["w"])
                t = srv.submit_mesh(m["mesh"], m["out"].numpy(), iters, 0.01)
                n_repart += 1
                npts = m["p"].shape[0]
            else:
                hp, hw, ho = host[k]
                t = srv.submit(ho.numpy(), hp.numpy(), hw.numpy(), iters, 0.01)
                npts = hp.shape[0]
            with lock:
                tickets.append((t, npts))
        done_evt.set()
        th.join()
        wall = time.perf_counter() - start
        tot = np.array([x["total_ms"] for x in lat])
        run = np.array([x["run_ms"] for x in lat])
        bsz = np.array([x["batch_requests"] for x in lat])
        pts_total = float(np.sum([x["points"] for x in lat]))
        return {"rate": rate, "requests": count, "repartitions": n_repart,
                "p50_ms": round(float(np.percentile(tot, 50)), 3), "p99_ms": round(float(np.percentile(tot, 99)), 3),
                "p50_run_ms": round(float(np.percentile(run, 50)), 3),
                "p99_run_ms": round(float(np.percentile(run, 99)), 3),
                "mean_batch_requests": round(float(bsz.mean()), 2),
                "achieved_req_s": round(count / wall, 1),
                "Mpts_s": round(pts_total / wall / 1e6, 1), "wall_s": round(wall, 2)}

    main = stream(rate, n_requests, np.random.default_rng(seed))
    log(f"served at {rate:.0f} req/s: p50 {main['p50_ms']} ms, p99 {main['p99_ms']} ms, {main['Mpts_s']} Mpts/s")
    # Saturation: rising rates until the server no longer keeps up (achieved
    # rate under 90% of the offered one) or p99 exceeds 4x the nominal p99.
    sweep, capacity = [], None
    for r in sat_rates:
        x = stream(r, sat_requests, np.random.default_rng(seed + int(r)))
        sweep.append(x)
        log(f"served at {r:.0f} req/s offered: {x['achieved_req_s']} req/s, p99 {x['p99_ms']} ms, "
            f"{x['Mpts_s']} Mpts/s, mean batch {x['mean_batch_requests']}")
        ok = x["achieved_req_s"] >= 0.9 * r and x["p99_ms"] <= 4 * main["p99_ms"]
        if ok:
            capacity = x
        else:
            break
    peak = max([main] + sweep, key=lambda x: x["Mpts_s"])
    st = srv.stats()
    for m in res:
        m["mesh"].destroy()
    srv.close()
    # Per-instance baseline: the same kind of requests, one coupe_rcb call at a time.
    prng = np.random.default_rng(seed + 99)
    pts_plain = 0
    t1 = time.perf_counter()
    for _ in range(plain_requests):
        hp, hw, ho = host[int(prng.integers(0, pool))]
        coupe_amd.Rcb(int(6 + np.floor(7 * prng.random())), 0.01).partition(ho.numpy().view(np.uint64),
                                                                             (hp.numpy(), hw.numpy()))
        pts_plain += hp.shape[0]
    plain_s = time.perf_counter() - t1
    plain = {"requests": plain_requests, "Mpts_s": round(pts_plain / plain_s / 1e6, 1),
             "req_s": round(plain_requests / plain_s, 1)}
    log(f"plain coupe_rcb one request at a time: {plain['Mpts_s']} Mpts/s")
    out = dict(main)
    out.update({
        "workload": f"Poisson {rate:.0f} req/s, sizes log-uniform in [1e4, 1e7] 3D points, weights 1..100, "
                    f"iter_count 6..12, tolerance 0.01, new requests from {pool} pinned host meshes, 30% "
                    f"re-partitions of {resident} meshes resident in HBM through mesh handles (N(0,0.01) drift "
                    f"and 5% of the weights resampled on the device, then the working copy updated), seed {seed}",
        "saturation": sweep,
        "capacity": ({"req_s": capacity["achieved_req_s"], "Mpts_s": capacity["Mpts_s"], "p99_ms": capacity["p99_ms"]}
                     if capacity else None),
        "peak_Mpts_s": peak["Mpts_s"],
        "per_instance_plain": plain,
        "peak_vs_per_instance_plain": round(peak["Mpts_s"] / plain["Mpts_s"], 2) if plain["Mpts_s"] else None,
        "peak_device_bytes": st["peak_device_bytes"], "memory_budget": st["memory_budget"],
    })
    return out


if __name__ == "__main__":
    main()
