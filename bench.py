"""RCB throughput on one MI355X.

Default workload: 512^3 = 2^27 jittered hex-cell centroids (f64), lognormal
integer-valued f64 weights, iter_count 12 (4096 parts), tolerance 0.01.
A step is one full RCB of the whole point set with inputs already resident in
HBM (coupe_hip_rcb_device): prologue, 12 levels, part ids written to HBM.

Prints one JSON line: value = points / second (Mpts/s) for the step, the
dominant kernel's achieved bandwidth against the 8 TB/s HBM peak (HIP events
on the launch stream), and the CPU oracle (fork-join C restatement of coupe's
Rcb) timed on the same input on this host, with an id-for-id parity check.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                       [--no-cpu] [--cpu-threads T]
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
    """Whole-job algorithmic bytes per point with coupe's working types
    (f32 coordinates, f64 weights, u32 index, u64 ids):
    prologue (12D+20) + one split sweep per level 12 L + partition of levels
    0..L-2 (L-1)(8D+24) + ids 12."""
    return (12 * dim + 20) + 12 * levels + (levels - 1) * (8 * dim + 24) + 12


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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the partitioner runs on a "
              f"single device, multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import torch

    import coupe_amd
    from coupe_amd import workloads

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
    roofline = {
        "kernel": dominant,
        "bound": "hbm",
        "achieved": round(dk["GBps"], 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(dk["GBps"] / HBM_PEAK_GBS, 4),
        "traffic": None,
        "avg_launch_ms": round(dk["avg_launch_ms"], 4),
        "algo_bytes_per_launch": dk["algo_bytes_per_launch"],
    }
    value = n / (ms_per_step * 1e-3) / 1e6
    bpp = algo_bytes_per_point(dim, iters)
    whole_job = {
        "algo_bytes_per_point": bpp,
        "achieved_GBps": round(n * bpp / (ms_per_step * 1e-3) / 1e9, 1),
        "frac_of_hbm_peak": round(n * bpp / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }

    # CPU baseline and parity on the same input.
    cpu = None
    parity = None
    gpu_ids = d_out.cpu().numpy().astype(np.uint64)
    if not args.no_cpu:
        from oracle import oracle

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu oracle on {threads} threads ...")
        t1 = time.perf_counter()
        err, cpu_ids, cst = oracle.rcb(pts, w, iters, tol, parallel=True, nthreads=threads)
        cpu_s = time.perf_counter() - t1
        assert err == 0
        cpu = {"value": round(n / cpu_s / 1e6, 2), "unit": "Mpts/s", "cores": threads, "kind": "port",
               "sample": f"the full workload ({n} points), one run, fork-join C restatement of coupe::Rcb "
                         f"(OpenMP tasks for rayon::join, chunked parallel split passes), {cpu_s:.2f} s"}
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
        "step_ms": {"median": round(step_ms[len(step_ms) // 2], 4), "min": round(step_ms[0], 4),
                    "max": round(step_ms[-1], 4), "profiled_median": round(sorted(prof_ms)[len(prof_ms) // 2], 4)},
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
