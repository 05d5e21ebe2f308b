"""Benchmark: effective HBM GB/s of plain plan->exec() on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]

One step = one exec() of the plan over the 6-D fp32 workload
{16,24,32,16,24,32}, perm {5,2,0,4,1,3}, alpha=2.3, beta=4.2 (A and B
resident in HBM before the timed region).  Effective bytes per step are
N*4*(2 + [beta used]) = read A + write B + read B.  The JSON line also carries
the same figure for fp64 and the complex<double> 4-D workloads (tile and
stride-1-kept paths), the 3-D fp32 copy workload (cache-flushed), the roofline
fraction, and a CPU baseline: the oracle (a port of the reference's OpenMP
RefTrans loop, test/inc/hptt/test_util.tcc:143-185) timed on host cores on
the same fp32 workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_GBPS = 8000.0  # MI355X HBM3E
METRIC = "effective GB/s & % HBM peak; batched transposes/s, p50/p99 vs plain exec"

WORKLOADS = {
    "6d-f32": dict(size=[16, 24, 32, 16, 24, 32], perm=[5, 2, 0, 4, 1, 3], dtype="float32",
                   alpha=2.3, beta=4.2, flush=False),
    "6d-f64": dict(size=[16, 24, 32, 16, 24, 32], perm=[5, 2, 0, 4, 1, 3], dtype="float64",
                   alpha=2.3, beta=4.2, flush=False),
    "4d-c128-tile": dict(size=[96, 96, 96, 96], perm=[3, 2, 1, 0], dtype="complex128",
                         alpha=complex(1.1, 0.3), beta=complex(0.7, -0.2), flush=False),
    "4d-c128-wide": dict(size=[96, 96, 96, 96], perm=[0, 3, 2, 1], dtype="complex128",
                         alpha=complex(1.1, 0.3), beta=complex(0.7, -0.2), flush=False),
    "3d-f32-copy": dict(size=[64, 128, 256], perm=[2, 1, 0], dtype="float32",
                        alpha=1.0, beta=0.0, flush=True),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_workload(name, steps, warmup):
    """Returns a dict with GB/s, ms per exec and plan info for one workload."""
    import torch

    import hptt_amd
    w = WORKLOADS[name]
    tdt = getattr(torch, w["dtype"])
    n = int(np.prod(w["size"]))
    g = torch.Generator(device="cuda").manual_seed(1234)
    A = (torch.rand(n, device="cuda", generator=g, dtype=tdt) * 2 - 1)
    B = (torch.rand(n, device="cuda", generator=g, dtype=tdt) * 2 - 1)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    plan = hptt_amd.create_plan_checked(A, B, w["size"], w["perm"], w["alpha"], w["beta"],
                                        stream=stream)
    info = plan.info()
    nbytes = info.algorithmic_bytes
    flush = torch.empty(512 * 2**20 // 4, device="cuda", dtype=torch.float32) if w["flush"] else None
    for _ in range(warmup):
        plan.exec()
    plan.sync()
    torch.cuda.synchronize()
    if not w["flush"]:
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            plan.exec()
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = e0.elapsed_time(e1) / steps
    else:
        # Data (16.8 MB) would sit in the 256 MB last-level cache: overwrite a
        # 512 MB buffer before each launch and time each launch on its own.
        evs = []
        t0 = time.perf_counter()
        for _ in range(steps):
            with torch.cuda.stream(stream):
                flush.fill_(1.0)
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            plan.exec()
            b.record(stream)
            evs.append((a, b))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        times = sorted(a.elapsed_time(b) for a, b in evs)
        ms = times[len(times) // 2]
    gbps = nbytes / (ms * 1e-3) / 1e9
    res = {
        "gbps": round(gbps, 1), "pct_hbm_peak": round(100 * gbps / PEAK_GBPS, 2),
        "ms_per_exec": round(ms, 5), "bytes_per_exec": nbytes, "elements": n,
        "reference_metric_gibps": round(hptt_amd.reference_gibps(n, info.elem_bytes, ms * 1e-3), 1),
        "path": hptt_amd.PATHS[info.path], "tile_r": info.tile_r, "tile_c": info.tile_c,
        "tile_o": info.tile_o, "vec": [info.vec_r, info.vec_w], "grid": info.grid,
        "wall_s": round(wall, 3), "steps": steps,
    }
    del plan, A, B, flush
    torch.cuda.empty_cache()
    return res


def cpu_baseline(budget_s=12.0):
    """Oracle (port of the reference's OpenMP RefTrans) on the fp32 workload."""
    import oracle
    w = WORKLOADS["6d-f32"]
    n = int(np.prod(w["size"]))
    threads = min(os.cpu_count() or 1, 16)
    A = oracle.uniform(n, np.float32, 1, nthreads=threads)
    B = oracle.uniform(n, np.float32, 2, nthreads=threads)
    times = []
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < budget_s and len(times) < 10:
        t0 = time.perf_counter()
        rc = oracle.transpose(A, B, w["size"], w["perm"], w["alpha"], w["beta"], nthreads=threads)
        times.append(time.perf_counter() - t0)
        assert rc == 0
    best = min(times)
    nbytes = n * 4 * 3
    return {"value": round(nbytes / best / 1e9, 2), "unit": "GB/s", "cores": threads,
            "kind": "port",
            "sample": "full 6-D fp32 workload (151M elements), %d runs, best of; "
                      "oracle OpenMP loop (reference RefTrans order), %d threads"
                      % (len(times), threads)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="main workload only")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(see DESIGN.md for how a transposition would shard)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm device", file=sys.stderr)
        sys.exit(1)
    import hptt_amd
    hptt_amd.lib()
    log("bench: main workload 6d-f32, %d steps" % args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    main_res = time_workload("6d-f32", args.steps, args.warmup)
    torch.cuda.synchronize()
    log("bench: 6d-f32 %.1f GB/s (%.1f%% of peak)" % (main_res["gbps"], main_res["pct_hbm_peak"]))
    secondary = {}
    if not args.quick:
        for name in ("6d-f64", "4d-c128-tile", "4d-c128-wide", "3d-f32-copy"):
            secondary[name] = time_workload(name, max(10, args.steps // 2), args.warmup)
            log("bench: %s %.1f GB/s" % (name, secondary[name]["gbps"]))
    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (oracle, OpenMP)")
        cpu = cpu_baseline()
        log("bench: CPU %.2f GB/s on %d threads" % (cpu["value"], cpu["cores"]))
    gbps = main_res["gbps"]
    line = {
        "metric": METRIC,
        "value": gbps,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_exec"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: A, B uniform[-1,1) from a fixed seed, generated on the device",
        "config": {"workload": "plain single transpose fp32 6-D {16,24,32,16,24,32} "
                               "perm {5,2,0,4,1,3} alpha=2.3 beta=4.2",
                   "elements": main_res["elements"], "bytes_per_step": main_res["bytes_per_exec"],
                   "path": main_res["path"]},
        "pct_hbm_peak": main_res["pct_hbm_peak"],
        "reference_metric_gibps": main_res["reference_metric_gibps"],
        "roofline": {"bound": "hbm", "achieved": gbps, "peak": PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbps / PEAK_GBPS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
