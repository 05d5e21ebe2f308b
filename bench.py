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
    # identity perm (axpby path): a plain streaming copy of 2 x 512 MiB, the
    # streaming reference the transposes are compared with
    "copy-f32-512m": dict(size=[131072, 1024], perm=[0, 1], dtype="float32",
                          alpha=1.0, beta=0.0, flush=False),
    # identity perm with beta on the main workload's element count: the same
    # bytes and the same 2:1 read:write mix as the 6-D transposes, with no
    # transposition -- their streaming reference
    "axpby-f32-151m": dict(size=[12288, 12288], perm=[0, 1], dtype="float32",
                           alpha=2.3, beta=4.2, flush=False),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


DTYPE_TAG = {"float32": "f32", "float64": "f64", "complex64": "c64 (f32 pairs)",
             "complex128": "c128 (f64 pairs)"}


def describe_workload(name):
    w = WORKLOADS[name]
    fmt = lambda v: ("%g%+gi" % (v.real, v.imag)) if isinstance(v, complex) else "%g" % v
    return "plain single transpose %s %d-D {%s} perm {%s} alpha=%s beta=%s" % (
        w["dtype"], len(w["size"]), ",".join(map(str, w["size"])), ",".join(map(str, w["perm"])),
        fmt(w["alpha"]), fmt(w["beta"]))


TRAFFIC_FILE = os.path.join("profiles", "hbm_traffic_pmc.md")


def pmc_traffic(name):
    """HBM bytes per launch of the workload's kernel as measured by rocprofv3
    FETCH_SIZE / WRITE_SIZE (scaled by a plain copy of known size), read from
    the committed summary; None when it holds no figure for the workload."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as f:
            for ln in f:
                if ln.startswith("traffic[%s] = " % name):
                    return int(ln.split("=")[1])
    except OSError:
        pass
    return None


def time_workload(name, steps, warmup, tune=0.0):
    """Returns a dict with GB/s, ms per exec and plan info for one workload
    (tune > 0: autotune the plan for that many seconds first)."""
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
                                        stream=stream, tuning_timeout=tune)
    info = plan.info()
    log(plan.describe().rstrip())
    nbytes = info.algorithmic_bytes
    flush = (torch.ones(512 * 2**20 // 4, device="cuda", dtype=torch.float32) if w["flush"]
             else None)
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
        # Data (16.8 MB) would sit in the 256 MB last-level cache: read a
        # 512 MB buffer before each launch and time each launch on its own.
        # A read leaves the cache cold and clean; writing the buffer instead
        # would leave it full of dirty lines, whose write-back the timed
        # launch would then pay for (1.38 against 1.46 TB/s).
        evs = []
        t0 = time.perf_counter()
        for _ in range(steps):
            with torch.cuda.stream(stream):
                flush.sum()
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


def time_batched(dtype_name, reps=5, seq_reps=3, tune=0.0, one_launch=False):
    """The batched workload: 4096 tensors (seeded manifest, rank 2..6, dims
    log-uniform in [2,128], <= 4 Mi elements each, random perms, 25% beta=0)
    as one Batch (work-queue launches, graph replay) against one plan->exec()
    per tensor in sequence on the same stream; identical bits required."""
    import torch

    import hptt_amd
    from hptt_amd.workloads import batch_manifest, layout
    man = batch_manifest()
    tdt = getattr(torch, dtype_name)
    esz = torch.empty(0, dtype=tdt).element_size()
    offs, total = layout(man, esz)
    real = torch.float32 if dtype_name in ("float32", "complex64") else torch.float64
    ncomp = 2 if "complex" in dtype_name else 1
    g = torch.Generator(device="cuda").manual_seed(4096)
    A = torch.empty(total * ncomp, dtype=real, device="cuda").uniform_(-1, 1, generator=g)
    B0 = torch.empty(total * ncomp, dtype=real, device="cuda").uniform_(-1, 1, generator=g)
    A = A.view(tdt) if ncomp == 2 else A
    B0 = B0.view(tdt) if ncomp == 2 else B0
    B = B0.clone()
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    items, plans = [], []
    for m, o in zip(man, offs):
        n = int(np.prod(m["in_size"]))
        a, b = A[o:o + n], B[o:o + n]
        items.append(hptt_amd.make_batch_item(a, b, m["in_size"], m["perm"], m["alpha"], m["beta"]))
        plans.append(hptt_amd.create_plan_checked(a, b, m["in_size"], m["perm"], m["alpha"],
                                                  m["beta"], stream=stream))
    t_build = time.perf_counter()
    batch = hptt_amd.Batch(items, stream=stream, tuning_timeout=tune, one_launch=one_launch)
    t_build = time.perf_counter() - t_build
    info = batch.info()
    exec_fn = hptt_amd.lib().hpttx_plan_exec
    handles = [p._h for p in plans]

    def run_seq():
        for h in handles:
            if exec_fn(h) != 0:
                raise RuntimeError(hptt_amd.last_error())

    # warm up both (the first batch exec captures the graph)
    run_seq()
    batch.exec()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    seq_ms = []
    for _ in range(seq_reps):
        e0.record(stream)
        run_seq()
        e1.record(stream)
        torch.cuda.synchronize()
        seq_ms.append(e0.elapsed_time(e1))
    bat_ms = []
    for _ in range(reps):
        e0.record(stream)
        batch.exec()
        e1.record(stream)
        torch.cuda.synchronize()
        bat_ms.append(e0.elapsed_time(e1))
    # identical output: sequential plans vs the batch from the same B0
    B.copy_(B0)
    torch.cuda.synchronize()
    run_seq()
    torch.cuda.synchronize()
    Bseq = B.clone()
    B.copy_(B0)
    torch.cuda.synchronize()
    batch.exec()
    torch.cuda.synchronize()
    iv = torch.int32 if esz == 4 else torch.int64
    same = bool(torch.equal(B.view(real).view(iv), Bseq.view(real).view(iv)))
    seq = float(np.median(seq_ms))
    bat = float(np.median(bat_ms))
    res = {"tensors": len(man), "batched_ms": round(bat, 3), "sequential_ms": round(seq, 3),
           "batched_transposes_per_s": round(len(man) / (bat * 1e-3), 1),
           "sequential_transposes_per_s": round(len(man) / (seq * 1e-3), 1),
           "speedup": round(seq / bat, 2),
           "batched_gbps": round(info.algorithmic_bytes / (bat * 1e-3) / 1e9, 1),
           "sequential_gbps": round(info.algorithmic_bytes / (seq * 1e-3) / 1e9, 1),
           "algorithmic_bytes": info.algorithmic_bytes, "tiles": info.tiles,
           "grid": list(info.grid)[:info.groups], "build_s": round(t_build, 3),
           "tuning_timeout": tune, "groups": info.groups,
           "layout": "one persistent launch per element type over a dynamic queue" if one_launch
                     else "one launch per (element type, variant, LDS class), one block per tile",
           "identical_to_sequential": same}
    del batch, plans, A, B, B0, Bseq
    torch.cuda.empty_cache()
    return res


def time_serving_manifest(dtype_name, passes=3, seq_reps=3, max_batch=4096, slots=3):
    """The served mode on the batched workload's mix: the 4096-tensor manifest
    handed to the continuous-batching server in one hpttx_server_submit_many
    call, closed loop, against one plain plan->exec() per tensor in sequence
    (plans created beforehand, back to back on one stream).  Wall time is the
    host clock from before the submission to the last completion; the first
    pass plans every request (cold plan cache), later passes find the plans
    cached.  Served results must equal the sequential ones bit for bit."""
    import torch

    import hptt_amd
    from hptt_amd.workloads import batch_manifest, layout
    man = batch_manifest()
    tdt = getattr(torch, dtype_name)
    esz = torch.empty(0, dtype=tdt).element_size()
    offs, total = layout(man, esz)
    g = torch.Generator(device="cuda").manual_seed(4096)
    A = torch.empty(total, dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    B0 = torch.empty(total, dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    B = B0.clone()
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    items, plans = [], []
    for m, o in zip(man, offs):
        n = int(np.prod(m["in_size"]))
        a, b = A[o:o + n], B[o:o + n]
        items.append(hptt_amd.make_batch_item(a, b, m["in_size"], m["perm"], m["alpha"], m["beta"]))
        plans.append(hptt_amd.create_plan_checked(a, b, m["in_size"], m["perm"], m["alpha"],
                                                  m["beta"], stream=stream))
    exec_fn = hptt_amd.lib().hpttx_plan_exec
    handles = [p._h for p in plans]

    def run_seq():
        for h in handles:
            if exec_fn(h) != 0:
                raise RuntimeError(hptt_amd.last_error())

    run_seq()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    seq_ms = []
    for _ in range(seq_reps):
        e0.record(stream)
        run_seq()
        e1.record(stream)
        torch.cuda.synchronize()
        seq_ms.append(e0.elapsed_time(e1))
    B.copy_(B0)
    torch.cuda.synchronize()
    run_seq()
    torch.cuda.synchronize()
    Bseq = B.clone()
    srv = hptt_amd.Server(slots=slots, max_batch=max_batch)
    arr = (hptt_amd.BatchItem * len(items))(*items)   # the request array a client hands over
    walls, launches = [], []
    same = True
    iv = torch.int32 if esz == 4 else torch.int64
    for k in range(passes):
        B.copy_(B0)
        torch.cuda.synchronize()
        n0 = srv.stats().launches
        t0 = srv.now()
        tickets = srv.submit_many(arr)
        srv.drain()
        t_last = 0.0
        for t in tickets:
            c = srv.wait(t, 10.0)
            if c.state != 2:
                raise RuntimeError("served request failed (%d)" % c.status)
            t_last = max(t_last, c.t_done)
        walls.append(t_last - t0)
        launches.append(srv.stats().launches - n0)
        torch.cuda.synchronize()
        same = same and bool(torch.equal(B.view(iv), Bseq.view(iv)))
    srv.close()
    seq = float(np.median(seq_ms)) * 1e-3
    warm = float(np.median(walls[1:])) if passes > 1 else walls[0]
    res = {"tensors": len(man), "sequential_ms": round(seq * 1e3, 3),
           "served_cold_ms": round(walls[0] * 1e3, 3), "served_warm_ms": round(warm * 1e3, 3),
           "sequential_transposes_per_s": round(len(man) / seq, 1),
           "served_transposes_per_s": round(len(man) / warm, 1),
           "speedup_warm": round(seq / warm, 2), "speedup_cold": round(seq / walls[0], 2),
           "launches_per_pass": launches, "max_batch": max_batch, "slots": slots,
           "identical_to_sequential": same,
           "note": "one hpttx_server_submit_many call per pass, closed loop; wall = host clock "
                   "from before the call to the last completion; cold = plans made in the "
                   "call, warm = plans cached"}
    del srv, plans, A, B, B0, Bseq
    torch.cuda.empty_cache()
    return res


def time_sweep(reps=10, tuned=True):
    """The reference benchmark's 102 workloads (fp32, alpha=2.3, beta=4.2,
    test/benchmark/inc/hptt/benchmark/benchmark_trans.h:24-140): effective
    GB/s of plain plan->exec() per shape (median of reps event-timed execs
    run back to back, like the other single-plan workloads), with the minimum
    and median over shapes and the slowest shapes, for model
    plans (tuning_timeout=0) and, with tuned, for plans autotuned exhaustively
    (tuning_timeout=-1, which is how the reference's benchmark creates its
    plans, test/benchmark/inc/hptt/perf-test/test_perf_util.tcc:22-24)."""
    import torch

    import hptt_amd
    from hptt_amd.workloads import (REFERENCE_SWEEP_ALPHA, REFERENCE_SWEEP_BETA,
                                    reference_sweep)
    cases = reference_sweep()
    nmax = max(int(np.prod(s)) for _, s, _ in cases)
    g = torch.Generator(device="cuda").manual_seed(102)
    Abuf = torch.empty(nmax, device="cuda").uniform_(-1, 1, generator=g)
    Bbuf = torch.empty(nmax, device="cuda").uniform_(-1, 1, generator=g)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()

    def rate(plan, info):
        plan.exec()
        evs = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            plan.exec()
            b.record(stream)
            evs.append((a, b))
        torch.cuda.synchronize()
        ms = float(np.median([a.elapsed_time(b) for a, b in evs]))
        return info.algorithmic_bytes / (ms * 1e-3) / 1e9

    def summary(rates):
        rates = sorted(rates)
        return {"min_gbps": rates[0], "median_gbps": float(np.median(rates)),
                "max_gbps": rates[-1], "min_pct": round(100 * rates[0] / PEAK_GBPS, 1),
                "median_pct": round(100 * float(np.median(rates)) / PEAK_GBPS, 1)}

    rows = {}
    for name, size, perm in cases:
        n = int(np.prod(size))
        plan = hptt_amd.create_plan_checked(Abuf[:n], Bbuf[:n], size, perm, REFERENCE_SWEEP_ALPHA,
                                            REFERENCE_SWEEP_BETA, stream=stream)
        info = plan.info()
        gbps = rate(plan, info)
        rows[name] = {"size": size, "perm": perm, "gbps": round(gbps, 1),
                      "pct_hbm_peak": round(100 * gbps / PEAK_GBPS, 1),
                      "path": hptt_amd.PATHS[info.path]}
        del plan
    # Tuned plans second: a tuned shape's parameters stay in the process's
    # tune cache and would be picked up by later model plans of the same shape.
    t_tune = 0.0
    for name, size, perm in (cases if tuned else []):
        n = int(np.prod(size))
        t0 = time.perf_counter()
        plan = hptt_amd.create_plan_checked(Abuf[:n], Bbuf[:n], size, perm, REFERENCE_SWEEP_ALPHA,
                                            REFERENCE_SWEEP_BETA, stream=stream,
                                            tuning_timeout=-1.0)
        t_tune += time.perf_counter() - t0
        info = plan.info()
        rows[name]["gbps_tuned"] = round(rate(plan, info), 1)
        rows[name]["path_tuned"] = hptt_amd.PATHS[info.path]
        del plan
    worst = sorted(rows.items(), key=lambda kv: kv[1]["gbps"])[:5]
    del Abuf, Bbuf
    torch.cuda.empty_cache()
    res = {"shapes": len(rows), **summary([r["gbps"] for r in rows.values()]),
           "note": "execs timed back to back; most of these shapes have A and B of about 200 MB "
                   "each, and where each buffer fits in the 256 MB last-level cache the planner "
                   "leaves A's loads unstreamed (no cost with a cold cache), so A stays in the "
                   "cache between runs: those rates count algorithmic bytes, part of which the "
                   "cache serves",
           "slowest": [dict(name=k, **v) for k, v in worst], "per_shape": rows}
    if tuned:
        res["tuned"] = {"tuning_timeout": -1.0, "tuning_s_total": round(t_tune, 1),
                        **summary([r["gbps_tuned"] for r in rows.values()]),
                        "note": "timed back to back like the model's plans; A and B of most "
                                "of these shapes are about 200 MB each, and the tuner (which "
                                "times launches back to back too) often picks access hints "
                                "that leave one of them in the 256 MB last-level cache "
                                "between runs, a gain that does not carry over to larger "
                                "problems or a cold cache"}
    return res


def cpu_batched_baseline(dtype_name, budget_s=5.0):
    """The oracle (port of the reference's RefTrans, OpenMP) over the batched
    manifest in order, one tensor after another, for about budget_s seconds:
    the CPU's sequential transposes/s on the same tensors."""
    import oracle
    from hptt_amd.workloads import batch_manifest
    man = batch_manifest()
    dt = np.dtype(dtype_name)
    threads, aff, physical = host_cores()
    done, nbytes, t_cpu = 0, 0, 0.0
    t_start = time.perf_counter()
    for k, m in enumerate(man):
        if time.perf_counter() - t_start > budget_s:
            break
        n = int(np.prod(m["in_size"]))
        A = oracle.uniform(n, dt, 2 * k + 1, nthreads=threads)
        B = oracle.uniform(n, dt, 2 * k + 2, nthreads=threads)
        t0 = time.perf_counter()
        assert oracle.transpose(A, B, m["in_size"], m["perm"], m["alpha"], m["beta"],
                                nthreads=threads) == 0
        t_cpu += time.perf_counter() - t0
        done += 1
        nbytes += n * dt.itemsize * (3 if abs(m["beta"]) >= 1e-16 else 2)
    return {"value": round(done / t_cpu, 1), "unit": "transposes/s", "cores": threads,
            "affinity_cpus": aff, "physical_cores_host": physical, "kind": "port",
            "gbps": round(nbytes / t_cpu / 1e9, 2),
            "sample": "the first %d of 4096 manifest tensors (%s) in order, timed per transpose "
                      "(data generation excluded); oracle OpenMP loop, %d threads"
                      % (done, dtype_name, threads)}


def serving_trace(count, seed=5150):
    """Serving requests: total elements log-uniform in [1K, 64M], rank 2..6
    with dims splitting that count, random perm, dtype uniform over the four,
    alpha/beta uniform in [-2, 2] (complex for the complex dtypes), beta = 0
    for 25%."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = float(np.exp(rng.uniform(np.log(1024), np.log(64 * 2**20))))
        rank = int(rng.integers(2, 7))
        w = rng.dirichlet(np.ones(rank))
        dims = [max(1, int(round(n ** wi))) for wi in w]
        while int(np.prod(dims)) > 64 * 2**20:
            dims[int(np.argmax(dims))] -= 1
        code = int(rng.integers(0, 4))
        cplx = code >= 2
        al = complex(rng.uniform(-2, 2), rng.uniform(-2, 2)) if cplx else float(rng.uniform(-2, 2))
        if rng.random() < 0.25:
            be = 0.0
        else:
            be = complex(rng.uniform(-2, 2), rng.uniform(-2, 2)) if cplx else float(rng.uniform(-2, 2))
        out.append({"code": code, "in_size": dims, "perm": [int(v) for v in rng.permutation(rank)],
                    "alpha": al, "beta": be, "n": int(np.prod(dims))})
    return out


def time_serving(count=600, rates=(0.5, 1.0, 2.0), inflight_gb=8.0):
    """Serving workload: sequential plain exec() (create plan, exec, wait, one
    request at a time, in arrival order) against the continuous-batching
    server, closed loop (all requests at once: capacity) and open loop with
    Poisson arrivals at 0.5x, 1x and 2x the sequential capacity."""
    import torch

    import hptt_amd
    trace = serving_trace(count)
    tdt = {0: torch.float32, 1: torch.float64, 2: torch.complex64, 3: torch.complex128}
    real = {0: torch.float32, 1: torch.float64, 2: torch.float32, 3: torch.float64}
    g = torch.Generator(device="cuda").manual_seed(55)
    maxn = max(r["n"] for r in trace)
    pools_a, pools_b, offs = {}, {}, []
    per_code = {c: 0 for c in range(4)}
    for r in trace:
        offs.append(per_code[r["code"]])
        per_code[r["code"]] += (r["n"] + 63) // 64 * 64
    for c in range(4):
        k = 2 if c >= 2 else 1
        pools_a[c] = torch.empty(maxn * k, dtype=real[c], device="cuda").uniform_(-1, 1, generator=g)
        pools_b[c] = torch.empty(max(1, per_code[c]) * k, dtype=real[c], device="cuda").uniform_(
            -1, 1, generator=g)
        if c >= 2:
            pools_a[c] = pools_a[c].view(tdt[c])
            pools_b[c] = pools_b[c].view(tdt[c])
    torch.cuda.synchronize()
    items, args = [], []
    for r, o in zip(trace, offs):
        a = pools_a[r["code"]][:r["n"]]
        b = pools_b[r["code"]][o:o + r["n"]]
        items.append(hptt_amd.make_batch_item(a, b, r["in_size"], r["perm"], r["alpha"], r["beta"]))
        args.append((a, b))
    total_bytes = 0
    L = hptt_amd.lib()
    stream = torch.cuda.Stream()
    sp = int(stream.cuda_stream)

    # Plain exec() baseline: every request's plan created beforehand (plans
    # kept resident, as the server keeps its own), then exec + wait per
    # request in arrival order.
    plans = [hptt_amd.create_plan_checked(a, b, r["in_size"], r["perm"], r["alpha"], r["beta"],
                                          stream=sp) for r, (a, b) in zip(trace, args)]
    exec_fn, sync_fn = L.hpttx_plan_exec, L.hpttx_plan_sync
    handles = [p._h for p in plans]

    def seq_run(arrivals, create=False):
        """One request at a time: exec, wait (create=True: also create the
        plan per request, inside the timed region)."""
        t0 = time.perf_counter()
        lat = []
        for k, (r, (a, b)) in enumerate(zip(trace, args)):
            due = t0 + arrivals[k]
            while time.perf_counter() < due:
                pass
            if create:
                p = hptt_amd.create_plan_checked(a, b, r["in_size"], r["perm"], r["alpha"],
                                                 r["beta"], stream=sp)
                p.exec()
                p.sync()
            else:
                h = handles[k]
                if exec_fn(h) != 0 or sync_fn(h) != 0:
                    raise RuntimeError(hptt_amd.last_error())
            lat.append(time.perf_counter() - due)
        return time.perf_counter() - t0, lat

    def serve_run(arrivals):
        srv = hptt_amd.Server(inflight_bytes=int(inflight_gb * 2**30), slots=3, max_batch=256)
        t0 = time.perf_counter()
        s0 = srv.now()
        tickets = []
        for k, it in enumerate(items):
            due = t0 + arrivals[k]
            while time.perf_counter() < due:
                pass
            tickets.append(srv.submit(it))
        srv.drain()
        st = srv.stats()
        lat = []
        lag = []
        for k, t in enumerate(tickets):
            c = srv.wait(t, 10.0)
            if c.state != 2:
                raise RuntimeError("serving request %d failed (%d)" % (k, c.status))
            lat.append(c.t_done - (s0 + arrivals[k]))
            lag.append(c.t_submit - (s0 + arrivals[k]))
        wall = st.last_done - s0
        launches = st.launches
        srv.close()
        return wall, lat, launches, float(np.max(lag))

    def summary(wall, lat):
        lat = np.sort(np.array(lat))
        return {"transposes_per_s": round(len(lat) / wall, 1), "wall_ms": round(wall * 1e3, 2),
                "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)}

    zero = [0.0] * len(trace)
    seq_run(zero)                     # warm-up (kernels)
    seq_wall, seq_lat = seq_run(zero)
    seqc_wall, seqc_lat = seq_run(zero, create=True)
    cap = len(trace) / seq_wall
    serve_run(zero)                   # warm-up
    sw, sl, launches, _ = serve_run(zero)
    nbytes = int(sum(r["n"] * [4, 8, 8, 16][r["code"]] * (2 + (abs(r["beta"]) >= 1e-16))
                     for r in trace))
    bound_s = nbytes / (PEAK_GBPS * 1e9)
    out = {"requests": len(trace), "bytes": nbytes,
           "byte_bound_ms": round(bound_s * 1e3, 3),
           "sequential_closed": dict(summary(seq_wall, seq_lat),
                                     bound_frac=round(bound_s / seq_wall, 4)),
           "sequential_create_exec_closed": summary(seqc_wall, seqc_lat),
           "served_closed": dict(summary(sw, sl), launches=launches,
                                 gbps=round(nbytes / sw / 1e9, 1),
                                 bound_frac=round(bound_s / sw, 4)),
           "inflight_budget_gb": inflight_gb,
           "note": "sequential = plans created beforehand, then exec + wait per request; "
                   "sequential_create_exec also creates each plan inside the timed loop; "
                   "bound_frac = (bytes / 8 TB/s) / wall time"}
    out["served_vs_sequential_throughput"] = round(
        out["served_closed"]["transposes_per_s"] / out["sequential_closed"]["transposes_per_s"], 2)
    out["served_vs_create_exec_throughput"] = round(
        out["served_closed"]["transposes_per_s"] /
        out["sequential_create_exec_closed"]["transposes_per_s"], 2)
    rng = np.random.default_rng(99)
    for f in rates:
        lam = f * cap
        arr = np.cumsum(rng.exponential(1.0 / lam, size=len(trace)))
        arr -= arr[0]
        w1, l1 = seq_run(arr)
        w2, l2, n2, lag = serve_run(arr)
        out["rate_%.1fx" % f] = {"offered_per_s": round(lam, 1),
                                 "sequential": summary(w1, l1),
                                 "served": dict(summary(w2, l2), launches=n2,
                                                max_submit_lag_ms=round(lag * 1e3, 3))}
    del pools_a, pools_b, items, args, plans, handles
    torch.cuda.empty_cache()
    return out


def host_cores():
    """(threads used, CPUs this process may run on, physical cores of the
    host per lscpu).  Threads: every CPU in the affinity mask, capped by
    OMP_NUM_THREADS when that is set (the share of the host a job gets)."""
    aff = len(os.sched_getaffinity(0))
    threads = aff
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
    physical = None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {l.split(":", 1)[0].strip(): l.split(":", 1)[1].strip() for l in out.splitlines()
              if ":" in l}
        physical = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except Exception:
        pass
    return threads, aff, physical


def cpu_baseline(budget_s=10.0, name="6d-f32"):
    """Oracle (port of the reference's OpenMP RefTrans) on one workload:
    repeated full transposes for about budget_s seconds of CPU time."""
    import oracle
    w = WORKLOADS[name]
    n = int(np.prod(w["size"]))
    dt = np.dtype(w["dtype"])
    threads, aff, physical = host_cores()
    A = oracle.uniform(n, dt, 1, nthreads=threads)
    B = oracle.uniform(n, dt, 2, nthreads=threads)
    cc = isinstance(w["alpha"], complex) or isinstance(w["beta"], complex)
    upd = abs(w["beta"]) >= 1e-16
    times = []
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < budget_s and len(times) < 1000:
        t0 = time.perf_counter()
        rc = oracle.transpose(A, B, w["size"], w["perm"], w["alpha"], w["beta"], nthreads=threads,
                              complex_coef=cc)
        times.append(time.perf_counter() - t0)
        assert rc == 0
    med = float(np.median(times))
    nbytes = n * dt.itemsize * (3 if upd else 2)
    return {"value": round(nbytes / med / 1e9, 2), "unit": "GB/s", "cores": threads,
            "affinity_cpus": aff, "physical_cores_host": physical,
            "kind": "port", "best_gbps": round(nbytes / min(times) / 1e9, 2),
            "sample": "%s (%s) repeated for %.1f s: %d runs, median; oracle OpenMP loop "
                      "(reference RefTrans order), %d threads"
                      % (name, describe_workload(name), time.perf_counter() - t_start,
                         len(times), threads)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="main workload only")
    ap.add_argument("--tune", type=float, default=0.0,
                    help="autotune each plan for this many seconds before timing")
    ap.add_argument("--workload", default="6d-f32",
                    help="main workload (%s), batched-f32 / batched-f64 (suffix -one-launch "
                         "for that layout), serving or sweep" % ", ".join(WORKLOADS))
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(a transposition is measured on one GPU's share)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm device", file=sys.stderr)
        sys.exit(1)
    import hptt_amd
    hptt_amd.lib()
    if args.workload == "sweep":
        print(json.dumps({"workload": "reference-sweep", **time_sweep()}), flush=True)
        return
    if args.workload == "serving":
        res = time_serving()
        print(json.dumps({"workload": "serving", **res}), flush=True)
        return
    if args.workload.startswith("serving-manifest-"):
        res = time_serving_manifest({"f32": "float32", "f64": "float64"}[args.workload.split("-")[2]])
        print(json.dumps({"workload": args.workload, **res}), flush=True)
        return
    if args.workload.startswith("batched-"):
        res = time_batched({"f32": "float32", "f64": "float64"}[args.workload.split("-")[1]],
                           reps=max(1, args.steps), tune=args.tune,
                           one_launch=args.workload.endswith("one-launch"))
        print(json.dumps({"workload": args.workload, **res}), flush=True)
        return
    log("bench: main workload %s, %d steps" % (args.workload, args.steps))
    torch.cuda.synchronize()
    main_res = time_workload(args.workload, args.steps, args.warmup, args.tune)
    torch.cuda.synchronize()
    log("bench: %s %.1f GB/s (%.1f%% of peak)" % (args.workload, main_res["gbps"],
                                                 main_res["pct_hbm_peak"]))
    secondary = {}
    if not args.quick:
        for name in ("6d-f64", "4d-c128-tile", "4d-c128-wide", "3d-f32-copy", "copy-f32-512m",
                     "axpby-f32-151m"):
            if name == args.workload:
                continue
            secondary[name] = time_workload(name, max(10, args.steps // 2), args.warmup, args.tune)
            log("bench: %s %.1f GB/s" % (name, secondary[name]["gbps"]))
        secondary["reference-sweep"] = time_sweep()
        sw = secondary["reference-sweep"]
        log("bench: reference sweep (%d shapes) min %.1f median %.1f GB/s; slowest %s; "
            "tuned (%.0f s of tuning) min %.1f median %.1f GB/s"
            % (sw["shapes"], sw["min_gbps"], sw["median_gbps"], sw["slowest"][0]["name"],
               sw["tuned"]["tuning_s_total"], sw["tuned"]["min_gbps"], sw["tuned"]["median_gbps"]))
        for name in ("batched-f32", "batched-f64", "batched-f32-one-launch",
                     "batched-f64-one-launch"):
            secondary[name] = time_batched(name.split("-")[1].replace("f", "float"),
                                           one_launch=name.endswith("one-launch"))
            log("bench: %s %s transposes/s batched vs %s sequential (x%.2f), identical=%s"
                % (name, secondary[name]["batched_transposes_per_s"],
                   secondary[name]["sequential_transposes_per_s"], secondary[name]["speedup"],
                   secondary[name]["identical_to_sequential"]))
        for name in ("serving-manifest-f32", "serving-manifest-f64"):
            secondary[name] = time_serving_manifest(
                {"f32": "float32", "f64": "float64"}[name.split("-")[2]])
            log("bench: %s served %s transposes/s (warm plans) vs %s sequential (x%.2f), "
                "identical=%s" % (name, secondary[name]["served_transposes_per_s"],
                                  secondary[name]["sequential_transposes_per_s"],
                                  secondary[name]["speedup_warm"],
                                  secondary[name]["identical_to_sequential"]))
        secondary["serving"] = time_serving()
        sv = secondary["serving"]
        log("bench: serving closed loop %s/s vs sequential %s/s (x%s); " % (
            sv["served_closed"]["transposes_per_s"], sv["sequential_closed"]["transposes_per_s"],
            sv["served_vs_sequential_throughput"]) + "; ".join(
            "%s: served p50 %s p99 %s ms, seq p50 %s p99 %s ms" % (
                k, v["served"]["p50_ms"], v["served"]["p99_ms"], v["sequential"]["p50_ms"],
                v["sequential"]["p99_ms"]) for k, v in sv.items() if k.startswith("rate_")))
    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (oracle, OpenMP)")
        cpu = cpu_baseline()
        log("bench: CPU %.2f GB/s on %d threads" % (cpu["value"], cpu["cores"]))
        if not args.quick:
            # the same port on the other single-transpose workloads, briefly
            for name in ("3d-f32-copy", "4d-c128-tile", "6d-f64"):
                c2 = cpu_baseline(2.0, name)
                secondary.setdefault("cpu_baseline", {})[name] = c2
                log("bench: CPU %s %.2f GB/s" % (name, c2["value"]))
            for name in ("batched-f32", "batched-f64"):
                c3 = cpu_batched_baseline(name.split("-")[1].replace("f", "float"), 3.0)
                secondary.setdefault("cpu_baseline", {})[name] = c3
                log("bench: CPU %s %.1f transposes/s" % (name, c3["value"]))
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
        "dtype": DTYPE_TAG[WORKLOADS[args.workload]["dtype"]],
        "data": "synthetic: A, B uniform[-1,1) from a fixed seed, generated on the device",
        "config": {"workload": describe_workload(args.workload),
                   "elements": main_res["elements"], "bytes_per_step": main_res["bytes_per_exec"],
                   "path": main_res["path"]},
        "pct_hbm_peak": main_res["pct_hbm_peak"],
        "reference_metric_gibps": main_res["reference_metric_gibps"],
        "roofline": {"bound": "hbm", "achieved": gbps, "peak": PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbps / PEAK_GBPS, 4), "traffic": pmc_traffic(args.workload),
                     "traffic_source": TRAFFIC_FILE},
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
