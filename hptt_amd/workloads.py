"""Synthetic workload generators for the batched and serving modes.

Batched manifest: rank uniform in 2..6, each dim floor(exp(U[ln 2, ln 128]))
clipped to [2, 128] and the whole shape rejection-resampled until it has at
most 4 Mi elements, a uniformly random perm, alpha and beta uniform in
[-2, 2] with beta = 0 for a seeded 25% of the tensors.
"""
from __future__ import annotations

import numpy as np


def batch_manifest(count: int = 4096, seed: int = 12345, max_elems: int = 4 * 2**20,
                   lo: int = 2, hi: int = 128):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        rank = int(rng.integers(2, 7))
        while True:
            dims = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=rank)))
            dims = np.clip(dims.astype(np.int64), lo, hi)
            if int(np.prod(dims)) <= max_elems:
                break
        perm = [int(v) for v in rng.permutation(rank)]
        alpha = float(rng.uniform(-2.0, 2.0))
        beta = 0.0 if rng.random() < 0.25 else float(rng.uniform(-2.0, 2.0))
        out.append({"in_size": [int(v) for v in dims], "perm": perm,
                    "alpha": alpha, "beta": beta})
    return out


def layout(manifest, elem_bytes: int, align_bytes: int = 256):
    """Element offsets of each tensor in one flat buffer (each start aligned
    to align_bytes) and the total element count."""
    step = max(1, align_bytes // elem_bytes)
    offs = []
    cur = 0
    for m in manifest:
        offs.append(cur)
        n = int(np.prod(m["in_size"]))
        cur += (n + step - 1) // step * step
    return offs, cur


# The reference benchmark's workload list: fp32, alpha=2.3, beta=4.2
# (test/benchmark/inc/hptt/perf-test/test_perf_util.h:20-22), each entry
# (in_size, perm).  The table is the reference's own data, kept as it is so
# that the sweep measures the same shapes: its first part follows
# test/benchmark/inc/hptt/benchmark/benchmark_trans.h:24-82 (57 shapes: small
# leading dims of 4-16 elements, perms that keep or move the stride-1 dim,
# ranks 2-6), its second part :85-140 (45 shapes grouped by the reference as
# one original transposition and four variants that merge to it).
REFERENCE_SWEEP_ALPHA = 2.3
REFERENCE_SWEEP_BETA = 4.2
REFERENCE_SWEEP = [
    ([7248, 7248], [1, 0]),
    ([1216, 43408], [1, 0]),
    ([43408, 1216], [1, 0]),
    ([368, 384, 384], [0, 2, 1]),
    ([2144, 64, 384], [0, 2, 1]),
    ([368, 64, 2307], [0, 2, 1]),
    ([384, 384, 355], [1, 0, 2]),
    ([2320, 384, 59], [1, 0, 2]),
    ([384, 2320, 59], [1, 0, 2]),
    ([384, 355, 384], [2, 1, 0]),
    ([2320, 59, 384], [2, 1, 0]),
    ([384, 59, 2320], [2, 1, 0]),
    ([80, 96, 75, 96], [0, 3, 2, 1]),
    ([464, 16, 75, 96], [0, 3, 2, 1]),
    ([80, 16, 75, 582], [0, 3, 2, 1]),
    ([96, 75, 96, 75], [2, 1, 3, 0]),
    ([608, 12, 96, 75], [2, 1, 3, 0]),
    ([96, 12, 608, 75], [2, 1, 3, 0]),
    ([96, 75, 96, 75], [2, 0, 3, 1]),
    ([608, 12, 96, 75], [2, 0, 3, 1]),
    ([96, 12, 608, 75], [2, 0, 3, 1]),
    ([96, 96, 75, 75], [1, 0, 3, 2]),
    ([608, 96, 12, 75], [1, 0, 3, 2]),
    ([96, 608, 12, 75], [1, 0, 3, 2]),
    ([96, 75, 75, 96], [3, 2, 1, 0]),
    ([608, 12, 75, 96], [3, 2, 1, 0]),
    ([96, 12, 75, 608], [3, 2, 1, 0]),
    ([32, 48, 28, 28, 48], [0, 4, 2, 1, 3]),
    ([176, 8, 28, 28, 48], [0, 4, 2, 1, 3]),
    ([32, 8, 28, 28, 298], [0, 4, 2, 1, 3]),
    ([48, 28, 28, 48, 28], [3, 2, 1, 4, 0]),
    ([352, 4, 28, 48, 28], [3, 2, 1, 4, 0]),
    ([48, 4, 28, 352, 28], [3, 2, 1, 4, 0]),
    ([48, 28, 48, 28, 28], [2, 0, 4, 1, 3]),
    ([352, 4, 48, 28, 28], [2, 0, 4, 1, 3]),
    ([48, 4, 352, 28, 28], [2, 0, 4, 1, 3]),
    ([48, 48, 28, 28, 28], [1, 3, 0, 4, 2]),
    ([352, 48, 4, 28, 28], [1, 3, 0, 4, 2]),
    ([48, 352, 4, 28, 28], [1, 3, 0, 4, 2]),
    ([48, 28, 28, 28, 48], [4, 3, 2, 1, 0]),
    ([352, 4, 28, 28, 48], [4, 3, 2, 1, 0]),
    ([48, 4, 28, 28, 352], [4, 3, 2, 1, 0]),
    ([16, 32, 15, 32, 15, 15], [0, 3, 2, 5, 4, 1]),
    ([48, 10, 15, 32, 15, 15], [0, 3, 2, 5, 4, 1]),
    ([16, 10, 15, 103, 15, 15], [0, 3, 2, 5, 4, 1]),
    ([32, 15, 15, 32, 15, 15], [3, 2, 0, 5, 1, 4]),
    ([112, 5, 15, 32, 15, 15], [3, 2, 0, 5, 1, 4]),
    ([32, 5, 15, 112, 15, 15], [3, 2, 0, 5, 1, 4]),
    ([32, 15, 32, 15, 15, 15], [2, 0, 4, 1, 5, 3]),
    ([112, 5, 32, 15, 15, 15], [2, 0, 4, 1, 5, 3]),
    ([32, 5, 112, 15, 15, 15], [2, 0, 4, 1, 5, 3]),
    ([32, 15, 15, 32, 15, 15], [3, 2, 5, 1, 0, 4]),
    ([112, 5, 15, 32, 15, 15], [3, 2, 5, 1, 0, 4]),
    ([32, 5, 15, 112, 15, 15], [3, 2, 5, 1, 0, 4]),
    ([32, 15, 15, 15, 15, 32], [5, 4, 3, 2, 1, 0]),
    ([112, 5, 15, 15, 15, 32], [5, 4, 3, 2, 1, 0]),
    ([32, 5, 15, 15, 15, 112], [5, 4, 3, 2, 1, 0]),
]
REFERENCE_SWEEP_MERGE = [
    ([7248, 7248], [1, 0]),
    ([7248, 151, 48], [1, 2, 0]),
    ([48, 151, 7248], [2, 0, 1]),
    ([48, 151, 151, 48], [2, 3, 0, 1]),
    ([1, 7248, 1, 7248], [2, 3, 0, 1]),
    ([368, 384, 384], [0, 2, 1]),
    ([24, 16, 384, 384], [0, 1, 3, 2]),
    ([368, 16, 24, 384], [0, 3, 1, 2]),
    ([23, 16, 16, 24, 384], [0, 1, 3, 4, 2]),
    ([368, 16, 24, 24, 16], [0, 3, 4, 1, 2]),
    ([384, 384, 355], [1, 0, 2]),
    ([24, 16, 384, 355], [2, 0, 1, 3]),
    ([384, 384, 5, 71], [1, 0, 2, 3]),
    ([16, 24, 384, 71, 5], [2, 0, 1, 3, 4]),
    ([384, 24, 16, 71, 5], [1, 2, 0, 3, 4]),
    ([384, 355, 384], [2, 1, 0]),
    ([16, 24, 355, 384], [3, 2, 0, 1]),
    ([384, 71, 5, 384], [3, 1, 2, 0]),
    ([16, 24, 71, 5, 384], [4, 2, 3, 0, 1]),
    ([384, 5, 71, 16, 24], [3, 4, 1, 2, 0]),
    ([80, 96, 75, 96], [0, 3, 2, 1]),
    ([5, 16, 96, 75, 96], [0, 1, 4, 3, 2]),
    ([80, 16, 6, 75, 96], [0, 4, 3, 1, 2]),
    ([5, 16, 16, 6, 75, 96], [0, 1, 5, 4, 2, 3]),
    ([80, 96, 5, 15, 6, 16], [0, 4, 5, 2, 3, 1]),
    ([96, 75, 96, 75], [2, 1, 3, 0]),
    ([6, 16, 75, 96, 75], [3, 2, 4, 0, 1]),
    ([96, 75, 96, 25, 3], [2, 1, 3, 4, 0]),
    ([6, 16, 5, 15, 96, 75], [4, 2, 3, 5, 0, 1]),
    ([96, 75, 16, 6, 25, 3], [2, 3, 1, 4, 5, 0]),
    ([96, 75, 96, 75], [2, 0, 3, 1]),
    ([8, 12, 75, 96, 75], [3, 0, 1, 4, 2]),
    ([96, 15, 5, 96, 75], [3, 0, 4, 1, 2]),
    ([6, 16, 75, 6, 16, 75], [3, 4, 0, 1, 5, 2]),
    ([96, 5, 15, 96, 15, 5], [3, 0, 4, 5, 1, 2]),
    ([96, 96, 75, 75], [1, 0, 3, 2]),
    ([6, 16, 96, 75, 75], [2, 0, 1, 4, 3]),
    ([96, 96, 75, 15, 5], [1, 0, 3, 4, 2]),
    ([12, 8, 96, 15, 5, 75], [2, 0, 1, 5, 3, 4]),
    ([96, 96, 3, 25, 15, 5], [1, 0, 4, 5, 2, 3]),
    ([96, 75, 75, 96], [3, 2, 1, 0]),
    ([6, 16, 75, 75, 96], [4, 3, 2, 0, 1]),
    ([96, 75, 75, 16, 6], [3, 4, 2, 1, 0]),
    ([1, 96, 75, 1, 75, 96], [5, 4, 2, 3, 0, 1]),
    ([96, 75, 1, 75, 1, 96], [4, 5, 3, 1, 2, 0]),
]


def reference_sweep():
    """All 102 (name, in_size, perm) entries of the reference's workload list."""
    out = []
    for tag, rows in (("s", REFERENCE_SWEEP), ("m", REFERENCE_SWEEP_MERGE)):
        for k, (size, perm) in enumerate(rows):
            out.append(("%s%02d-%dd" % (tag, k, len(size)), list(size), list(perm)))
    return out
