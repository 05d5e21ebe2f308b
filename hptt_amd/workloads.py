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
