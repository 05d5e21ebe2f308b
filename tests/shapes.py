"""Seeded shape generators shared by the CPU emulator tests and GPU tests."""
from __future__ import annotations

import numpy as np

DTYPES = [np.float32, np.float64, np.complex64, np.complex128]

# The workloads the benchmark and the parity tests are defined on.
CONFIG1 = ([64, 128, 256], [2, 1, 0])
CONFIG2 = ([16, 24, 32, 16, 24, 32], [5, 2, 0, 4, 1, 3])
CONFIG3_TILE = ([96, 96, 96, 96], [3, 2, 1, 0])
CONFIG3_WIDE = ([96, 96, 96, 96], [0, 3, 2, 1])


def random_case(rng, max_elems=20000, max_dim=40, ranks=(2, 3, 4, 5, 6)):
    """A random (in_size, perm) with a bounded element count; dims log-uniform
    in [1, max_dim] (size-1 dims included)."""
    rank = int(rng.choice(ranks))
    while True:
        dims = np.floor(np.exp(rng.uniform(0.0, np.log(max_dim + 1), size=rank))).astype(int)
        dims = np.clip(dims, 1, max_dim)
        if int(np.prod(dims)) <= max_elems:
            break
    perm = list(rng.permutation(rank))
    return [int(v) for v in dims], [int(v) for v in perm]


def random_outer(rng, size, perm):
    """Random padded outer sizes for A and B (about half the dims padded)."""
    io = [s + (int(rng.integers(0, 3)) if rng.random() < 0.5 else 0) for s in size]
    out_size = [size[p] for p in perm]
    oo = [s + (int(rng.integers(0, 3)) if rng.random() < 0.5 else 0) for s in out_size]
    return io, oo
