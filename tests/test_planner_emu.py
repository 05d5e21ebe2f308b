"""The planner and the kernels' index math, checked on the CPU.

tests/emu/libemu.so runs the library's own block bodies (hptt_amd/csrc/
hpttx_core.h) with the library's own planner on the host: every global and LDS
access is bounds-checked, every store into B is counted (each element of the
tensor exactly once, padding never), and the result is compared bit for bit
with the oracle.  This is how launch shapes are shown to stay in bounds before
they run on a GPU.
"""
import numpy as np
import pytest

import emu_runner as E
import oracle
import shapes


def test_fastdiv_exact():
    lib = E.lib()
    for d in [1, 2, 3, 5, 7, 24, 32, 96, 384, 1000, 65537, 2**31 - 1, 2**31, 2**31 + 1,
              2**32 - 1]:
        assert lib.emu_check_fastdiv(d, 0, 200000) == -1, d
        assert lib.emu_check_fastdiv(d, 2**32 - 200000, 2**32 - 1) == -1, d
    rng = np.random.default_rng(0)
    for d in rng.integers(1, 2**32 - 1, size=40):
        lo = int(rng.integers(0, 2**32 - 20000))
        assert lib.emu_check_fastdiv(int(d), lo, lo + 20000) == -1


def _check(dt, size, perm, alpha, beta, io=None, oo=None, tune=None, cc=False,
           offset=0, force_scalar=False):
    n_a = int(np.prod(io if io else size))
    out_size = [size[p] for p in perm]
    n_b = int(np.prod(oo if oo else out_size))
    Abuf = oracle.uniform(n_a + offset, dt, 101)
    Bbuf = oracle.uniform(n_b + offset, dt, 202)
    A = Abuf[offset:]
    B0 = Bbuf[offset:]
    exp = B0.copy()
    assert oracle.transpose(A, exp, size, perm, alpha, beta, io, oo, complex_coef=cc) == 0
    got = B0.copy() if offset == 0 else Bbuf.copy()[offset:]
    rc, err, desc = E.transpose(A, got, size, perm, alpha, beta, io, oo, tune=tune,
                                complex_coef=cc, force_scalar=force_scalar)
    assert rc == 0, (size, perm, err, desc)
    assert oracle.first_mismatch(got, exp) == -1, (size, perm, desc)
    return desc


@pytest.mark.parametrize("dt", shapes.DTYPES)
def test_emulated_random_shapes(dt):
    rng = np.random.default_rng(1234 + shapes.DTYPES.index(dt))
    paths = set()
    for k in range(60):
        size, perm = shapes.random_case(rng, max_elems=6000)
        al, be = [(1.0, 0.0), (2.3, 4.2), (-0.5, 1e-17), (1.25, -0.75)][k % 4]
        desc = _check(dt, size, perm, al, be)
        paths.add(desc.split("path: ")[1].split("\n")[0])
    assert {"tile", "tile-shared", "axpby"} <= paths, paths


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_emulated_complex_coefficients(dt):
    rng = np.random.default_rng(77)
    for k in range(20):
        size, perm = shapes.random_case(rng, max_elems=4000)
        be = complex(0.7, -0.2) if k % 2 else 0j
        _check(dt, size, perm, complex(1.1, 0.3), be, cc=True)


@pytest.mark.parametrize("dt", shapes.DTYPES)
def test_emulated_padded_and_misaligned(dt):
    rng = np.random.default_rng(4321)
    for k in range(25):
        size, perm = shapes.random_case(rng, max_elems=3000, max_dim=16)
        io, oo = shapes.random_outer(rng, size, perm)
        _check(dt, size, perm, 2.3, 4.2 if k % 2 else 0.0, io, oo)
    # pointers one element off a 16-byte boundary force the unaligned-vector
    # kernels (scalars for 16-byte elements)
    for k in range(10):
        size, perm = shapes.random_case(rng, max_elems=3000)
        desc = _check(dt, size, perm, 1.5, 0.25, offset=1)
        if np.dtype(dt).itemsize < 16 and "path: tile" in desc:
            assert "vec(r,w)=%du,%du" % ((16 // np.dtype(dt).itemsize,) * 2) in desc, desc


@pytest.mark.parametrize("dt", shapes.DTYPES)
def test_emulated_unaligned_runs(dt):
    """Odd extents (runs start at any element offset and end in partial
    vectors) and the forced unaligned mode on every path, with and without
    beta, against the oracle; every access in bounds and dword-aligned, every
    element of B written once."""
    rng = np.random.default_rng(99)
    odd = [[7, 5, 3], [3, 9, 11, 2], [5, 7, 6, 3, 2], [13, 3], [3, 5, 7, 2, 3, 2],
           [31, 33], [6, 10, 14]]
    for size in odd:
        for _ in range(3):
            perm = [int(v) for v in rng.permutation(len(size))]
            for be in (0.0, 4.2):
                _check(dt, size, perm, 2.3, be)
                _check(dt, size, perm, 2.3, be, force_scalar=True)
                _check(dt, size, perm, 2.3, be, offset=1)


@pytest.mark.parametrize("tune", [[8, 8, 256, 64], [16, 64, 1024, 128], [128, 16, 4096, 1024],
                                  [256, 256, 8192, 4096], [4, 4, 64, 16]])
def test_emulated_tile_shapes(tune):
    rng = np.random.default_rng(sum(tune))
    for dt in shapes.DTYPES:
        for k in range(8):
            size, perm = shapes.random_case(rng, max_elems=5000)
            _check(dt, size, perm, 2.3, 4.2 if k % 2 else 0.0, tune=tune)


@pytest.mark.parametrize("dt", shapes.DTYPES)
def test_emulated_wide_blocks(dt):
    # tiles over 32 KiB of LDS run 512-thread blocks: full, ragged and
    # unaligned tiles, with and without beta
    tune = {4: [64, 128, 8192, 256], 8: [32, 64, 4096, 256],
            16: [32, 64, 2048, 256]}[np.dtype(dt).itemsize]
    wide = 0
    for size, perm in [([64, 3, 128], [2, 1, 0]), ([70, 5, 130], [2, 1, 0]),
                       ([66, 2, 3, 129], [3, 1, 0, 2]), ([128, 2, 64], [2, 1, 0])]:
        for be in (0.0, 4.2):
            d = _check(dt, size, perm, 2.3, be, tune=tune)
            lds = int(d.split(" lds=")[1].split("B")[0])
            assert (" x 512 " in d) == (lds > 32768), d
            wide += " x 512 " in d
            _check(dt, size, perm, 2.3, be, tune=tune, offset=1)
    assert wide >= 4
    rc, err, d = E.plan(np.float32, *shapes.CONFIG2)
    assert rc == 0 and "grid: 18432 x 512 " in d, d
    rc, err, d = E.plan(np.float64, *shapes.CONFIG2)
    assert rc == 0 and "C=1024/128 " in d and "grid: 36864 x 512 " in d, d
    rc, err, d = E.plan(np.complex128, *shapes.CONFIG3_TILE)
    assert rc == 0 and "grid: 165888 x 256 " in d, d


@pytest.mark.parametrize("dt", shapes.DTYPES)
def test_emulated_long_rows_in_segments(dt):
    # kept rows longer than a block's elements are cut into equal segments
    # (2144 -> 1088 + 1056 for fp32), exact and unaligned, with and without beta
    for size, perm in (([2144, 3, 4], [0, 2, 1]), ([4099, 2, 3], [0, 2, 1]), ([2100, 5], [0, 1]),
                       ([1025, 3, 3], [0, 2, 1])):
        for be in (0.0, 4.2):
            for off in (0, 1):
                for tune in (None, [0, 0, 512, 0], [0, 0, 1000, 0]):
                    d = _check(dt, size, perm, 2.3, be, tune=tune, offset=off)
                    assert "wide-copy" in d or "axpby" in d, d
    rc, err, d = E.plan(np.float32, [2144, 64, 384], [0, 2, 1])
    assert rc == 0 and "seg=1088 " in d, d


def test_emulated_benchmark_shapes_reduced():
    # the benchmark workloads' perms on reduced extents (same fusion classes)
    _check(np.float32, [16, 12, 8, 4, 6, 8], shapes.CONFIG2[1], 2.3, 4.2)
    _check(np.float64, [16, 12, 8, 4, 6, 8], shapes.CONFIG2[1], 2.3, 4.2)
    _check(np.complex128, [12, 10, 9, 8], [3, 2, 1, 0], 1.1, 0.7)
    _check(np.complex128, [12, 10, 9, 8], [0, 3, 2, 1], 1.1, 0.7)
    _check(np.float32, [16, 32, 64], [2, 1, 0], 1.0, 0.0)


def test_plans_for_benchmark_shapes():
    rc, err, d = E.plan(np.float32, *shapes.CONFIG2)
    assert rc == 0 and "path: tile\n" in d and "vec(r,w)=4,4" in d and "swizzled" in d, d
    rc, err, d = E.plan(np.float64, *shapes.CONFIG2)
    assert rc == 0 and "vec(r,w)=2,2" in d, d
    rc, err, d = E.plan(np.complex128, *shapes.CONFIG3_TILE)
    assert rc == 0 and "path: tile\n" in d, d
    rc, err, d = E.plan(np.complex128, *shapes.CONFIG3_WIDE)
    assert rc == 0 and "path: wide-copy" in d, d
    rc, err, d = E.plan(np.float32, *shapes.CONFIG1)
    assert rc == 0 and "path: tile\n" in d, d


def test_streaming_and_tile_order_rules():
    """Problems beyond the last-level cache (A and B together): XCD-aware
    tile order always.  Runs of whole 128-byte lines: every access streamed,
    except A's loads where each buffer fits in the cache.  Runs that are not
    all whole lines: where each buffer fits, loads only (4-byte elements) or
    none (wider); larger buffers, everything for 4-byte elements on 64-byte
    half lines, none otherwise.  Problems the cache holds stream nothing
    and keep the plain order."""
    cases = [
        # whole lines, buffers of 604 MB to 1.4 GB
        (np.float32, shapes.CONFIG2, "on"),
        (np.float64, shapes.CONFIG2, "on"),
        (np.complex128, shapes.CONFIG3_TILE, "on"),
        (np.complex128, ([64, 128, 128, 64], [3, 2, 1, 0]), "on"),
        (np.float32, ([192, 75, 75, 96], [3, 2, 1, 0]), "on"),
        # whole lines, buffers of 198 and 200 MB
        (np.float32, ([96, 75, 75, 96], [3, 2, 1, 0]), "stores+b-loads"),
        (np.float32, ([384, 355, 384], [2, 1, 0]), "stores+b-loads"),
        # partial lines, buffers of 200 MB
        (np.float32, ([2320, 384, 59], [1, 0, 2]), "loads"),
        (np.float32, ([7248, 7248], [1, 0]), "loads"),
        # partial lines, larger buffers: half lines (rows of 9280 bytes) ...
        (np.float32, ([2320, 384, 236], [1, 0, 2]), "on"),
        (np.float32, ([7248, 24000], [1, 0]), "on"),
        # ... and less (388, 400 and 19248 bytes)
        (np.float32, ([97, 97, 97, 97], [3, 2, 1, 0]), "off"),
        (np.float32, ([100, 100, 100, 200], [3, 2, 1, 0]), "off"),
        (np.float32, ([4812, 36000], [1, 0]), "off"),
        (np.complex128, ([97, 97, 97, 97], [3, 2, 1, 0]), "off"),
        (np.float64, ([100, 100, 100, 100], [3, 2, 1, 0]), "off"),
    ]
    for dt, (size, perm), mode in cases:
        rc, err, d = E.plan(dt, size, perm)
        assert rc == 0 and "path: tile\n" in d, d
        assert ("streaming: %s " % mode) in d and "xcd-order" in d, (size, perm, d)
    rc, err, d = E.plan(np.float32, *shapes.CONFIG1)
    assert rc == 0 and "streaming: off " in d and "xcd-order" not in d, d
    # kept rows stream every access (1.4 GB per buffer), except A's loads
    # where each buffer fits in the cache (200 MB)
    rc, err, d = E.plan(np.complex128, *shapes.CONFIG3_WIDE)
    assert rc == 0 and "streaming: on " in d, d
    rc, err, d = E.plan(np.float32, [368, 384, 384], [0, 2, 1])
    assert rc == 0 and "path: wide-copy" in d and "streaming: stores+b-loads " in d, d
    # a tuned choice is kept: loads only (3), none (2), all (1), stores only
    # (4), stores and B loads (5)
    for st, mode in ((3, "loads"), (2, "off"), (1, "on"), (4, "stores"), (5, "stores+b-loads")):
        rc, err, d = E.plan(np.float32, *shapes.CONFIG2, tune=[0, 0, 0, 0, st])
        assert rc == 0 and ("streaming: %s " % mode) in d, (st, d)


def test_short_r_extents_split_evenly():
    """Beyond the last-level cache, an R extent of at most 4 tiles that the
    tile does not divide is cut into equal runs of whole lines (96 fp32 as
    3 x 32, 160 as 5 x 32); longer extents keep the model's tile (352 as
    5 x 64 + 32), and so do extents the tile divides (the 6-D workload's
    384)."""
    for size, perm, r in [([96, 75, 75, 96], [3, 2, 1, 0], "R=96/32 "),
                          ([32, 5, 15, 112, 15, 15], [3, 2, 0, 5, 1, 4], "R=160/32 "),
                          ([352, 4, 48, 28, 28], [2, 0, 4, 1, 3], "R=352/64 "),
                          (shapes.CONFIG2[0], shapes.CONFIG2[1], "R=384/64 ")]:
        rc, err, d = E.plan(np.float32, size, perm)
        assert rc == 0 and r in d, (size, perm, d)


def test_fusion_and_validation():
    # adjacent ascending pairs fuse (reference parameter_trans.tcc:235-250)
    rc, err, d = E.plan(np.float32, [4, 5, 6, 7], [2, 3, 0, 1])
    assert rc == 0 and "fused dims (A order, size:strideA:strideB): 20:1:42 42:20:1" in d, d
    # identity fuses to one dim -> axpby (the reference crashes here)
    rc, err, d = E.plan(np.float32, [4, 5, 6], [0, 1, 2])
    assert rc == 0 and "path: axpby" in d
    # padding blocks fusion
    rc, err, d = E.plan(np.float32, [4, 5, 6, 7], [2, 3, 0, 1], [5, 5, 6, 7])
    assert rc == 0 and d.count(":") >= 12
    # invalid arguments
    assert E.plan(np.float32, [4], [0])[0] == 1
    assert E.plan(np.float32, [4, 4], [0, 0])[0] == 1
    assert E.plan(np.float32, [4, 0], [1, 0])[0] == 1
    assert E.plan(np.float32, [2] * 7, list(range(7)))[0] == 4
    assert E.plan(np.float32, [4, 4], [1, 0], [3, 4])[0] == 1


def test_huge_fused_dims_are_split():
    """A fused dim of 2^31 or more elements (the kernels index with 32 bits)
    is split into two below 2^31; one that cannot be split is refused."""
    rc, err, d = E.plan(np.float32, [65536, 131072], [0, 1])     # 2^33, identity
    assert rc == 0, err
    assert "fused dims (A order, size:strideA:strideB): 1073741824:1:1 8:1073741824:1073741824" in d
    rc, err, d = E.plan(np.float32, [2**16, 2**16, 3], [1, 0, 2])
    assert rc == 0 and "path: tile" in d
    rc, err, d = E.plan(np.float32, [5, 4294967291], [0, 1])     # prime > 2^31
    assert rc == 4 and "split" in err

