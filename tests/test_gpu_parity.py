"""GPU parity: the HIP kernels through the C ABI against the oracle and the
golden vectors, bit for bit (pure data movement plus the reference's
two-rounding axpby, contraction off -- the tolerance is zero ulp)."""
import json
import os

import numpy as np
import pytest

import hptt_amd
import oracle
import shapes
from gpu_util import dev, host, run_plan

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _case(c):
    al = complex(*c["alpha"])
    be = complex(*c["beta"])
    if not c["complex_coef"]:
        al, be = al.real, be.real
    return al, be


def test_golden_vectors(device):
    man = json.load(open(os.path.join(HERE, "golden", "golden_small.json")))["cases"]
    data = np.load(os.path.join(HERE, "golden", "golden_small.npz"))
    for c in man:
        al, be = _case(c)
        A, B = data[c["name"] + "_A"], data[c["name"] + "_B"]
        got, _ = run_plan(A, B, c["in_size"], c["perm"], al, be, c["in_outer"],
                          c["out_outer"], c["complex_coef"])
        assert oracle.first_mismatch(got, data[c["name"] + "_out"]) == -1, c["name"]


def _vs_oracle(dt, size, perm, al, be, io=None, oo=None, cc=False, offset=0, tt=0.0,
               seed=0):
    na = int(np.prod(io if io else size))
    nb = int(np.prod(oo if oo else [size[p] for p in perm]))
    A = oracle.uniform(na, dt, 1000 + seed)
    B0 = oracle.uniform(nb, dt, 2000 + seed)
    exp = B0.copy()
    assert oracle.transpose(A, exp, size, perm, al, be, io, oo, nthreads=16,
                            complex_coef=cc) == 0
    got, info = run_plan(A, B0, size, perm, al, be, io, oo, cc, offset, tt)
    bad = oracle.first_mismatch(got, exp)
    assert bad == -1, (np.dtype(dt).name, size, perm, al, be, bad)
    return info


@pytest.mark.parametrize("dt", shapes.DTYPES)
def test_random_shapes(device, dt):
    rng = np.random.default_rng(99 + shapes.DTYPES.index(dt))
    for k in range(40):
        size, perm = shapes.random_case(rng, max_elems=200000, max_dim=64)
        al, be = [(1.0, 0.0), (2.3, 4.2), (-0.5, 1e-17), (1.25, -0.75)][k % 4]
        _vs_oracle(dt, size, perm, al, be, seed=k)


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_complex_coefficients(device, dt):
    rng = np.random.default_rng(5)
    for k in range(16):
        size, perm = shapes.random_case(rng, max_elems=100000, max_dim=64)
        be = complex(0.7, -0.2) if k % 2 else 0j
        _vs_oracle(dt, size, perm, complex(1.1, 0.3), be, cc=True, seed=k)


@pytest.mark.parametrize("dt", shapes.DTYPES)
def test_padded_and_misaligned(device, dt):
    rng = np.random.default_rng(7)
    for k in range(12):
        size, perm = shapes.random_case(rng, max_elems=50000, max_dim=32)
        io, oo = shapes.random_outer(rng, size, perm)
        _vs_oracle(dt, size, perm, 2.3, 4.2 if k % 2 else 0.0, io, oo, seed=k)
    for k in range(8):
        size, perm = shapes.random_case(rng, max_elems=50000, max_dim=64)
        _vs_oracle(dt, size, perm, 1.5, 0.25, offset=1, seed=k)


@pytest.mark.parametrize("dt", shapes.DTYPES)
def test_unaligned_runs(device, dt):
    """Odd extents: runs start at any element offset and end in partial
    vectors (the unaligned-vector kernels), with and without beta, on aligned
    and one-element-offset pointers."""
    rng = np.random.default_rng(31)
    odd = [[7, 5, 3], [3, 9, 11, 2], [5, 7, 6, 3, 2], [13, 3], [3, 5, 7, 2, 3, 2],
           [31, 33], [127, 65, 33], [31, 33, 35, 7], [1001, 999], [30, 30, 30, 30]]
    seen_unaligned = False
    for k, size in enumerate(odd):
        perm = [int(v) for v in rng.permutation(len(size))]
        if perm == list(range(len(size))):
            perm = perm[::-1]
        for be in (0.0, 4.2):
            for off in (0, 1):
                info = _vs_oracle(dt, size, perm, 2.3, be, offset=off, seed=10 * k + off)
                seen_unaligned |= bool(info.unaligned_r or info.unaligned_w)
    assert seen_unaligned or np.dtype(dt).itemsize == 16


@pytest.mark.parametrize("dt", shapes.DTYPES)
def test_random_large_shapes(device, dt):
    """Random ranks 2-4 with every extent in [24, 400) and up to 6M elements:
    many full-size tiles next to ragged edges, on aligned and
    one-element-offset pointers, with and without beta."""
    rng = np.random.default_rng(500 + shapes.DTYPES.index(dt))
    for k in range(10):
        rank = int(rng.integers(2, 5))
        while True:
            size = [int(v) for v in rng.integers(24, 400, size=rank)]
            if int(np.prod(size)) <= 6 * 2**20:
                break
        perm = [int(v) for v in rng.permutation(rank)]
        al, be = [(1.0, 0.0), (2.3, 4.2)][k % 2]
        _vs_oracle(dt, size, perm, al, be, offset=(k // 2) % 2, seed=300 + k)


@pytest.mark.parametrize("dt", shapes.DTYPES)
def test_wide_block_tiles(device, dt):
    """Tiles over 32 KiB of LDS (the fp32 64x128 and 8-byte 32x128 tiles) run
    512-thread blocks: full and ragged tiles, exact and unaligned runs, with
    and without beta; complex<double> tiles stay at 256 threads."""
    wide = np.dtype(dt).itemsize <= 8
    for k, (size, perm) in enumerate([([256, 64, 512], [2, 1, 0]),
                                      ([251, 64, 509], [2, 1, 0]),
                                      ([254, 8, 8, 508], [3, 1, 2, 0])]):
        for be in (0.0, 4.2):
            for off in (0, 1):
                info = _vs_oracle(dt, size, perm, 2.3, be, offset=off, seed=70 + k)
                assert (info.threads == 512) == wide, (size, info.threads)
                assert (info.lds_bytes > 32768) == wide, (size, info.lds_bytes)
                if wide and off == 0:
                    assert bool(info.unaligned_r or info.unaligned_w) == (k == 1 or (
                        k == 2 and np.dtype(dt).itemsize == 4))


def test_benchmark_workloads_full_size(device):
    """The benchmark workloads at their full sizes, against the oracle."""
    info = _vs_oracle(np.float32, *shapes.CONFIG1, 1.0, 0.0)
    assert hptt_amd.PATHS[info.path] == "tile"
    info = _vs_oracle(np.float32, *shapes.CONFIG2, 2.3, 4.2)
    assert info.update == 1 and info.vec_r == 4 and info.vec_w == 4
    assert info.unaligned_r == 0 and info.unaligned_w == 0
    _vs_oracle(np.float64, *shapes.CONFIG2, 2.3, 4.2)


def test_complex_workloads_full_size(device):
    for perm in (shapes.CONFIG3_TILE[1], shapes.CONFIG3_WIDE[1]):
        _vs_oracle(np.complex128, shapes.CONFIG3_TILE[0], perm, complex(1.1, 0.3),
                   complex(0.7, -0.2), cc=True)
        # the real-coefficient twin the reference can express
        _vs_oracle(np.complex128, shapes.CONFIG3_TILE[0], perm, 1.1, 0.7)


def test_round_trip_and_write_once(device):
    """Size-independent properties at a full benchmark size: perm then inverse
    restores A exactly; with beta = 0 every element of B is overwritten (B
    starts as NaN) and the result is a permutation of A's bits."""
    import torch
    size, perm = shapes.CONFIG2
    n = int(np.prod(size))
    g = torch.Generator(device="cuda").manual_seed(3)
    A = torch.rand(n, device="cuda", generator=g, dtype=torch.float32) * 2 - 1
    B = torch.full((n,), float("nan"), device="cuda", dtype=torch.float32)
    C = torch.empty_like(A)
    p1 = hptt_amd.create_plan_checked(A, B, size, perm, 1.0, 0.0)
    inv = list(np.argsort(perm))
    p2 = hptt_amd.create_plan_checked(B, C, [size[p] for p in perm], inv, 1.0, 0.0)
    p1.exec()
    p2.exec()
    torch.cuda.synchronize()
    assert not torch.isnan(B).any().item()
    assert torch.equal(C.view(torch.int32), A.view(torch.int32))
    assert torch.equal(torch.sort(B.view(torch.int32))[0], torch.sort(A.view(torch.int32))[0])


def test_plan_interface(device):
    """reset_data, exec on another stream, autotuning, describe."""
    import torch
    size, perm = [40, 33, 20], [2, 0, 1]
    n = int(np.prod(size))
    A1 = oracle.uniform(n, np.float64, 1)
    A2 = oracle.uniform(n, np.float64, 2)
    dA1, dA2 = dev(A1), dev(A2)
    dB1, dB2 = torch.zeros(n, dtype=torch.float64, device="cuda"), torch.zeros(
        n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    plan = hptt_amd.create_plan_checked(dA1, dB1, size, perm, 2.0, 0.0, stream=s)
    plan.exec()
    plan.reset_data(dA2, dB2)
    plan.exec()
    plan.sync()
    for A, dB in ((A1, dB1), (A2, dB2)):
        exp = np.zeros(n)
        oracle.transpose(A, exp, size, perm, 2.0, 0.0)
        assert oracle.first_mismatch(host(dB), exp) == -1
    text = plan.describe()
    assert "path: tile" in text and "grid:" in text
    # autotuned plan: same results, never touches the caller's B while tuning
    dB3 = torch.full((n,), 7.0, dtype=torch.float64, device="cuda")
    tuned = hptt_amd.create_plan_checked(dA1, dB3, size, perm, 2.0, 1.0, tuning_timeout=2.0)
    torch.cuda.synchronize()
    assert torch.all(dB3 == 7.0).item()
    tuned.exec()
    tuned.sync()
    exp = np.full(n, 7.0)
    oracle.transpose(A1, exp, size, perm, 2.0, 1.0)
    assert oracle.first_mismatch(host(dB3), exp) == -1


def test_tensors_beyond_2_31_elements(device):
    """2^31 + 2^16 fp32 elements (8.6 GB per tensor): an identity copy whose
    one fused dim is split for 32-bit indexing, and a transpose round trip
    (perm then its inverse) with 64-bit offsets, compared bit for bit."""
    import torch
    size = [65536, 32769]
    n = size[0] * size[1]
    assert n > 2**31
    A = torch.empty(n, device="cuda", dtype=torch.float32).uniform_(-1, 1)
    B = torch.empty_like(A)
    hptt_amd.create_plan_checked(A, B, size, [0, 1], 1.0, 0.0).exec()
    torch.cuda.synchronize()
    assert torch.equal(A.view(torch.int32), B.view(torch.int32))
    p1 = hptt_amd.create_plan_checked(A, B, size, [1, 0], 1.0, 0.0)
    C = torch.empty_like(A)
    p2 = hptt_amd.create_plan_checked(B, C, size[::-1], [1, 0], 1.0, 0.0)
    p1.exec()
    p2.exec()
    torch.cuda.synchronize()
    assert torch.equal(A.view(torch.int32), C.view(torch.int32))
    # the transposed layout: A is {65536, 32769} column-major, so a row-major
    # view (32769, 65536) holds A(i0, i1) at [i1, i0]; B is {32769, 65536}
    # with B(i1, i0) = A(i0, i1) at row-major [i0, i1]
    Av = A.view(size[1], size[0])
    Bv = B.view(size[0], size[1])
    for i0, i1 in [(0, 0), (65535, 32768), (12345, 23456), (1, 32767)]:
        assert Bv[i0, i1].item() == Av[i1, i0].item()
    del A, B, C
    torch.cuda.empty_cache()



def test_tuning_cache_file(device, tmp_path):
    """tuning_timeout != 0 records the chosen tile parameters in the file
    named by HPTTX_TUNE_CACHE (key: arch, dtype, update flag, fused problem;
    value: run_r run_c block wide stream), and the tuned plan is exact."""
    import os
    path = tmp_path / "tune.txt"
    os.environ["HPTTX_TUNE_CACHE"] = str(path)
    try:
        size, perm = [48, 20, 36], [1, 2, 0]
        info = _vs_oracle(np.float32, size, perm, 1.5, 0.5, tt=1.0, seed=77)
    finally:
        del os.environ["HPTTX_TUNE_CACHE"]
    assert info.update == 1
    lines = path.read_text().splitlines()
    assert lines, "no tuning record written"
    key, value = lines[-1].split("\t")
    # {48, 20, 36} with perm {1, 2, 0} fuses A's dims 1 and 2 (contiguous in both)
    assert key == "gfx950|0|1|2|48:1:720|720:48:1", key
    assert len(value.split()) == 5


def test_concurrent_exec_from_host_threads(device):
    """Plans are immutable after creation: several host threads exec their own
    plans, and one shared plan, each on its own stream, at the same time; every
    result is exact (the C ABI keeps no per-exec host state)."""
    import threading

    import torch
    size, perm = [40, 36, 50], [2, 0, 1]
    n = int(np.prod(size))
    A = oracle.uniform(n, np.float32, 5)
    exp = np.zeros(n, np.float32)
    assert oracle.transpose(A, exp, size, perm, 1.0, 0.0) == 0
    dA = dev(A)
    outs = [torch.zeros(n, dtype=torch.float32, device="cuda") for _ in range(8)]
    streams = [torch.cuda.Stream() for _ in range(8)]
    torch.cuda.synchronize()
    # threads 0 and 1 share one plan (same idempotent output, alpha=1,
    # beta=0), threads 2..7 each own one
    shared = hptt_amd.create_plan_checked(dA, outs[0], size, perm, 1.0, 0.0)
    own = {k: hptt_amd.create_plan_checked(dA, outs[k], size, perm, 1.0, 0.0, stream=streams[k])
           for k in range(2, 8)}
    errors = []

    def worker(k):
        try:
            for _ in range(20):
                if k < 2:
                    shared.exec_on(streams[k])
                else:
                    own[k].exec()
            streams[k].synchronize()
        except Exception as e:  # noqa: BLE001 -- reported below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for k in [0] + list(range(2, 8)):
        assert oracle.first_mismatch(host(outs[k]), exp) == -1, k


def test_every_streaming_mode_is_exact(device, tmp_path):
    """Tuning-cache records force each streaming mode (1 loads and stores, 2
    none, 3 loads only, 4 stores only, 5 stores and B's loads) on tile plans,
    and mode 5 on a row plan, with beta != 0, in fp32 and c128; every result
    equals the oracle.  Runs in a fresh process, which reads the cache
    file."""
    import subprocess
    import sys
    root = os.path.dirname(HERE)
    cache = tmp_path / "tune.txt"
    # {n0, 24, 20, 16} perm {3, 2, 1, 0}: fused dims n0:1:7680, 24:n0:320,
    # 20:24*n0:16, 16:480*n0:1 (fp32 code 0, c128 code 3, update flag 1)
    lines = []
    for mode, n0 in zip(range(1, 6), (64, 32, 96, 128, 160)):
        fd = "%d:1:7680|24:%d:320|20:%d:16|16:%d:1" % (n0, n0, 24 * n0, 480 * n0)
        for code in (0, 3):
            lines.append("gfx950|%d|1|4|%s\t0 0 0 0 %d" % (code, fd, mode))
    # {64, 24, 20, 16} perm {0, 3, 2, 1}: rows of 64 along the kept dim
    for code in (0, 3):
        lines.append("gfx950|%d|1|4|64:1:1|24:64:20480|20:1536:1024|16:30720:64\t0 0 0 0 5" % code)
    cache.write_text("\n".join(lines) + "\n")
    script = (
        "import sys; sys.path.insert(0, 'tests')\n"
        "import numpy as np, torch, hptt_amd, oracle\n"
        "names = {1: 'on', 2: 'off', 3: 'loads', 4: 'stores', 5: 'stores+b-loads'}\n"
        "for mode, n0 in zip(range(1, 6), (64, 32, 96, 128, 160)):\n"
        "    for dt in (np.float32, np.complex128):\n"
        "        size, perm = [n0, 24, 20, 16], [3, 2, 1, 0]\n"
        "        n = int(np.prod(size))\n"
        "        A = oracle.uniform(n, dt, mode); B = oracle.uniform(n, dt, 10 + mode)\n"
        "        dA = torch.from_numpy(A).cuda(); dB = torch.from_numpy(B).cuda()\n"
        "        p = hptt_amd.create_plan_checked(dA, dB, size, perm, 1.5, -0.75)\n"
        "        d = p.describe()\n"
        "        assert 'path: tile' in d and ('streaming: %s ' % names[mode]) in d, (mode, d)\n"
        "        p.exec(); p.sync(); torch.cuda.synchronize()\n"
        "        exp = B.copy()\n"
        "        assert oracle.transpose(A, exp, size, perm, 1.5, -0.75) == 0\n"
        "        assert oracle.first_mismatch(dB.cpu().numpy(), exp) == -1, (mode, dt)\n"
        "# rows (the stride-1 dim kept) under stores and B's loads\n"
        "size, perm = [64, 24, 20, 16], [0, 3, 2, 1]\n"
        "n = int(np.prod(size))\n"
        "for dt in (np.float32, np.complex128):\n"
        "    A = oracle.uniform(n, dt, 21); B = oracle.uniform(n, dt, 22)\n"
        "    dA = torch.from_numpy(A).cuda(); dB = torch.from_numpy(B).cuda()\n"
        "    p = hptt_amd.create_plan_checked(dA, dB, size, perm, 1.5, -0.75)\n"
        "    d = p.describe()\n"
        "    assert 'path: wide-copy' in d and 'streaming: stores+b-loads ' in d, d\n"
        "    p.exec(); p.sync(); torch.cuda.synchronize()\n"
        "    exp = B.copy()\n"
        "    assert oracle.transpose(A, exp, size, perm, 1.5, -0.75) == 0\n"
        "    assert oracle.first_mismatch(dB.cpu().numpy(), exp) == -1, ('rows', dt)\n"
        "print('MODES OK')\n")
    env = dict(os.environ, HPTTX_TUNE_CACHE=str(cache))
    r = subprocess.run([sys.executable, "-c", script], cwd=root, env=env, capture_output=True,
                       text=True, timeout=110)
    assert r.returncode == 0 and "MODES OK" in r.stdout, r.stdout + r.stderr


def test_bad_tuning_cache_record_falls_back(device, tmp_path):
    """A tuning-cache record that does not plan (here tile runs of 4096
    elements, which need far more than 64 KiB of LDS) is replaced by the
    model's pick instead of failing create_plan; the plan is exact.  Runs in
    a fresh process, which reads the cache file named by HPTTX_TUNE_CACHE."""
    import subprocess
    import sys
    root = os.path.dirname(HERE)
    cache = tmp_path / "tune.txt"
    key = "gfx950|0|0|2|8192:1:8192|8192:8192:1"
    cache.write_text(key + "\t4096 4096 1048576 256 0\n")
    script = (
        "import torch, hptt_amd\n"
        "A = torch.rand(8192 * 8192, device='cuda')\n"
        "B = torch.zeros_like(A)\n"
        "p = hptt_amd.create_plan_checked(A, B, [8192, 8192], [1, 0], 1.0, 0.0)\n"
        "p.exec(); p.sync()\n"
        "assert torch.equal(B.view(8192, 8192), A.view(8192, 8192).t())\n"
        "print('FALLBACK OK')\n")
    env = dict(os.environ, HPTTX_TUNE_CACHE=str(cache))
    r = subprocess.run([sys.executable, "-c", script], cwd=root, env=env, capture_output=True,
                       text=True, timeout=110)
    assert r.returncode == 0 and "FALLBACK OK" in r.stdout, r.stdout + r.stderr
    assert cache.read_text().splitlines()[-1] == key + "\t0 0 0 0 0"
