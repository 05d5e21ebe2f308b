"""Batched mode on the GPU: work-queue launches per element type and kernel
variant, replayed from a hipGraph, against the oracle and against sequential
plans -- bit for bit."""
import numpy as np
import pytest

import hptt_amd
import oracle
import shapes
from gpu_util import dev, host
from hptt_amd.workloads import batch_manifest

pytestmark = pytest.mark.gpu

NP = {hptt_amd.F32: np.float32, hptt_amd.F64: np.float64, hptt_amd.C64: np.complex64,
      hptt_amd.C128: np.complex128}


def _random_items(rng, count, dtypes, max_elems=60000, padded=False):
    items = []
    for k in range(count):
        code = int(rng.choice(dtypes))
        size, perm = shapes.random_case(rng, max_elems=max_elems, max_dim=48)
        io = oo = None
        if padded and k % 3 == 0:
            io, oo = shapes.random_outer(rng, size, perm)
        cc = NP[code] in (np.complex64, np.complex128) and k % 4 == 1
        al = complex(rng.uniform(-2, 2), rng.uniform(-1, 1)) if cc else float(rng.uniform(-2, 2))
        if k % 4 == 0:
            be = 0.0
        else:
            be = complex(rng.uniform(-2, 2), rng.uniform(-1, 1)) if cc else float(rng.uniform(-2, 2))
        items.append(dict(code=code, size=size, perm=perm, io=io, oo=oo, alpha=al, beta=be,
                          cc=cc))
    return items


def _materialize(items, seed):
    """Host and device buffers for every item; device A/B per item."""
    out = []
    for k, it in enumerate(items):
        dt = NP[it["code"]]
        na = int(np.prod(it["io"] or it["size"]))
        nb = int(np.prod(it["oo"] or [it["size"][p] for p in it["perm"]]))
        A = oracle.uniform(na, dt, seed + 2 * k)
        B = oracle.uniform(nb, dt, seed + 2 * k + 1)
        out.append(dict(A=A, B=B, dA=dev(A), dB=dev(B)))
    return out


def _expected(it, buf, B0=None):
    exp = (B0 if B0 is not None else buf["B"]).copy()
    rc = oracle.transpose(buf["A"], exp, it["size"], it["perm"], it["alpha"], it["beta"],
                          it["io"], it["oo"], nthreads=8, complex_coef=it["cc"])
    assert rc == 0
    return exp


def _batch(items, bufs, tuning_timeout=0.0):
    bi = [hptt_amd.make_batch_item(b["dA"], b["dB"], it["size"], it["perm"], it["alpha"],
                                   it["beta"], in_outer_size=it["io"] or (),
                                   out_outer_size=it["oo"] or ())
          for it, b in zip(items, bufs)]
    return hptt_amd.Batch(bi, tuning_timeout=tuning_timeout)


def test_mixed_batch_vs_oracle(device):
    import torch
    rng = np.random.default_rng(31)
    items = _random_items(rng, 240, [0, 1, 2, 3], padded=True)
    bufs = _materialize(items, 500)
    torch.cuda.synchronize()
    batch = _batch(items, bufs)
    info = batch.info()
    assert info.count == 240 and info.groups >= 4   # one launch per (dtype, variant)
    batch.exec()
    batch.sync()
    for k, (it, b) in enumerate(zip(items, bufs)):
        got = host(b["dB"])
        exp = _expected(it, b)
        assert oracle.first_mismatch(got, exp) == -1, (k, it, batch.item_path(k))


def test_graph_replays_with_eager_work_between(device):
    """Replay the captured graph several times; before each replay B is reset
    by eager torch work; every replay must match the oracle."""
    import torch
    rng = np.random.default_rng(8)
    items = _random_items(rng, 120, [0, 3], max_elems=30000)
    bufs = _materialize(items, 900)
    torch.cuda.synchronize()
    batch = _batch(items, bufs)
    exps = [_expected(it, b) for it, b in zip(items, bufs)]
    B0s = [b["dB"].clone() for b in bufs]
    for rep in range(4):
        for b, b0 in zip(bufs, B0s):
            b["dB"].copy_(b0)
        torch.cuda.synchronize()
        junk = torch.randn(1 << 20, device="cuda").sum()   # eager work between replays
        batch.exec()
        batch.sync()
        torch.cuda.synchronize()
        assert torch.isfinite(junk).item()
        for k, (b, e) in enumerate(zip(bufs, exps)):
            assert oracle.first_mismatch(host(b["dB"]), e) == -1, (rep, k)


def test_direct_launch_equals_graph(device):
    import torch
    rng = np.random.default_rng(9)
    items = _random_items(rng, 80, [1, 2])
    bufs = _materialize(items, 1300)
    torch.cuda.synchronize()
    batch = _batch(items, bufs)
    batch.use_graph(False)
    batch.exec()
    batch.sync()
    for it, b in zip(items, bufs):
        assert oracle.first_mismatch(host(b["dB"]), _expected(it, b)) == -1


def test_persistent_grid_mode(device):
    """HPTTX_BATCH_GRID=persistent launches the same kernels on a grid of
    resident blocks that walk the tile queue; results match the oracle and
    the grid is smaller than the tile count."""
    import os

    import torch
    rng = np.random.default_rng(13)
    items = _random_items(rng, 120, [0, 1, 2, 3])
    bufs = _materialize(items, 1400)
    torch.cuda.synchronize()
    os.environ["HPTTX_BATCH_GRID"] = "persistent"
    try:
        batch = _batch(items, bufs)
    finally:
        del os.environ["HPTTX_BATCH_GRID"]
    info = batch.info()
    assert info.tiles > sum(info.grid[:min(4, info.groups)]) or info.groups > 4
    batch.exec()
    batch.sync()
    for it, b in zip(items, bufs):
        assert oracle.first_mismatch(host(b["dB"]), _expected(it, b)) == -1


def test_manifest_batch_equals_sequential_plans(device):
    """The first 512 tensors of the batched workload manifest (fp32): the
    batch and one plan->exec() per tensor give identical bits, and a sample
    matches the oracle."""
    import torch
    man = batch_manifest(count=512)
    items = [dict(code=0, size=m["in_size"], perm=m["perm"], io=None, oo=None,
                  alpha=m["alpha"], beta=m["beta"], cc=False) for m in man]
    g = torch.Generator(device="cuda").manual_seed(4)
    As = [torch.rand(int(np.prod(it["size"])), device="cuda", generator=g) * 2 - 1 for it in items]
    B0 = [torch.rand(int(np.prod(it["size"])), device="cuda", generator=g) * 2 - 1 for it in items]
    Bb = [b.clone() for b in B0]
    Bs = [b.clone() for b in B0]
    bufs = [dict(dA=a, dB=b) for a, b in zip(As, Bb)]
    batch = _batch(items, bufs)
    plans = [hptt_amd.create_plan_checked(a, b, it["size"], it["perm"], it["alpha"], it["beta"])
             for a, b, it in zip(As, Bs, items)]
    torch.cuda.synchronize()
    batch.exec()
    for p in plans:
        p.exec()
    torch.cuda.synchronize()
    for k in range(len(items)):
        assert torch.equal(Bb[k].view(torch.int32), Bs[k].view(torch.int32)), k
    for k in range(0, len(items), 17):
        A = host(As[k])
        exp = host(B0[k])
        oracle.transpose(A, exp, items[k]["size"], items[k]["perm"], items[k]["alpha"],
                         items[k]["beta"], nthreads=8)
        assert oracle.first_mismatch(host(Bb[k]), exp) == -1, k


def test_set_data_and_tuned_classes(device):
    import torch
    rng = np.random.default_rng(10)
    items = _random_items(rng, 40, [0])
    bufs = _materialize(items, 1700)
    torch.cuda.synchronize()
    batch = _batch(items, bufs, tuning_timeout=3.0)
    # re-point item 5 at fresh buffers
    it = items[5]
    n_a, n_b = bufs[5]["A"].size, bufs[5]["B"].size
    A2 = oracle.uniform(n_a, np.float32, 77)
    B2 = oracle.uniform(n_b, np.float32, 78)
    dA2, dB2 = dev(A2), dev(B2)
    torch.cuda.synchronize()
    batch.set_data(5, dA2, dB2)
    batch.exec()
    batch.sync()
    exp = B2.copy()
    oracle.transpose(A2, exp, it["size"], it["perm"], it["alpha"], it["beta"], it["io"],
                     it["oo"])
    assert oracle.first_mismatch(host(dB2), exp) == -1
    for k, (it, b) in enumerate(zip(items, bufs)):
        if k == 5:
            continue
        assert oracle.first_mismatch(host(b["dB"]), _expected(it, b)) == -1, k


def test_single_persistent_launch_per_type(device):
    """HPTTX_BATCH_MODE=merged with HPTTX_BATCH_GRID=persistent: one persistent
    work-queue launch per element type (the kernel dispatches each entry's
    variant), replayed from the graph; results match the oracle."""
    import os

    import torch
    rng = np.random.default_rng(17)
    items = _random_items(rng, 90, [0, 1])
    bufs = _materialize(items, 1500)
    torch.cuda.synchronize()
    os.environ["HPTTX_BATCH_MODE"] = "merged"
    os.environ["HPTTX_BATCH_GRID"] = "persistent"
    try:
        batch = _batch(items, bufs)
    finally:
        del os.environ["HPTTX_BATCH_MODE"]
        del os.environ["HPTTX_BATCH_GRID"]
    info = batch.info()
    assert info.groups == len({it["code"] for it in items})
    assert all(g < info.tiles for g in info.grid[:info.groups])
    batch.exec()
    batch.sync()
    for it, b in zip(items, bufs):
        assert oracle.first_mismatch(host(b["dB"]), _expected(it, b)) == -1


def test_one_launch_dynamic_queue(device):
    """The one-launch layout (hpttx_batch_create_ex with HPTTX_BATCH_ONE_LAUNCH):
    one persistent launch per element type whose blocks take chunks of tiles
    off a device counter.  The counter is cleared by a kernel node, so graph
    replays (with eager work between them) and direct launches all match the
    oracle."""
    import torch
    rng = np.random.default_rng(23)
    items = _random_items(rng, 200, [0, 1, 2, 3], padded=True)
    bufs = _materialize(items, 2100)
    torch.cuda.synchronize()
    bi = [hptt_amd.make_batch_item(b["dA"], b["dB"], it["size"], it["perm"], it["alpha"],
                                   it["beta"], in_outer_size=it["io"] or (),
                                   out_outer_size=it["oo"] or ())
          for it, b in zip(items, bufs)]
    batch = hptt_amd.Batch(bi, one_launch=True)
    info = batch.info()
    assert info.groups == len({it["code"] for it in items})
    assert all(g < info.tiles for g in info.grid[:info.groups])
    exps = [_expected(it, b) for it, b in zip(items, bufs)]
    B0s = [b["dB"].clone() for b in bufs]
    for rep in range(4):
        for b, b0 in zip(bufs, B0s):
            b["dB"].copy_(b0)
        torch.cuda.synchronize()
        batch.use_graph(rep != 3)
        junk = torch.randn(1 << 20, device="cuda").sum()
        batch.exec()
        batch.sync()
        torch.cuda.synchronize()
        assert torch.isfinite(junk).item()
        for k, (b, e) in enumerate(zip(bufs, exps)):
            assert oracle.first_mismatch(host(b["dB"]), e) == -1, (rep, k)
