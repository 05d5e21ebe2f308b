"""hpttx_batch_run_noalloc through the C ABI: the same bits as a batch built
by hpttx_batch_create on the same inputs, on two calls in a row (the second
with fewer and smaller items, so stale table contents from the first call
would show), and a clear refusal of a batch whose tables exceed the static
scratch.  The refusal happens before any launch."""
import ctypes

import numpy as np
import pytest

import hptt_amd
import oracle
from gpu_util import dev, host
from test_gpu_batch import NP, _random_items

pytestmark = pytest.mark.gpu


def _buffers(items, seed):
    out = []
    for k, it in enumerate(items):
        dt = NP[it["code"]]
        na = int(np.prod(it["io"] or it["size"]))
        nb = int(np.prod(it["oo"] or [it["size"][p] for p in it["perm"]]))
        A = oracle.uniform(na, dt, seed + 2 * k)
        B = oracle.uniform(nb, dt, seed + 2 * k + 1)
        out.append(dict(A=A, B=B, dA=dev(A), dB_ref=dev(B), dB_na=dev(B)))
    return out


def _items(items, bufs, key):
    return [hptt_amd.make_batch_item(b["dA"], b[key], it["size"], it["perm"], it["alpha"],
                                     it["beta"], in_outer_size=it["io"] or (),
                                     out_outer_size=it["oo"] or ())
            for it, b in zip(items, bufs)]


def _run_and_compare(items, bufs, stream):
    import torch
    ref = hptt_amd.Batch(_items(items, bufs, "dB_ref"))
    ref.exec()
    ref.sync()
    hptt_amd.batch_run_noalloc(_items(items, bufs, "dB_na"), stream=stream)
    stream.synchronize()
    torch.cuda.synchronize()
    for k, (it, b) in enumerate(zip(items, bufs)):
        got, exp = host(b["dB_na"]), host(b["dB_ref"])
        assert oracle.first_mismatch(got, exp) == -1, (k, it)
    # and a sample against the oracle
    for k in range(0, len(items), 11):
        it, b = items[k], bufs[k]
        exp = b["B"].copy()
        assert oracle.transpose(b["A"], exp, it["size"], it["perm"], it["alpha"], it["beta"],
                                it["io"], it["oo"], nthreads=8, complex_coef=it["cc"]) == 0
        assert oracle.first_mismatch(host(b["dB_na"]), exp) == -1, (k, it)
    del ref


def test_noalloc_batch_equals_allocating_batch(device):
    import torch
    stream = torch.cuda.Stream()
    rng = np.random.default_rng(707)
    first = _random_items(rng, 300, [0, 1, 2, 3], max_elems=80000, padded=True)
    bufs1 = _buffers(first, 31000)
    torch.cuda.synchronize()
    _run_and_compare(first, bufs1, stream)
    second = _random_items(rng, 37, [0, 1, 2, 3], max_elems=4000)
    bufs2 = _buffers(second, 47000)
    torch.cuda.synchronize()
    _run_and_compare(second, bufs2, stream)


def test_noalloc_refuses_tables_beyond_the_scratch(device):
    import torch
    A = torch.zeros(64, device="cuda")
    B = torch.zeros(64, device="cuda")
    torch.cuda.synchronize()
    item = hptt_amd.make_batch_item(A, B, [8, 8], [1, 0], 1.0, 0.0)
    many = [item] * 12000   # 12000 table entries of 872 bytes > 8 MiB
    with pytest.raises(hptt_amd.HpttError) as ei:
        hptt_amd.batch_run_noalloc(many)
    assert ei.value.status == hptt_amd.EINVAL
    assert "static scratch" in str(ei.value)
    assert torch.count_nonzero(B).item() == 0
