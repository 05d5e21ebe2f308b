"""The batched workload's whole manifest on the GPU: 4096 tensors (rank 2..6,
dims log-uniform in [2, 128], at most 4 Mi elements each, random perms,
alpha/beta uniform in [-2, 2], beta = 0 for 25%), all fp32 and all fp64, in
the default launch layout (one launch per element type, kernel variant and
LDS class), as merged persistent launches (one static work-queue launch per
element type) and in the one-launch layout (one persistent launch per
element type over a dynamic queue).  Every tensor must equal its own plain
plan->exec() bit for bit, nothing outside the tensors may be written, and at
least one tensor of every (variant, LDS class) group -- the largest -- is
checked against the oracle (the reference's RefTrans,
test/inc/hptt/test_util.tcc:143-185)."""
import os

import numpy as np
import pytest

import hptt_amd
import oracle
from hptt_amd.workloads import batch_manifest, layout

pytestmark = pytest.mark.gpu

# layout: (environment while the batch is built, one_launch flag)
LAYOUTS = {
    "default": ({}, False),
    "merged-persistent": ({"HPTTX_BATCH_MODE": "merged", "HPTTX_BATCH_GRID": "persistent"}, False),
    "one-launch": ({}, True),
}


def _build(items, layout):
    env, one_launch = layout
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return hptt_amd.Batch(items, one_launch=one_launch)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("lay", list(LAYOUTS))
@pytest.mark.parametrize("dtname", ["float32", "float64"])
def test_full_manifest(device, dtname, lay):
    import torch
    man = batch_manifest()
    assert len(man) == 4096
    tdt = getattr(torch, dtname)
    npdt = np.float32 if dtname == "float32" else np.float64
    iv = torch.int32 if dtname == "float32" else torch.int64
    esz = 4 if dtname == "float32" else 8
    offs, total = layout(man, esz)
    print("manifest %s %s: %d tensors, %d elements" % (dtname, lay, len(man), total), flush=True)
    g = torch.Generator(device="cuda").manual_seed(77 if dtname == "float32" else 78)
    A = torch.empty(total, dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    B0 = torch.empty(total, dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    Bb, Bs = B0.clone(), B0.clone()
    covered = torch.zeros(total, dtype=torch.bool, device="cuda")
    items, plans = [], []
    for m, o in zip(man, offs):
        n = int(np.prod(m["in_size"]))
        covered[o:o + n] = True
        items.append(hptt_amd.make_batch_item(A[o:o + n], Bb[o:o + n], m["in_size"], m["perm"],
                                              m["alpha"], m["beta"]))
        plans.append(hptt_amd.create_plan_checked(A[o:o + n], Bs[o:o + n], m["in_size"],
                                                  m["perm"], m["alpha"], m["beta"]))
    torch.cuda.synchronize()
    batch = _build(items, LAYOUTS[lay])
    info = batch.info()
    if lay != "default":
        assert info.groups == 1
    batch.exec()
    batch.sync()
    for p in plans:
        p.exec()
    torch.cuda.synchronize()
    # every tensor: batch == sequential plans, bit for bit
    if not torch.equal(Bb.view(iv), Bs.view(iv)):
        for k, (m, o) in enumerate(zip(man, offs)):
            n = int(np.prod(m["in_size"]))
            assert torch.equal(Bb[o:o + n].view(iv), Bs[o:o + n].view(iv)), (k, m)
        raise AssertionError("a gap between tensors differs")
    # nothing written between the tensors
    assert torch.equal(Bb[~covered].view(iv), B0[~covered].view(iv))
    # the largest tensor of every (variant, LDS class) group against the oracle
    pick = {}
    for k, m in enumerate(man):
        inf = batch.item_info(k)
        key = (inf.variant, (inf.lds_bytes + 8191) // 8192)
        n = int(np.prod(m["in_size"]))
        if key not in pick or n > pick[key][1]:
            pick[key] = (k, n)
    print("oracle checks: %d groups" % len(pick), flush=True)
    assert len(pick) >= 8
    for key, (k, n) in sorted(pick.items()):
        m, o = man[k], offs[k]
        a = A[o:o + n].cpu().numpy()
        exp = B0[o:o + n].cpu().numpy()
        assert oracle.transpose(a, exp, m["in_size"], m["perm"], m["alpha"], m["beta"],
                                nthreads=8) == 0
        got = Bb[o:o + n].cpu().numpy()
        assert got.dtype == npdt
        assert oracle.first_mismatch(got, exp) == -1, (k, key, m)
    del batch, plans, A, B0, Bb, Bs, covered
    torch.cuda.empty_cache()
