"""Serving mode on the GPU: requests admitted as they arrive, batched into
work-queue launches on several streams; every result must equal, bit for bit,
the same transposition run alone through a plain plan and the oracle."""
import threading
import time

import numpy as np
import pytest

import hptt_amd
import oracle
import shapes
from gpu_util import dev, host, read_device

pytestmark = pytest.mark.gpu

NP = {0: np.float32, 1: np.float64, 2: np.complex64, 3: np.complex128}


def _requests(rng, count, max_elems=150000):
    out = []
    for k in range(count):
        code = int(rng.integers(0, 4))
        size, perm = shapes.random_case(rng, max_elems=max_elems, max_dim=64)
        cc = code >= 2 and k % 3 == 0
        al = complex(rng.uniform(-2, 2), rng.uniform(-1, 1)) if cc else float(rng.uniform(-2, 2))
        if k % 4 == 0:
            be = 0.0
        else:
            be = complex(rng.uniform(-2, 2), rng.uniform(-1, 1)) if cc else float(rng.uniform(-2, 2))
        n = int(np.prod(size))
        A = oracle.uniform(n, NP[code], 10 * k + 1)
        B = oracle.uniform(n, NP[code], 10 * k + 2)
        out.append(dict(code=code, size=size, perm=perm, alpha=al, beta=be, cc=cc, A=A, B=B,
                        dA=dev(A), dB=dev(B)))
    return out


def _item(r, out=True):
    return hptt_amd.make_batch_item(r["dA"], r["dB"] if out else None, r["size"], r["perm"],
                                    r["alpha"], r["beta"])


def _alone(r):
    """The same transposition through a plain plan, on a fresh copy of B."""
    import torch
    dB = dev(r["B"])
    p = hptt_amd.create_plan_checked(r["dA"], dB, r["size"], r["perm"], r["alpha"], r["beta"])
    p.exec()
    p.sync()
    torch.cuda.synchronize()
    return host(dB)


def test_served_equals_alone_and_oracle(device):
    """Requests arrive from a second thread while earlier ones run."""
    import torch
    rng = np.random.default_rng(2024)
    reqs = _requests(rng, 160)
    torch.cuda.synchronize()
    srv = hptt_amd.Server(slots=3, max_batch=64)
    tickets = [None] * len(reqs)

    def client():
        for k, r in enumerate(reqs):
            tickets[k] = srv.submit(_item(r))
            if k % 7 == 0:
                time.sleep(0.0005)

    th = threading.Thread(target=client)
    th.start()
    th.join()
    srv.drain()
    st = srv.stats()
    assert st.completed == len(reqs) and st.failed == 0
    assert 1 <= st.launches <= len(reqs)
    assert 0 < st.p50_latency <= st.p99_latency
    for k, r in enumerate(reqs):
        c = srv.wait(tickets[k], 5.0)
        assert c.state == 2, (k, c.status)
        got = host(r["dB"])
        exp = r["B"].copy()
        oracle.transpose(r["A"], exp, r["size"], r["perm"], r["alpha"], r["beta"], nthreads=8,
                         complex_coef=r["cc"])
        assert oracle.first_mismatch(got, exp) == -1, k
        assert oracle.first_mismatch(got, _alone(r)) == -1, k
    srv.close()


def test_submit_many_equals_alone_and_oracle(device):
    """Requests handed over in groups (hpttx_server_submit_many): each result
    equals its plain plan and the oracle; a rejected request stops the
    submission at that request."""
    import torch
    rng = np.random.default_rng(4096)
    reqs = _requests(rng, 120)
    torch.cuda.synchronize()
    srv = hptt_amd.Server(slots=3, max_batch=256)
    tickets = []
    for lo in range(0, len(reqs), 50):
        tickets += srv.submit_many([_item(r) for r in reqs[lo:lo + 50]])
    assert len(set(tickets)) == len(reqs) and 0 not in tickets
    srv.drain()
    st = srv.stats()
    assert st.completed == len(reqs) and st.failed == 0
    for k, r in enumerate(reqs):
        c = srv.wait(tickets[k], 5.0)
        assert c.state == 2, (k, c.status)
        got = host(r["dB"])
        exp = r["B"].copy()
        oracle.transpose(r["A"], exp, r["size"], r["perm"], r["alpha"], r["beta"], nthreads=8,
                         complex_coef=r["cc"])
        assert oracle.first_mismatch(got, exp) == -1, k
        assert oracle.first_mismatch(got, _alone(r)) == -1, k
    # the third request needs a server-allocated output with beta != 0 (no
    # arena): the first two are queued, the rest are not
    extra = _requests(rng, 4)
    items = [_item(r) for r in extra]
    items[2] = _item(extra[2], out=False)
    L = hptt_amd.lib()
    import ctypes
    arr = (hptt_amd.BatchItem * 4)(*items)
    tk = (ctypes.c_uint64 * 4)()
    q = ctypes.c_uint32()
    rc = L.hpttx_server_submit_many(srv._h, 4, arr, tk, ctypes.byref(q))
    assert rc != hptt_amd.OK and q.value == 2 and tk[0] and tk[1] and not tk[2] and not tk[3]
    srv.drain()
    for i in range(2):
        c = srv.wait(tk[i], 5.0)
        assert c.state == 2
        exp = extra[i]["B"].copy()
        oracle.transpose(extra[i]["A"], exp, extra[i]["size"], extra[i]["perm"], extra[i]["alpha"],
                         extra[i]["beta"], complex_coef=extra[i]["cc"])
        assert oracle.first_mismatch(host(extra[i]["dB"]), exp) == -1, i
    srv.close()


def test_submit_many_across_chunks(device):
    """One submission of 420 requests (planned and queued in chunks of 128,
    256, ...) with a rejected request at index 401, in the third chunk: the
    401 before it are queued, run and equal the oracle; none after it is
    queued."""
    import ctypes

    import torch
    rng = np.random.default_rng(420)
    reqs = _requests(rng, 420, max_elems=20000)
    torch.cuda.synchronize()
    srv = hptt_amd.Server(slots=3, max_batch=256)
    items = [_item(r) for r in reqs]
    items[401] = _item(reqs[401], out=False)   # server output with beta != 0 and no arena
    assert reqs[401]["beta"] != 0
    arr = (hptt_amd.BatchItem * len(items))(*items)
    tk = (ctypes.c_uint64 * len(items))()
    q = ctypes.c_uint32()
    rc = hptt_amd.lib().hpttx_server_submit_many(srv._h, len(items), arr, tk, ctypes.byref(q))
    assert rc != hptt_amd.OK and q.value == 401
    assert all(tk[i] for i in range(401)) and not any(tk[i] for i in range(401, len(items)))
    assert len(set(tk[i] for i in range(401))) == 401
    srv.drain()
    st = srv.stats()
    assert st.completed == 401 and st.failed == 0 and st.launches >= 2
    for k in range(401):
        r = reqs[k]
        c = srv.wait(tk[k], 5.0)
        assert c.state == 2, (k, c.status)
        exp = r["B"].copy()
        oracle.transpose(r["A"], exp, r["size"], r["perm"], r["alpha"], r["beta"], nthreads=8,
                         complex_coef=r["cc"])
        assert oracle.first_mismatch(host(r["dB"]), exp) == -1, k
    for k in range(401, len(reqs)):   # not queued: B untouched
        assert oracle.first_mismatch(host(reqs[k]["dB"]), reqs[k]["B"]) == -1, k
    srv.close()


def test_submit_many_from_two_threads(device):
    """Two threads hand over 300 requests each in one submit_many call at the
    same time: their chunks interleave in the queue, every ticket is unique,
    each thread's tickets rise in its own order, and every result equals the
    oracle."""
    import torch
    rng = np.random.default_rng(2600)
    reqs = _requests(rng, 600, max_elems=20000)
    torch.cuda.synchronize()
    srv = hptt_amd.Server(slots=3, max_batch=128)
    out = [None, None]
    barrier = threading.Barrier(2)

    def client(j):
        items = (hptt_amd.BatchItem * 300)(*[_item(r) for r in reqs[300 * j:300 * (j + 1)]])
        barrier.wait()
        out[j] = srv.submit_many(items)

    ths = [threading.Thread(target=client, args=(j,)) for j in range(2)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    tickets = out[0] + out[1]
    assert len(set(tickets)) == 600 and 0 not in tickets
    assert out[0] == sorted(out[0]) and out[1] == sorted(out[1])
    srv.drain()
    st = srv.stats()
    assert st.completed == 600 and st.failed == 0
    for k, r in enumerate(reqs):
        c = srv.wait(tickets[k], 5.0)
        assert c.state == 2, (k, c.status)
        exp = r["B"].copy()
        oracle.transpose(r["A"], exp, r["size"], r["perm"], r["alpha"], r["beta"], nthreads=8,
                         complex_coef=r["cc"])
        assert oracle.first_mismatch(host(r["dB"]), exp) == -1, k
    srv.close()


def test_server_allocated_outputs(device):
    import torch
    rng = np.random.default_rng(77)
    reqs = _requests(rng, 40)
    for r in reqs:
        r["beta"] = 0.0
    torch.cuda.synchronize()
    srv = hptt_amd.Server(arena_bytes=64 << 20, slots=2, max_batch=16)
    tickets = [srv.submit(_item(r, out=False)) for r in reqs]
    srv.drain()
    for k, (t, r) in enumerate(zip(tickets, reqs)):
        c = srv.wait(t, 5.0)
        assert c.state == 2 and c.out
        got = read_device(c.out, r["A"].size, NP[r["code"]])
        exp = np.zeros_like(r["B"])
        oracle.transpose(r["A"], exp, r["size"], r["perm"], r["alpha"], 0.0,
                         complex_coef=r["cc"])
        assert oracle.first_mismatch(got, exp) == -1, k
        srv.release(t)
    # a beta != 0 request cannot use a server-allocated (uninitialised) B
    bad = _item(reqs[0], out=False)
    bad.beta_re = 1.0
    with pytest.raises(hptt_amd.HpttError):
        srv.submit(bad)
    srv.close()


def test_admission_budget_serialises(device):
    """With an in-flight budget smaller than any request, requests still all
    complete, each admitted alone (one launch per request)."""
    import torch
    rng = np.random.default_rng(5)
    reqs = _requests(rng, 12, max_elems=50000)
    torch.cuda.synchronize()
    srv = hptt_amd.Server(inflight_bytes=1, slots=3)
    tickets = [srv.submit(_item(r)) for r in reqs]
    srv.drain()
    st = srv.stats()
    assert st.completed == len(reqs) and st.launches == len(reqs)
    for t, r in zip(tickets, reqs):
        assert srv.wait(t, 5.0).state == 2
        assert oracle.first_mismatch(host(r["dB"]), _alone(r)) == -1
    srv.close()


# One request of at least 16 Mi elements per element type (shape, perm).
BIG = {0: ([256, 256, 256], [2, 0, 1]),
       1: ([64, 64, 64, 64], [3, 1, 0, 2]),
       2: ([4096, 4096], [1, 0]),
       3: ([32, 16, 32, 16, 64], [4, 2, 0, 3, 1])}


def _poisson_requests(rng, count):
    reqs = []
    big_at = set(int(v) for v in rng.choice(np.arange(4, count), size=4, replace=False))
    codes = iter(range(4))
    for k in range(count):
        if k in big_at:
            code = next(codes)
            size, perm = BIG[code]
        else:
            code = int(rng.integers(0, 4))
            n_target = float(np.exp(rng.uniform(np.log(1024), np.log(2 * 2**20))))
            size, perm = shapes.random_case(rng, max_elems=int(n_target), max_dim=128)
        cc = code >= 2 and k % 3 == 0
        al = complex(rng.uniform(-2, 2), rng.uniform(-1, 1)) if cc else float(rng.uniform(-2, 2))
        if k % 4 == 0:
            be = 0.0
        else:
            be = complex(rng.uniform(-2, 2), rng.uniform(-1, 1)) if cc else float(rng.uniform(-2, 2))
        n = int(np.prod(size))
        A = oracle.uniform(n, NP[code], 50 * k + 1, nthreads=16)
        B = oracle.uniform(n, NP[code], 50 * k + 2, nthreads=16)
        reqs.append(dict(code=code, size=size, perm=perm, alpha=al, beta=be, cc=cc, A=A, B=B,
                         dA=dev(A), dB=dev(B)))
    return reqs


def test_poisson_stream_with_large_requests(device):
    """A submitter thread issues requests at Poisson arrival times (mean gap
    0.2 ms), among them one of >= 16 Mi elements per element type, while
    earlier requests run; every result equals the oracle and the same
    transposition run alone through a plain plan."""
    import torch
    rng = np.random.default_rng(4242)
    reqs = _poisson_requests(rng, 48)
    assert sum(int(np.prod(r["size"])) >= 16 * 2**20 for r in reqs) == 4
    torch.cuda.synchronize()
    gaps = rng.exponential(0.2e-3, size=len(reqs))
    srv = hptt_amd.Server(slots=3, max_batch=64)
    tickets = [None] * len(reqs)

    def client():
        t0 = time.perf_counter()
        due = 0.0
        for k, r in enumerate(reqs):
            due += gaps[k]
            while time.perf_counter() - t0 < due:
                pass
            tickets[k] = srv.submit(_item(r))

    th = threading.Thread(target=client)
    th.start()
    th.join()
    srv.drain()
    st = srv.stats()
    assert st.completed == len(reqs) and st.failed == 0
    print("poisson serving: %d requests in %d launches" % (len(reqs), st.launches), flush=True)
    for k, r in enumerate(reqs):
        c = srv.wait(tickets[k], 10.0)
        assert c.state == 2, (k, c.status)
        got = host(r["dB"])
        exp = r["B"].copy()
        assert oracle.transpose(r["A"], exp, r["size"], r["perm"], r["alpha"], r["beta"],
                                nthreads=16, complex_coef=r["cc"]) == 0
        assert oracle.first_mismatch(got, exp) == -1, (k, r["size"], r["perm"])
        assert oracle.first_mismatch(got, _alone(r)) == -1, k
    srv.close()


def test_admission_follows_free_hbm(device):
    """With the default budget (a quarter of the HBM free at each admission
    round), an allocation made outside the server shrinks the budget the next
    admission uses by about a quarter of its size, and freeing it restores
    the budget."""
    import torch
    rng = np.random.default_rng(6)
    reqs = _requests(rng, 3, max_elems=20000)
    torch.cuda.synchronize()
    srv = hptt_amd.Server(slots=2)

    def admitted_budget(r):
        srv.wait(srv.submit(_item(r)), 10.0)
        return srv.stats().inflight_budget

    b0 = admitted_budget(reqs[0])
    hog_bytes = 64 << 30
    hog = torch.empty(hog_bytes // 4, dtype=torch.float32, device="cuda")
    b1 = admitted_budget(reqs[1])
    del hog
    torch.cuda.empty_cache()
    b2 = admitted_budget(reqs[2])
    srv.close()
    drop = b0 - b1
    assert 0.8 * hog_bytes / 4 <= drop <= 1.2 * hog_bytes / 4, (b0, b1, drop)
    assert b2 > b1 + 0.8 * hog_bytes / 4, (b1, b2)
