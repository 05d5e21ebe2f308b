"""The CPU oracle, pinned against checks that do not share its code.

* golden vectors written by tests/golden/gen_golden.py (numpy.transpose of
  Fortran-ordered arrays + numpy's once-rounded elementwise arithmetic);
* brute force over every index on tiny shapes (pure Python loops);
* closed forms: alpha=1, beta=0 is an exact permutation; integer-valued data
  with integer coefficients is exact; perm followed by its inverse restores
  A; the identity perm is axpby;
* the UPDATE threshold |beta| >= 1e-16 (reference hptt_trans_impl.h:75-78);
* validation returns the reference's failure cases (src/hptt.cc:25-56);
* planted mistakes (oracle built with ORACLE_MUTANT=1..5) are each caught.
"""
import ctypes
import itertools
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from golden.gen_golden import numpy_transpose

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def load_golden():
    with open(os.path.join(GOLD, "golden_small.json")) as f:
        man = json.load(f)
    data = np.load(os.path.join(GOLD, "golden_small.npz"))
    return man["cases"], data


def run_case(c, data, fn=oracle.transpose):
    A = data[c["name"] + "_A"]
    B = data[c["name"] + "_B"].copy()
    al = complex(*c["alpha"])
    be = complex(*c["beta"])
    if not c["complex_coef"]:
        al, be = al.real, be.real
    rc = fn(A, B, c["in_size"], c["perm"], al, be, c["in_outer"], c["out_outer"],
            complex_coef=c["complex_coef"])
    return rc, B


def test_golden_vectors_bit_exact(oracle_lib):
    cases, data = load_golden()
    assert len(cases) > 150
    for c in cases:
        rc, B = run_case(c, data)
        assert rc == 0, c
        ref = data[c["name"] + "_out"]
        assert oracle.first_mismatch(B, ref) == -1, c["name"]


def _brute(A, B, in_size, perm, alpha, beta):
    """Pure-Python loops over every index: B[j] with j_d = i_{perm[d]}."""
    out = B.copy()
    rank = len(in_size)
    out_size = [in_size[p] for p in perm]
    sa = [int(np.prod(in_size[:d])) for d in range(rank)]
    sb = [int(np.prod(out_size[:d])) for d in range(rank)]
    R = np.float32 if A.dtype == np.float32 else np.float64
    upd = abs(float(beta)) >= 1e-16
    for i in itertools.product(*[range(s) for s in in_size]):
        ia = sum(i[d] * sa[d] for d in range(rank))
        ib = sum(i[perm[d]] * sb[d] for d in range(rank))
        ax = R(alpha) * A[ia]
        out[ib] = (R(beta) * B[ib] + ax) if upd else ax
    return out


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("size,perm", [([3, 4], [1, 0]), ([2, 3, 4], [2, 0, 1]),
                                       ([2, 2, 3, 2], [3, 1, 0, 2]),
                                       ([2, 1, 3, 2, 2], [4, 2, 0, 1, 3]),
                                       ([2, 2, 2, 2, 2, 2], [5, 4, 3, 2, 1, 0])])
def test_brute_force_tiny(oracle_lib, dt, size, perm):
    n = int(np.prod(size))
    A = oracle.uniform(n, dt, 11)
    B = oracle.uniform(n, dt, 12)
    for al, be in [(1.0, 0.0), (2.3, 4.2), (-1.5, 1e-17)]:
        got = B.copy()
        assert oracle.transpose(A, got, size, perm, al, be) == 0
        exp = _brute(A, B, size, perm, al, be)
        assert oracle.first_mismatch(got, exp) == -1


@pytest.mark.parametrize("dt", [np.float32, np.float64, np.complex64, np.complex128])
def test_closed_forms(oracle_lib, dt):
    rng = np.random.default_rng(5)
    for rank in range(2, 7):
        size = list(rng.integers(2, 6, size=rank))
        perm = list(rng.permutation(rank))
        n = int(np.prod(size))
        A = oracle.uniform(n, dt, 3)
        # alpha = 1, beta = 0: an exact permutation (the multiset is unchanged)
        B = np.zeros(n, dtype=dt)
        assert oracle.transpose(A, B, size, perm, 1.0, 0.0) == 0
        assert np.array_equal(np.sort(B.view(np.uint8).reshape(n, -1), axis=0),
                              np.sort(A.view(np.uint8).reshape(n, -1), axis=0))
        # perm then its inverse restores A bit for bit
        inv = list(np.argsort(perm))
        out_size = [size[p] for p in perm]
        C = np.zeros(n, dtype=dt)
        assert oracle.transpose(B, C, out_size, inv, 1.0, 0.0) == 0
        assert oracle.first_mismatch(C, A) == -1
        # identity perm with beta: axpby elementwise
        Y = oracle.uniform(n, dt, 4)
        Z = Y.copy()
        assert oracle.transpose(A, Z, size, list(range(rank)), 2.0, -0.5) == 0
        R = np.float32 if dt in (np.float32, np.complex64) else np.float64
        exp = (R(-0.5) * Y.view(R) + R(2.0) * A.view(R)).view(dt)
        assert oracle.first_mismatch(Z, exp) == -1


def test_integer_valued_exact(oracle_lib):
    rng = np.random.default_rng(9)
    for dt in (np.float32, np.float64):
        size, perm = [6, 5, 4, 3], [2, 0, 3, 1]
        n = int(np.prod(size))
        A = rng.integers(-1000, 1000, n).astype(dt)
        B = rng.integers(-1000, 1000, n).astype(dt)
        got = B.copy()
        assert oracle.transpose(A, got, size, perm, 3.0, -2.0) == 0
        Af = A.reshape(size, order="F").astype(np.int64)
        Bf = B.reshape([size[p] for p in perm], order="F").astype(np.int64)
        exp = (3 * np.transpose(Af, perm) - 2 * Bf).reshape(-1, order="F").astype(dt)
        assert np.array_equal(got, exp)


def test_complex_coefficients_integer_exact(oracle_lib):
    """Complex alpha/beta on integer-valued complex data: every product and
    sum is exact, so the result must equal exact complex arithmetic
    (alpha*x + beta*y in Python integers), whatever the rounding order.  This
    fixes which components multiply and every sign; alpha = i additionally
    reduces to (-x.im, x.re)."""
    rng = np.random.default_rng(21)
    for dt in (np.complex64, np.complex128):
        size, perm = [5, 3, 4], [1, 2, 0]
        n = int(np.prod(size))
        re = lambda: rng.integers(-300, 300, n).astype(np.float64)
        A = (re() + 1j * re()).astype(dt)
        B = (re() + 1j * re()).astype(dt)
        Af = A.reshape(size, order="F")
        Bf = B.reshape([size[p] for p in perm], order="F")
        for al, be in [(complex(3, -2), complex(-1, 4)), (1j, 0.0), (complex(0, -2), 1j)]:
            got = B.copy()
            assert oracle.transpose(A, got, size, perm, al, be, complex_coef=True) == 0
            exp = np.empty(n, dtype=dt)
            x = np.transpose(Af, perm).reshape(-1, order="F")
            y = Bf.reshape(-1, order="F")
            for k in range(n):
                xr, xi = int(x[k].real), int(x[k].imag)
                yr, yi = int(y[k].real), int(y[k].imag)
                ar, ai, br, bi = int(al.real), int(al.imag), int(complex(be).real), int(complex(be).imag)
                r = (ar * xr - ai * xi, ar * xi + ai * xr)
                if abs(complex(be)) >= 1e-16:
                    r = (r[0] + br * yr - bi * yi, r[1] + br * yi + bi * yr)
                exp[k] = complex(*r)
            assert np.array_equal(got, exp), (dt, al, be)
            if al == 1j and be == 0.0:
                assert np.array_equal(got, (-x.imag + 1j * x.real).astype(dt))


def test_update_threshold(oracle_lib):
    assert oracle.update_real(1e-16)
    assert not oracle.update_real(9.999e-17)
    assert oracle.update_real(-1e-16)
    assert not oracle.update_real(0.0)
    # float beta compared after promotion to double, as std::fabs(float) >= 1e-16
    assert oracle.update_real(float(np.float32(1e-16))) == (float(np.float32(1e-16)) >= 1e-16)
    assert oracle.update_cplx(complex(0.0, 2e-16)) and not oracle.update_cplx(complex(5e-17, -5e-17))
    # below the threshold B is never read: NaN in B does not leak
    A = np.ones(6, np.float32)
    B = np.full(6, np.nan, np.float32)
    assert oracle.transpose(A, B, [2, 3], [1, 0], 2.0, 5e-17) == 0
    assert np.all(B == 2.0)


def test_validation(oracle_lib):
    A = np.zeros(256, np.float32)
    B = np.zeros(256, np.float32)
    t = oracle.transpose
    assert t(A, B, [4, 4], [1, 0], 1.0, 0.0) == 0
    assert t(A, B, [4], [0], 1.0, 0.0) == 1                   # order <= 1
    assert t(A, B, [4, 4], [1, 1], 1.0, 0.0) == 1             # not a permutation
    assert t(A, B, [4, 4], [2, 0], 1.0, 0.0) == 1             # out of range
    assert t(A, B, [4, 0], [1, 0], 1.0, 0.0) == 1             # zero size
    assert t(A, B, [4, 4], [1, 0], 1.0, 0.0, [3, 4]) == 1     # outer < size
    assert t(A, B, [4, 4], [1, 0], 1.0, 0.0, None, [4, 3]) == 1
    assert t(A, B, [2] * 7, list(range(7)), 1.0, 0.0) == 4    # rank > 6


def test_numpy_restatement_matches_brute_force():
    """The golden generator's numpy form against pure loops (tiny, padded)."""
    rng = np.random.default_rng(1)
    size, perm, io, oo = [3, 2, 4], [2, 0, 1], [4, 3, 4], [5, 3, 3]
    A = rng.uniform(-1, 1, int(np.prod(io))).astype(np.float64)
    B = rng.uniform(-1, 1, int(np.prod(oo))).astype(np.float64)
    got = numpy_transpose(A, B, size, perm, 2.5, 0.5, io, oo)
    exp = B.copy()
    out_size = [size[p] for p in perm]
    for i in itertools.product(*[range(s) for s in size]):
        ia = i[0] + io[0] * (i[1] + io[1] * i[2])
        j = [i[perm[d]] for d in range(3)]
        ib = j[0] + oo[0] * (j[1] + oo[1] * j[2])
        exp[ib] = 0.5 * B[ib] + 2.5 * A[ia]
    assert np.array_equal(got, exp)
    assert out_size == [4, 3, 2]


@pytest.mark.parametrize("mutant", [1, 2, 3, 4, 5])
def test_planted_mistakes_are_caught(mutant):
    """Each planted mistake in the oracle fails at least one golden vector."""
    d = os.path.join(os.path.dirname(HERE), "oracle")
    subprocess.run(["make", "-s", "-C", d, "mutants"], check=True)
    L = ctypes.CDLL(os.path.join(d, "liboracle_mutant%d.so" % mutant))
    L.oracle_transpose.restype = ctypes.c_int
    L.oracle_transpose.argtypes = oracle.lib().oracle_transpose.argtypes

    def fn(A, B, in_size, perm, alpha, beta, in_outer=None, out_outer=None,
           complex_coef=False):
        a, b = complex(alpha), complex(beta)
        return L.oracle_transpose(
            oracle.DTYPE_CODE[A.dtype], A.ctypes.data, B.ctypes.data, len(perm),
            oracle._u32(in_size), oracle._u32(perm), a.real, a.imag if complex_coef else 0.0,
            b.real, b.imag if complex_coef else 0.0, 1 if complex_coef else 0,
            oracle._u32(in_outer), oracle._u32(out_outer), 1)

    cases, data = load_golden()
    failures = 0
    for c in cases:
        rc, B = run_case(c, data, fn)
        if rc != 0 or oracle.first_mismatch(B, data[c["name"] + "_out"]) != -1:
            failures += 1
    assert failures > 0, "mutant %d not caught" % mutant
