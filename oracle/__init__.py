"""CPU oracle for HPTT's out-of-place transposition -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker.  The product library
(``hptt_amd``) never imports, links or calls it.

The arithmetic lives in ``hptt_oracle.c`` (built with ``-ffp-contract=off``);
this module is a thin numpy/ctypes wrapper around it.  The C file's header
lists the reference clauses it restates (file:line).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

DTYPES = {
    "s": (0, np.float32),
    "d": (1, np.float64),
    "c": (2, np.complex64),
    "z": (3, np.complex128),
}
DTYPE_CODE = {np.dtype(v[1]): v[0] for v in DTYPES.values()}

_lib = None


def build() -> None:
    """Compile liboracle.so from hptt_oracle.c (gcc, contraction off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.oracle_transpose.restype = ctypes.c_int
        L.oracle_transpose.argtypes = [
            ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
            u32p, u32p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_int, u32p, u32p, ctypes.c_int]
        for name in ("oracle_fill_uniform_f32", "oracle_fill_uniform_f64"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                          ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.oracle_first_mismatch.restype = ctypes.c_int64
        L.oracle_first_mismatch.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_uint64, ctypes.c_uint32]
        L.oracle_update_real.restype = ctypes.c_int
        L.oracle_update_real.argtypes = [ctypes.c_double]
        L.oracle_update_cplx.restype = ctypes.c_int
        L.oracle_update_cplx.argtypes = [ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


def _u32(seq):
    if seq is None:
        return None
    arr = (ctypes.c_uint32 * len(seq))(*[int(v) for v in seq])
    return arr


def out_shape(in_size, perm):
    """out_size[d] = in_size[perm[d]] (hptt_trans_gen.py:196-197)."""
    return [int(in_size[p]) for p in perm]


def storage_len(size, outer=None):
    """Number of elements a column-major tensor with these (outer) sizes spans."""
    dims = outer if outer is not None else size
    n = 1
    for v in dims:
        n *= int(v)
    return n


def transpose(A: np.ndarray, B: np.ndarray, in_size, perm, alpha, beta,
              in_outer=None, out_outer=None, nthreads: int = 1,
              complex_coef: bool = False) -> int:
    """Run the oracle in place on B.  A and B are flat (1-D) numpy arrays of
    the same dtype holding column-major data.  Returns the C status code
    (0 = ok, 1 = invalid argument, 4 = unsupported rank)."""
    assert A.dtype == B.dtype
    assert A.flags.c_contiguous and B.flags.c_contiguous
    code = DTYPE_CODE[A.dtype]
    rank = len(perm)
    ar, ai = (complex(alpha).real, complex(alpha).imag)
    br, bi = (complex(beta).real, complex(beta).imag)
    if not complex_coef:
        ai = bi = 0.0
    if rank == len(in_size) and rank >= 1:
        need_a = storage_len(in_size, in_outer)
        need_b = storage_len(out_shape(in_size, perm) if all(
            p < len(in_size) for p in perm) else in_size, out_outer)
        if A.size < need_a or B.size < need_b:
            raise ValueError("buffer too small for the described tensor")
    return lib().oracle_transpose(
        code, A.ctypes.data, B.ctypes.data, rank, _u32(in_size), _u32(perm),
        ar, ai, br, bi, 1 if complex_coef else 0, _u32(in_outer),
        _u32(out_outer), nthreads)


def uniform(n: int, dtype, seed: int, lo=-1.0, hi=1.0, nthreads=8) -> np.ndarray:
    """Deterministic uniform[lo,hi) data; complex dtypes draw re/im as an
    interleaved scalar stream."""
    dt = np.dtype(dtype)
    if dt in (np.dtype(np.float32), np.dtype(np.complex64)):
        base, fn = np.float32, lib().oracle_fill_uniform_f32
    else:
        base, fn = np.float64, lib().oracle_fill_uniform_f64
    ncomp = 2 if dt.kind == "c" else 1
    buf = np.empty(n * ncomp, dtype=base)
    fn(buf.ctypes.data, buf.size, seed, lo, hi, nthreads)
    return buf.view(dt)


def update_real(beta: float) -> bool:
    return bool(lib().oracle_update_real(float(beta)))


def update_cplx(beta: complex) -> bool:
    b = complex(beta)
    return bool(lib().oracle_update_cplx(b.real, b.imag))


def first_mismatch(x: np.ndarray, y: np.ndarray) -> int:
    """Bitwise comparison; index of the first differing element or -1."""
    assert x.dtype == y.dtype and x.size == y.size
    return int(lib().oracle_first_mismatch(x.ctypes.data, y.ctypes.data,
                                           x.size, x.dtype.itemsize))
