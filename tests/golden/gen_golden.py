"""Generate tests/golden/golden_small.npz -- small input/output vectors for the
transposition B[perm(i)] = alpha*A[i] + beta*B[perm(i)].

The reference commits no golden vectors and its tests do not build (its unit
tests target a removed API and its benchmark's check runs after a reset, so it
compares a buffer with itself: test/benchmark/inc/hptt/perf-test/
test_perf_util.tcc:40-48).  These vectors are therefore computed by an
independent restatement that shares no code with the C oracle:

  * index map: numpy.transpose(A.reshape(in_size, order="F"), axes=perm),
    numpy's own permutation of a Fortran-ordered array (result axis d is input
    axis perm[d], the reference's j_d = i_{perm[d]},
    inc/hptt/operations/operation_trans.tcc:47-51), written into B viewed
    Fortran-ordered over its outer sizes;
  * arithmetic: numpy elementwise float32/float64 products and sums, each
    rounded once (numpy never fuses them), applied to real and imaginary parts
    separately for real coefficients; complex coefficients follow the fixed
    order of the C ABI (fl(fl(ar*xr) - fl(ai*xi)), ...);
  * UPDATE iff |beta| >= 1e-16 (inc/hptt/impl/hptt_trans_impl.h:75-78); when
    not updating, B is never read (a NaN planted in B must not leak).

Run:  python tests/golden/gen_golden.py   (rewrites the .npz and .json)
"""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_NPZ = os.path.join(HERE, "golden_small.npz")
OUT_JSON = os.path.join(HERE, "golden_small.json")

REAL = {np.dtype(np.float32): np.float32, np.dtype(np.float64): np.float64,
        np.dtype(np.complex64): np.float32, np.dtype(np.complex128): np.float64}


def _components(x: np.ndarray) -> np.ndarray:
    """(…, ncomp) real view."""
    if x.dtype.kind == "c":
        return x.view(REAL[x.dtype]).reshape(x.shape + (2,))
    return x.reshape(x.shape + (1,))


def numpy_transpose(A, B, in_size, perm, alpha, beta, in_outer=None, out_outer=None,
                    complex_coef=False):
    """Return a new B after the operation (A, B flat column-major buffers)."""
    R = REAL[A.dtype]
    in_size = [int(v) for v in in_size]
    out_size = [in_size[p] for p in perm]
    in_outer = list(in_outer) if in_outer is not None else in_size
    out_outer = list(out_outer) if out_outer is not None else out_size
    a_full = A[: int(np.prod(in_outer))].reshape(in_outer, order="F")
    a_view = a_full[tuple(slice(0, s) for s in in_size)]
    x = np.transpose(a_view, axes=perm)            # shape out_size
    Bn = B.copy()
    b_full = Bn[: int(np.prod(out_outer))].reshape(out_outer, order="F")
    b_view = b_full[tuple(slice(0, s) for s in out_size)]
    al, be = complex(alpha), complex(beta)
    if complex_coef:
        update = abs(be.real) >= 1e-16 or abs(be.imag) >= 1e-16
    else:
        update = abs(be.real) >= 1e-16
    xc = _components(np.ascontiguousarray(x))
    res = np.empty_like(xc)
    if not complex_coef:
        a = R(al.real)
        ax = (a * xc).astype(R)
        if update:
            yc = _components(np.ascontiguousarray(b_view))
            b = R(be.real)
            by = (b * yc).astype(R)
            res[...] = (by + ax).astype(R)
        else:
            res[...] = ax
    else:
        ar, ai, br, bi = R(al.real), R(al.imag), R(be.real), R(be.imag)
        xr, xi = xc[..., 0], xc[..., 1]
        axr = (ar * xr) - (ai * xi)
        axi = (ar * xi) + (ai * xr)
        if update:
            yc = _components(np.ascontiguousarray(b_view))
            yr, yi = yc[..., 0], yc[..., 1]
            byr = (br * yr) - (bi * yi)
            byi = (br * yi) + (bi * yr)
            res[..., 0] = byr + axr
            res[..., 1] = byi + axi
        else:
            res[..., 0] = axr
            res[..., 1] = axi
    if A.dtype.kind == "c":
        b_view[...] = res.reshape(-1).view(A.dtype).reshape(out_size)
    else:
        b_view[...] = res[..., 0]
    return Bn


def _data(rng, n, dtype, kind):
    R = REAL[np.dtype(dtype)]
    ncomp = 2 if np.dtype(dtype).kind == "c" else 1
    if kind == "int":
        v = rng.integers(-40, 41, size=n * ncomp).astype(R)
    else:
        v = rng.uniform(-1.0, 1.0, size=n * ncomp).astype(R)
    return v.view(np.dtype(dtype)) if ncomp == 2 else v


def cases():
    """(name, dtype, in_size, perm, alpha, beta, complex_coef, in_outer, out_outer, data)"""
    out = []
    shapes = [
        ([7, 5], [1, 0]),
        ([2, 3], [1, 0]),
        ([9, 4, 6], [2, 1, 0]),
        ([5, 8, 3], [0, 2, 1]),
        ([4, 4, 9], [1, 0, 2]),
        ([3, 5, 2, 4], [3, 2, 1, 0]),
        ([6, 3, 4, 5], [0, 3, 2, 1]),
        ([4, 6, 3, 5], [2, 0, 3, 1]),
        ([3, 4, 2, 3, 2], [4, 1, 3, 0, 2]),
        ([2, 3, 2, 2, 3, 2], [5, 2, 0, 4, 1, 3]),
        ([3, 1, 4, 2], [2, 0, 3, 1]),      # size-1 dim
        ([5, 4, 3], [0, 1, 2]),            # identity (reference crashes, D1)
        ([2, 3, 4, 5], [1, 2, 3, 0]),
        ([33, 17, 5], [1, 2, 0]),          # odd extents, edge tiles
    ]
    coefs = [(1.0, 0.0, False), (2.3, 4.2, False), (-0.75, 1e-17, False),
             (2.0, -3.0, False)]
    cplx = [(complex(1.1, 0.3), complex(0.7, -0.2)), (complex(0.5, -2.0), 0j)]
    k = 0
    for dt in (np.float32, np.float64, np.complex64, np.complex128):
        for si, (size, perm) in enumerate(shapes):
            for ci, (al, be, _) in enumerate(coefs):
                if (si + ci) % 2 and ci != 0:
                    continue
                kind = "int" if ci == 3 else "rand"
                out.append(("c%03d" % k, dt, size, perm, al, be, False, None, None, kind))
                k += 1
            if np.dtype(dt).kind == "c":
                for al, be in cplx:
                    out.append(("c%03d" % k, dt, size, perm, al, be, True, None, None, "rand"))
                    k += 1
    # padded (outer-size) cases
    for dt in (np.float32, np.complex128):
        out.append(("c%03d" % k, dt, [5, 4, 3], [2, 0, 1], 2.3, 4.2, False,
                    [7, 4, 5], [4, 6, 6], "rand"))
        k += 1
        out.append(("c%03d" % k, dt, [6, 5], [1, 0], 1.0, 0.0, False, [8, 5], None, "rand"))
        k += 1
        out.append(("c%03d" % k, dt, [4, 3, 5], [0, 2, 1], 1.5, -0.5, False,
                    [4, 4, 5], [6, 5, 3], "rand"))
        k += 1
    return out


def main():
    rng = np.random.default_rng(20261203)
    arrays = {}
    manifest = []
    for (name, dt, size, perm, al, be, cc, io, oo, kind) in cases():
        out_size = [size[p] for p in perm]
        na = int(np.prod(io if io else size))
        nb = int(np.prod(oo if oo else out_size))
        A = _data(rng, na, dt, kind)
        B = _data(rng, nb, dt, kind)
        upd = (abs(complex(be).real) >= 1e-16 or (cc and abs(complex(be).imag) >= 1e-16))
        if not upd:
            # B must not be read when beta is below threshold: plant NaNs
            B = B.copy()
            B.view(REAL[np.dtype(dt)])[::3] = np.nan
        Bout = numpy_transpose(A, B, size, perm, al, be, io, oo, cc)
        arrays[name + "_A"] = A
        arrays[name + "_B"] = B
        arrays[name + "_out"] = Bout
        manifest.append({
            "name": name, "dtype": np.dtype(dt).name, "in_size": size, "perm": perm,
            "alpha": [complex(al).real, complex(al).imag],
            "beta": [complex(be).real, complex(be).imag], "complex_coef": cc,
            "in_outer": io, "out_outer": oo, "data": kind})
    np.savez_compressed(OUT_NPZ, **arrays)
    with open(OUT_JSON, "w") as f:
        json.dump({"generator": "tests/golden/gen_golden.py", "seed": 20261203,
                   "cases": manifest}, f, indent=1)
    print("wrote %d cases to %s" % (len(manifest), OUT_NPZ))


if __name__ == "__main__":
    main()
