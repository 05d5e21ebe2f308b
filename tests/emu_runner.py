"""TEST INFRASTRUCTURE: ctypes wrapper of tests/emu/libemu.so (the host
emulator of the library's kernels) -- see tests/emu/emu.cpp."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_EMU_DIR = os.path.join(_HERE, "emu")
_LIB = os.path.join(_EMU_DIR, "libemu.so")
_lib = None

CODE = {np.dtype(np.float32): 0, np.dtype(np.float64): 1,
        np.dtype(np.complex64): 2, np.dtype(np.complex128): 3}


def build():
    subprocess.run(["make", "-s", "-C", _EMU_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.emu_transpose.restype = ctypes.c_int
        L.emu_transpose.argtypes = [
            ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
            ctypes.c_uint32, u32p, u32p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_int, u32p, u32p, u32p, ctypes.c_int,
            ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
        L.emu_plan.restype = ctypes.c_int
        L.emu_plan.argtypes = [ctypes.c_int, ctypes.c_uint32, u32p, u32p, u32p, u32p, u32p,
                               ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
        L.emu_check_fastdiv.restype = ctypes.c_int64
        L.emu_check_fastdiv.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64]
        _lib = L
    return _lib


def _u32(seq):
    if seq is None:
        return None
    return (ctypes.c_uint32 * len(seq))(*[int(v) for v in seq])


def _tune(seq):
    """Tune parameters (run_r, run_c, block_elems, wide_min_bytes[, stream]),
    passed to the emulator as five values."""
    if seq is None:
        return None
    return _u32(list(seq) + [0] * (5 - len(seq)))


def transpose(A, B, in_size, perm, alpha, beta, in_outer=None, out_outer=None,
              tune=None, complex_coef=False, force_scalar=False):
    """Emulate the library's launch on host arrays A, B (flat numpy arrays).
    Returns (status, error text, plan description)."""
    err = ctypes.create_string_buffer(512)
    desc = ctypes.create_string_buffer(2048)
    a, b = complex(alpha), complex(beta)
    rc = lib().emu_transpose(
        CODE[A.dtype], A.ctypes.data, A.size, B.ctypes.data, B.size, len(perm),
        _u32(in_size), _u32(perm), a.real, a.imag if complex_coef else 0.0,
        b.real, b.imag if complex_coef else 0.0, 1 if complex_coef else 0,
        _u32(in_outer), _u32(out_outer), _tune(tune), 1 if force_scalar else 0,
        err, 512, desc, 2048)
    return rc, err.value.decode(), desc.value.decode()


def plan(dtype, in_size, perm, in_outer=None, out_outer=None, tune=None):
    err = ctypes.create_string_buffer(512)
    desc = ctypes.create_string_buffer(2048)
    rc = lib().emu_plan(CODE[np.dtype(dtype)], len(perm), _u32(in_size), _u32(perm),
                        _u32(in_outer), _u32(out_outer), _tune(tune), err, 512, desc, 2048)
    return rc, err.value.decode(), desc.value.decode()
