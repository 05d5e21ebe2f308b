"""Helpers for GPU tests: move numpy buffers to the device, run a plan through
the C ABI, bring B back."""
from __future__ import annotations

import numpy as np

import hptt_amd


def dev(x: np.ndarray):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def host(t) -> np.ndarray:
    return t.cpu().numpy()


def run_plan(A: np.ndarray, B0: np.ndarray, size, perm, alpha, beta, io=None, oo=None,
             complex_coef=False, offset=0, tuning_timeout=0.0, stream=None):
    """Execute one plan on device copies of A and B0; returns (B, plan info).
    offset > 0 shifts both device pointers by that many elements (misaligned)."""
    import torch
    a_full = np.concatenate([np.zeros(offset, A.dtype), A]) if offset else A
    b_full = np.concatenate([np.zeros(offset, B0.dtype), B0]) if offset else B0
    dA, dB = dev(a_full), dev(b_full)
    a_view, b_view = dA[offset:], dB[offset:]
    if complex_coef:
        alpha, beta = complex(alpha), complex(beta)
    plan = hptt_amd.create_plan_checked(a_view, b_view, size, perm, alpha, beta, 0,
                                        tuning_timeout, io or (), oo or (), stream)
    plan.exec()
    plan.sync()
    torch.cuda.synchronize()
    info = plan.info()
    out = host(b_view)
    del plan
    return out, info


def read_device(addr: int, n: int, dtype) -> np.ndarray:
    """Copy n elements at a raw device address to the host (hipMemcpy)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    out = np.empty(n, dtype=dtype)
    rc = hip.hipMemcpy(out.ctypes.data, addr, out.nbytes, 2)  # device to host
    assert rc == 0, rc
    return out
