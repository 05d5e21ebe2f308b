"""hptt_amd -- MI355X (gfx950) tensor transposition with HPTT's plan interface.

    B[perm(i)] = alpha * A[i] + beta * B[perm(i)]

Python mirror of the reference's C++ interface (inc/hptt/hptt.h:85-93,
inc/hptt/impl/hptt_trans.h:12-36) over the C ABI in include/hptt_amd.h::

    plan = hptt_amd.create_plan(A, B, in_size, perm, alpha, beta, num_threads)
    plan.exec()                 # enqueued on the plan's HIP stream
    plan.reset_data(A2, B2)     # same sizes / perm / coefficients

A and B are device buffers: torch tensors on a ROCm device, or raw device
addresses (int) together with ``dtype=``.  Layout is column-major (dim 0 is
stride-1), as in the reference.  ``create_plan`` returns None on invalid
arguments, like the reference's nullptr (inc/hptt/hptt.h:71-83);
``last_error()`` says why.

The library (hptt_amd/libhptt_amd.so) is required: there is no fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# HPTTX_LIB selects another build of the library (e.g. one compiled with a
# different streaming policy); the default is the in-tree build.
LIB_PATH = os.environ.get("HPTTX_LIB") or os.path.join(_HERE, "libhptt_amd.so")

OK, EINVAL, ENOMEM, EHIP, EUNSUPPORTED = 0, 1, 2, 3, 4
F32, F64, C64, C128 = 0, 1, 2, 3
ELEM_BYTES = {F32: 4, F64: 8, C64: 8, C128: 16}
PATHS = {0: "tile", 1: "tile-shared", 2: "wide-copy", 3: "axpby", 4: "gather"}
BATCH_ONE_LAUNCH = 1

_lib = None


class HpttError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__("hptt_amd status %d: %s" % (status, msg))
        self.status = status


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("elements", ctypes.c_uint64), ("elem_bytes", ctypes.c_uint32),
        ("update", ctypes.c_int32), ("path", ctypes.c_int32),
        ("fused_rank", ctypes.c_int32), ("algorithmic_bytes", ctypes.c_uint64),
        ("grid", ctypes.c_uint32), ("lds_bytes", ctypes.c_uint32),
        ("tile_r", ctypes.c_uint32), ("tile_c", ctypes.c_uint32),
        ("tile_o", ctypes.c_uint32), ("vec_r", ctypes.c_uint32),
        ("vec_w", ctypes.c_uint32), ("unaligned_r", ctypes.c_uint32),
        ("unaligned_w", ctypes.c_uint32), ("threads", ctypes.c_uint32)]


def lib() -> ctypes.CDLL:
    """Load libhptt_amd.so (built by ``python -m hptt_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("hptt_amd: %s is missing -- build it with "
                           "`python -m hptt_amd.build`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u32p = ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)
    pp = ctypes.POINTER(ctypes.c_void_p)
    f, d = ctypes.c_float, ctypes.c_double
    for name, real in (("s", f), ("d", d), ("c", f), ("z", d)):
        fn = getattr(L, "hpttx_plan_create_" + name)
        fn.restype = ctypes.c_int
        fn.argtypes = [vp, vp, u32, u32p, u32p, real, real, u32, d, u32p, u32p, vp, pp]
    for name, real in (("c", f), ("z", d)):
        fn = getattr(L, "hpttx_plan_create_%s_cplx" % name)
        fn.restype = ctypes.c_int
        fn.argtypes = [vp, vp, u32, u32p, u32p, real, real, real, real, u32, d, u32p,
                       u32p, vp, pp]
    for name in ("hpttx_plan_exec", "hpttx_plan_sync"):
        getattr(L, name).restype = ctypes.c_int
        getattr(L, name).argtypes = [vp]
    L.hpttx_plan_exec_on.restype = ctypes.c_int
    L.hpttx_plan_exec_on.argtypes = [vp, vp]
    L.hpttx_plan_reset_data.restype = ctypes.c_int
    L.hpttx_plan_reset_data.argtypes = [vp, vp, vp]
    L.hpttx_plan_set_stream.restype = ctypes.c_int
    L.hpttx_plan_set_stream.argtypes = [vp, vp]
    L.hpttx_plan_describe.restype = ctypes.c_int
    L.hpttx_plan_describe.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    L.hpttx_plan_get_info.restype = ctypes.c_int
    L.hpttx_plan_get_info.argtypes = [vp, ctypes.POINTER(PlanInfo)]
    L.hpttx_plan_destroy.restype = None
    L.hpttx_plan_destroy.argtypes = [vp]
    L.hpttx_last_error.restype = ctypes.c_char_p
    L.hpttx_last_error.argtypes = []
    L.hpttx_version.restype = ctypes.c_int
    L.hpttx_version.argtypes = []
    _bind_batch(L)
    _bind_server(L)
    _lib = L
    return L


def last_error() -> str:
    return lib().hpttx_last_error().decode()


def _check(rc: int) -> None:
    if rc != OK:
        raise HpttError(rc, last_error())


def _u32(seq):
    if seq is None or len(seq) == 0:
        return None
    return (ctypes.c_uint32 * len(seq))(*[int(v) for v in seq])


def _torch_dtype_code(t):
    import torch
    return {torch.float32: F32, torch.float64: F64, torch.complex64: C64,
            torch.complex128: C128}.get(t.dtype)


def _addr_and_dtype(x, dtype):
    """Device address and dtype code of a torch tensor or a raw address."""
    if isinstance(x, int):
        if dtype is None:
            raise TypeError("dtype= is required when passing raw device addresses")
        return x, dtype
    code = _torch_dtype_code(x)
    if code is None:
        raise TypeError("unsupported tensor dtype %s" % (x.dtype,))
    if not x.is_cuda:
        raise TypeError("hptt_amd plans take device tensors")
    return int(x.data_ptr()), code


def dtype_code(dtype) -> int:
    """Map 's'/'d'/'c'/'z', numpy or torch dtypes to a dtype code."""
    if isinstance(dtype, int):
        return dtype
    if isinstance(dtype, str) and dtype in "sdcz" and len(dtype) == 1:
        return "sdcz".index(dtype)
    name = str(dtype).replace("torch.", "")
    return {"float32": F32, "float64": F64, "complex64": C64, "complex128": C128}[name]


class Plan:
    """A transposition plan (reference CGraphTransPackBase<T>,
    inc/hptt/impl/hptt_trans.h:12-36).  exec() enqueues and returns."""

    def __init__(self, handle: int, dtype: int, keep=()):
        self._h = ctypes.c_void_p(handle)
        self.dtype = dtype
        self._keep = list(keep)

    @property
    def handle(self) -> int:
        return self._h.value

    def exec(self) -> None:
        _check(lib().hpttx_plan_exec(self._h))

    __call__ = exec

    def exec_on(self, stream) -> None:
        _check(lib().hpttx_plan_exec_on(self._h, _stream_ptr(stream)))

    def sync(self) -> None:
        _check(lib().hpttx_plan_sync(self._h))

    def reset_data(self, in_data, out_data, dtype=None) -> None:
        a, _ = _addr_and_dtype(in_data, dtype if dtype is not None else self.dtype)
        b, _ = _addr_and_dtype(out_data, dtype if dtype is not None else self.dtype)
        self._keep = [in_data, out_data]
        _check(lib().hpttx_plan_reset_data(self._h, a, b))

    def set_stream(self, stream) -> None:
        _check(lib().hpttx_plan_set_stream(self._h, _stream_ptr(stream)))

    def describe(self) -> str:
        n = lib().hpttx_plan_describe(self._h, None, 0)
        buf = ctypes.create_string_buffer(max(n, 1))
        lib().hpttx_plan_describe(self._h, buf, n)
        return buf.value.decode()

    def print_plan(self) -> None:
        print(self.describe(), end="")

    def info(self) -> PlanInfo:
        inf = PlanInfo()
        _check(lib().hpttx_plan_get_info(self._h, ctypes.byref(inf)))
        return inf

    # The reference pins OpenMP thread ids (cgraph_trans.tcc:71-89); a GPU plan
    # has no host threads to pin, so these are accepted no-ops.
    def set_thread_ids(self, thread_ids) -> None:
        return None

    def unset_thread_ids(self) -> None:
        return None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.hpttx_plan_destroy(h)
            self._h = ctypes.c_void_p(0)


def _stream_ptr(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)  # torch.cuda.Stream


def _check_shape_args(in_size, perm, in_outer_size, out_outer_size):
    """Array lengths the C ABI relies on (it reads rank entries of each):
    in_size as long as perm, each outer size empty or as long as perm
    (the reference returns nullptr on a mismatch, src/hptt.cc:35-39)."""
    rank = len(perm)
    if len(in_size) != rank:
        raise HpttError(EINVAL, "in_size has %d entries but perm has %d" % (len(in_size), rank))
    for name, o in (("in_outer_size", in_outer_size), ("out_outer_size", out_outer_size)):
        if o is not None and len(o) not in (0, rank):
            raise HpttError(EINVAL, "%s has %d entries; expected 0 or %d" % (name, len(o), rank))


def _spans(in_size, perm, in_outer_size, out_outer_size):
    """Elements A and B must hold (1 + sum (extent-1)*stride, column-major
    over the outer sizes), or None when the shape is invalid (the C ABI then
    says why)."""
    rank = len(perm)
    try:
        size = [int(v) for v in in_size]
        pm = [int(v) for v in perm]
    except (TypeError, ValueError):
        return None
    if sorted(pm) != list(range(rank)) or any(v < 1 for v in size):
        return None
    out_size = [size[p] for p in pm]
    io = [int(v) for v in in_outer_size] if in_outer_size else size
    oo = [int(v) for v in out_outer_size] if out_outer_size else out_size
    span = []
    for ext, outer in ((size, io), (out_size, oo)):
        s, stride = 1, 1
        for d in range(rank):
            s += (ext[d] - 1) * stride
            stride *= outer[d]
        span.append(s)
    return span


def _check_tensor_spans(in_data, out_data, in_size, perm, in_outer_size, out_outer_size):
    sp = _spans(in_size, perm, in_outer_size, out_outer_size)
    if sp is None:
        return
    for name, t, need in (("A", in_data, sp[0]), ("B", out_data, sp[1])):
        if t is not None and not isinstance(t, int) and t.numel() < need:
            raise HpttError(EINVAL, "%s holds %d elements; the shape needs %d" % (
                name, t.numel(), need))


def create_plan_checked(in_data, out_data, in_size, perm, alpha, beta,
                        num_threads: int = 0, tuning_timeout: float = 0.0,
                        in_outer_size=(), out_outer_size=(), stream=None,
                        dtype=None) -> Plan:
    """create_plan that raises HpttError (with the status code) on failure."""
    code = dtype_code(dtype) if dtype is not None else None
    a, code_a = _addr_and_dtype(in_data, code)
    b, code_b = _addr_and_dtype(out_data, code)
    if code_a != code_b:
        raise TypeError("A and B must have the same dtype")
    code = code_a
    _check_shape_args(in_size, perm, in_outer_size, out_outer_size)
    _check_tensor_spans(in_data, out_data, in_size, perm, in_outer_size, out_outer_size)
    rank = len(perm)
    out = ctypes.c_void_p()
    L = lib()
    args_tail = (_u32(in_outer_size), _u32(out_outer_size), _stream_ptr(stream),
                 ctypes.byref(out))
    cplx_coef = isinstance(alpha, complex) or isinstance(beta, complex)
    if cplx_coef and code in (C64, C128):
        al, be = complex(alpha), complex(beta)
        fn = L.hpttx_plan_create_c_cplx if code == C64 else L.hpttx_plan_create_z_cplx
        rc = fn(a, b, rank, _u32(in_size), _u32(perm), al.real, al.imag, be.real, be.imag,
                int(num_threads), float(tuning_timeout), *args_tail)
    else:
        if cplx_coef:
            raise TypeError("complex coefficients need a complex dtype")
        fn = getattr(L, "hpttx_plan_create_" + "sdcz"[code])
        rc = fn(a, b, rank, _u32(in_size), _u32(perm), float(alpha), float(beta),
                int(num_threads), float(tuning_timeout), *args_tail)
    _check(rc)
    return Plan(out.value, code, keep=(in_data, out_data))


def create_plan(in_data, out_data, in_size, perm, alpha, beta,
                num_threads: int = 0, tuning_timeout: float = 0.0,
                in_outer_size=(), out_outer_size=(), stream=None, dtype=None):
    """Reference hptt::create_plan (inc/hptt/hptt.h:85-93): a Plan, or None
    when the arguments are invalid (last_error() gives the reason)."""
    try:
        return create_plan_checked(in_data, out_data, in_size, perm, alpha, beta,
                                   num_threads, tuning_timeout, in_outer_size,
                                   out_outer_size, stream, dtype)
    except HpttError:
        return None


class BatchItem(ctypes.Structure):
    """include/hptt_amd.h hpttx_batch_item."""
    _fields_ = [
        ("dtype", ctypes.c_int32), ("rank", ctypes.c_uint32),
        ("in_", ctypes.c_void_p), ("out", ctypes.c_void_p),
        ("in_size", ctypes.c_uint32 * 6), ("perm", ctypes.c_uint32 * 6),
        ("in_outer", ctypes.c_uint32 * 6), ("out_outer", ctypes.c_uint32 * 6),
        ("alpha_re", ctypes.c_double), ("alpha_im", ctypes.c_double),
        ("beta_re", ctypes.c_double), ("beta_im", ctypes.c_double),
        ("complex_coef", ctypes.c_int32), ("reserved_", ctypes.c_int32)]


class BatchItemInfo(ctypes.Structure):
    """include/hptt_amd.h hpttx_batch_item_info."""
    _fields_ = [("group", ctypes.c_uint32), ("kernel", ctypes.c_uint32),
                ("variant", ctypes.c_uint32), ("lds_bytes", ctypes.c_uint32),
                ("tiles", ctypes.c_uint32), ("reserved_", ctypes.c_uint32)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint32), ("groups", ctypes.c_uint32),
                ("algorithmic_bytes", ctypes.c_uint64), ("tiles", ctypes.c_uint64),
                ("grid", ctypes.c_uint32 * 4), ("lds_bytes", ctypes.c_uint32 * 4)]


def _bind_batch(L):
    vp, u32, d = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double
    L.hpttx_batch_create.restype = ctypes.c_int
    L.hpttx_batch_create.argtypes = [u32, ctypes.POINTER(BatchItem), d, vp,
                                     ctypes.POINTER(ctypes.c_void_p)]
    L.hpttx_batch_create_ex.restype = ctypes.c_int
    L.hpttx_batch_create_ex.argtypes = [u32, ctypes.POINTER(BatchItem), d, vp, u32,
                                        ctypes.POINTER(ctypes.c_void_p)]
    for name in ("hpttx_batch_exec", "hpttx_batch_sync"):
        getattr(L, name).restype = ctypes.c_int
        getattr(L, name).argtypes = [vp]
    L.hpttx_batch_exec_on.restype = ctypes.c_int
    L.hpttx_batch_exec_on.argtypes = [vp, vp]
    L.hpttx_batch_set_data.restype = ctypes.c_int
    L.hpttx_batch_set_data.argtypes = [vp, u32, vp, vp]
    L.hpttx_batch_use_graph.restype = ctypes.c_int
    L.hpttx_batch_use_graph.argtypes = [vp, ctypes.c_int]
    L.hpttx_batch_get_info.restype = ctypes.c_int
    L.hpttx_batch_get_info.argtypes = [vp, ctypes.POINTER(BatchInfo)]
    L.hpttx_batch_get_item_info.restype = ctypes.c_int
    L.hpttx_batch_get_item_info.argtypes = [vp, u32, ctypes.POINTER(BatchItemInfo)]
    L.hpttx_batch_item_path.restype = ctypes.c_int
    L.hpttx_batch_item_path.argtypes = [vp, u32, ctypes.c_char_p, ctypes.c_size_t]
    L.hpttx_batch_destroy.restype = None
    L.hpttx_batch_destroy.argtypes = [vp]
    L.hpttx_batch_run_noalloc.restype = ctypes.c_int
    L.hpttx_batch_run_noalloc.argtypes = [u32, ctypes.POINTER(BatchItem), vp]


def make_batch_item(in_data, out_data, in_size, perm, alpha, beta, dtype=None,
                    in_outer_size=(), out_outer_size=()) -> BatchItem:
    """One hpttx_batch_item; in/out are device tensors or raw addresses.
    out_data=None (server requests only): the server allocates B."""
    code = dtype_code(dtype) if dtype is not None else None
    a, ca = _addr_and_dtype(in_data, code)
    if out_data is None:
        b, cb = 0, ca
    else:
        b, cb = _addr_and_dtype(out_data, code)
    if ca != cb:
        raise TypeError("A and B must have the same dtype")
    _check_shape_args(in_size, perm, in_outer_size, out_outer_size)
    if len(perm) > 6:
        raise HpttError(EUNSUPPORTED, "rank above 6 is not supported")
    _check_tensor_spans(in_data, out_data, in_size, perm, in_outer_size, out_outer_size)
    it = BatchItem()
    it.dtype = ca
    it.rank = len(perm)
    it.in_ = a
    it.out = b
    for k, v in enumerate(in_size):
        it.in_size[k] = int(v)
    for k, v in enumerate(perm):
        it.perm[k] = int(v)
    for k, v in enumerate(in_outer_size or ()):
        it.in_outer[k] = int(v)
    for k, v in enumerate(out_outer_size or ()):
        it.out_outer[k] = int(v)
    cc = isinstance(alpha, complex) or isinstance(beta, complex)
    al, be = complex(alpha), complex(beta)
    it.alpha_re, it.alpha_im = al.real, al.imag
    it.beta_re, it.beta_im = be.real, be.imag
    it.complex_coef = 1 if cc else 0
    return it


class Batch:
    """Many transpositions executed by work-queue launches over one tile
    queue (one launch per element type, kernel variant and LDS class, forked
    over side streams; or, with one_launch, one persistent launch per element
    type over a dynamic queue), replayed from a captured hipGraph
    (hpttx_batch_* in include/hptt_amd.h).  items: a sequence of BatchItem."""

    def __init__(self, items, tuning_timeout: float = 0.0, stream=None, keep=(),
                 one_launch: bool = False):
        L = lib()
        arr = (BatchItem * len(items))(*items)
        h = ctypes.c_void_p()
        flags = BATCH_ONE_LAUNCH if one_launch else 0
        _check(L.hpttx_batch_create_ex(len(items), arr, float(tuning_timeout),
                                       _stream_ptr(stream), flags, ctypes.byref(h)))
        self._h = h
        self.count = len(items)
        self._keep = list(keep)

    def exec(self) -> None:
        _check(lib().hpttx_batch_exec(self._h))

    __call__ = exec

    def exec_on(self, stream) -> None:
        _check(lib().hpttx_batch_exec_on(self._h, _stream_ptr(stream)))

    def sync(self) -> None:
        _check(lib().hpttx_batch_sync(self._h))

    def use_graph(self, enable: bool) -> None:
        _check(lib().hpttx_batch_use_graph(self._h, 1 if enable else 0))

    def set_data(self, index: int, in_data, out_data, dtype=None) -> None:
        a, _ = _addr_and_dtype(in_data, dtype)
        b, _ = _addr_and_dtype(out_data, dtype)
        _check(lib().hpttx_batch_set_data(self._h, index, a, b))

    def info(self) -> BatchInfo:
        inf = BatchInfo()
        _check(lib().hpttx_batch_get_info(self._h, ctypes.byref(inf)))
        return inf

    def item_info(self, index: int) -> BatchItemInfo:
        inf = BatchItemInfo()
        _check(lib().hpttx_batch_get_item_info(self._h, index, ctypes.byref(inf)))
        return inf

    def item_path(self, index: int) -> str:
        buf = ctypes.create_string_buffer(32)
        _check(lib().hpttx_batch_item_path(self._h, index, buf, 32))
        return buf.value.decode()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.hpttx_batch_destroy(h)
            self._h = ctypes.c_void_p(0)


def batch_run_noalloc(items, stream=None) -> None:
    """Enqueue the transpositions of `items` (BatchItems) on `stream` without
    allocating device memory: their tables go to the library's static device
    scratch (hpttx_batch_run_noalloc in include/hptt_amd.h).  Calls whose
    work overlaps on one GPU race on that scratch."""
    arr = (BatchItem * len(items))(*items)
    _check(lib().hpttx_batch_run_noalloc(len(items), arr, _stream_ptr(stream)))


class ServerConfigC(ctypes.Structure):
    _fields_ = [("inflight_bytes", ctypes.c_uint64), ("arena_bytes", ctypes.c_uint64),
                ("slots", ctypes.c_uint32), ("max_batch", ctypes.c_uint32)]


class Completion(ctypes.Structure):
    """include/hptt_amd.h hpttx_completion (state 2 = done, 3 = failed)."""
    _fields_ = [("ticket", ctypes.c_uint64), ("state", ctypes.c_int32),
                ("status", ctypes.c_int32), ("out", ctypes.c_void_p),
                ("t_submit", ctypes.c_double), ("t_launch", ctypes.c_double),
                ("t_done", ctypes.c_double), ("bytes", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("failed", ctypes.c_uint64), ("launches", ctypes.c_uint64),
                ("bytes_done", ctypes.c_uint64), ("first_submit", ctypes.c_double),
                ("last_done", ctypes.c_double), ("p50_latency", ctypes.c_double),
                ("p99_latency", ctypes.c_double), ("mean_latency", ctypes.c_double),
                ("inflight_budget", ctypes.c_uint64)]


def _bind_server(L):
    vp, u32, u64, d = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double
    L.hpttx_server_create.restype = ctypes.c_int
    L.hpttx_server_create.argtypes = [ctypes.POINTER(ServerConfigC), ctypes.POINTER(vp)]
    L.hpttx_server_submit.restype = ctypes.c_int
    L.hpttx_server_submit.argtypes = [vp, ctypes.POINTER(BatchItem), ctypes.POINTER(u64)]
    L.hpttx_server_submit_many.restype = ctypes.c_int
    L.hpttx_server_submit_many.argtypes = [vp, u32, ctypes.POINTER(BatchItem),
                                           ctypes.POINTER(u64), ctypes.POINTER(u32)]
    L.hpttx_server_poll.restype = ctypes.c_int
    L.hpttx_server_poll.argtypes = [vp, ctypes.POINTER(Completion), u32, ctypes.POINTER(u32)]
    L.hpttx_server_wait.restype = ctypes.c_int
    L.hpttx_server_wait.argtypes = [vp, u64, d, ctypes.POINTER(Completion)]
    L.hpttx_server_drain.restype = ctypes.c_int
    L.hpttx_server_drain.argtypes = [vp]
    L.hpttx_server_release.restype = ctypes.c_int
    L.hpttx_server_release.argtypes = [vp, u64]
    L.hpttx_server_get_stats.restype = ctypes.c_int
    L.hpttx_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    L.hpttx_server_now.restype = d
    L.hpttx_server_now.argtypes = [vp]
    L.hpttx_server_destroy.restype = None
    L.hpttx_server_destroy.argtypes = [vp]


class Server:
    """Continuous-batching transposition server (hpttx_server_* in
    include/hptt_amd.h).  Requests are BatchItems; an item whose `out` is 0
    gets its B allocated in the server's arena (beta must be 0)."""

    def __init__(self, inflight_bytes: int = 0, arena_bytes: int = 0, slots: int = 3,
                 max_batch: int = 256):
        L = lib()
        cfg = ServerConfigC(int(inflight_bytes), int(arena_bytes), int(slots), int(max_batch))
        h = ctypes.c_void_p()
        _check(L.hpttx_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def submit(self, item: BatchItem) -> int:
        t = ctypes.c_uint64()
        _check(lib().hpttx_server_submit(self._h, ctypes.byref(item), ctypes.byref(t)))
        return t.value

    def submit_many(self, items) -> list:
        """Submit a sequence of BatchItems, or a ctypes array of them (built
        once and reused), in one call (planned and queued in chunks of
        doubling size, so the first run while the rest are planned);
        returns their tickets."""
        n = len(items)
        arr = items if isinstance(items, ctypes.Array) else (BatchItem * n)(*items)
        tickets = (ctypes.c_uint64 * max(1, n))()
        queued = ctypes.c_uint32()
        _check(lib().hpttx_server_submit_many(self._h, n, arr, tickets, ctypes.byref(queued)))
        return [tickets[i] for i in range(n)]

    def poll(self, max_n: int = 256):
        buf = (Completion * max_n)()
        n = ctypes.c_uint32()
        _check(lib().hpttx_server_poll(self._h, buf, max_n, ctypes.byref(n)))
        return [buf[i] for i in range(n.value)]

    def wait(self, ticket: int, timeout: float = -1.0) -> Completion:
        c = Completion()
        _check(lib().hpttx_server_wait(self._h, ticket, float(timeout), ctypes.byref(c)))
        return c

    def drain(self) -> None:
        _check(lib().hpttx_server_drain(self._h))

    def release(self, ticket: int) -> None:
        _check(lib().hpttx_server_release(self._h, ticket))

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(lib().hpttx_server_get_stats(self._h, ctypes.byref(st)))
        return st

    def now(self) -> float:
        return lib().hpttx_server_now(self._h)

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.hpttx_server_destroy(h)
            self._h = ctypes.c_void_p(0)

    __del__ = close


def algorithmic_bytes(n_elements: int, elem_bytes: int, update: bool) -> int:
    """Bytes one exec must move: read A, write B, read B iff beta is used."""
    return n_elements * elem_bytes * (3 if update else 2)


def reference_gibps(n_elements: int, elem_bytes: int, seconds: float) -> float:
    """The reference's own throughput figure, 3*N*sizeof(T) bytes in GiB per
    second whatever beta is (src/util/util_trans.cc:12-27)."""
    return 3.0 * n_elements * elem_bytes / 2.0 ** 30 / seconds
