"""Loader for libstntt.so (the HIP product) and its C-ABI signatures.

The library is built in-tree by stntt_amd.build.build(); importing this
module never builds it, and a missing library is a hard error: there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libstntt.so")

STNTT_OK = 0
STNTT_EINVAL = 1
STNTT_ERANGE = 2
STNTT_ENOMEM = 3
STNTT_EDEVICE = 4
STNTT_EAGAIN = 5
STNTT_ENOTFOUND = 6

FIELD_U32 = 0
FIELD_GOLDILOCKS = 1
FIELD_U64 = 2
FORWARD = 0
INVERSE = 1
SIGNED_INPUT = 1
BITREV_IN = 2
BITREV_OUT = 4

_c = ctypes
_vp = _c.c_void_p
_u32 = _c.c_uint32
_u64 = _c.c_uint64
_int = _c.c_int
_sz = _c.c_size_t

# (name, restype, argtypes) for every symbol declared in include/stntt.h
SIGNATURES = [
    ("stntt_version", _c.c_char_p, []),
    ("stntt_last_error", _c.c_char_p, []),
    ("stntt_ctx_create", _int, [_int, _c.POINTER(_vp)]),
    ("stntt_ctx_destroy", _int, [_vp]),
    ("stntt_ctx_stream", _vp, [_vp]),
    ("stntt_find_params", _int, [_u64, _u64, _c.POINTER(_u64), _c.POINTER(_u64)]),
    ("stntt_twiddles_u32", _int, [_vp, _u64, _u32, _u32, _int, _vp]),
    ("stntt_twiddles_goldilocks", _int, [_vp, _u32, _int, _vp]),
    ("stntt_ntt_u32", _int, [_vp, _vp, _u32, _u64, _u32, _u32, _int, _u32, _vp]),
    ("stntt_ntt_goldilocks", _int, [_vp, _vp, _u32, _u64, _int, _vp]),
    ("stntt_ntt_goldilocks_ex", _int, [_vp, _vp, _u32, _u64, _int, _u32, _vp]),
    ("stntt_ntt_goldilocks_noalloc", _int, [_vp, _vp, _u32, _u64, _int, _u32, _vp]),
    ("stntt_ntt_u64", _int, [_vp, _vp, _u32, _u64, _u64, _u64, _int, _u32, _vp]),
    ("stntt_twiddles_u64", _int, [_vp, _u32, _u64, _u64, _int, _vp]),
    ("stntt_pointwise_u64", _int, [_vp, _vp, _vp, _u64, _u64, _vp]),
    ("stntt_polymul_u64", _int, [_vp, _vp, _vp, _u32, _u64, _u64, _u64, _vp]),
    ("stntt_pointwise_u32", _int, [_vp, _vp, _vp, _u64, _u32, _vp]),
    ("stntt_pointwise_goldilocks", _int, [_vp, _vp, _vp, _u64, _vp]),
    ("stntt_polymul_u32", _int, [_vp, _vp, _vp, _u32, _u64, _u32, _u32, _u32, _vp]),
    ("stntt_polymul_goldilocks", _int, [_vp, _vp, _vp, _u32, _u64, _vp]),
    ("stntt_auto_ntt_i32", _int, [_vp, _vp, _sz, _u32, _u32, _u32, _u32, _int, _c.POINTER(_c.c_float)]),
    ("stntt_auto_ntt_goldilocks", _int, [_vp, _vp, _sz, _u32, _u32, _int, _c.POINTER(_c.c_float)]),
    ("stntt_auto_ntt_i32_ex", _int, [_vp, _vp, _sz, _u32, _u32, _u32, _u32, _int, _u32, _c.POINTER(_c.c_float)]),
    ("stntt_auto_polymul_i32", _int, [_vp, _vp, _vp, _u32, _u32, _u32, _c.POINTER(_c.c_float)]),
    ("stntt_stockham_steps_u32", _int, [_vp, _vp, _u32, _u64, _u32, _u32, _int, _u32, _vp]),
    ("stntt_stockham_idx_u32", _int, [_vp, _vp, _u32, _u64, _u32, _int, _vp]),
    ("stntt_auto_stockham_steps_i32", _int, [_vp, _vp, _sz, _u32, _u32, _u32, _u32, _int,
                                             _c.POINTER(_c.c_float)]),
    ("stntt_auto_stockham_idx_i32", _int, [_vp, _vp, _sz, _u32, _u32, _u32, _int, _c.POINTER(_c.c_float)]),
    ("stntt_batch_create", _int, [_vp, _vp, _sz, _c.POINTER(_vp)]),
    ("stntt_batch_run", _int, [_vp, _vp]),
    ("stntt_batch_capture", _int, [_vp, _vp]),
    ("stntt_batch_run_graph", _int, [_vp, _vp]),
    ("stntt_batch_run_plain", _int, [_vp, _vp]),
    ("stntt_batch_destroy", _int, [_vp]),
    ("stntt_set_batch_split", _int, [_vp, _int, _u32, _u32, _u32]),
    ("stntt_set_pass_split", _int, [_vp, _int, _u32, _u32, _c.POINTER(_u32)]),
    ("stntt_set_pass_shape", _int, [_vp, _int, _u32, _u32, _c.POINTER(_u32), _c.POINTER(_c.c_int32)]),
    ("stntt_server_create", _int, [_int, _u64, _u32, _c.POINTER(_vp)]),
    ("stntt_server_destroy", _int, [_vp]),
    ("stntt_server_host_alloc", _int, [_vp, _sz, _c.POINTER(_vp)]),
    ("stntt_server_host_free", _int, [_vp, _vp]),
    ("stntt_server_submit", _int, [_vp, _vp, _c.POINTER(_u64)]),
    ("stntt_server_wait", _int, [_vp, _u64, _c.POINTER(_c.c_double)]),
    ("stntt_server_poll", _int, [_vp, _u64, _c.POINTER(_int)]),
    ("stntt_server_stats_get", _int, [_vp, _vp]),
    ("stntt_server_reset_stats", _int, [_vp]),
    ("stntt_dev_alloc", _int, [_vp, _sz, _c.POINTER(_vp)]),
    ("stntt_dev_free", _int, [_vp, _vp]),
    ("stntt_memcpy_htod", _int, [_vp, _vp, _vp, _sz, _vp]),
    ("stntt_memcpy_dtoh", _int, [_vp, _vp, _vp, _sz, _vp]),
    ("stntt_memcpy_dtod", _int, [_vp, _vp, _vp, _sz, _vp]),
    ("stntt_stream_sync", _int, [_vp, _vp]),
    ("stntt_timer_start", _int, [_vp, _vp]),
    ("stntt_timer_stop", _int, [_vp, _vp, _c.POINTER(_c.c_float)]),
    ("stntt_set_profiling", _int, [_vp, _int]),
    ("stntt_profile_read", _int, [_vp, _c.POINTER(_c.c_float), _int, _c.POINTER(_int)]),
]

_lib = None


class Desc(_c.Structure):
    """stntt_desc (include/stntt.h)"""
    _fields_ = [("field", _u32), ("direction", _u32), ("log2n", _u32), ("flags", _u32),
                ("root", _u32), ("mod", _u32), ("src", _vp), ("dst", _vp)]


class Request(_c.Structure):
    """stntt_request (include/stntt.h)"""
    _fields_ = [("field", _u32), ("direction", _u32), ("log2n", _u32), ("flags", _u32),
                ("root", _u32), ("mod", _u32), ("host_in", _vp), ("host_out", _vp)]


class ServerStats(_c.Structure):
    """stntt_server_stats (include/stntt.h)"""
    _fields_ = [("completed", _u64), ("p50_ms", _c.c_double), ("p99_ms", _c.c_double), ("mean_ms", _c.c_double),
                ("max_ms", _c.c_double), ("span_s", _c.c_double), ("peak_bytes", _u64), ("rounds", _u64),
                ("batched", _u64), ("held_bytes", _u64), ("grouped", _u64)]


class NativeLibraryMissing(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is not built; run __graft_entry__.build() (the HIP library has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class StnttError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"stntt error {code}: {msg}")
        self.code = code


def check(rc: int) -> None:
    """Raise the reference's exception kinds: invalid_argument -> ValueError,
    out_of_range -> IndexError, anything else -> StnttError."""
    if rc == STNTT_OK:
        return
    msg = (lib().stntt_last_error() or b"").decode(errors="replace")
    if rc == STNTT_EINVAL:
        raise ValueError(msg)
    if rc == STNTT_ERANGE:
        raise IndexError(msg)
    if rc == STNTT_ENOMEM:
        raise MemoryError(msg)
    raise StnttError(rc, msg)
