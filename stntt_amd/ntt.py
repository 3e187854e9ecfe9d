"""Python host side of the MI355X Stockham NTT, over the C ABI (include/stntt.h).

Mirrors the reference's host interfaces for this path:
  * find_params(n, min_mod)                  <- findParams   (src/ntt/ntt_util.cpp:10-14)
  * NttArgs / auto_ntt(args)                 <- cuda::ntt::NttArgs / autoNtt<n>
                                                (src/cuda/ntt/ntt_kernel.cuh:41-82,152-201)
  * ntt_stockham(vec, size, root, mod, b)    <- nttStockham(span, size, root, mod, batches)
                                                (src/ntt/ntt_cpu.cpp:87-90), run on the GPU
plus the inverse and the Goldilocks field, which the reference leaves
unimplemented (src/cuda/ntt/implementations/goldilocks/arithmetic.cuh:32-35).

Errors keep the reference's kinds: a non-power-of-two size raises ValueError
(std::invalid_argument), a short host buffer raises IndexError
(std::out_of_range).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib

GL_P = 0xFFFFFFFF00000001
DEFAULT_BLOCK_SIZE = 1024  # src/cuda/ntt/ntt_kernel.cuh:25


def version() -> str:
    return lib().stntt_version().decode()


def find_params(n: int, min_mod: int = 11) -> tuple[int, int]:
    """(root, mod): smallest prime mod = k n + 1 >= min_mod, root of order n."""
    r, m = ctypes.c_uint64(), ctypes.c_uint64()
    check(lib().stntt_find_params(n, min_mod, ctypes.byref(r), ctypes.byref(m)))
    return r.value, m.value


def _log2_exact(n: int) -> int:
    if n <= 0 or n & (n - 1):
        raise ValueError(f"size ({n}) must be a power of 2")
    return n.bit_length() - 1


class Context:
    """One device: its stream, resident twiddle plans and scratch pool."""

    def __init__(self, device: int = 0, tuned: bool = True):
        h = ctypes.c_void_p()
        check(lib().stntt_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device
        if tuned:
            from .autotune import apply_table
            apply_table(self)

    def close(self) -> None:
        if self.handle:
            lib().stntt_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return lib().stntt_ctx_stream(self.handle)

    # ---- memory plumbing ------------------------------------------------
    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def upload(self, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = DeviceBuffer(self, arr.nbytes)
        buf.copy_from(arr)
        return buf

    def copy(self, dst: "DeviceBuffer", src: "DeviceBuffer", nbytes: int, stream: int | None = None) -> None:
        check(lib().stntt_memcpy_dtod(self.handle, dst.ptr, src.ptr, nbytes, stream))

    def sync(self, stream: int | None = None) -> None:
        check(lib().stntt_stream_sync(self.handle, stream))

    def timer_start(self, stream: int | None = None) -> None:
        check(lib().stntt_timer_start(self.handle, stream))

    def timer_stop(self, stream: int | None = None) -> float:
        ms = ctypes.c_float()
        check(lib().stntt_timer_stop(self.handle, stream, ctypes.byref(ms)))
        return ms.value

    def set_profiling(self, on: bool) -> None:
        check(lib().stntt_set_profiling(self.handle, int(on)))

    def profile_read(self) -> list[float]:
        """Per-kernel durations (ms) of the last transform run with profiling on."""
        buf = (ctypes.c_float * 16)()
        cnt = ctypes.c_int()
        check(lib().stntt_profile_read(self.handle, buf, 16, ctypes.byref(cnt)))
        return [buf[i] for i in range(min(cnt.value, 16))]

    # ---- transforms on device buffers (asynchronous) ----------------------
    def ntt_u32(self, buf: "DeviceBuffer | int", log2n: int, batches: int, root: int, mod: int,
                inverse: bool = False, signed: bool = False, stream: int | None = None) -> None:
        ptr = buf.ptr if isinstance(buf, DeviceBuffer) else buf
        check(lib().stntt_ntt_u32(self.handle, ptr, log2n, batches, root, mod,
                                  _lib.INVERSE if inverse else _lib.FORWARD,
                                  _lib.SIGNED_INPUT if signed else 0, stream))

    def ntt_goldilocks(self, buf: "DeviceBuffer | int", log2n: int, batches: int, inverse: bool = False,
                       stream: int | None = None) -> None:
        ptr = buf.ptr if isinstance(buf, DeviceBuffer) else buf
        check(lib().stntt_ntt_goldilocks(self.handle, ptr, log2n, batches,
                                         _lib.INVERSE if inverse else _lib.FORWARD, stream))

    # ---- twiddle tables in the reference layout ---------------------------
    def twiddles_u32(self, n: int, root: int, mod: int, inverse: bool = False) -> np.ndarray:
        out = np.zeros(max(n // 2, 1), dtype=np.uint32)
        check(lib().stntt_twiddles_u32(self.handle, n, root, mod, int(inverse), out.ctypes.data))
        return out[: n // 2]

    def twiddles_goldilocks(self, log2n: int, inverse: bool = False) -> np.ndarray:
        out = np.zeros(max((1 << log2n) // 2, 1), dtype=np.uint64)
        check(lib().stntt_twiddles_goldilocks(self.handle, log2n, int(inverse), out.ctypes.data))
        return out[: (1 << log2n) // 2]


class DeviceBuffer:
    """Raw device allocation owned by a Context (hipMalloc / hipFree)."""

    def __init__(self, ctx: Context, nbytes: int):
        p = ctypes.c_void_p()
        check(lib().stntt_dev_alloc(ctx.handle, nbytes, ctypes.byref(p)))
        self.ctx, self.ptr, self.nbytes = ctx, p.value, nbytes

    def copy_from(self, arr: np.ndarray, stream: int | None = None) -> None:
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise IndexError("host array larger than device buffer")
        check(lib().stntt_memcpy_htod(self.ctx.handle, self.ptr, arr.ctypes.data, arr.nbytes, stream))

    def copy_to(self, arr: np.ndarray, stream: int | None = None) -> np.ndarray:
        if not arr.flags.c_contiguous or arr.nbytes > self.nbytes:
            raise IndexError("host array must be contiguous and no larger than the device buffer")
        check(lib().stntt_memcpy_dtoh(self.ctx.handle, arr.ctypes.data, self.ptr, arr.nbytes, stream))
        return arr

    def free(self) -> None:
        if self.ptr:
            lib().stntt_dev_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


# --------------------------------------------------------------------------
# reference-shaped API
# --------------------------------------------------------------------------
@dataclass
class NttArgs:
    """cuda::ntt::NttArgs (src/cuda/ntt/ntt_kernel.cuh:41-82).

    radix and block_size are accepted for call compatibility; the tile shape
    of every pass is chosen by the library.
    """
    vec: np.ndarray
    n: int
    batches: int
    root: int
    mod: int
    radix: int = 2
    block_size: int = DEFAULT_BLOCK_SIZE
    inverse: bool = False
    field: str = "u32"  # "u32" or "goldilocks"
    is_global: bool = False  # NttArgs::isGlobal, informational

    def __post_init__(self):
        self.is_global = (self.n // max(self.radix, 1)) > min(self.n // max(self.radix, 1), self.block_size)


def auto_ntt(args: NttArgs, ctx: Context | None = None) -> float:
    """autoNtt<n>(args): transform args.vec in place, return kernel-only ms.

    args.vec must be a C-contiguous numpy array: int32 (the reference's int
    storage, any value) or uint32 for the 32-bit field, uint64 for Goldilocks.
    """
    ctx = ctx or default_context()
    vec = args.vec
    if not isinstance(vec, np.ndarray) or not vec.flags.c_contiguous:
        raise ValueError("vec must be a C-contiguous numpy array")
    ms = ctypes.c_float()
    direction = _lib.INVERSE if args.inverse else _lib.FORWARD
    if args.field == "goldilocks":
        if vec.dtype != np.uint64:
            raise ValueError("Goldilocks vectors are uint64")
        check(lib().stntt_auto_ntt_goldilocks(ctx.handle, vec.ctypes.data, vec.size, args.n, args.batches,
                                              direction, ctypes.byref(ms)))
    else:
        if vec.dtype not in (np.int32, np.uint32):
            raise ValueError("32-bit field vectors are int32 or uint32")
        if vec.dtype == np.int32:
            check(lib().stntt_auto_ntt_i32(ctx.handle, vec.ctypes.data, vec.size, args.n, args.batches,
                                           args.root, args.mod, direction, ctypes.byref(ms)))
        else:
            _log2_exact(args.n)
            if vec.size < args.n * args.batches:
                raise IndexError("Host buffer is smaller than device buffer")
            buf = ctx.upload(vec[: args.n * args.batches])
            ctx.timer_start()
            ctx.ntt_u32(buf, _log2_exact(args.n), args.batches, args.root, args.mod, inverse=args.inverse)
            ms.value = ctx.timer_stop()
            buf.copy_to(vec[: args.n * args.batches])
            buf.free()
    return ms.value


def ntt_stockham(vec: np.ndarray, size: int, root: int, mod: int, batches: int = 1,
                 inverse: bool = False, ctx: Context | None = None) -> np.ndarray:
    """nttStockham(vec, size, root, mod, batches) on the GPU, in place."""
    auto_ntt(NttArgs(vec, size, batches, root, mod, inverse=inverse), ctx)
    return vec


def ntt_goldilocks(vec: np.ndarray, size: int, batches: int = 1, inverse: bool = False,
                   ctx: Context | None = None) -> np.ndarray:
    """Goldilocks transform of vec[batches][size] (uint64) on the GPU, in place."""
    auto_ntt(NttArgs(vec, size, batches, 0, GL_P, field="goldilocks", inverse=inverse), ctx)
    return vec


# --------------------------------------------------------------------------
# batched mixed transforms
# --------------------------------------------------------------------------
class Batch:
    """Thousands of independent transforms in one persistent launch.

    descs: iterable of (field, inverse, log2n, buf, root, mod) where field is
    "u32" or "goldilocks", buf a DeviceBuffer (transformed in place) and
    root/mod the 32-bit field parameters (ignored for Goldilocks).
    """

    def __init__(self, ctx: Context, descs):
        self.ctx = ctx
        arr = (_lib.Desc * max(len(descs), 1))()
        for i, (field, inverse, log2n, buf, root, mod) in enumerate(descs):
            ptr = buf.ptr if isinstance(buf, DeviceBuffer) else buf
            arr[i] = _lib.Desc(_lib.FIELD_GOLDILOCKS if field == "goldilocks" else _lib.FIELD_U32,
                               _lib.INVERSE if inverse else _lib.FORWARD, log2n, 0,
                               root or 0, mod or 0, ptr, ptr)
        h = ctypes.c_void_p()
        check(lib().stntt_batch_create(ctx.handle, arr, len(descs), ctypes.byref(h)))
        self.handle = h
        self.count = len(descs)

    def run(self, stream: int | None = None) -> None:
        check(lib().stntt_batch_run(self.handle, stream))

    def capture(self, stream: int | None = None) -> None:
        check(lib().stntt_batch_capture(self.handle, stream))

    def run_graph(self, stream: int | None = None) -> None:
        check(lib().stntt_batch_run_graph(self.handle, stream))

    def run_plain(self, stream: int | None = None) -> None:
        check(lib().stntt_batch_run_plain(self.handle, stream))

    def close(self) -> None:
        if self.handle:
            lib().stntt_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_pass_split(ctx: Context, field: str, log2n: int, logr: list[int] | None) -> None:
    """Install an autotuned stages-per-pass split for the multi-pass path."""
    f = _lib.FIELD_GOLDILOCKS if field == "goldilocks" else _lib.FIELD_U32
    arr = (ctypes.c_uint32 * max(len(logr or []), 1))(*(logr or []))
    check(lib().stntt_set_pass_split(ctx.handle, f, log2n, len(logr or []), arr))


def set_batch_split(ctx: Context, field: str, log2n: int, radix_logs: list[int] | None) -> None:
    """Install an autotuned sub-step split for the batched kernel."""
    f = _lib.FIELD_GOLDILOCKS if field == "goldilocks" else _lib.FIELD_U32
    if not radix_logs:
        check(lib().stntt_set_batch_split(ctx.handle, f, log2n, 0, 0))
        return
    packed = 0
    for i, lr in enumerate(radix_logs):
        packed |= lr << (4 * i)
    check(lib().stntt_set_batch_split(ctx.handle, f, log2n, len(radix_logs), packed))


# --------------------------------------------------------------------------
# serving
# --------------------------------------------------------------------------
class PinnedArray:
    """numpy view of pinned host memory owned by a Server."""

    def __init__(self, server: "Server", count: int, dtype):
        self.server = server
        nbytes = max(int(count) * np.dtype(dtype).itemsize, 1)
        p = ctypes.c_void_p()
        check(lib().stntt_server_host_alloc(server.handle, nbytes, ctypes.byref(p)))
        self.ptr = p.value
        buf = (ctypes.c_char * nbytes).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=dtype, count=int(count))

    def free(self) -> None:
        if self.ptr:
            lib().stntt_server_host_free(self.server.handle, self.ptr)
            self.ptr = None
            self.array = None


class Server:
    """Continuous-batching transform server (include/stntt.h, serving)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 64 << 30, workers: int = 4):
        h = ctypes.c_void_p()
        check(lib().stntt_server_create(device, hbm_budget_bytes, workers, ctypes.byref(h)))
        self.handle = h

    def pinned(self, count: int, dtype) -> PinnedArray:
        return PinnedArray(self, count, dtype)

    def submit(self, field: str, log2n: int, data: PinnedArray, inverse: bool = False, root: int = 0,
               mod: int = 0, out: PinnedArray | None = None) -> int:
        req = _lib.Request(_lib.FIELD_GOLDILOCKS if field == "goldilocks" else _lib.FIELD_U32,
                           _lib.INVERSE if inverse else _lib.FORWARD, log2n, 0, root, mod, data.ptr,
                           (out or data).ptr)
        rid = ctypes.c_uint64()
        check(lib().stntt_server_submit(self.handle, ctypes.byref(req), ctypes.byref(rid)))
        return rid.value

    def wait(self, rid: int) -> float:
        ms = ctypes.c_double()
        check(lib().stntt_server_wait(self.handle, rid, ctypes.byref(ms)))
        return ms.value

    def poll(self, rid: int) -> bool:
        d = ctypes.c_int()
        check(lib().stntt_server_poll(self.handle, rid, ctypes.byref(d)))
        return bool(d.value)

    def stats(self) -> dict:
        st = _lib.ServerStats()
        check(lib().stntt_server_stats_get(self.handle, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def reset_stats(self) -> None:
        check(lib().stntt_server_reset_stats(self.handle))

    def close(self) -> None:
        if self.handle:
            lib().stntt_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
