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
import weakref
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
    """One device: its stream, resident twiddle plans and scratch pool.

    Every context starts with the splits compiled into the library
    (csrc/tuned_gfx950.hpp, generated from tuning/gfx950.json at build
    time); ``tuned=True`` re-applies tuning/gfx950.json itself, so a table
    re-tuned since the last build takes effect without rebuilding."""

    def __init__(self, device: int = 0, tuned: bool = True):
        h = ctypes.c_void_p()
        check(lib().stntt_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device
        # buffers and batches made on this context: released before it is
        # destroyed (a batch must not outlive its context)
        self._owned: weakref.WeakSet = weakref.WeakSet()
        if tuned:
            from .autotune import apply_table
            apply_table(self)

    def _own(self, obj) -> None:
        self._owned.add(obj)

    def close(self) -> None:
        if self.handle:
            for obj in list(self._owned):
                obj._release()
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
                inverse: bool = False, signed: bool = False, stream: int | None = None,
                bitrev_in: bool = False, bitrev_out: bool = False) -> None:
        """In place.  bitrev_out: output in bit-reversed order (nttCt's);
        bitrev_in: input in bit-reversed order (inttGs's)."""
        ptr = buf.ptr if isinstance(buf, DeviceBuffer) else buf
        flags = (_lib.SIGNED_INPUT if signed else 0) | _order_flags(bitrev_in, bitrev_out)
        check(lib().stntt_ntt_u32(self.handle, ptr, log2n, batches, root, mod,
                                  _lib.INVERSE if inverse else _lib.FORWARD, flags, stream))

    def ntt_goldilocks(self, buf: "DeviceBuffer | int", log2n: int, batches: int, inverse: bool = False,
                       stream: int | None = None, bitrev_in: bool = False, bitrev_out: bool = False) -> None:
        ptr = buf.ptr if isinstance(buf, DeviceBuffer) else buf
        check(lib().stntt_ntt_goldilocks_ex(self.handle, ptr, log2n, batches,
                                            _lib.INVERSE if inverse else _lib.FORWARD,
                                            _order_flags(bitrev_in, bitrev_out), stream))

    def ntt_u64(self, buf: "DeviceBuffer | int", log2n: int, batches: int, root: int, mod: int,
                inverse: bool = False, stream: int | None = None, bitrev_in: bool = False,
                bitrev_out: bool = False) -> None:
        """Any odd prime mod < 2^64 with a root of order 2^log2n (e.g. from
        find_params with a 64-bit min_mod); in place, any input word."""
        ptr = buf.ptr if isinstance(buf, DeviceBuffer) else buf
        check(lib().stntt_ntt_u64(self.handle, ptr, log2n, batches, root, mod,
                                  _lib.INVERSE if inverse else _lib.FORWARD, _order_flags(bitrev_in, bitrev_out),
                                  stream))

    # ---- pointwise products and polynomial multiplication ----------------
    def pointwise_u32(self, a: "DeviceBuffer", b: "DeviceBuffer", count: int, mod: int,
                      stream: int | None = None) -> None:
        """a[i] = a[i] * b[i] mod p (elemWiseMul)."""
        check(lib().stntt_pointwise_u32(self.handle, a.ptr, b.ptr, count, mod, stream))

    def pointwise_goldilocks(self, a: "DeviceBuffer", b: "DeviceBuffer", count: int,
                             stream: int | None = None) -> None:
        check(lib().stntt_pointwise_goldilocks(self.handle, a.ptr, b.ptr, count, stream))

    def polymul_u32(self, a: "DeviceBuffer", b: "DeviceBuffer", log2n: int, batches: int, root: int, mod: int,
                    signed: bool = False, stream: int | None = None) -> None:
        """pwcNtt on device buffers: a <- a (*) b (cyclic), b <- nttCt(b)."""
        check(lib().stntt_polymul_u32(self.handle, a.ptr, b.ptr, log2n, batches, root, mod,
                                      _lib.SIGNED_INPUT if signed else 0, stream))

    def polymul_goldilocks(self, a: "DeviceBuffer", b: "DeviceBuffer", log2n: int, batches: int,
                           stream: int | None = None) -> None:
        check(lib().stntt_polymul_goldilocks(self.handle, a.ptr, b.ptr, log2n, batches, stream))

    def pointwise_u64(self, a: "DeviceBuffer", b: "DeviceBuffer", count: int, mod: int,
                      stream: int | None = None) -> None:
        check(lib().stntt_pointwise_u64(self.handle, a.ptr, b.ptr, count, mod, stream))

    def polymul_u64(self, a: "DeviceBuffer", b: "DeviceBuffer", log2n: int, batches: int, root: int, mod: int,
                    stream: int | None = None) -> None:
        check(lib().stntt_polymul_u64(self.handle, a.ptr, b.ptr, log2n, batches, root, mod, stream))

    # ---- twiddle tables in the reference layout ---------------------------
    def twiddles_u32(self, n: int, root: int, mod: int, inverse: bool = False) -> np.ndarray:
        out = np.zeros(max(n // 2, 1), dtype=np.uint32)
        check(lib().stntt_twiddles_u32(self.handle, n, root, mod, int(inverse), out.ctypes.data))
        return out[: n // 2]

    def twiddles_u64(self, log2n: int, root: int, mod: int, inverse: bool = False) -> np.ndarray:
        out = np.zeros(max((1 << log2n) // 2, 1), dtype=np.uint64)
        check(lib().stntt_twiddles_u64(self.handle, log2n, root, mod, int(inverse), out.ctypes.data))
        return out[: (1 << log2n) // 2]

    def twiddles_goldilocks(self, log2n: int, inverse: bool = False) -> np.ndarray:
        out = np.zeros(max((1 << log2n) // 2, 1), dtype=np.uint64)
        check(lib().stntt_twiddles_goldilocks(self.handle, log2n, int(inverse), out.ctypes.data))
        return out[: (1 << log2n) // 2]


_FIELDS = {"u32": _lib.FIELD_U32, "goldilocks": _lib.FIELD_GOLDILOCKS, "u64": _lib.FIELD_U64}


def _order_flags(bitrev_in: bool, bitrev_out: bool) -> int:
    return (_lib.BITREV_IN if bitrev_in else 0) | (_lib.BITREV_OUT if bitrev_out else 0)


class DeviceBuffer:
    """Raw device allocation owned by a Context (hipMalloc / hipFree)."""

    def __init__(self, ctx: Context, nbytes: int):
        p = ctypes.c_void_p()
        check(lib().stntt_dev_alloc(ctx.handle, nbytes, ctypes.byref(p)))
        self.ctx, self.ptr, self.nbytes = ctx, p.value, nbytes
        ctx._own(self)

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
        """Release the allocation (a no-op once freed or once the context is closed)."""
        if self.ptr and self.ctx.handle:
            lib().stntt_dev_free(self.ctx.handle, self.ptr)
        self.ptr = None

    _release = free

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
                 inverse: bool = False, ctx: Context | None = None, nsteps: int | None = None) -> np.ndarray:
    """nttStockham(vec, size, root, mod, batches, nsteps) (src/ntt/ntt_cpu.cpp:46-90)
    on the GPU, in place.  nsteps=None runs the whole transform on the
    multi-pass path; a value runs the reference's first nsteps radix-2
    stages one by one (stntt_stockham_steps_u32), forward only."""
    if nsteps is None:
        auto_ntt(NttArgs(vec, size, batches, root, mod, inverse=inverse), ctx)
        return vec
    if inverse:
        raise ValueError("nsteps applies to the forward stage loop only")
    _stage_loop(vec, size, batches, root, mod, nsteps, True, ctx)
    return vec


def ntt_stockham_idx(vec: np.ndarray, size: int, root: int, mod: int, batches: int = 1,
                     nsteps: int | None = None, ctx: Context | None = None) -> np.ndarray:
    """nttStockhamIdx(vec, size, root, mod, batches, nsteps) (src/ntt/ntt_cpu.cpp:92-127)
    on the GPU, in place: the Stockham data movement without arithmetic
    (with every stage, the bit-reversal permutation).  root is unused, as in
    the reference; mod only selects the stage strides pow(2, i, mod)."""
    _stage_loop(vec, size, batches, root, mod, (_log2_exact(size) if nsteps is None else nsteps), False, ctx)
    return vec


def _stage_loop(vec: np.ndarray, size: int, batches: int, root: int, mod: int, nsteps: int, arith: bool,
                ctx: Context | None) -> None:
    ctx = ctx or default_context()
    if not isinstance(vec, np.ndarray) or not vec.flags.c_contiguous or vec.dtype not in (np.int32, np.uint32):
        raise ValueError("vec must be a C-contiguous int32 or uint32 numpy array")
    ms = ctypes.c_float()
    root = int(root) % int(mod) if int(mod) > 0 else 0
    if vec.dtype == np.int32:
        if arith:
            check(lib().stntt_auto_stockham_steps_i32(ctx.handle, vec.ctypes.data, vec.size, size, batches, root,
                                                      mod, nsteps, ctypes.byref(ms)))
        else:
            check(lib().stntt_auto_stockham_idx_i32(ctx.handle, vec.ctypes.data, vec.size, size, batches, mod,
                                                    nsteps, ctypes.byref(ms)))
        return
    L = _log2_exact(size)
    if vec.size < size * batches:
        raise IndexError("Host buffer is smaller than device buffer")
    buf = ctx.upload(vec[: size * batches])
    if arith:
        check(lib().stntt_stockham_steps_u32(ctx.handle, buf.ptr, L, batches, root, mod, nsteps, 0, None))
    else:
        check(lib().stntt_stockham_idx_u32(ctx.handle, buf.ptr, L, batches, mod, nsteps, None))
    buf.copy_to(vec[: size * batches])
    buf.free()


def ntt_goldilocks(vec: np.ndarray, size: int, batches: int = 1, inverse: bool = False,
                   ctx: Context | None = None) -> np.ndarray:
    """Goldilocks transform of vec[batches][size] (uint64) on the GPU, in place."""
    auto_ntt(NttArgs(vec, size, batches, 0, GL_P, field="goldilocks", inverse=inverse), ctx)
    return vec


def _host_transform(vec: np.ndarray, inverse: bool, flags: int, root: int, mod: int, ctx: Context | None) -> float:
    """One transform of the whole vector (length = power of two) in place."""
    ctx = ctx or default_context()
    if not isinstance(vec, np.ndarray) or not vec.flags.c_contiguous:
        raise ValueError("vec must be a C-contiguous numpy array")
    n = vec.size
    L = _log2_exact(n)
    direction = _lib.INVERSE if inverse else _lib.FORWARD
    ms = ctypes.c_float()
    if vec.dtype == np.int32:
        check(lib().stntt_auto_ntt_i32_ex(ctx.handle, vec.ctypes.data, n, n, 1, root, mod, direction, flags,
                                          ctypes.byref(ms)))
        return ms.value
    if vec.dtype not in (np.uint32, np.uint64):
        raise ValueError("vectors are int32 / uint32 (32-bit field) or uint64 (Goldilocks)")
    buf = ctx.upload(vec)
    ctx.timer_start()
    if vec.dtype == np.uint64:
        check(lib().stntt_ntt_goldilocks_ex(ctx.handle, buf.ptr, L, 1, direction, flags, None))
    else:
        check(lib().stntt_ntt_u32(ctx.handle, buf.ptr, L, 1, root, mod, direction, flags, None))
    ms.value = ctx.timer_stop()
    buf.copy_to(vec)
    buf.free()
    return ms.value


def ntt_ct(vec: np.ndarray, root: int = 0, mod: int = GL_P, ctx: Context | None = None) -> np.ndarray:
    """nttCt(vec, root, mod) (src/ntt/ntt_cpu.cpp:11-42) on the GPU, in place:
    natural order in, the transform in bit-reversed order out.  uint64
    vectors are Goldilocks (root and mod ignored)."""
    _host_transform(vec, False, _lib.BITREV_OUT, root, mod, ctx)
    return vec


def intt_gs(vec: np.ndarray, root: int = 0, mod: int = GL_P, ctx: Context | None = None) -> np.ndarray:
    """inttGs(vec, root, mod) (src/ntt/ntt_cpu.cpp:130-164) on the GPU, in
    place: bit-reversed order in, the inverse (scaled by n^-1) in natural
    order out.  root is the forward root, as in the reference."""
    _host_transform(vec, True, _lib.BITREV_IN, root, mod, ctx)
    return vec


def elem_wise_mul(s0: np.ndarray, s1: np.ndarray, mod: int = GL_P, ctx: Context | None = None) -> np.ndarray:
    """elemWiseMul(s0, s1, mod) (src/ntt/ntt_util.h:28-33): s0[i] = s0[i] s1[i]
    mod p for i < len(s0), in place; len(s0) <= len(s1)."""
    ctx = ctx or default_context()
    n = s0.size
    if s1.size < n:
        raise IndexError("s1 is shorter than s0")
    if s0.dtype != s1.dtype or s0.dtype not in (np.uint32, np.uint64):
        raise ValueError("s0 and s1 are both uint32 (32-bit field) or both uint64 (Goldilocks)")
    if n == 0:
        return s0
    a, b = ctx.upload(s0), ctx.upload(np.ascontiguousarray(s1[:n]))
    if s0.dtype == np.uint64:
        ctx.pointwise_goldilocks(a, b, n)
    else:
        ctx.pointwise_u32(a, b, n, mod)
    a.copy_to(s0)
    a.free()
    b.free()
    return s0


def pwc_ntt(vec0: np.ndarray, vec1: np.ndarray, root: int = 0, mod: int = GL_P,
            ctx: Context | None = None) -> np.ndarray:
    """pwcNtt(vec0, vec1, root, mod) (src/ntt/ntt_cpu.cpp:167-174) on the GPU:
    vec0 <- cyclic convolution of vec0 and vec1, vec1 <- nttCt(vec1); both in
    place.  int32 vectors follow the reference's int storage."""
    ctx = ctx or default_context()
    for v in (vec0, vec1):
        if not isinstance(v, np.ndarray) or not v.flags.c_contiguous:
            raise ValueError("vectors must be C-contiguous numpy arrays")
    if vec0.size != vec1.size or vec0.dtype != vec1.dtype:
        raise ValueError("vec0 and vec1 must have the same length and dtype")
    n = vec0.size
    L = _log2_exact(n)
    if vec0.dtype == np.int32:
        ms = ctypes.c_float()
        check(lib().stntt_auto_polymul_i32(ctx.handle, vec0.ctypes.data, vec1.ctypes.data, n, root, mod,
                                           ctypes.byref(ms)))
        return vec0
    if vec0.dtype not in (np.uint32, np.uint64):
        raise ValueError("vectors are int32 / uint32 (32-bit field) or uint64 (Goldilocks)")
    a, b = ctx.upload(vec0), ctx.upload(vec1)
    if vec0.dtype == np.uint64:
        ctx.polymul_goldilocks(a, b, L, 1)
    else:
        ctx.polymul_u32(a, b, L, 1, root, mod)
    a.copy_to(vec0)
    b.copy_to(vec1)
    a.free()
    b.free()
    return vec0


def pwc_convolution(s0: np.ndarray, s1: np.ndarray, ctx: Context | None = None) -> tuple[np.ndarray, int, int]:
    """pwc_convolution(s0, s1) (src/ntt/ntt_cpu.cpp:180-192): the cyclic
    convolution of two positive int vectors of equal power-of-two length,
    with (root, mod) from findParams(len, max^2 len + 1) so that no output
    wraps the modulus.  The inputs are not modified.  Returns (vec, root, mod).
    The reference forms max^2 len + 1 in int; here it is exact and a value
    of 2^31 or more is an error (the 32-bit field holds p < 2^31)."""
    a = np.array(s0, dtype=np.int64)
    b = np.array(s1, dtype=np.int64)
    if a.size == 0 or a.size != b.size:
        raise ValueError("s0 and s1 must be non-empty and of equal length")
    max_val = int(max(a.max(), b.max()))
    min_mod = max_val * max_val * a.size + 1
    if min_mod >= 1 << 31:
        raise ValueError(f"max^2 * len + 1 = {min_mod} needs a modulus of 2^31 or more")
    root, mod = find_params(a.size, min_mod)
    v0, v1 = a.astype(np.int32), b.astype(np.int32)
    pwc_ntt(v0, v1, root, mod, ctx)
    return v0, root, mod


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
            arr[i] = _lib.Desc(_FIELDS[field],
                               _lib.INVERSE if inverse else _lib.FORWARD, log2n, 0,
                               root or 0, mod or 0, ptr, ptr)
        h = ctypes.c_void_p()
        check(lib().stntt_batch_create(ctx.handle, arr, len(descs), ctypes.byref(h)))
        self.handle = h
        self.count = len(descs)
        ctx._own(self)

    def run(self, stream: int | None = None) -> None:
        check(lib().stntt_batch_run(self.handle, stream))

    def capture(self, stream: int | None = None) -> None:
        check(lib().stntt_batch_capture(self.handle, stream))

    def run_graph(self, stream: int | None = None) -> None:
        check(lib().stntt_batch_run_graph(self.handle, stream))

    def run_plain(self, stream: int | None = None) -> None:
        check(lib().stntt_batch_run_plain(self.handle, stream))

    def close(self) -> None:
        """Destroy the batch (a no-op once closed or once the context is closed)."""
        if self.handle and self.ctx.handle:
            lib().stntt_batch_destroy(self.handle)
        self.handle = None

    _release = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_pass_split(ctx: Context, field: str, log2n: int, logr: list[int] | None) -> None:
    """Install an autotuned stages-per-pass split for the multi-pass path."""
    f = _FIELDS[field]
    arr = (ctypes.c_uint32 * max(len(logr or []), 1))(*(logr or []))
    check(lib().stntt_set_pass_split(ctx.handle, f, log2n, len(logr or []), arr))


def set_pass_shape(ctx: Context, field: str, log2n: int, logr: list[int], logc: list[int]) -> None:
    """A pass split with the tile columns (log2) of each pass; -1 = the
    library's choice.  Passes of radix 2^10..2^12 run the large-radix kernel."""
    f = _FIELDS[field]
    r = (ctypes.c_uint32 * len(logr))(*logr)
    c = (ctypes.c_int32 * len(logc))(*logc)
    check(lib().stntt_set_pass_shape(ctx.handle, f, log2n, len(logr), r, c))


def restore_pass_split(ctx: Context, field: str, log2n: int) -> None:
    """Put back the tuned split of (field, log2n) from the tuning table (or
    the fewest-pass default where the table has none), undoing a
    set_pass_split override."""
    from .autotune import table_pass_split
    t = table_pass_split(field, log2n)
    if t is None:
        set_pass_split(ctx, field, log2n, None)
    else:
        set_pass_shape(ctx, field, log2n, t[0], t[1])


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
        server._owned.add(self)

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
        self._owned: weakref.WeakSet = weakref.WeakSet()  # pinned arrays, freed after the server

    def pinned(self, count: int, dtype) -> PinnedArray:
        return PinnedArray(self, count, dtype)

    def submit(self, field: str, log2n: int, data: PinnedArray, inverse: bool = False, root: int = 0,
               mod: int = 0, out: PinnedArray | None = None) -> int:
        req = _lib.Request(_FIELDS[field],
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
        """Finish every request, stop the server, then free its pinned arrays."""
        if self.handle:
            lib().stntt_server_destroy(self.handle)
            self.handle = None
            for arr in list(self._owned):
                arr.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
