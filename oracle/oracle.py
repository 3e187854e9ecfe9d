"""ctypes view of the CPU oracle (liboracle.so) and of the compiled reference
(oracle/_ref/libstntt_ref.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product package (stntt_amd)
never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libstntt_ref.so")
GL_P = 0xFFFFFFFF00000001

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_c = ctypes
_U64 = _c.c_uint64
_U32 = _c.c_uint32


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and, when the reference tree exists, _ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/src/ntt"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_find_params.argtypes = [_U64, _U64, _c.POINTER(_U64), _c.POINTER(_U64)]
        L.orc_powmod.argtypes = [_U64, _U64, _U64]
        L.orc_powmod.restype = _U64
        L.orc_invmod.argtypes = [_U64, _U64]
        L.orc_invmod.restype = _U64
        L.orc_is_prime.argtypes = [_U64]
        L.orc_bit_reverse.argtypes = [_c.c_size_t, _c.c_uint]
        L.orc_bit_reverse.restype = _c.c_size_t
        L.orc_twiddles_u32.argtypes = [_u32p, _U64, _U32, _U32]
        L.orc_stockham_u32.argtypes = [_u32p, _U64, _U32, _U32, _c.c_int]
        L.orc_stockham_u32_batch.argtypes = [_u32p, _U64, _U32, _U32, _U64, _c.c_int]
        L.orc_inverse_u32.argtypes = [_u32p, _U64, _U32, _U32]
        L.orc_inverse_u32_batch.argtypes = [_u32p, _U64, _U32, _U32, _U64]
        L.orc_naive_dft_u32.argtypes = [_u32p, _u32p, _U64, _U32, _U32]
        L.orc_ntt_ct_u32.argtypes = [_u32p, _U64, _U32, _U32]
        L.orc_intt_gs_u32.argtypes = [_u32p, _U64, _U32, _U32]
        L.orc_pwc_ntt_u32.argtypes = [_u32p, _u32p, _U64, _U32, _U32]
        L.orc_stockham_refcompat.argtypes = [_i32p, _U64, _c.c_int32, _c.c_int32, _c.c_int]
        L.orc_stockham_idx.argtypes = [_i32p, _U64, _c.c_int]
        L.orc_stockham_idx_mod.argtypes = [_i32p, _U64, _U32, _c.c_int]
        for f in ("orc_gl_mul", "orc_gl_add", "orc_gl_sub", "orc_gl_pow"):
            getattr(L, f).argtypes = [_U64, _U64]
            getattr(L, f).restype = _U64
        L.orc_gl_root.argtypes = [_c.c_uint]
        L.orc_gl_root.restype = _U64
        L.orc_twiddles_gl.argtypes = [_u64p, _c.c_uint, _c.c_int]
        L.orc_stockham_gl.argtypes = [_u64p, _c.c_uint, _c.c_int]
        L.orc_stockham_gl_batch.argtypes = [_u64p, _c.c_uint, _U64, _c.c_int]
        L.orc_inverse_gl.argtypes = [_u64p, _c.c_uint]
        L.orc_inverse_gl_batch.argtypes = [_u64p, _c.c_uint, _U64]
        L.orc_naive_dft_gl.argtypes = [_u64p, _u64p, _c.c_uint, _c.c_int]
        L.orc_ntt_ct_gl.argtypes = [_u64p, _c.c_uint]
        L.orc_intt_gs_gl.argtypes = [_u64p, _c.c_uint]
        L.orc_pwc_ntt_gl.argtypes = [_u64p, _u64p, _c.c_uint]
        L.orc_twiddles_u64.argtypes = [_u64p, _U64, _U64, _U64]
        L.orc_stockham_u64_batch.argtypes = [_u64p, _U64, _U64, _U64, _U64, _c.c_int]
        L.orc_inverse_u64_batch.argtypes = [_u64p, _U64, _U64, _U64, _U64]
        L.orc_naive_dft_u64.argtypes = [_u64p, _u64p, _U64, _U64, _U64]
        L.orc_ntt_ct_u64.argtypes = [_u64p, _U64, _U64, _U64]
        L.orc_intt_gs_u64.argtypes = [_u64p, _U64, _U64, _U64]
        L.orc_pwc_ntt_u64.argtypes = [_u64p, _u64p, _U64, _U64, _U64]
        L.orc_dft_points_gl.argtypes = [_u64p, _c.c_uint, _c.c_int, _u64p, _U64, _u64p]
        L.orc_dft_points_u32.argtypes = [_u32p, _U64, _U32, _U32, _c.c_int, _u64p, _U64, _u32p]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref() -> ctypes.CDLL:
    """The reference's own CPU NTT compiled in place (see Makefile)."""
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_PATH)
        R.ref_find_params.argtypes = [_U64, _U64, _c.POINTER(_U64), _c.POINTER(_U64)]
        R.ref_ntt_stockham.argtypes = [_i32p, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int]
        R.ref_ntt_stockham_idx.argtypes = [_i32p, _c.c_int, _c.c_int, _c.c_int]
        R.ref_ntt_stockham_idx_mod.argtypes = [_i32p, _c.c_int, _c.c_int, _c.c_int, _c.c_int]
        R.ref_elem_wise_mul.argtypes = [_i32p, _i32p, _c.c_int, _U64]
        R.ref_ntt_ct.argtypes = [_i32p, _c.c_int, _c.c_int, _c.c_int]
        R.ref_intt_gs.argtypes = [_i32p, _c.c_int, _c.c_int, _c.c_int]
        R.ref_pwc_convolution.argtypes = [_i32p, _i32p, _c.c_int, _i32p, _c.POINTER(_c.c_int),
                                          _c.POINTER(_c.c_int)]
        _ref = R
    return _ref


def _check(rc: int, what: str) -> None:
    if rc == -1:
        raise ValueError(f"{what}: size must be a power of 2")
    if rc != 0:
        raise RuntimeError(f"{what}: oracle error {rc}")


# --------------------------------------------------------------------------
# 32-bit prime field
# --------------------------------------------------------------------------
def find_params(n: int, min_mod: int = 11) -> tuple[int, int]:
    r, m = _U64(), _U64()
    _check(lib().orc_find_params(n, min_mod, _c.byref(r), _c.byref(m)), "find_params")
    return r.value, m.value


def twiddles_u32(n: int, root: int, mod: int) -> np.ndarray:
    out = np.zeros(max(n // 2, 1), dtype=np.uint32)
    _check(lib().orc_twiddles_u32(out, n, root, mod), "twiddles_u32")
    return out[: n // 2]


def stockham_u32(x: np.ndarray, n: int, root: int, mod: int, batches: int = 1, nsteps: int = -1) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.uint32).copy()
    _check(lib().orc_stockham_u32_batch(v, n, root, mod, batches, nsteps), "stockham_u32")
    return v


def inverse_u32(x: np.ndarray, n: int, root: int, mod: int, batches: int = 1) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.uint32).copy()
    _check(lib().orc_inverse_u32_batch(v, n, root, mod, batches), "inverse_u32")
    return v


def naive_dft_u32(x: np.ndarray, root: int, mod: int) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.zeros_like(v)
    _check(lib().orc_naive_dft_u32(v, out, v.size, root, mod), "naive_dft_u32")
    return out


def ntt_ct_u32(x: np.ndarray, root: int, mod: int) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.uint32).copy()
    _check(lib().orc_ntt_ct_u32(v, v.size, root, mod), "ntt_ct_u32")
    return v


def intt_gs_u32(x: np.ndarray, root: int, mod: int) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.uint32).copy()
    _check(lib().orc_intt_gs_u32(v, v.size, root, mod), "intt_gs_u32")
    return v


def pwc_ntt_u32(a: np.ndarray, b: np.ndarray, root: int, mod: int) -> np.ndarray:
    va = np.ascontiguousarray(a, dtype=np.uint32).copy()
    vb = np.ascontiguousarray(b, dtype=np.uint32).copy()
    _check(lib().orc_pwc_ntt_u32(va, vb, va.size, root, mod), "pwc_ntt_u32")
    return va


def stockham_refcompat(x: np.ndarray, n: int, root: int, mod: int, batches: int = 1, nsteps: int = -1) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.int32).copy()
    for b in range(batches):
        seg = v[b * n:(b + 1) * n]
        _check(lib().orc_stockham_refcompat(seg, n, root, mod, nsteps), "stockham_refcompat")
    return v


def stockham_idx(x: np.ndarray, n: int, nsteps: int = -1, mod: int = 2**31 - 1) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.int32).copy()
    _check(lib().orc_stockham_idx_mod(v, n, mod, nsteps), "stockham_idx")
    return v


# --------------------------------------------------------------------------
# any odd modulus below 2^64
# --------------------------------------------------------------------------
def _u64(x):
    return np.ascontiguousarray(x, dtype=np.uint64).copy()


def twiddles_u64(n: int, root: int, mod: int) -> np.ndarray:
    out = np.zeros(max(n // 2, 1), dtype=np.uint64)
    _check(lib().orc_twiddles_u64(out, n, root, mod), "twiddles_u64")
    return out[: n // 2]


def stockham_u64(x: np.ndarray, n: int, root: int, mod: int, batches: int = 1, nsteps: int = -1) -> np.ndarray:
    v = _u64(x)
    _check(lib().orc_stockham_u64_batch(v, n, root, mod, batches, nsteps), "stockham_u64")
    return v


def inverse_u64(x: np.ndarray, n: int, root: int, mod: int, batches: int = 1) -> np.ndarray:
    v = _u64(x)
    _check(lib().orc_inverse_u64_batch(v, n, root, mod, batches), "inverse_u64")
    return v


def naive_dft_u64(x: np.ndarray, root: int, mod: int) -> np.ndarray:
    v = _u64(x)
    out = np.zeros_like(v)
    _check(lib().orc_naive_dft_u64(v, out, v.size, root, mod), "naive_dft_u64")
    return out


def ntt_ct_u64(x: np.ndarray, root: int, mod: int) -> np.ndarray:
    v = _u64(x)
    _check(lib().orc_ntt_ct_u64(v, v.size, root, mod), "ntt_ct_u64")
    return v


def intt_gs_u64(x: np.ndarray, root: int, mod: int) -> np.ndarray:
    v = _u64(x)
    _check(lib().orc_intt_gs_u64(v, v.size, root, mod), "intt_gs_u64")
    return v


def pwc_ntt_u64(a: np.ndarray, b: np.ndarray, root: int, mod: int) -> np.ndarray:
    va, vb = _u64(a), _u64(b)
    _check(lib().orc_pwc_ntt_u64(va, vb, va.size, root, mod), "pwc_ntt_u64")
    return va


# --------------------------------------------------------------------------
# Goldilocks
# --------------------------------------------------------------------------
def gl_root(log2n: int) -> int:
    return lib().orc_gl_root(log2n)


def twiddles_gl(log2n: int, inverse: bool = False) -> np.ndarray:
    out = np.zeros(max((1 << log2n) // 2, 1), dtype=np.uint64)
    _check(lib().orc_twiddles_gl(out, log2n, int(inverse)), "twiddles_gl")
    return out[: (1 << log2n) // 2]


def stockham_gl(x: np.ndarray, log2n: int, batches: int = 1, nsteps: int = -1) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.uint64).copy()
    _check(lib().orc_stockham_gl_batch(v, log2n, batches, nsteps), "stockham_gl")
    return v


def ntt_ct_gl(x: np.ndarray, log2n: int) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.uint64).copy()
    _check(lib().orc_ntt_ct_gl(v, log2n), "ntt_ct_gl")
    return v


def intt_gs_gl(x: np.ndarray, log2n: int) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.uint64).copy()
    _check(lib().orc_intt_gs_gl(v, log2n), "intt_gs_gl")
    return v


def pwc_ntt_gl(a: np.ndarray, b: np.ndarray, log2n: int) -> np.ndarray:
    va = np.ascontiguousarray(a, dtype=np.uint64).copy()
    vb = np.ascontiguousarray(b, dtype=np.uint64).copy()
    _check(lib().orc_pwc_ntt_gl(va, vb, log2n), "pwc_ntt_gl")
    return va


def inverse_gl(x: np.ndarray, log2n: int, batches: int = 1) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.uint64).copy()
    _check(lib().orc_inverse_gl_batch(v, log2n, batches), "inverse_gl")
    return v


def naive_dft_gl(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.uint64)
    out = np.zeros_like(v)
    _check(lib().orc_naive_dft_gl(v, out, int(v.size).bit_length() - 1, int(inverse)), "naive_dft_gl")
    return out


def dft_points_gl(x: np.ndarray, log2n: int, ks, inverse: bool = False) -> np.ndarray:
    """Selected outputs X[k] of the textbook DFT (times n^-1 for the inverse)."""
    ks = np.ascontiguousarray(ks, dtype=np.uint64)
    out = np.zeros(ks.size, dtype=np.uint64)
    _check(lib().orc_dft_points_gl(np.ascontiguousarray(x, dtype=np.uint64), log2n, int(inverse), ks, ks.size, out),
           "dft_points_gl")
    return out


def dft_points_u32(x: np.ndarray, n: int, root: int, mod: int, ks, inverse: bool = False) -> np.ndarray:
    ks = np.ascontiguousarray(ks, dtype=np.uint64)
    out = np.zeros(ks.size, dtype=np.uint32)
    _check(lib().orc_dft_points_u32(np.ascontiguousarray(x, dtype=np.uint32), n, root, mod, int(inverse), ks,
                                    ks.size, out), "dft_points_u32")
    return out


def ntt(field: str, x: np.ndarray, log2n: int, batches: int = 1, inverse: bool = False,
        root: int | None = None, mod: int | None = None) -> np.ndarray:
    """Dispatch helper used by the parity tests and the bench baseline."""
    n = 1 << log2n
    if field == "u32":
        if root is None or mod is None:
            root, mod = find_params(n, 11)
        return inverse_u32(x, n, root, mod, batches) if inverse else stockham_u32(x, n, root, mod, batches)
    if field == "gl":
        return inverse_gl(x, log2n, batches) if inverse else stockham_gl(x, log2n, batches)
    raise ValueError(field)
