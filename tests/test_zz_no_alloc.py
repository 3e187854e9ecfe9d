"""The allocation-free Goldilocks entry point (stntt_ntt_goldilocks_noalloc)
against stntt_ntt_goldilocks_ex, through the C ABI: two calls in a row, the
second on a smaller input, each result bit-identical to the allocating
entry point's and to the oracle; both directions and the order flags;
sizes past the fixed scratch are rejected."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from stntt_amd import _lib

pytestmark = pytest.mark.gpu
P = 0xFFFFFFFF00000001


def run(ctx, fn, x, L, batches, direction, flags=0):
    L_ = _lib.lib()
    d = _lib._c.c_void_p()
    _lib.check(L_.stntt_dev_alloc(ctx, x.nbytes, _lib._c.byref(d)))
    try:
        _lib.check(L_.stntt_memcpy_htod(ctx, d, x.ctypes.data, x.nbytes, None))
        _lib.check(getattr(L_, fn)(ctx, d, L, batches, direction, flags, None))
        _lib.check(L_.stntt_stream_sync(ctx, None))
        out = np.empty_like(x)
        _lib.check(L_.stntt_memcpy_dtoh(ctx, out.ctypes.data, d, x.nbytes, None))
    finally:
        L_.stntt_dev_free(ctx, d)
    return out


@pytest.fixture(scope="module")
def ctx():
    h = _lib._c.c_void_p()
    _lib.check(_lib.lib().stntt_ctx_create(0, _lib._c.byref(h)))
    yield h
    _lib.lib().stntt_ctx_destroy(h)


@pytest.mark.parametrize("sizes", [[(23, 1), (18, 16)], [(22, 1), (12, 3)], [(18, 16), (16, 2)], [(20, 2), (9, 5)],
                                   [(13, 1), (5, 2)]])
def test_noalloc_matches_allocating_entry_point(ctx, sizes):
    rng = np.random.default_rng(sum(L for L, _ in sizes))
    for L, batches in sizes:  # two calls in a row, the second smaller
        x = rng.integers(0, 2**64, size=batches << L, dtype=np.uint64)
        for direction in (0, 1):
            print(f"noalloc L={L} batches={batches} dir={direction}", flush=True)
            got = run(ctx, "stntt_ntt_goldilocks_noalloc", x, L, batches, direction)
            want = run(ctx, "stntt_ntt_goldilocks_ex", x, L, batches, direction)
            assert np.array_equal(got, want), (L, batches, direction)
            if L <= 20:
                ref = O.inverse_gl(x, L, batches) if direction else O.stockham_gl(x, L, batches)
                assert np.array_equal(got, ref), (L, batches, direction)


def test_noalloc_orders(ctx):
    rng = np.random.default_rng(7)
    for L, flags in [(17, _lib.BITREV_OUT), (15, _lib.BITREV_IN), (11, _lib.BITREV_IN | _lib.BITREV_OUT)]:
        x = rng.integers(0, P, size=1 << L, dtype=np.uint64)
        got = run(ctx, "stntt_ntt_goldilocks_noalloc", x, L, 1, 0, flags)
        want = run(ctx, "stntt_ntt_goldilocks_ex", x, L, 1, 0, flags)
        assert np.array_equal(got, want), (L, flags)


def test_noalloc_rejects_sizes_past_its_scratch(ctx):
    L_ = _lib.lib()
    # 2^24 words need 128 MiB or more of scratch: rejected before any device work
    assert L_.stntt_ntt_goldilocks_noalloc(ctx, _lib._c.c_void_p(1 << 20), 24, 1, 0, 0, None) == _lib.STNTT_EINVAL
    assert b"allocation-free" in L_.stntt_last_error()
    assert L_.stntt_ntt_goldilocks_noalloc(ctx, _lib._c.c_void_p(1 << 20), 29, 1, 0, 0, None) == _lib.STNTT_EINVAL
    # 2^23 fits in natural order (three passes over the data and one 64 MiB
    # scratch) but not with bit-reversed output, whose odd pass count needs
    # two scratch buffers
    assert L_.stntt_ntt_goldilocks_noalloc(ctx, _lib._c.c_void_p(1 << 20), 23, 1, 0, _lib.BITREV_OUT,
                                           None) == _lib.STNTT_EINVAL
