"""Bit-reversed element orders (nttCt / inttGs) and polynomial multiply
(pwcNtt, pwc_convolution, elemWiseMul) on the GPU, against the oracle.

The 32-bit oracle functions for these orders are pinned to the compiled
reference (tests/test_oracle.py); the Goldilocks ones to big-integer DFTs.
Sizes cover the single-pass kernel, two- and three-pass schedules, passes of
odd radix (two sub-step groups per thread on the permuted pass) and batches.
"""
from __future__ import annotations

import numpy as np
import pytest

import stntt_amd as S
from oracle import oracle as O
from stntt_amd import _lib

pytestmark = pytest.mark.gpu
P = S.GL_P


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


def bitrev_perm(L: int) -> np.ndarray:
    n = 1 << L
    idx = np.arange(n, dtype=np.uint64)
    r = np.zeros(n, dtype=np.uint64)
    for b in range(L):
        r |= ((idx >> np.uint64(b)) & np.uint64(1)) << np.uint64(L - 1 - b)
    return r.astype(np.int64)


def gl_input(L, batches, seed):
    return np.random.default_rng(seed).integers(0, P, size=batches << L, dtype=np.uint64)


def u32_input(L, batches, seed):
    root, mod = S.find_params(1 << L, 11)
    x = np.random.default_rng(seed).integers(0, mod, size=batches << L, dtype=np.uint64).astype(np.uint32)
    return x, root, mod


def gl_dft(x, L, batches, inverse=False):
    return O.inverse_gl(x, L, batches) if inverse else O.stockham_gl(x, L, batches)


def u32_dft(x, L, batches, root, mod, inverse=False):
    n = 1 << L
    return O.inverse_u32(x, n, root, mod, batches) if inverse else O.stockham_u32(x, n, root, mod, batches)


def permute(v, L, batches):
    perm = bitrev_perm(L)
    return v.reshape(batches, 1 << L)[:, perm].reshape(-1).copy()


ORDER_CASES = [(1, 1), (5, 2), (11, 3), (13, 1), (14, 2), (16, 1), (18, 2), (21, 1)]


@pytest.mark.parametrize("L,batches", ORDER_CASES)
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("rin,rout", [(False, True), (True, False), (True, True)])
def test_goldilocks_orders(ctx, L, batches, inverse, rin, rout):
    print(f"goldilocks L={L} batches={batches} inverse={inverse} rin={rin} rout={rout}", flush=True)
    x = gl_input(L, batches, 100 + L)
    src = permute(x, L, batches) if rin else x     # x sits at bit-reversed positions
    want = gl_dft(x, L, batches, inverse)
    if rout:
        want = permute(want, L, batches)
    buf = ctx.upload(src)
    ctx.ntt_goldilocks(buf, L, batches, inverse=inverse, bitrev_in=rin, bitrev_out=rout)
    assert np.array_equal(buf.copy_to(np.empty_like(src)), want)


@pytest.mark.parametrize("L,batches", ORDER_CASES)
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("rin,rout", [(False, True), (True, False), (True, True)])
def test_u32_orders(ctx, L, batches, inverse, rin, rout):
    print(f"u32 L={L} batches={batches} inverse={inverse} rin={rin} rout={rout}", flush=True)
    x, root, mod = u32_input(L, batches, 200 + L)
    src = permute(x, L, batches) if rin else x
    want = u32_dft(x, L, batches, root, mod, inverse)
    if rout:
        want = permute(want, L, batches)
    buf = ctx.upload(src)
    ctx.ntt_u32(buf, L, batches, root, mod, inverse=inverse, bitrev_in=rin, bitrev_out=rout)
    assert np.array_equal(buf.copy_to(np.empty_like(src)), want)


@pytest.mark.parametrize("field,L,split", [("goldilocks", 16, [5, 6, 5]), ("goldilocks", 17, [9, 8]),
                                          ("goldilocks", 19, [7, 5, 7]), ("u32", 17, [5, 7, 5]),
                                          ("u32", 19, [9, 5, 5]), ("u32", 20, [7, 6, 7])])
def test_orders_with_odd_radix_splits(ctx, field, L, split):
    """Odd radices put two sub-step groups on a thread in the permuted pass."""
    print(f"{field} L={L} split={split}", flush=True)
    S.set_pass_split(ctx, field, L, split)
    try:
        if field == "goldilocks":
            x = gl_input(L, 1, 300 + L)
            buf = ctx.upload(x)
            ctx.ntt_goldilocks(buf, L, 1, bitrev_out=True)
            X = buf.copy_to(np.empty_like(x))
            assert np.array_equal(X, permute(O.stockham_gl(x, L), L, 1))
            ctx.ntt_goldilocks(buf, L, 1, inverse=True, bitrev_in=True)
        else:
            x, root, mod = u32_input(L, 1, 300 + L)
            buf = ctx.upload(x)
            ctx.ntt_u32(buf, L, 1, root, mod, bitrev_out=True)
            X = buf.copy_to(np.empty_like(x))
            assert np.array_equal(X, O.ntt_ct_u32(x, root, mod))
            ctx.ntt_u32(buf, L, 1, root, mod, inverse=True, bitrev_in=True)
        assert np.array_equal(buf.copy_to(np.empty_like(x)), x)
    finally:
        S.restore_pass_split(ctx, field, L)


@pytest.mark.parametrize("L", [1, 4, 10, 14, 17])
def test_ntt_ct_intt_gs_dropins_match_reference_semantics(L):
    """int32 vectors through the reference-shaped calls, against the oracle's
    nttCt / inttGs (pinned to the compiled reference)."""
    n = 1 << L
    root, mod = S.find_params(n, 11)
    x = np.random.default_rng(L).integers(-1000, 1000, size=n).astype(np.int32)
    xm = (x.astype(np.int64) % mod).astype(np.uint32)
    ct = S.ntt_ct(x.copy(), root, mod)
    assert np.array_equal(ct.astype(np.uint32), O.ntt_ct_u32(xm, root, mod))
    gs = S.intt_gs(ct.copy(), root, mod)
    assert np.array_equal(gs.astype(np.uint32), O.intt_gs_u32(ct.astype(np.uint32), root, mod))
    assert np.array_equal(gs.astype(np.uint32), xm)


@pytest.mark.parametrize("L,batches", [(1, 1), (3, 2), (8, 1), (13, 3), (14, 1), (17, 2), (20, 1)])
def test_polymul_u32(ctx, L, batches):
    print(f"polymul u32 L={L} batches={batches}", flush=True)
    n = 1 << L
    root, mod = S.find_params(n, 11)
    rng = np.random.default_rng(40 + L)
    a = rng.integers(0, mod, size=batches * n, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, mod, size=batches * n, dtype=np.uint64).astype(np.uint32)
    da, db = ctx.upload(a), ctx.upload(b)
    ctx.polymul_u32(da, db, L, batches, root, mod)
    got_a, got_b = da.copy_to(np.empty_like(a)), db.copy_to(np.empty_like(b))
    for i in range(batches):
        sl = slice(i * n, (i + 1) * n)
        assert np.array_equal(got_a[sl], O.pwc_ntt_u32(a[sl], b[sl], root, mod)), i
        assert np.array_equal(got_b[sl], O.ntt_ct_u32(b[sl], root, mod)), i


@pytest.mark.parametrize("L,batches", [(1, 1), (3, 2), (8, 1), (12, 3), (13, 1), (16, 2), (20, 1)])
def test_polymul_goldilocks(ctx, L, batches):
    print(f"polymul goldilocks L={L} batches={batches}", flush=True)
    n = 1 << L
    rng = np.random.default_rng(50 + L)
    a = rng.integers(0, P, size=batches * n, dtype=np.uint64)
    b = rng.integers(0, P, size=batches * n, dtype=np.uint64)
    da, db = ctx.upload(a), ctx.upload(b)
    ctx.polymul_goldilocks(da, db, L, batches)
    got_a, got_b = da.copy_to(np.empty_like(a)), db.copy_to(np.empty_like(b))
    for i in range(batches):
        sl = slice(i * n, (i + 1) * n)
        assert np.array_equal(got_a[sl], O.pwc_ntt_gl(a[sl], b[sl], L)), i
        assert np.array_equal(got_b[sl], O.ntt_ct_gl(b[sl], L)), i


def test_polymul_side_stream_limit_and_two_caller_streams(ctx):
    """The largest multiplies whose forward pair overlaps (2^22 words per
    operand: Goldilocks 2^21 x 2, 32-bit 64 x 2^16), issued back to back on
    two callers' streams of one context (they share its side stream and its
    scratch), then a repeat on the first stream; every result exact."""
    other = S.Context(0)
    try:
        st = [ctx.stream, other.stream]
        L1, B1 = 21, 2
        n1 = 1 << L1
        rng = np.random.default_rng(91)
        a1 = rng.integers(0, P, size=B1 * n1, dtype=np.uint64)
        b1 = rng.integers(0, P, size=B1 * n1, dtype=np.uint64)
        L2, B2 = 16, 64
        n2 = 1 << L2
        root, mod = S.find_params(n2, 11)
        a2 = rng.integers(0, mod, size=B2 * n2, dtype=np.uint64).astype(np.uint32)
        b2 = rng.integers(0, mod, size=B2 * n2, dtype=np.uint64).astype(np.uint32)
        da1, db1, da2, db2 = ctx.upload(a1), ctx.upload(b1), ctx.upload(a2), ctx.upload(b2)
        ctx.polymul_goldilocks(da1, db1, L1, B1, stream=st[0])
        ctx.polymul_u32(da2, db2, L2, B2, root, mod, stream=st[1])
        ctx.sync(st[0])
        ctx.sync(st[1])
        want_a1 = np.concatenate([O.pwc_ntt_gl(a1[i * n1:(i + 1) * n1], b1[i * n1:(i + 1) * n1], L1)
                                  for i in range(B1)])
        want_b1 = np.concatenate([O.ntt_ct_gl(b1[i * n1:(i + 1) * n1], L1) for i in range(B1)])
        want_a2 = np.concatenate([O.pwc_ntt_u32(a2[i * n2:(i + 1) * n2], b2[i * n2:(i + 1) * n2], root, mod)
                                  for i in range(B2)])
        want_b2 = np.concatenate([O.ntt_ct_u32(b2[i * n2:(i + 1) * n2], root, mod) for i in range(B2)])
        assert np.array_equal(da1.copy_to(np.empty_like(a1)), want_a1)
        assert np.array_equal(db1.copy_to(np.empty_like(b1)), want_b1)
        assert np.array_equal(da2.copy_to(np.empty_like(a2)), want_a2)
        assert np.array_equal(db2.copy_to(np.empty_like(b2)), want_b2)
        # again on the first stream, from fresh inputs
        da1.copy_from(a1)
        db1.copy_from(b1)
        ctx.polymul_goldilocks(da1, db1, L1, B1, stream=st[0])
        ctx.sync(st[0])
        assert np.array_equal(da1.copy_to(np.empty_like(a1)), want_a1)
        for d in (da1, db1, da2, db2):
            d.free()
    finally:
        other.close()


def test_polymul_goldilocks_full_size_shift(ctx):
    """At 2^24, multiplying by z^s is a cyclic shift by s: a size-independent
    exact check of forward, fused pointwise product and inverse."""
    L, s = 24, 12345
    n = 1 << L
    a = np.random.default_rng(7).integers(0, P, size=n, dtype=np.uint64)
    b = np.zeros(n, dtype=np.uint64)
    b[s] = 1
    da, db = ctx.upload(a), ctx.upload(b)
    ctx.polymul_goldilocks(da, db, L, 1)
    assert np.array_equal(da.copy_to(np.empty_like(a)), np.roll(a, s))


def test_pwc_dropins_match_schoolbook():
    rng = np.random.default_rng(3)
    for L in (1, 5, 9):
        n = 1 << L
        s0 = rng.integers(1, 50, size=n).astype(np.int32)
        s1 = rng.integers(1, 50, size=n).astype(np.int32)
        vec, root, mod = S.pwc_convolution(s0, s1)
        want = np.zeros(n, dtype=np.int64)
        for i in range(n):
            want += s0[i].astype(np.int64) * np.roll(s1.astype(np.int64), i)
        assert np.array_equal(vec.astype(np.int64), want)           # no output wraps the modulus
        assert (root, mod) == S.find_params(n, int(max(s0.max(), s1.max())) ** 2 * n + 1)
    v0 = rng.integers(0, 1000, size=64).astype(np.int32)
    v1 = rng.integers(0, 1000, size=64).astype(np.int32)
    root, mod = S.find_params(64, 11)
    w0, w1 = v0.copy(), v1.copy()
    S.pwc_ntt(w0, w1, root, mod)
    assert np.array_equal(w0.astype(np.uint32), O.pwc_ntt_u32(v0.astype(np.uint32), v1.astype(np.uint32), root, mod))
    assert np.array_equal(w1.astype(np.uint32), O.ntt_ct_u32(v1.astype(np.uint32), root, mod))


def test_elem_wise_mul(ctx):
    rng = np.random.default_rng(4)
    for mod in (17, 7340033, 2013265921):
        a = rng.integers(0, mod, size=5000, dtype=np.uint64).astype(np.uint32)
        b = rng.integers(0, mod, size=6000, dtype=np.uint64).astype(np.uint32)
        want = (a.astype(object) * b[:5000].astype(object) % mod).astype(np.uint32)
        assert np.array_equal(S.elem_wise_mul(a.copy(), b, mod, ctx), want)
    a = rng.integers(0, P, size=3000, dtype=np.uint64)
    b = rng.integers(0, P, size=3000, dtype=np.uint64)
    want = np.array([int(x) * int(y) % P for x, y in zip(a, b)], dtype=np.uint64)
    assert np.array_equal(S.elem_wise_mul(a.copy(), b, ctx=ctx), want)


def test_order_and_polymul_errors(ctx):
    buf = ctx.alloc(8 << 10)
    arr = (_lib.Desc * 1)()
    arr[0] = _lib.Desc(1, 0, 10, _lib.BITREV_OUT, 0, 0, buf.ptr, buf.ptr)
    import ctypes
    h = ctypes.c_void_p()
    with pytest.raises(ValueError):
        _lib.check(_lib.lib().stntt_batch_create(ctx.handle, arr, 1, ctypes.byref(h)))
    with pytest.raises(ValueError):
        ctx.polymul_goldilocks(buf, buf, 10, 1)            # a and b alias
    with pytest.raises(ValueError):
        _lib.check(_lib.lib().stntt_ntt_goldilocks_ex(ctx.handle, buf.ptr, 10, 1, 0, 64, None))  # unknown flag
    with pytest.raises(ValueError):
        S.pwc_convolution(np.array([60000, 1], dtype=np.int32), np.array([1, 1], dtype=np.int32))  # p >= 2^31
    buf.free()
