"""The 64-bit prime field (STNTT_FIELD_U64, Montgomery arithmetic) on the
GPU against the oracle's generic-modulus restatements (pinned in
tests/test_field64_host.py), and against the Goldilocks kernels word for
word when the modulus is P.  Primes below 2^32, between 2^32 and 2^63, above
2^63 and just below 2^64; inputs are arbitrary 64-bit words (reduced exactly).
"""
from __future__ import annotations

import numpy as np
import pytest

import stntt_amd as S
from oracle import oracle as O
from stntt_amd import _lib

pytestmark = pytest.mark.gpu
P = S.GL_P
MIN_MODS = [2**31, 2**40 + 7, 2**63 + 1, 2**64 - 2**36]


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


def words(n, seed):
    return np.random.default_rng(seed).integers(0, 2**64, size=n, dtype=np.uint64)


@pytest.mark.parametrize("L,batches", [(1, 1), (4, 3), (9, 2), (13, 1), (14, 2), (17, 1), (20, 1)])
@pytest.mark.parametrize("min_mod", MIN_MODS)
@pytest.mark.parametrize("inverse", [False, True])
def test_u64_matches_oracle(ctx, L, batches, min_mod, inverse):
    n = 1 << L
    root, mod = S.find_params(n, min_mod)
    print(f"u64 L={L} batches={batches} mod={mod} inverse={inverse}", flush=True)
    x = words(batches * n, L + batches)
    want = O.inverse_u64(x, n, root, mod, batches) if inverse else O.stockham_u64(x, n, root, mod, batches)
    buf = ctx.upload(x)
    ctx.ntt_u64(buf, L, batches, root, mod, inverse=inverse)
    assert np.array_equal(buf.copy_to(np.empty_like(x)), want)


@pytest.mark.parametrize("split", [[9, 5, 6], [5, 9, 6], [6, 5, 9]])
@pytest.mark.parametrize("inverse", [False, True])
def test_u64_radix512_passes(ctx, split, inverse):
    """Radix-512 tiles (exchange halves, the per-tile LDS root table) as the
    first, a middle and the last pass, above 2^63."""
    L = sum(split)
    n = 1 << L
    root, mod = S.find_params(n, 2**63 + 1)
    x = words(n, 77 + split[0])
    want = O.inverse_u64(x, n, root, mod) if inverse else O.stockham_u64(x, n, root, mod)
    S.set_pass_split(ctx, "u64", L, split)
    try:
        buf = ctx.upload(x)
        ctx.ntt_u64(buf, L, 1, root, mod, inverse=inverse)
        assert np.array_equal(buf.copy_to(np.empty_like(x)), want), split
        buf.free()
    finally:
        S.restore_pass_split(ctx, "u64", L)


@pytest.mark.parametrize("split", [[9, 5, 6], [6, 5, 9]])
def test_u64_radix512_orders_and_polymul(ctx, split):
    """Radix-512 first / last passes that stage the tile for bit-reversed
    input (with and without the fused pointwise operand) and output."""
    L = sum(split)
    n = 1 << L
    root, mod = S.find_params(n, 2**63 + 1)
    print(f"u64 radix-512 orders/polymul split={split}", flush=True)
    x = words(n, 300 + split[0]) % np.uint64(mod)
    perm = bitrev_perm(L)
    want = O.stockham_u64(x, n, root, mod)
    S.set_pass_split(ctx, "u64", L, split)
    try:
        buf = ctx.upload(x[perm].copy())
        ctx.ntt_u64(buf, L, 1, root, mod, bitrev_in=True)
        assert np.array_equal(buf.copy_to(np.empty_like(x)), want), "bit-reversed input"
        buf.copy_from(x)
        ctx.ntt_u64(buf, L, 1, root, mod, bitrev_out=True)
        assert np.array_equal(buf.copy_to(np.empty_like(x)), want[perm]), "bit-reversed output"
        buf.free()
        a, b = words(n, 11) % np.uint64(mod), words(n, 12) % np.uint64(mod)
        da, db = ctx.upload(a), ctx.upload(b)
        ctx.polymul_u64(da, db, L, 1, root, mod)
        assert np.array_equal(da.copy_to(np.empty_like(a)), O.pwc_ntt_u64(a, b, root, mod)), "polymul"
        da.free()
        db.free()
    finally:
        S.restore_pass_split(ctx, "u64", L)


@pytest.mark.parametrize("L", [12, 20, 24])
def test_u64_at_goldilocks_equals_goldilocks_kernels(ctx, L):
    """Montgomery arithmetic at p = P against the special-form Goldilocks
    kernels: identical words at every size, including 2^24."""
    print(f"u64 vs goldilocks L={L}", flush=True)
    n = 1 << L
    root, mod = S.find_params(n, P)
    assert mod == P
    x = np.random.default_rng(L).integers(0, P, size=n, dtype=np.uint64)
    a, b = ctx.upload(x), ctx.upload(x)
    ctx.ntt_u64(a, L, 1, root, mod)
    ctx.ntt_goldilocks(b, L, 1)
    got = a.copy_to(np.empty_like(x))
    assert np.array_equal(got, b.copy_to(np.empty_like(x)))
    ctx.ntt_u64(a, L, 1, root, mod, inverse=True)
    assert np.array_equal(a.copy_to(np.empty_like(x)), x)


@pytest.mark.parametrize("L", [5, 13, 16, 19])
@pytest.mark.parametrize("rin,rout", [(False, True), (True, False)])
def test_u64_orders(ctx, L, rin, rout):
    n = 1 << L
    root, mod = S.find_params(n, 2**63 + 1)
    print(f"u64 orders L={L} rin={rin} rout={rout}", flush=True)
    x = words(n, 100 + L) % np.uint64(mod)
    perm = bitrev_perm(L)
    src = x[perm].copy() if rin else x
    want = O.stockham_u64(x, n, root, mod)
    if rout:
        want = want[perm]
    buf = ctx.upload(src)
    ctx.ntt_u64(buf, L, 1, root, mod, bitrev_in=rin, bitrev_out=rout)
    assert np.array_equal(buf.copy_to(np.empty_like(src)), want)


@pytest.mark.parametrize("L,min_mod", [(3, 2**40), (12, 2**63 + 1), (17, 2**50), (20, 2**62)])
def test_u64_polymul(ctx, L, min_mod):
    n = 1 << L
    root, mod = S.find_params(n, min_mod)
    print(f"u64 polymul L={L} mod={mod}", flush=True)
    a, b = words(n, 1) % np.uint64(mod), words(n, 2) % np.uint64(mod)
    da, db = ctx.upload(a), ctx.upload(b)
    ctx.polymul_u64(da, db, L, 1, root, mod)
    assert np.array_equal(da.copy_to(np.empty_like(a)), O.pwc_ntt_u64(a, b, root, mod))
    assert np.array_equal(db.copy_to(np.empty_like(b)), O.ntt_ct_u64(b, root, mod))


def test_u64_pointwise_any_words(ctx):
    mod = S.find_params(1 << 10, 2**63 + 1)[1]
    a, b = words(4000, 5), words(4000, 6)
    da, db = ctx.upload(a), ctx.upload(b)
    ctx.pointwise_u64(da, db, a.size, mod)
    want = np.array([int(x) * int(y) % mod for x, y in zip(a, b)], dtype=np.uint64)
    assert np.array_equal(da.copy_to(np.empty_like(a)), want)


@pytest.mark.parametrize("L,min_mod", [(10, 2**40 + 7), (18, 2**63 + 1), (23, 2**62)])
def test_device_generated_twiddles(ctx, L, min_mod):
    """Plans are generated on the device; the exported logical table equals
    the oracle's host power loop (and the inverse table its inverse)."""
    root, mod = S.find_params(1 << L, min_mod)
    assert np.array_equal(ctx.twiddles_u64(L, root, mod), O.twiddles_u64(1 << L, root, mod))
    iroot = pow(root, -1, mod)
    assert np.array_equal(ctx.twiddles_u64(L, root, mod, inverse=True), O.twiddles_u64(1 << L, iroot, mod))


def test_u64_errors(ctx):
    buf = ctx.alloc(8 << 10)
    root, mod = S.find_params(1 << 10, 2**40)
    with pytest.raises(ValueError):
        ctx.ntt_u64(buf, 10, 1, root, mod + 1)                 # even modulus
    with pytest.raises(ValueError):
        ctx.ntt_u64(buf, 10, 1, 3, mod)                        # 3 is not of order 2^10
    with pytest.raises(ValueError):
        ctx.ntt_u64(buf, 33, 1, root, mod)                     # above 2^32
    with pytest.raises(ValueError):
        S.Batch(ctx, [("u64", False, 10, buf, root, mod)])     # batches take the two named fields
    buf.free()
