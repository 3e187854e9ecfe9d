"""The large-radix pass (ntt_big.hpp: radix 2^10..2^12, 128 KiB tiles of
Goldilocks words or 64 KiB of 32-bit words, persistent workgroups in
XCD-aware tile order) against the oracle: both fields, both directions,
batches, every tile width it accepts, and schedules that mix it with the
128-byte-row tile passes as first, middle and last pass."""
from __future__ import annotations

import numpy as np
import pytest

import stntt_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
P = S.GL_P


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


def inputs(field, L, batches, seed):
    n = 1 << L
    rng = np.random.default_rng(seed)
    if field == "goldilocks":
        return rng.integers(0, 2**64, size=batches * n, dtype=np.uint64), 0, P
    root, mod = S.find_params(n, 11)
    return rng.integers(0, 2**32, size=batches * n, dtype=np.uint64).astype(np.uint32), root, mod


def run(ctx, field, x, L, batches, root, mod, inverse):
    buf = ctx.upload(x)
    if field == "goldilocks":
        ctx.ntt_goldilocks(buf, L, batches, inverse=inverse)
    else:
        ctx.ntt_u32(buf, L, batches, root, mod, inverse=inverse)
    out = buf.copy_to(np.empty_like(x))
    buf.free()
    return out


def want(field, x, L, batches, root, mod, inverse):
    if field == "goldilocks":
        return O.inverse_gl(x, L, batches) if inverse else O.stockham_gl(x, L, batches)
    n = 1 << L
    x = x % np.uint32(mod)
    return O.inverse_u32(x, n, root, mod, batches) if inverse else O.stockham_u32(x, n, root, mod, batches)


CASES = [  # (L, split, logc per pass: -1 = the library's choice)
    (20, [10, 10], [-1, -1]), (21, [10, 11], [-1, -1]), (21, [11, 10], [-1, -1]), (22, [11, 11], [-1, -1]),
    (23, [11, 12], [-1, -1]), (24, [12, 12], [-1, -1]), (22, [11, 11], [1, 2]), (24, [12, 12], [0, 1]),
    (20, [10, 10], [2, 4]), (22, [6, 12, 4], [-1, -1, -1]), (22, [10, 6, 6], [-1, -1, -1]),
    (21, [5, 6, 10], [-1, -1, -1]), (20, [4, 12, 4], [-1, 1, -1]),
]


@pytest.mark.parametrize("field", ["goldilocks", "u32"])
@pytest.mark.parametrize("L,split,logc", CASES)
def test_big_passes_match_oracle(ctx, field, L, split, logc):
    if field == "u32":  # 64 KiB tiles: two fewer columns at the same radix
        logc = [c if c < 0 else min(c, 14 - r) for c, r in zip(logc, split)]
    print(f"{field} L={L} split={split} logc={logc}", flush=True)
    S.set_pass_shape(ctx, field, L, split, logc)
    try:
        for inverse in (False, True):
            x, root, mod = inputs(field, L, 1, 1000 + L + inverse)
            got = run(ctx, field, x, L, 1, root, mod, inverse)
            assert np.array_equal(got, want(field, x, L, 1, root, mod, inverse)), (inverse, split, logc)
    finally:
        S.restore_pass_split(ctx, field, L)


@pytest.mark.parametrize("field", ["goldilocks", "u32"])
def test_big_passes_in_batches(ctx, field):
    """batches > 1 take the fewest-pass default; forcing the split through a
    batch-sized call is done with a two-pass schedule picked per size"""
    for L, batches in [(20, 3), (22, 2)]:
        x, root, mod = inputs(field, L, batches, 77 + L)
        got = run(ctx, field, x, L, batches, root, mod, False)
        assert np.array_equal(got, want(field, x, L, batches, root, mod, False)), (L, batches)


def test_shape_validation(ctx):
    with pytest.raises(ValueError):
        S.set_pass_shape(ctx, "goldilocks", 24, [12, 12], [5, -1])   # wider than the 128 KiB tile
    with pytest.raises(ValueError):
        S.set_pass_shape(ctx, "goldilocks", 14, [12, 2], [-1, -1])   # radix 4 pass
    with pytest.raises(ValueError):
        S.set_pass_shape(ctx, "goldilocks", 14, [4, 10], [-1, 5])    # 32 columns of 1024 rows: over 2^14 words
    with pytest.raises(ValueError):
        S.set_pass_shape(ctx, "goldilocks", 22, [11, 11], [3, 0])    # a quarter of the tile is the narrowest
