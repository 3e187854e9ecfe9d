"""nttStockham with nsteps and nttStockhamIdx (src/ntt/ntt_cpu.cpp:46-127) on
the device: the reference's radix-2 stage loop, truncated after nsteps
stages, against the oracle's restatement (pinned to the reference's own code
for nsteps past log2 n in tests/test_oracle.py) and the reference's fixtures.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import stntt_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("L", list(range(0, 14)) + [16, 20])
def test_truncated_stockham_matches_oracle(ctx, L):
    n = 1 << L
    root, mod = S.find_params(n, 11)
    rng = np.random.default_rng(50 + L)
    for batches in (1, 3):
        x = rng.integers(0, mod, size=batches * n, dtype=np.int64).astype(np.int32)
        for k in sorted({0, 1, L // 2, L, L + 1, L + 3}):
            got = S.ntt_stockham(x.copy(), n, root, mod, batches, ctx=ctx, nsteps=k)
            want = O.stockham_u32(x.astype(np.uint32), n, root, mod, batches, k)
            assert np.array_equal(got.astype(np.uint32), want), (L, batches, k)
    # every stage: the full transform of the multi-pass path
    full = S.ntt_stockham(x.copy(), n, root, mod, batches, ctx=ctx)
    assert np.array_equal(S.ntt_stockham(x.copy(), n, root, mod, batches, ctx=ctx, nsteps=L), full)


def test_truncated_stockham_fixture_and_signed(ctx):
    ref = np.load(os.path.join(GOLD, "ref_u32.npz"))
    n = 256
    root, mod = S.find_params(n, 11)
    x = ref["nsteps_in"]
    for s, want in enumerate(ref["nsteps_out"]):
        got = S.ntt_stockham(x.copy(), n, root, mod, ctx=ctx, nsteps=s)
        assert np.array_equal(got, want), s
    # signed ints take the reference's modulo() semantics
    v = np.random.default_rng(3).integers(-2**31, 2**31, size=4 * n, dtype=np.int64).astype(np.int32)
    red = (v.astype(np.int64) % mod).astype(np.int32)
    for k in (1, 5, 8):
        assert np.array_equal(S.ntt_stockham(v.copy(), n, root, mod, 4, ctx=ctx, nsteps=k),
                              S.ntt_stockham(red.copy(), n, root, mod, 4, ctx=ctx, nsteps=k)), k


@pytest.mark.parametrize("L", [0, 1, 4, 7, 10, 13, 18])
def test_stockham_idx(ctx, L):
    n = 1 << L
    x = np.arange(2 * n, dtype=np.int32)
    got = S.ntt_stockham_idx(x.copy(), n, 0, 2**31 - 1, 2, ctx=ctx)
    brev = np.array([O.lib().orc_bit_reverse(i, L) for i in range(n)], dtype=np.int64)
    want = np.concatenate([brev, brev + n]).astype(np.int32)
    assert np.array_equal(got, want)
    if L in (1, 4, 7, 10):
        ref = np.load(os.path.join(GOLD, "ref_u32.npz"))
        assert np.array_equal(got[:n], ref[f"idx_L{L}"])
    for mod in (17, 257, 7681, 2**31 - 1):
        for k in sorted({0, 1, L // 2, L + 2, L + 9}):
            g = S.ntt_stockham_idx(x.copy(), n, 0, mod, 2, nsteps=k, ctx=ctx)
            w = np.concatenate([O.stockham_idx(x[:n], n, k, mod), O.stockham_idx(x[n:], n, k, mod)])
            assert np.array_equal(g, w), (L, mod, k)


def test_stage_loop_errors(ctx):
    v = np.zeros(16, dtype=np.int32)
    with pytest.raises(ValueError):
        S.ntt_stockham(v, 16, 3, 17, ctx=ctx, nsteps=31)        # above the stage limit
    with pytest.raises(ValueError):
        S.ntt_stockham_idx(v, 16, 0, 2, nsteps=3, ctx=ctx)      # pow(2, 1, 2) = 0: the reference divides by it
    with pytest.raises(ValueError):
        S.ntt_stockham(np.zeros(12, dtype=np.int32), 12, 3, 13, ctx=ctx, nsteps=2)  # not a power of two
    w = np.arange(16, dtype=np.int32)
    assert np.array_equal(S.ntt_stockham(w.copy(), 16, 3, 17, ctx=ctx, nsteps=0), w)  # no stage: untouched
