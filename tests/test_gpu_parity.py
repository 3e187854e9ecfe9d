"""Parity of the HIP path (libstntt.so through its C ABI) with the oracle.

Bit-exact for every integer output.  Small and medium sizes compare whole
outputs with the oracle; 2^23..2^27 use size-independent properties: round
trip, X[0] = sum x, sampled outputs against the textbook DFT.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import stntt_amd as S
from oracle import oracle as O
from tests.golden.gen import digits, splitmix64

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
P = S.GL_P


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


def u32_input(seed: int, count: int, mod: int) -> np.ndarray:
    return (splitmix64(seed, count) % np.uint64(mod)).astype(np.uint32)


def gl_input(seed: int, count: int) -> np.ndarray:
    # full 64-bit words: values >= P are valid (non-canonical) inputs
    return splitmix64(seed, count)


def run_u32(ctx, x: np.ndarray, L: int, batches: int, root: int, mod: int, inverse: bool = False) -> np.ndarray:
    buf = ctx.upload(x)
    ctx.ntt_u32(buf, L, batches, root, mod, inverse=inverse)
    out = buf.copy_to(np.empty_like(x))
    buf.free()
    return out


def run_gl(ctx, x: np.ndarray, L: int, batches: int, inverse: bool = False) -> np.ndarray:
    buf = ctx.upload(x)
    ctx.ntt_goldilocks(buf, L, batches, inverse=inverse)
    out = buf.copy_to(np.empty_like(x))
    buf.free()
    return out


# ----------------------------------------------------------------------------
# whole-output parity
# ----------------------------------------------------------------------------
SMALL = [(L, b) for L in range(1, 13) for b in (1, 2, 5, 10)] + [(L, b) for L in range(13, 23) for b in (1, 2)]


@pytest.mark.parametrize("L,batches", SMALL)
def test_u32_forward_inverse(ctx, L, batches):
    n = 1 << L
    root, mod = S.find_params(n, 11)
    x = u32_input(L * 17 + batches, n * batches, mod)
    X = run_u32(ctx, x, L, batches, root, mod)
    assert np.array_equal(X, O.stockham_u32(x, n, root, mod, batches)), f"forward L={L} b={batches}"
    y = run_u32(ctx, X, L, batches, root, mod, inverse=True)
    assert np.array_equal(y, x), f"round trip L={L} b={batches}"


@pytest.mark.parametrize("L,batches", SMALL)
def test_goldilocks_forward_inverse(ctx, L, batches):
    n = 1 << L
    x = gl_input(L * 31 + batches, n * batches)
    X = run_gl(ctx, x, L, batches)
    assert np.array_equal(X, O.stockham_gl(x, L, batches)), f"forward L={L} b={batches}"
    assert np.all(X < np.uint64(P))
    y = run_gl(ctx, X, L, batches, inverse=True)
    assert np.array_equal(y, x % np.uint64(P)), f"round trip L={L} b={batches}"
    Z = run_gl(ctx, x, L, batches, inverse=True)
    assert np.array_equal(Z, O.inverse_gl(x, L, batches)), f"inverse L={L} b={batches}"


@pytest.mark.parametrize("L", [3, 9, 14, 17, 20])
def test_u32_inverse_of_arbitrary_input(ctx, L):
    n = 1 << L
    root, mod = S.find_params(n, 11)
    x = u32_input(L + 900, 2 * n, mod)
    assert np.array_equal(run_u32(ctx, x, L, 2, root, mod, inverse=True), O.inverse_u32(x, n, root, mod, 2))


@pytest.mark.parametrize("L", range(1, 14))
def test_reference_golden_vectors(ctx, L):
    """The reference's own nttStockham outputs (tests/golden/ref_u32.npz)."""
    g = np.load(os.path.join(GOLD, "ref_u32.npz"))
    n = 1 << L
    root, mod = S.find_params(n, 11)
    for b in (1, 2):
        x = g[f"st_in_L{L}_b{b}"]
        want = g[f"st_out_L{L}_b{b}"]
        v = x.copy()
        S.ntt_stockham(v, n, root, mod, batches=b, ctx=ctx)
        assert np.array_equal(v, want), (L, b)


def test_goldilocks_bigint_fixtures(ctx):
    g = np.load(os.path.join(GOLD, "gl_naive.npz"))
    for L in range(1, 7):
        assert np.array_equal(run_gl(ctx, g[f"x_L{L}"], L, 1), g[f"X_L{L}"]), L


# ----------------------------------------------------------------------------
# the reference's test protocol: st_ntt_test.cu:16-97 (inputs 0..9,
# batches {1,2,5,10}, block sizes {64..1024}) through NttArgs / auto_ntt
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("L", list(range(2, 21)))
def test_reference_protocol_auto_ntt(ctx, L):
    n = 1 << L
    root, mod = S.find_params(n, 11)
    vec = digits(42, n * 10)
    for batches in (1, 2, 5, 10):
        for bs in (64, 1024) if L > 6 else (64, 128, 256, 512, 1024):
            cpu = O.stockham_u32(vec[: n * batches].astype(np.uint32), n, root, mod, batches)
            gpu = vec[: n * batches].copy()
            ms = S.auto_ntt(S.NttArgs(gpu, n, batches, root, mod, 2, bs), ctx)
            assert ms >= 0.0
            assert np.array_equal(gpu.astype(np.uint32), cpu), (L, batches, bs)
            if L <= 13:
                assert np.array_equal(gpu, O.stockham_refcompat(vec[: n * batches], n, root, mod, batches))


def test_signed_inputs_follow_reference_modulo(ctx):
    """int inputs, negatives included, reduce like modulo() (src/ntt/ntt_util.cpp:88-94)."""
    for L in (4, 11, 16):
        n = 1 << L
        root, mod = S.find_params(n, 11)
        raw = splitmix64(L + 5, 2 * n).view(np.int64)[: 2 * n]
        x = (raw >> np.int64(32)).astype(np.int32)  # full int32 range
        canon = np.mod(x.astype(np.int64), mod).astype(np.uint32)
        want = O.stockham_u32(canon, n, root, mod, 2)
        got = x.copy()
        S.ntt_stockham(got, n, root, mod, batches=2, ctx=ctx)
        assert np.array_equal(got.astype(np.uint32), want), L
        if L <= 13 and O.ref_available():
            # the reference's first stage adds raw ints (src/ntt/ntt_cpu.cpp:73), which wraps for
            # full-range inputs; with |x| < 2^29 it is exact and must agree
            small = (x >> np.int32(3)).astype(np.int32)
            r = small.copy()
            assert O.ref().ref_ntt_stockham(r, n, root, mod, 2, -1) == 0
            g2 = small.copy()
            S.ntt_stockham(g2, n, root, mod, batches=2, ctx=ctx)
            assert np.array_equal(g2, r), L


def test_size_one_is_identity(ctx):
    v = np.array([7, -3, 2**31 - 1], dtype=np.int32)
    w = v.copy()
    S.ntt_stockham(w, 1, 1, 11, batches=3, ctx=ctx)
    assert np.array_equal(w, v)
    g = np.array([P + 5, 3], dtype=np.uint64)
    h = g.copy()
    S.ntt_goldilocks(h, 1, batches=2, ctx=ctx)
    assert np.array_equal(h, g)


# ----------------------------------------------------------------------------
# twiddle table in the reference layout
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("L", [1, 5, 12, 16, 20])
def test_twiddle_layout(ctx, L):
    n = 1 << L
    root, mod = S.find_params(n, 11)
    assert np.array_equal(ctx.twiddles_u32(n, root, mod), O.twiddles_u32(n, root, mod))
    assert np.array_equal(ctx.twiddles_goldilocks(L), O.twiddles_gl(L))
    assert np.array_equal(ctx.twiddles_goldilocks(L, inverse=True), O.twiddles_gl(L, inverse=True))


# ----------------------------------------------------------------------------
# large sizes: properties that do not need a full oracle run
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("L", [23, 24, 26])
def test_goldilocks_large(ctx, L):
    n = 1 << L
    x = gl_input(1000 + L, n)
    X = run_gl(ctx, x, L, 1)
    ks = np.array([0, 1, 2, 3, n // 2, n - 1, 12345 % n, (n * 5) // 7], dtype=np.uint64)
    assert np.array_equal(X[ks], O.dft_points_gl(x, L, ks))
    assert np.all(X < np.uint64(P))
    y = run_gl(ctx, X, L, 1, inverse=True)
    assert np.array_equal(y, x % np.uint64(P))
    if L <= 24:
        assert np.array_equal(X, O.stockham_gl(x, L))


def test_goldilocks_beyond_4gib(ctx):
    """2^30 Goldilocks words (8 GiB): tile offsets past 2^32 bytes.  Sampled
    outputs against the O(n)-per-point DFT, and the round trip."""
    L = 30
    n = 1 << L
    print("goldilocks 2^30: building input", flush=True)
    x = np.arange(n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    buf = ctx.upload(x)
    ctx.ntt_goldilocks(buf, L, 1)
    ks = np.array([1, n // 2 + 3, n - 1], dtype=np.uint64)
    X = buf.copy_to(np.empty_like(x))
    print("goldilocks 2^30: sampled points", flush=True)
    assert np.array_equal(X[ks], O.dft_points_gl(x, L, ks))
    del X
    ctx.ntt_goldilocks(buf, L, 1, inverse=True)
    y = buf.copy_to(np.empty_like(x))
    assert np.array_equal(y, x % np.uint64(P))
    buf.free()


@pytest.mark.parametrize("L", [23, 24, 27])
def test_u32_large(ctx, L):
    n = 1 << L
    root, mod = S.find_params(n, 11)
    x = u32_input(2000 + L, n, mod)
    X = run_u32(ctx, x, L, 1, root, mod)
    ks = np.array([0, 1, 5, n // 2, n - 1, (n * 3) // 11], dtype=np.uint64)
    assert np.array_equal(X[ks], O.dft_points_u32(x, n, root, mod, ks))
    assert np.array_equal(run_u32(ctx, X, L, 1, root, mod, inverse=True), x)
    if L <= 24:
        assert np.array_equal(X, O.stockham_u32(x, n, root, mod))


# The 32-bit field's twiddles are Montgomery words (REDC with -p^-1 mod 2^32):
# moduli at both ends of its range, through every kernel family (single-pass,
# tile passes, large-radix passes, the batch queue).  Below 2^30 the pass
# kernels keep words in [0, 2p) (lazy reduction), from 2^30 on in [0, p): the
# moduli just below and just above 2^30 cover both sides of that switch.
# root = None: findParams from that modulus up (each listed modulus is the
# prime it finds).
EXTREME_MODULI = [(1, 3, 2), (2, 5, 2), (4, 17, 3), (1, 2**31 - 1, 2**31 - 2),
                  (10, 2145389569, None), (14, 2013265921, None), (16, 2130706433, None),
                  (20, 1888485377, None),
                  (12, 1073479681, None), (18, 1073479681, None), (20, 1053818881, None),
                  (21, 1004535809, None), (20, 1085276161, None), (12, 1073750017, None)]


@pytest.mark.parametrize("L,mod,root", EXTREME_MODULI)
def test_u32_extreme_moduli(ctx, L, mod, root):
    n = 1 << L
    if root is None:
        root, found = S.find_params(n, mod)
        assert found == mod
    batches = 3
    x = u32_input(7000 + L, n * batches, mod)
    X = run_u32(ctx, x, L, batches, root, mod)
    assert np.array_equal(X, O.stockham_u32(x, n, root, mod, batches)), f"forward mod={mod}"
    assert np.array_equal(run_u32(ctx, X, L, batches, root, mod, inverse=True), x), f"round trip mod={mod}"
    # the same transforms through the batch queue / grouped passes
    bufs = [ctx.upload(x[i * n:(i + 1) * n].copy()) for i in range(batches)]
    b = S.Batch(ctx, [("u32", False, L, buf, root, mod) for buf in bufs])
    b.run()
    ctx.sync()
    for i, buf in enumerate(bufs):
        assert np.array_equal(buf.copy_to(np.empty(n, dtype=np.uint32)), X[i * n:(i + 1) * n]), f"batch mod={mod}"
        buf.free()
    b.close()


# ----------------------------------------------------------------------------
# error behaviour of the reference's API
# ----------------------------------------------------------------------------
def test_errors(ctx):
    v = np.zeros(12, dtype=np.int32)
    with pytest.raises(ValueError):
        S.ntt_stockham(v, 12, 3, 13, ctx=ctx)            # not a power of two
    with pytest.raises(IndexError):
        S.auto_ntt(S.NttArgs(np.zeros(8, dtype=np.int32), 8, 2, 9, 17), ctx)  # host span too short
    with pytest.raises(ValueError):
        S.ntt_stockham(np.zeros(16, dtype=np.int32), 16, 2, 17, ctx=ctx)  # 2 is not of order 16 mod 17
    with pytest.raises(ValueError):
        S.ntt_stockham(np.zeros(16, dtype=np.int32), 16, 3, 2**31 + 11, ctx=ctx)  # modulus too large
    # 9 has order 2 mod 10 (9 = -1), but twiddles are Montgomery words: odd moduli only
    with pytest.raises(ValueError):
        S.ntt_stockham(np.zeros(2, dtype=np.int32), 2, 9, 10, ctx=ctx)
    w = np.arange(1, 2, dtype=np.int32)
    S.ntt_stockham(w, 1, 9, 10, ctx=ctx)                  # n = 1: the identity, any modulus
    assert w[0] == 1


@pytest.mark.parametrize("field,L,split", [("goldilocks", 24, [9, 8, 7]), ("goldilocks", 16, [6, 5, 5]),
                                           ("u32", 20, [6, 7, 7]), ("u32", 22, [9, 9, 4]),
                                           ("goldilocks", 20, [5, 5, 5, 5]),
                                           # radix-512 last passes (row halves, final scaling)
                                           ("goldilocks", 20, [6, 5, 9]), ("u32", 21, [6, 6, 9])])
def test_tuned_pass_splits_same_result(ctx, field, L, split):
    """Any valid stages-per-pass split gives the same bits."""
    n = 1 << L
    S.set_pass_split(ctx, field, L, split)
    try:
        if field == "goldilocks":
            x = gl_input(3000 + L, n)
            assert np.array_equal(run_gl(ctx, x, L, 1), O.stockham_gl(x, L))
            assert np.array_equal(run_gl(ctx, x, L, 1, inverse=True), O.inverse_gl(x, L))
        else:
            root, mod = S.find_params(n, 11)
            x = u32_input(3000 + L, n, mod)
            assert np.array_equal(run_u32(ctx, x, L, 1, root, mod), O.stockham_u32(x, n, root, mod))
            assert np.array_equal(run_u32(ctx, x, L, 1, root, mod, inverse=True), O.inverse_u32(x, n, root, mod))
    finally:
        S.restore_pass_split(ctx, field, L)


def test_pass_split_validation(ctx):
    with pytest.raises(ValueError):
        S.set_pass_split(ctx, "goldilocks", 24, [8, 8, 7])     # covers 23 stages
    with pytest.raises(ValueError):
        S.set_pass_split(ctx, "goldilocks", 26, [13, 13])      # pass above 2^12
    with pytest.raises(ValueError):
        S.set_pass_split(ctx, "u32", 12, [4, 4, 4])            # J = 16 < 32 columns
