"""Pins the CPU oracle (oracle/ntt_oracle.c) before it is trusted as the checker.

Three kinds of evidence:
  * golden fixtures written from the reference's own CPU code
    (tests/golden/make_golden.py over src/ntt/ntt_cpu.cpp, ntt_util.cpp);
  * the compiled reference itself (oracle/_ref), when present;
  * mathematics that fixes the answer independently of both: the textbook
    DFT, impulse / constant closed forms, linearity, the convolution theorem,
    root-of-unity facts of Goldilocks, and Python big-integer DFTs.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.golden.gen import digits, splitmix64

GOLD = os.path.join(os.path.dirname(__file__), "golden")
P = O.GL_P


@pytest.fixture(scope="module")
def ref_u32():
    return np.load(os.path.join(GOLD, "ref_u32.npz"))


@pytest.fixture(scope="module")
def gl_naive():
    return np.load(os.path.join(GOLD, "gl_naive.npz"))


# ----------------------------------------------------------------------------
# against the reference's own outputs (fixtures)
# ----------------------------------------------------------------------------
def test_find_params_table(ref_u32):
    for L, root, mod in ref_u32["params"]:
        assert O.find_params(1 << int(L), 11) == (int(root), int(mod)), L


@pytest.mark.parametrize("L", range(1, 14))
@pytest.mark.parametrize("batches", [1, 2])
def test_stockham_u32_matches_reference(ref_u32, L, batches):
    n = 1 << L
    root, mod = O.find_params(n, 11)
    x = ref_u32[f"st_in_L{L}_b{batches}"]
    want = ref_u32[f"st_out_L{L}_b{batches}"]
    got = O.stockham_u32(x.astype(np.uint32), n, root, mod, batches)
    assert np.array_equal(got.astype(np.int32), want)
    # the int32-wrap restatement agrees too (no overflow below p = 46341)
    assert np.array_equal(O.stockham_refcompat(x, n, root, mod, batches), want)


def test_stockham_nsteps_matches_reference(ref_u32):
    n = 256
    root, mod = O.find_params(n, 11)
    x = ref_u32["nsteps_in"]
    for s, want in enumerate(ref_u32["nsteps_out"]):
        got = O.stockham_u32(x.astype(np.uint32), n, root, mod, 1, s)
        assert np.array_equal(got.astype(np.int32), want), s


@pytest.mark.skipif(not O.ref_available(), reason="the reference's CPU code is not built here")
@pytest.mark.parametrize("L", range(0, 8))
def test_stage_loop_past_log2n_matches_reference(L):
    """nsteps beyond log2 n: the reference's strides J = pow(2, i, mod) and
    its zeroed ping-pong buffer (src/ntt/ntt_cpu.cpp:52,64-84,98-122), with
    moduli small enough that its int products do not wrap; the arithmetic
    form and the index-only form, two rows each."""
    R = O.ref()
    n = 1 << L
    for mod in (11, 13, 17, 97, 193, 257, 7681):
        for nsteps in range(0, 14):
            x = ((np.arange(2 * n, dtype=np.int64) * 7 + 3) % mod).astype(np.int32)
            want = x.copy()
            assert R.ref_ntt_stockham(want, n, 3, mod, 2, nsteps) == 0
            got = O.stockham_u32(x.astype(np.uint32), n, 3, mod, 2, nsteps)
            assert np.array_equal(got.astype(np.int32), want), (L, mod, nsteps)
            xi = np.arange(2 * n, dtype=np.int32)
            want = xi.copy()
            assert R.ref_ntt_stockham_idx_mod(want, n, 2, mod, nsteps) == 0
            got = np.concatenate([O.stockham_idx(xi[:n], n, nsteps, mod), O.stockham_idx(xi[n:], n, nsteps, mod)])
            assert np.array_equal(got, want), (L, mod, nsteps)


def test_stage_loop_past_log2n_closed_forms():
    """Without the reference: n = 1 writes nothing, so the result is the
    zeroed buffer after an odd number of stages and the input after an even
    number; with mod = 17, pow(2, 8, 17) = 1 makes stage 8 of an 8-point
    transform an ordinary J = 1 stage again."""
    one = np.array([5], dtype=np.int32)
    assert O.stockham_idx(one, 1, 3, 17)[0] == 0 and O.stockham_idx(one, 1, 2, 17)[0] == 5
    assert O.stockham_u32(np.array([5], dtype=np.uint32), 1, 3, 17, 1, 1)[0] == 0
    x = np.arange(8, dtype=np.int32)
    nine = O.stockham_idx(x, 8, 9, 17)  # stages 0-2, then 3..7 (J > 4, nothing written), then J = 1
    # after stages 0..7 the words are those of stage 2 or of the buffer two swaps back; stage 8 is
    # a J = 1 stage applied to the buffer current after stage 7
    eight = O.stockham_idx(x, 8, 8, 17)
    assert np.array_equal(nine, O.stockham_idx(eight, 8, 1, 17))


@pytest.mark.parametrize("L", [1, 4, 7, 10])
def test_stockham_idx_is_bit_reversal(ref_u32, L):
    n = 1 << L
    got = O.stockham_idx(np.arange(n, dtype=np.int32), n)
    assert np.array_equal(got, ref_u32[f"idx_L{L}"])
    assert np.array_equal(got, [O.lib().orc_bit_reverse(i, L) for i in range(n)])


@pytest.mark.parametrize("L", [2, 5, 8, 11])
def test_ct_gs_match_reference(ref_u32, L):
    n = 1 << L
    root, mod = O.find_params(n, 11)
    x = ref_u32[f"ct_in_L{L}"]
    ct = O.ntt_ct_u32(x.astype(np.uint32), root, mod)
    assert np.array_equal(ct.astype(np.int32), ref_u32[f"ct_out_L{L}"])
    gs = O.intt_gs_u32(ct, root, mod)
    assert np.array_equal(gs.astype(np.int32), ref_u32[f"gs_out_L{L}"])
    assert np.array_equal(gs.astype(np.int32), x)


@pytest.mark.parametrize("L", [2, 4, 6])
def test_pwc_matches_reference(ref_u32, L):
    n = 1 << L
    a, b = ref_u32[f"pwc_a_L{L}"], ref_u32[f"pwc_b_L{L}"]
    root, mod = (int(v) for v in ref_u32[f"pwc_params_L{L}"])
    max_val = int(max(a.max(), b.max()))
    assert (root, mod) == O.find_params(n, max_val * max_val * n + 1)  # src/ntt/ntt_cpu.cpp:184-186
    got = O.pwc_ntt_u32(a.astype(np.uint32), b.astype(np.uint32), root, mod)
    assert np.array_equal(got.astype(np.int32), ref_u32[f"pwc_out_L{L}"])
    # cyclic convolution, schoolbook
    want = np.zeros(n, dtype=np.int64)
    for i in range(n):
        for j in range(n):
            want[(i + j) % n] += int(a[i]) * int(b[j])
    assert np.array_equal(got.astype(np.int64), want % mod)


def test_refcompat_wrapped_digests():
    """Above n = 2^13 the reference's int products wrap; the ref-compat mode
    reproduces its literal outputs (digests written from the reference)."""
    wrap = json.load(open(os.path.join(GOLD, "ref_wrap.json")))
    for L, rec in wrap.items():
        L = int(L)
        n = 1 << L
        x = digits(L, n)
        assert hashlib.sha256(x.tobytes()).hexdigest() == rec["input_sha256"]
        y = O.stockham_refcompat(x, n, rec["root"], rec["mod"])
        assert hashlib.sha256(y.tobytes()).hexdigest() == rec["output_sha256"], L


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("L", [3, 9, 13, 16, 20])
def test_live_reference_cross_check(L):
    n = 1 << L
    root, mod = O.find_params(n, 11)
    x = (splitmix64(100 + L, 2 * n) % np.uint64(mod)).astype(np.int32)
    want = x.copy()
    assert O.ref().ref_ntt_stockham(want, n, root, mod, 2, -1) == 0
    assert np.array_equal(O.stockham_refcompat(x, n, root, mod, 2), want)
    if L <= 13:
        assert np.array_equal(O.stockham_u32(x.astype(np.uint32), n, root, mod, 2).astype(np.int32), want)


# ----------------------------------------------------------------------------
# against mathematics
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("L", [1, 2, 5, 8])
def test_u32_equals_textbook_dft(L):
    n = 1 << L
    root, mod = O.find_params(n, 11)
    x = (splitmix64(7, n) % np.uint64(mod)).astype(np.uint32)
    assert np.array_equal(O.stockham_u32(x, n, root, mod), O.naive_dft_u32(x, root, mod))


@pytest.mark.parametrize("L", [14, 17, 20])
def test_u32_exact_above_wrap_range(L):
    """Exact arithmetic where the reference wraps: closed forms and the
    first output X[0] = sum x, X[n/2] = alternating sum."""
    n = 1 << L
    root, mod = O.find_params(n, 11)
    d = np.zeros(n, dtype=np.uint32)
    d[0] = 1
    assert np.all(O.stockham_u32(d, n, root, mod) == 1)              # impulse -> ones
    ones = np.ones(n, dtype=np.uint32)
    want = np.zeros(n, dtype=np.uint32)
    want[0] = n % mod
    assert np.array_equal(O.stockham_u32(ones, n, root, mod), want)  # ones -> n * delta
    x = (splitmix64(L, n) % np.uint64(mod)).astype(np.uint32)
    X = O.stockham_u32(x, n, root, mod)
    assert int(X[0]) == int(x.astype(np.int64).sum()) % mod
    alt = int(x[0::2].astype(np.int64).sum() - x[1::2].astype(np.int64).sum()) % mod
    assert int(X[n // 2]) == alt
    # X[1] = sum x_j w^j, computed with Python integers
    w = [1]
    for _ in range(n - 1):
        w.append(w[-1] * root % mod)
    assert int(X[1]) == sum(int(a) * b for a, b in zip(x, w)) % mod
    assert np.array_equal(O.inverse_u32(X, n, root, mod), x)


def test_u32_linearity():
    n = 1 << 11
    root, mod = O.find_params(n, 11)
    a = (splitmix64(1, n) % np.uint64(mod)).astype(np.uint32)
    b = (splitmix64(2, n) % np.uint64(mod)).astype(np.uint32)
    s = ((a.astype(np.uint64) + 3 * b.astype(np.uint64)) % mod).astype(np.uint32)
    A, B, S = (O.stockham_u32(v, n, root, mod) for v in (a, b, s))
    assert np.array_equal(S, ((A.astype(np.uint64) + 3 * B.astype(np.uint64)) % mod).astype(np.uint32))


@pytest.mark.parametrize("L", [1, 4, 10, 13])
def test_inverse_is_reference_gs_of_bitreversed(L):
    """inverse(X) == inttGs(bitrev(X)), the reference's only inverse."""
    n = 1 << L
    root, mod = O.find_params(n, 11)
    x = (splitmix64(3, n) % np.uint64(mod)).astype(np.uint32)
    X = O.stockham_u32(x, n, root, mod)
    perm = [O.lib().orc_bit_reverse(i, L) for i in range(n)]
    gs = O.intt_gs_u32(X[perm].copy(), root, mod)
    assert np.array_equal(O.inverse_u32(X, n, root, mod), gs)
    assert np.array_equal(gs, x)
    if O.ref_available() and L <= 11:
        r = X[perm].astype(np.int32).copy()
        assert O.ref().ref_intt_gs(r, n, root, mod) == 0
        assert np.array_equal(r, x.astype(np.int32))


def test_zero_stages_is_identity():
    """nsteps = 0 and n = 1 return the input untouched (src/ntt/ntt_cpu.cpp:64-84)."""
    x = np.array([123456], dtype=np.uint32)
    assert np.array_equal(O.stockham_u32(x, 1, 1, 11), x)
    y = np.arange(8, dtype=np.uint32) + 40
    assert np.array_equal(O.stockham_u32(y, 8, 9, 17, 1, 0), y)


def test_goldilocks_roots():
    assert O.gl_root(32) == 1753635133440165772
    for L in (1, 5, 12, 24, 32):
        w = O.gl_root(L)
        assert O.lib().orc_gl_pow(w, 1 << (L - 1)) == P - 1
        assert O.lib().orc_gl_pow(w, 1 << L) == 1
    assert O.lib().orc_gl_pow(2, 96) == P - 1   # 2^96 == -1
    assert O.lib().orc_gl_pow(2, 64) == (1 << 32) - 1
    assert O.gl_root(6) == pow(2, 39, P)          # omega_64 = 2^39


def test_goldilocks_arithmetic_edges():
    L = O.lib()
    for a, b in [(P - 1, P - 1), (2**64 - 1, 2**64 - 1), (0, P - 1), (2**64 - 1, 1), (P, 5)]:
        assert L.orc_gl_mul(a, b) == (a * b) % P
        assert L.orc_gl_add(a, b) == (a + b) % P
        assert L.orc_gl_sub(a, b) == (a - b) % P


@pytest.mark.parametrize("L", range(1, 7))
def test_goldilocks_matches_bigint_dft(gl_naive, L):
    x, X = gl_naive[f"x_L{L}"], gl_naive[f"X_L{L}"]
    assert np.array_equal(O.stockham_gl(x, L), X)
    assert np.array_equal(O.naive_dft_gl(x), X)
    assert np.array_equal(O.inverse_gl(X, L), x)


@pytest.mark.parametrize("L", [8, 12, 16])
def test_goldilocks_closed_forms(L):
    n = 1 << L
    d = np.zeros(n, dtype=np.uint64)
    d[1] = 1
    X = O.stockham_gl(d, L)                         # delta_1 -> w^k
    w = O.gl_root(L)
    assert int(X[1]) == w and int(X[2]) == O.lib().orc_gl_pow(w, 2) and int(X[n - 1]) == O.lib().orc_gl_pow(w, n - 1)
    x = splitmix64(L, n)
    Y = O.stockham_gl(x, L)
    assert int(Y[0]) == sum(int(v) for v in x) % P
    assert np.all(Y < np.uint64(P))
    assert np.array_equal(O.inverse_gl(Y, L), x % np.uint64(P))


def test_goldilocks_convolution_theorem():
    L, n = 6, 64
    a = splitmix64(11, n) % np.uint64(P)
    b = splitmix64(12, n) % np.uint64(P)
    A, B = O.stockham_gl(a, L), O.stockham_gl(b, L)
    C = np.array([O.lib().orc_gl_mul(int(u), int(v)) for u, v in zip(A, B)], dtype=np.uint64)
    c = O.inverse_gl(C, L)
    want = [0] * n
    for i in range(n):
        for j in range(n):
            want[(i + j) % n] = (want[(i + j) % n] + int(a[i]) * int(b[j])) % P
    assert [int(v) for v in c] == want


@pytest.mark.parametrize("L", range(1, 7))
def test_goldilocks_ct_gs_match_bigint_dft(gl_naive, L):
    """The Goldilocks twins of nttCt / inttGs against the Python big-int DFT:
    CT gives the DFT in bit-reversed order, GS takes that order back to x."""
    x, X = gl_naive[f"x_L{L}"], gl_naive[f"X_L{L}"]
    perm = [O.lib().orc_bit_reverse(i, L) for i in range(1 << L)]
    assert np.array_equal(O.ntt_ct_gl(x, L), X[perm])
    assert np.array_equal(O.intt_gs_gl(X[perm], L), x % np.uint64(P))


@pytest.mark.parametrize("L", [3, 6])
def test_goldilocks_pwc_is_cyclic_convolution(L):
    n = 1 << L
    a = splitmix64(21 + L, n) % np.uint64(P)
    b = splitmix64(31 + L, n) % np.uint64(P)
    want = [0] * n
    for i in range(n):
        for j in range(n):
            want[(i + j) % n] = (want[(i + j) % n] + int(a[i]) * int(b[j])) % P
    assert [int(v) for v in O.pwc_ntt_gl(a, b, L)] == want


def test_fixtures_discriminate_plausible_mistakes(ref_u32):
    """A wrong root (the inverse root), dropped autosort (bit-reversed output)
    or a transposed batch would each fail the fixture check above."""
    L, n = 10, 1024
    root, mod = O.find_params(n, 11)
    x = ref_u32[f"st_in_L{L}_b2"].astype(np.uint32)
    want = ref_u32[f"st_out_L{L}_b2"].astype(np.uint32)
    inv_root = O.lib().orc_invmod(root, mod)
    assert not np.array_equal(O.stockham_u32(x, n, inv_root, mod, 2), want)
    perm = [O.lib().orc_bit_reverse(i, L) for i in range(n)]
    assert not np.array_equal(O.stockham_u32(x, n, root, mod, 2)[:n][perm], want[:n])
    swapped = np.concatenate([x[n:], x[:n]])
    assert not np.array_equal(O.stockham_u32(swapped, n, root, mod, 2), want)


@pytest.mark.parametrize("L", [1, 6, 11])
def test_sampled_dft_points_agree_with_full_transforms(L):
    n = 1 << L
    ks = np.array(sorted({0, 1, n // 2, n - 1, (3 * n) // 7}), dtype=np.uint64)
    x = splitmix64(70 + L, n)
    assert np.array_equal(O.dft_points_gl(x, L, ks), O.stockham_gl(x, L)[ks])
    assert np.array_equal(O.dft_points_gl(x, L, ks, inverse=True), O.inverse_gl(x, L)[ks])
    root, mod = O.find_params(n, 11)
    y = (x % np.uint64(mod)).astype(np.uint32)
    assert np.array_equal(O.dft_points_u32(y, n, root, mod, ks), O.stockham_u32(y, n, root, mod)[ks])
    assert np.array_equal(O.dft_points_u32(y, n, root, mod, ks, inverse=True), O.inverse_u32(y, n, root, mod)[ks])


@pytest.mark.parametrize("L", [1, 5, 10])
def test_twiddle_tables_are_root_powers(L):
    n = 1 << L
    root, mod = O.find_params(n, 11)
    assert [int(v) for v in O.twiddles_u32(n, root, mod)] == [pow(root, i, mod) for i in range(n // 2)]
    w = pow(7, (P - 1) >> L, P)
    assert [int(v) for v in O.twiddles_gl(L)] == [pow(w, i, P) for i in range(n // 2)]
    wi = pow(w, P - 2, P)
    assert [int(v) for v in O.twiddles_gl(L, inverse=True)] == [pow(wi, i, P) for i in range(n // 2)]
