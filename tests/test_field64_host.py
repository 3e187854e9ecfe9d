"""CPU checks of the 64-bit prime field's parameters and its oracle.

* The oracle's generic-modulus functions (oracle/ntt_oracle.c, *_u64) are
  pinned to the Goldilocks oracle at p = 2^64 - 2^32 + 1 (itself pinned to
  big-integer DFTs), to O(n^2) textbook DFTs for other primes (below and
  above 2^63), and to Python big-integer arithmetic.
* The library's findParams (Miller-Rabin + Pollard rho) equals the oracle's
  trial-division restatement of the reference's search
  (src/ntt/ntt_util.cpp:10-84) where that finishes quickly, and reproduces
  the published Goldilocks root of order 2^32.
No GPU call is made: stntt_find_params is host code.
"""
from __future__ import annotations

import numpy as np
import pytest

import stntt_amd as S
from oracle import oracle as O
from tests.golden.gen import splitmix64

P = O.GL_P
GL_ROOT_2_32 = 1753635133440165772  # 7^((P-1)/2^32)


def bitrev_perm(L):
    return [O.lib().orc_bit_reverse(i, L) for i in range(1 << L)]


@pytest.mark.parametrize("L", [1, 5, 8, 13])
def test_u64_oracle_equals_goldilocks_oracle(L):
    n = 1 << L
    x = splitmix64(70 + L, n)
    w = O.gl_root(L)
    assert np.array_equal(O.stockham_u64(x, n, w, P), O.stockham_gl(x, L))
    assert np.array_equal(O.inverse_u64(x, n, w, P), O.inverse_gl(x, L))
    assert np.array_equal(O.ntt_ct_u64(x, w, P), O.ntt_ct_gl(x, L))
    assert np.array_equal(O.intt_gs_u64(x, w, P), O.intt_gs_gl(x, L))
    y = splitmix64(80 + L, n)
    assert np.array_equal(O.pwc_ntt_u64(x, y, w, P), O.pwc_ntt_gl(x, y, L))


PRIMES = [(6, 2**40), (8, 2**62 + 12345), (7, 2**63 + 99), (4, 2**64 - 2**40)]


@pytest.mark.parametrize("L,min_mod", PRIMES)
def test_u64_oracle_matches_textbook_dft(L, min_mod):
    n = 1 << L
    root, mod = S.find_params(n, min_mod)
    x = splitmix64(L, n)  # any 64-bit words: reduced mod p first
    X = O.naive_dft_u64(x, root, mod)
    assert np.array_equal(O.stockham_u64(x, n, root, mod), X)
    assert np.array_equal(O.ntt_ct_u64(x, root, mod), X[bitrev_perm(L)])
    assert np.array_equal(O.inverse_u64(X, n, root, mod), x % np.uint64(mod))
    # one output by Python integers
    k = 3 % n
    assert int(X[k]) == sum(int(v) * pow(root, j * k, mod) for j, v in enumerate(x)) % mod


def test_u64_pwc_is_cyclic_convolution():
    L, n = 5, 32
    root, mod = S.find_params(n, 2**63)
    a, b = splitmix64(1, n) % np.uint64(mod), splitmix64(2, n) % np.uint64(mod)
    want = [0] * n
    for i in range(n):
        for j in range(n):
            want[(i + j) % n] = (want[(i + j) % n] + int(a[i]) * int(b[j])) % mod
    assert [int(v) for v in O.pwc_ntt_u64(a, b, root, mod)] == want


def test_find_params_goldilocks_root():
    """findParams(2^32, P) finds P itself (prime, P - 1 = (2^32 - 1) 2^32),
    generator 7 and the published root of order 2^32."""
    assert S.find_params(1 << 32, P) == (GL_ROOT_2_32, P)
    assert S.find_params(1 << 16, P) == (pow(7, (P - 1) >> 16, P), P)


@pytest.mark.parametrize("n,min_mod", [(1 << 10, 11), (1 << 20, 11), (1 << 27, 11), (1 << 12, 2**36 + 7),
                                       (1 << 16, 2**40), (1 << 3, 2**38 + 1), (6, 10**11)])
def test_find_params_equals_trial_division(n, min_mod):
    assert S.find_params(n, min_mod) == O.find_params(n, min_mod)


@pytest.mark.parametrize("n,min_mod", [(1 << 20, 2**62), (1 << 30, 2**63 + 1), (1 << 8, 2**64 - 2**33)])
def test_find_params_64bit_properties(n, min_mod):
    """Beyond trial division's reach: p = k n + 1 >= min_mod, root of order
    exactly n, and no prime k' n + 1 in [min_mod, p) (checked by Python's
    Fermat test on the candidates below p, which has no false negatives)."""
    root, mod = S.find_params(n, min_mod)
    assert mod >= min_mod and (mod - 1) % n == 0
    assert pow(root, n, mod) == 1 and pow(root, n // 2, mod) == mod - 1
    k = (min_mod - 1 + n - 1) // n
    for kk in range(k, (mod - 1) // n):
        c = kk * n + 1
        assert any(pow(a, c - 1, c) != 1 for a in (2, 3, 5, 7)), c  # composite
    assert all(pow(a, mod - 1, mod) == 1 for a in (2, 3, 5, 7, 11, 13))
