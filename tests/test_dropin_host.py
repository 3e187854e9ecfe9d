"""Host-only helpers of the C++ drop-in (include/stntt/ntt_kernel.hpp), run
without a GPU through tests/cpp/dropin_host_test (built by build()).

Each printed case is checked against exact integer arithmetic and, where the
reference's CPU code is built (oracle/_ref), against the reference's own
function on the same inputs: elemWiseMul on negative ints
(src/ntt/ntt_util.h:28-33), the parameter helpers (ntt_util.cpp:10-84) and
the launch geometry NttArgs carries (src/cuda/ntt/ntt_kernel.cuh:30-39).
"""
from __future__ import annotations

import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_host_test")


@pytest.fixture(scope="module")
def cases():
    if not os.path.exists(BIN):
        pytest.skip("build() first")
    out = subprocess.run([BIN], capture_output=True, text=True, check=True, timeout=60).stdout
    rows = []
    for line in out.splitlines():
        lhs, rhs = line.split(" = ")
        name, *args = lhs.split()
        rows.append((name, args, rhs.split()))
    return rows


def of(cases, name):
    return [(a, r) for n, a, r in cases if n == name]


def test_elem_wise_mul_negative_ints(cases):
    rows = of(cases, "elemWiseMul")
    assert len(rows) == 5 * 64
    for (a, b, m), (r,) in rows:
        a, b, m, r = int(a), int(b), int(m), int(r)
        assert abs(a * b) < 2**31
        assert r == (a * b) % m, (a, b, m, r)  # Python's % is the non-negative residue
    if O.ref_available():
        R = O.ref()
        for m in sorted({int(a[2]) for a, _ in rows}):
            sel = [(int(a[0]), int(a[1]), int(r[0])) for a, r in rows if int(a[2]) == m]
            s0 = np.array([x for x, _, _ in sel], dtype=np.int32)
            s1 = np.array([y for _, y, _ in sel], dtype=np.int32)
            assert R.ref_elem_wise_mul(s0, s1, len(sel), m) == 0
            assert np.array_equal(s0, np.array([r for _, _, r in sel], dtype=np.int32)), m


def test_modulo_pow_sqrt_is_prime(cases):
    for (x, m), (r,) in of(cases, "modulo"):
        assert int(r) == int(x) % int(m)
    for (b, e, m), (r,) in of(cases, "pow"):
        b, e, m = int(b), int(e), int(m)
        assert int(r) == pow(b % m, e, m)  # e < 0: the inverse, as the reference's Fermat branch
    for (v,), (r,) in of(cases, "sqrt"):
        assert int(r) == math.isqrt(int(v))
    for (v,), (r,) in of(cases, "isPrime"):
        v = int(v)
        # the reference's trial division over 2..isqrt(n) calls 0 and 1 prime
        assert bool(int(r)) == all(v % d for d in range(2, math.isqrt(v) + 1))


def test_parameter_helpers(cases):
    for (v,), fs in of(cases, "uniquePrimeFactors"):
        v, fs = int(v), [int(f) for f in fs]
        assert fs == sorted(fs)
        rest = v
        for f in fs:
            assert all(f % d for d in range(2, math.isqrt(f) + 1)) and rest % f == 0
            while rest % f == 0:
                rest //= f
        assert rest == 1
    (_, (g,)), = of(cases, "findGenerator")
    g = int(g)
    assert all(pow(g, 7680 // q, 7681) != 1 for q in (2, 3, 5)) and all(
        any(pow(c, 7680 // q, 7681) == 1 for q in (2, 3, 5)) for c in range(2, g))
    (_, (r,)), = of(cases, "findPrimitiveRoot")
    assert int(r) == pow(g, 7680 // 512, 7681) == 7146  # findParams(512, 11) = (7146, 7681)
    (_, (ok,)), = of(cases, "isPrimitiveRoot")
    assert ok == "1"
    (_, (m,)), = of(cases, "findModulus")
    assert int(m) == 12289
    for (n, mm), (root, mod) in of(cases, "findParams"):
        assert (int(root), int(mod)) == O.find_params(int(n), int(mm))
        if O.ref_available():
            import ctypes
            r, p = ctypes.c_uint64(), ctypes.c_uint64()
            assert O.ref().ref_find_params(int(n), int(mm), ctypes.byref(r), ctypes.byref(p)) == 0
            assert (r.value, p.value) == (int(root), int(mod))


def test_bit_reversal(cases):
    (_, (r,)), = of(cases, "bitReverse")
    assert int(r) == 3
    (_, perm), = of(cases, "bitReverseVector")
    assert [int(x) for x in perm] == [int(f"{i:04b}"[::-1], 2) for i in range(16)]


def test_launch_geometry(cases):
    rows = {tuple(int(x) for x in a): [int(x) for x in r] for a, r in of(cases, "NttArgs")}
    assert len(rows) == 16
    # hand-evaluated from src/cuda/ntt/ntt_kernel.cuh:30-39 (radix 2, warp 32)
    assert rows[(1024, 1, 1024)] == [1, 1, 1, 512, 1, 1, 2048, 0]
    assert rows[(4, 10, 64)] == [1, 1, 1, 2, 10, 1, 0, 0]
    assert rows[(1 << 20, 1, 1024)] == [512, 1, 1, 1024, 1, 1, 4096, 1]
    assert rows[(64, 10, 64)] == [5, 1, 1, 32, 2, 1, 0, 0]
    for (n, batches, bs), (gx, gy, gz, bx, by, bz, smem, glob) in rows.items():
        per = n // 2
        assert gx * bs >= per * batches and bx == min(per, bs) and glob == int(per > bx)
        assert bx * by <= max(bs, per)
