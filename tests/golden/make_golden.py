"""Write the golden fixtures in tests/golden/ from the compiled reference.

Run in a container that has the reference tree (it builds
oracle/_ref/libstntt_ref.so from the reference's own src/ntt/*.cpp through
oracle/Makefile):

    python tests/golden/make_golden.py

Fixtures written:
  ref_u32.npz        -- reference nttStockham outputs (src/ntt/ntt_cpu.cpp:46-90)
                        on seeded inputs, exact range log2 n <= 13, batches 1
                        and 2, all nsteps at log2 n = 8; findParams table;
                        nttStockhamIdx(iota); nttCt / inttGs pairs;
                        pwc_convolution cases.
  ref_wrap.json      -- SHA-256 digests of the reference's int32-wrapped
                        outputs for log2 n = 14..20; inputs are values 0..9
                        (the reference's own test distribution) from
                        gen.digits(seed = log2 n).
  gl_naive.npz       -- Goldilocks X[k] = sum_j x[j] w^{jk} computed with
                        Python integers (textbook DFT; the reference has no
                        Goldilocks NTT), log2 n = 1..6, forward and inverse.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests.golden.gen import digits  # noqa: E402

GL_P = O.GL_P


def ref_stockham(x: np.ndarray, n: int, root: int, mod: int, batches: int, nsteps: int = -1) -> np.ndarray:
    v = np.ascontiguousarray(x, dtype=np.int32).copy()
    assert O.ref().ref_ntt_stockham(v, n, root, mod, batches, nsteps) == 0
    return v


def main() -> None:
    O.build(ref=True)
    R = O.ref()
    rng = np.random.default_rng(20240611)
    out: dict[str, np.ndarray] = {}

    # findParams(n, 11) for log2 n = 1..27 (src/ntt/ntt_util.cpp:10-14)
    params = []
    for L in range(1, 28):
        r, m = ctypes.c_uint64(), ctypes.c_uint64()
        assert R.ref_find_params(1 << L, 11, ctypes.byref(r), ctypes.byref(m)) == 0
        params.append((L, r.value, m.value))
    out["params"] = np.array(params, dtype=np.uint64)

    # Stockham, exact range (no int32 overflow below p = 46341, i.e. n <= 2^13)
    for L in range(1, 14):
        n = 1 << L
        root, mod = O.find_params(n, 11)
        for batches in (1, 2):
            x = rng.integers(0, mod, size=n * batches, dtype=np.int64).astype(np.int32)
            out[f"st_in_L{L}_b{batches}"] = x
            out[f"st_out_L{L}_b{batches}"] = ref_stockham(x, n, root, mod, batches)

    # every truncation nsteps at log2 n = 8 (src/ntt/ntt_cpu.cpp:64-65)
    L, n = 8, 256
    root, mod = O.find_params(n, 11)
    x = rng.integers(0, mod, size=n, dtype=np.int64).astype(np.int32)
    out["nsteps_in"] = x
    out["nsteps_out"] = np.stack([ref_stockham(x, n, root, mod, 1, s) for s in range(L + 1)])

    # nttStockhamIdx(iota) (src/ntt/ntt_cpu.cpp:92-127)
    for L in (1, 4, 7, 10):
        v = np.arange(1 << L, dtype=np.int32)
        assert R.ref_ntt_stockham_idx(v, 1 << L, 1, -1) == 0
        out[f"idx_L{L}"] = v

    # nttCt / inttGs (src/ntt/ntt_cpu.cpp:11-42, 130-164), exact range
    for L in (2, 5, 8, 11):
        n = 1 << L
        root, mod = O.find_params(n, 11)
        x = rng.integers(0, mod, size=n, dtype=np.int64).astype(np.int32)
        ct = x.copy()
        assert R.ref_ntt_ct(ct, n, root, mod) == 0
        gs = ct.copy()
        assert R.ref_intt_gs(gs, n, root, mod) == 0
        out[f"ct_in_L{L}"] = x
        out[f"ct_out_L{L}"] = ct
        out[f"gs_out_L{L}"] = gs

    # pwc_convolution (src/ntt/ntt_cpu.cpp:180-192): values 1..9 keep
    # min_mod = 81*len+1 small enough for the int arithmetic
    for L in (2, 4, 6):
        n = 1 << L
        a = rng.integers(1, 10, size=n, dtype=np.int64).astype(np.int32)
        b = rng.integers(1, 10, size=n, dtype=np.int64).astype(np.int32)
        res = np.zeros(n, dtype=np.int32)
        r, m = ctypes.c_int(), ctypes.c_int()
        assert R.ref_pwc_convolution(a, b, n, res, ctypes.byref(r), ctypes.byref(m)) == 0
        out[f"pwc_a_L{L}"] = a
        out[f"pwc_b_L{L}"] = b
        out[f"pwc_out_L{L}"] = res
        out[f"pwc_params_L{L}"] = np.array([r.value, m.value], dtype=np.int64)

    np.savez_compressed(os.path.join(HERE, "ref_u32.npz"), **out)

    # int32-wrapped range: digests only.  Inputs are values 0..9 (the
    # reference's own test distribution) from gen.digits(seed = L).
    wrap = {}
    for L in range(14, 21):
        n = 1 << L
        root, mod = O.find_params(n, 11)
        x = digits(L, n)
        y = ref_stockham(x, n, root, mod, 1)
        wrap[str(L)] = {
            "root": root, "mod": mod, "input_sha256": hashlib.sha256(x.tobytes()).hexdigest(),
            "output_sha256": hashlib.sha256(y.tobytes()).hexdigest(),
        }
    with open(os.path.join(HERE, "ref_wrap.json"), "w") as f:
        json.dump(wrap, f, indent=1)

    # Goldilocks naive DFT with Python integers (textbook; not the reference)
    gl = {}
    for L in range(1, 7):
        n = 1 << L
        w = pow(7, (GL_P - 1) >> L, GL_P)
        winv = pow(w, GL_P - 2, GL_P)
        ninv = pow(n, GL_P - 2, GL_P)
        x = [int(v) for v in rng.integers(0, 2**63, size=n, dtype=np.int64)]
        x = [(v * 2 + (v & 1)) % GL_P for v in x]  # spread over [0, P)
        X = [sum(x[j] * pow(w, j * k, GL_P) for j in range(n)) % GL_P for k in range(n)]
        xi = [ninv * sum(X[j] * pow(winv, j * k, GL_P) for j in range(n)) % GL_P for k in range(n)]
        assert xi == x
        gl[f"x_L{L}"] = np.array(x, dtype=np.uint64)
        gl[f"X_L{L}"] = np.array(X, dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "gl_naive.npz"), **gl)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
