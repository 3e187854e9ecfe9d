"""Every context starts with the pass splits compiled into the library (the
tuning table), with no help from the Python package: Context(tuned=False)
creates a context through the C ABI only.  The pass count is read from the
library's per-pass profiling events; results stay exact either way."""
from __future__ import annotations

import json

import numpy as np
import pytest

import stntt_amd as S
from oracle import oracle as O
from stntt_amd import autotune as A
from tests.golden.gen import splitmix64

pytestmark = pytest.mark.gpu

# sizes whose tuned split has more passes than the fewest-pass default (2)
CASES = [("goldilocks", 17), ("u32", 18), ("u64", 18)]


def run(ctx, field: str, L: int, x: np.ndarray, root: int, mod: int) -> tuple[np.ndarray, int]:
    buf = ctx.upload(x)
    ctx.set_profiling(True)
    if field == "goldilocks":
        ctx.ntt_goldilocks(buf, L, 1)
    elif field == "u32":
        ctx.ntt_u32(buf, L, 1, root, mod)
    else:
        ctx.ntt_u64(buf, L, 1, root, mod)
    passes = len(ctx.profile_read())
    ctx.set_profiling(False)
    out = buf.copy_to(np.empty_like(x))
    buf.free()
    return out, passes


def expected(field: str, L: int, x: np.ndarray, root: int, mod: int) -> np.ndarray:
    n = 1 << L
    if field == "goldilocks":
        return O.stockham_gl(x, L)
    if field == "u32":
        return O.stockham_u32(x, n, root, mod)
    return O.stockham_u64(x, n, root, mod)


@pytest.mark.parametrize("field,L", CASES)
def test_context_starts_with_compiled_splits(field, L):
    with open(A.TABLE) as f:
        want = json.load(f)["pass_split"][field][str(L)]
    assert len(want) > 2, "pick a size whose tuned split differs from the default"
    n = 1 << L
    if field == "goldilocks":
        root, mod = 0, S.GL_P
        x = splitmix64(700 + L, n)
    else:
        root, mod = S.find_params(n, 11 if field == "u32" else 2**63 + 1)
        x = splitmix64(700 + L, n) % np.uint64(mod)
        if field == "u32":
            x = x.astype(np.uint32)
    ref = expected(field, L, x, root, mod)
    ctx = S.Context(0, tuned=False)
    try:
        got, passes = run(ctx, field, L, x, root, mod)
        assert passes == len(want), (field, L, passes, want)
        assert np.array_equal(got, ref)
        # removing the entry falls back to the fewest-pass default
        S.set_pass_split(ctx, field, L, None)
        got, passes = run(ctx, field, L, x, root, mod)
        assert passes == 2, (field, L, passes)
        assert np.array_equal(got, ref)
    finally:
        ctx.close()
