"""Transforms of one context on several streams.

Multi-pass sizes share the context's scratch buffer; a transform on one
stream must not run its passes through that buffer while a transform on
another stream is still using it.  Two streams interleave forward/inverse
pairs of multi-pass sizes with no host synchronisation in between (each
buffer stays on one stream); every buffer must come back to its input
exactly, and a third stream's forward result must equal the oracle.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import stntt_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
P = S.GL_P


def _hip():
    return ctypes.CDLL("libamdhip64.so")


class HipStream:
    """A plain HIP stream (default priority) for the test's own submissions."""

    def __init__(self):
        self.hip = _hip()
        self.handle = ctypes.c_void_p()
        assert self.hip.hipStreamCreate(ctypes.byref(self.handle)) == 0

    @property
    def ptr(self) -> int:
        return self.handle.value

    def sync(self):
        assert self.hip.hipStreamSynchronize(self.handle) == 0

    def close(self):
        if self.handle.value:
            self.hip.hipStreamDestroy(self.handle)
            self.handle = ctypes.c_void_p()


def test_multipass_transforms_on_two_streams():
    ctx = S.Context(0)
    s1, s2, s3 = HipStream(), HipStream(), HipStream()
    try:
        rng = np.random.default_rng(11)
        a = rng.integers(0, P, size=1 << 20, dtype=np.uint64)
        b = rng.integers(0, P, size=1 << 22, dtype=np.uint64)
        root, mod = S.find_params(1 << 21, 11)
        c = rng.integers(0, mod, size=1 << 21, dtype=np.uint32)
        ba, bb, bc = ctx.upload(a), ctx.upload(b), ctx.upload(c)
        ctx.sync()
        for _ in range(4):
            ctx.ntt_goldilocks(ba, 20, 1, stream=s1.ptr)
            ctx.ntt_goldilocks(bb, 22, 1, stream=s2.ptr)
            ctx.ntt_u32(bc, 21, 1, root, mod, stream=s1.ptr)
            ctx.ntt_goldilocks(ba, 20, 1, inverse=True, stream=s1.ptr)
            ctx.ntt_goldilocks(bb, 22, 1, inverse=True, stream=s2.ptr)
            ctx.ntt_u32(bc, 21, 1, root, mod, inverse=True, stream=s1.ptr)
        s1.sync()
        s2.sync()
        assert np.array_equal(ba.copy_to(np.empty_like(a)), a)
        assert np.array_equal(bb.copy_to(np.empty_like(b)), b)
        assert np.array_equal(bc.copy_to(np.empty_like(c)), c)
        # a forward result on a third stream while the first runs a round trip
        ctx.ntt_goldilocks(bb, 22, 1, stream=s3.ptr)
        ctx.ntt_goldilocks(ba, 20, 1, stream=s1.ptr)
        ctx.ntt_goldilocks(ba, 20, 1, inverse=True, stream=s1.ptr)
        s1.sync()
        s3.sync()
        assert np.array_equal(bb.copy_to(np.empty_like(b)), O.stockham_gl(b, 22))
        assert np.array_equal(ba.copy_to(np.empty_like(a)), a)
        for buf in (ba, bb, bc):
            buf.free()
    finally:
        for st in (s1, s2, s3):
            st.close()
        ctx.close()


def test_batch_runs_on_two_streams_in_order():
    """A batch's runs share its queue counter and group scratch: a run on a
    second stream starts after the first (here also the data dependency:
    the transforms are in place, so two runs apply each transform twice)."""
    ctx = S.Context(0)
    s1, s2 = HipStream(), HipStream()
    try:
        rng = np.random.default_rng(12)
        specs = [("goldilocks", 10), ("u32", 12), ("goldilocks", 16), ("u32", 17), ("goldilocks", 13)]
        xs, bufs, descs = [], [], []
        for field, L in specs:
            n = 1 << L
            if field == "goldilocks":
                x, root, mod = rng.integers(0, P, size=n, dtype=np.uint64), 0, P
            else:
                root, mod = S.find_params(n, 11)
                x = rng.integers(0, mod, size=n, dtype=np.uint32)
            xs.append((x, root, mod))
            buf = ctx.upload(x)
            bufs.append(buf)
            descs.append((field, False, L, buf, root, mod))
        ctx.sync()
        b = S.Batch(ctx, descs)
        b.run(stream=s1.ptr)
        b.run(stream=s2.ptr)
        s2.sync()
        for (field, L), (x, root, mod), buf in zip(specs, xs, bufs):
            n = 1 << L
            if field == "goldilocks":
                want = O.stockham_gl(O.stockham_gl(x, L), L)
            else:
                want = O.stockham_u32(O.stockham_u32(x, n, root, mod), n, root, mod)
            assert np.array_equal(buf.copy_to(np.empty_like(x)), want), (field, L)
        b.close()
        for buf in bufs:
            buf.free()
    finally:
        s1.close()
        s2.close()
        ctx.close()
