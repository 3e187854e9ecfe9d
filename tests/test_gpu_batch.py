"""Batched mixed transforms: one persistent launch over a descriptor queue.

Every transform in a mixed batch must equal the oracle bit for bit, equal
the same transform run alone (plain launches), and stay exact across
repeated hipGraph replays with other GPU work in between.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import stntt_amd as S
from oracle import oracle as O
from stntt_amd import _lib
from stntt_amd.workloads import mixed_batch

pytestmark = pytest.mark.gpu
P = S.GL_P


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


def make_inputs(specs, seed):
    rng = np.random.default_rng(seed)
    out = []
    for sp in specs:
        n = 1 << sp.log2n
        if sp.field == "goldilocks":
            out.append((rng.integers(0, 2**64, size=n, dtype=np.uint64), 0, P))
        else:
            root, mod = S.find_params(n, 11)
            out.append((rng.integers(0, mod, size=n, dtype=np.uint32), root, mod))
    return out


def oracle_apply(sp, x, root, mod, times=1):
    y = x
    for _ in range(times):
        if sp.field == "goldilocks":
            y = O.inverse_gl(y, sp.log2n) if sp.inverse else O.stockham_gl(y, sp.log2n)
        else:
            n = 1 << sp.log2n
            y = O.inverse_u32(y, n, root, mod) if sp.inverse else O.stockham_u32(y, n, root, mod)
    return y


def upload_all(ctx, inputs):
    return [ctx.upload(x) for x, _, _ in inputs]


def descs_of(specs, inputs, bufs):
    return [(sp.field, sp.inverse, sp.log2n, b, r, m) for sp, (_, r, m), b in zip(specs, inputs, bufs)]


def check_all(specs, inputs, bufs, times=1, tag=""):
    for i, (sp, (x, r, m), b) in enumerate(zip(specs, inputs, bufs)):
        got = b.copy_to(np.empty_like(x))
        want = oracle_apply(sp, x, r, m, times)
        assert np.array_equal(got, want), f"{tag} transform {i}: {sp}"


def test_mixed_batch_matches_oracle(ctx):
    specs = mixed_batch(seed=7, count=96, lo=1, hi=18)
    inputs = make_inputs(specs, 1)
    bufs = upload_all(ctx, inputs)
    b = S.Batch(ctx, descs_of(specs, inputs, bufs))
    print("batch of", len(specs), flush=True)
    b.run()
    ctx.sync()
    check_all(specs, inputs, bufs)
    b.close()


def test_batch_equals_plain_launches(ctx):
    """Each result is the same alone (per-transform launches) and in the mix."""
    specs = mixed_batch(seed=9, count=48, lo=10, hi=18)
    inputs = make_inputs(specs, 2)
    bufs_mix = upload_all(ctx, inputs)
    bufs_alone = upload_all(ctx, inputs)
    S.Batch(ctx, descs_of(specs, inputs, bufs_mix)).run()
    plain = S.Batch(ctx, descs_of(specs, inputs, bufs_alone))
    plain.run_plain()
    ctx.sync()
    for i, (bm, ba, (x, _, _)) in enumerate(zip(bufs_mix, bufs_alone, inputs)):
        assert np.array_equal(bm.copy_to(np.empty_like(x)), ba.copy_to(np.empty_like(x))), i


def test_graph_replays_stay_exact(ctx):
    """Replays of the captured launch, with eager work and syncs between
    them, each advance every transform exactly once more."""
    specs = mixed_batch(seed=11, count=40, lo=10, hi=16)
    inputs = make_inputs(specs, 3)
    bufs = upload_all(ctx, inputs)
    b = S.Batch(ctx, descs_of(specs, inputs, bufs))
    b.capture()
    side = ctx.upload(np.arange(1 << 12, dtype=np.uint64))
    for rep in range(1, 4):
        b.run_graph()
        ctx.ntt_goldilocks(side, 12, 1)  # unrelated eager work between replays
        ctx.sync()
        check_all(specs, inputs, bufs, times=rep, tag=f"replay {rep}")
    b.close()


@pytest.mark.parametrize("split", [[4, 4, 4, 4], [3, 3, 3, 3, 2, 2], [4, 4, 2, 2, 2, 2]])
def test_tuned_splits_same_result(ctx, split):
    L = 16
    S.set_batch_split(ctx, "goldilocks", L, split)
    try:
        x = np.random.default_rng(5).integers(0, 2**64, size=1 << L, dtype=np.uint64)
        buf = ctx.upload(x)
        S.Batch(ctx, [("goldilocks", False, L, buf, 0, P)]).run()
        assert np.array_equal(buf.copy_to(np.empty_like(x)), O.stockham_gl(x, L))
    finally:
        S.set_batch_split(ctx, "goldilocks", L, None)


def test_out_of_place_descriptors(ctx):
    rng = np.random.default_rng(13)
    xs, srcs, dsts, arr = [], [], [], (_lib.Desc * 3)()
    for i, (field, L) in enumerate([(1, 9), (0, 14), (1, 17)]):
        n = 1 << L
        if field:
            x, root, mod = rng.integers(0, 2**64, size=n, dtype=np.uint64), 0, 0
        else:
            root, mod = S.find_params(n, 11)
            x = rng.integers(0, mod, size=n, dtype=np.uint32)
        src, dst = ctx.upload(x), ctx.alloc(x.nbytes)
        xs.append((field, L, x, root, mod))
        srcs.append(src)
        dsts.append(dst)
        arr[i] = _lib.Desc(field, 0, L, 0, root, mod, src.ptr, dst.ptr)
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().stntt_batch_create(ctx.handle, arr, 3, ctypes.byref(h)))
    _lib.check(_lib.lib().stntt_batch_run(h, None))
    ctx.sync()
    for (field, L, x, root, mod), src, dst in zip(xs, srcs, dsts):
        got = dst.copy_to(np.empty_like(x))
        want = O.stockham_gl(x, L) if field else O.stockham_u32(x, 1 << L, root, mod)
        assert np.array_equal(got, want)
        assert np.array_equal(src.copy_to(np.empty_like(x)), x)  # source untouched
    _lib.lib().stntt_batch_destroy(h)


def test_batch_rejects_bad_descriptors(ctx):
    buf = ctx.alloc(8 << 19)
    with pytest.raises(ValueError):
        S.Batch(ctx, [("goldilocks", False, 25, buf, 0, P)])   # above the batched limit
    with pytest.raises(ValueError):
        S.Batch(ctx, [("u32", False, 4, buf, 2, 17)])          # 2 is not of order 16 mod 17
    with pytest.raises(ValueError):
        S.set_batch_split(ctx, "goldilocks", 16, [4, 4, 4, 3, 1])  # odd count on a streamed size


def test_grouped_single_pass_with_tables(ctx):
    """130 Goldilocks 2^13 transforms form one group whose schedule is the
    single-pass kernel over pointer tables; half of them out of place."""
    L, cnt = 13, 130
    rng = np.random.default_rng(17)
    xs = [rng.integers(0, 2**64, size=1 << L, dtype=np.uint64) for _ in range(cnt)]
    srcs = [ctx.upload(x) for x in xs]
    dsts = [s if i % 2 == 0 else ctx.alloc(x.nbytes) for i, (s, x) in enumerate(zip(srcs, xs))]
    arr = (_lib.Desc * cnt)()
    for i in range(cnt):
        arr[i] = _lib.Desc(1, i % 3 == 0, L, 0, 0, 0, srcs[i].ptr, dsts[i].ptr)
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().stntt_batch_create(ctx.handle, arr, cnt, ctypes.byref(h)))
    _lib.check(_lib.lib().stntt_batch_run(h, None))
    ctx.sync()
    for i, x in enumerate(xs):
        want = O.inverse_gl(x, L) if i % 3 == 0 else O.stockham_gl(x, L)
        assert np.array_equal(dsts[i].copy_to(np.empty_like(x)), want), i
        if i % 2:
            assert np.array_equal(srcs[i].copy_to(np.empty_like(x)), x), i  # source untouched
    _lib.lib().stntt_batch_destroy(h)


def test_large_descriptors_in_a_batch(ctx):
    """Sizes past the old batched limit, signed 32-bit inputs, both
    directions, graph capture and replay."""
    rng = np.random.default_rng(23)
    specs = [("goldilocks", False, 22), ("u32", True, 21), ("goldilocks", True, 20), ("u32", False, 19),
             ("goldilocks", False, 22)]
    items, descs = [], []
    for field, inv, L in specs:
        n = 1 << L
        if field == "goldilocks":
            x, root, mod = rng.integers(0, P, size=n, dtype=np.uint64), 0, P
        else:
            root, mod = S.find_params(n, 11)
            x = rng.integers(0, mod, size=n, dtype=np.uint64).astype(np.uint32)
        buf = ctx.upload(x)
        items.append((field, inv, L, x, root, mod, buf))
        descs.append((field, inv, L, buf, root, mod))
    b = S.Batch(ctx, descs)
    b.capture()
    for rep in (1, 2):
        b.run_graph()
        ctx.sync()
        for field, inv, L, x, root, mod, buf in items:
            y = x
            for _ in range(rep):
                if field == "goldilocks":
                    y = O.inverse_gl(y, L) if inv else O.stockham_gl(y, L)
                else:
                    y = O.inverse_u32(y, 1 << L, root, mod) if inv else O.stockham_u32(y, 1 << L, root, mod)
            assert np.array_equal(buf.copy_to(np.empty_like(x)), y), (field, L, rep)
    b.close()


def test_empty_batch_runs_captures_and_replays(ctx):
    b = S.Batch(ctx, [])
    b.run()
    b.capture()
    b.run_graph()
    b.run_graph()
    ctx.sync()
    b.close()


def test_lanes_from_a_callers_stream(ctx):
    """Two groups and queue work go out on three streams forked from the
    caller's stream and joined back into it: eager runs and graph replays
    on that stream, every result exact after each."""
    class Spec:
        def __init__(self, field, inverse, log2n):
            self.field, self.inverse, self.log2n = field, inverse, log2n
    specs = [Spec("goldilocks", False, 16), Spec("u32", True, 15), Spec("goldilocks", False, 16),
             Spec("u32", True, 15), Spec("goldilocks", False, 11), Spec("u32", False, 10), Spec("goldilocks", True, 12)]
    inputs = make_inputs(specs, 77)
    bufs = upload_all(ctx, inputs)
    other = S.Context(0)
    try:
        st = other.stream
        b = S.Batch(ctx, descs_of(specs, inputs, bufs))
        b.run(st)
        ctx.sync(st)
        check_all(specs, inputs, bufs, 1, "eager")
        b.capture(st)
        b.run_graph(st)
        ctx.sync(st)
        check_all(specs, inputs, bufs, 2, "replay 1")
        b.run(st)
        b.run_graph(st)
        ctx.sync(st)
        check_all(specs, inputs, bufs, 4, "eager + replay 2")
        b.close()
    finally:
        for x in bufs:
            x.free()
        other.close()


def test_closing_the_context_releases_its_batches_and_buffers():
    """A batch or buffer may outlive an explicit Context.close(): the context
    releases them first, and their own close/free are then no-ops."""
    c = S.Context(0)
    x = np.arange(1 << 12, dtype=np.uint64)
    buf = c.upload(x)
    b = S.Batch(c, [("goldilocks", False, 12, buf, 0, P)])
    b.run()
    c.sync()
    c.close()
    assert buf.ptr is None and b.handle is None
    b.close()
    buf.free()
