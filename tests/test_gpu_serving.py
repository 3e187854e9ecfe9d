"""Serving: continuous batching under an HBM budget.

Each result must be bit-identical to the same transform run alone (and to
the oracle) whatever else was in flight; admission must respect the budget;
bad requests fail with the reference's error kinds.
"""
from __future__ import annotations

import threading
import time

import numpy as np
import pytest

import stntt_amd as S
from oracle import oracle as O
from stntt_amd import _lib

pytestmark = pytest.mark.gpu
P = S.GL_P
RESERVE = 2 * (8 << 22) + 256 * (8 << 16)  # per-worker scratch reservation


def make_request(srv, rng, field, L, inverse):
    n = 1 << L
    if field == "goldilocks":
        data = srv.pinned(n, np.uint64)
        data.array[:] = rng.integers(0, 2**64, size=n, dtype=np.uint64)
        root = mod = 0
    else:
        root, mod = S.find_params(n, 11)
        data = srv.pinned(n, np.uint32)
        data.array[:] = rng.integers(0, mod, size=n, dtype=np.uint32)
    return data, root, mod


def expected(field, L, inverse, x, root, mod):
    n = 1 << L
    if field == "goldilocks":
        return O.inverse_gl(x, L) if inverse else O.stockham_gl(x, L)
    return O.inverse_u32(x, n, root, mod) if inverse else O.stockham_u32(x, n, root, mod)


def test_mixed_stream_matches_oracle_and_alone():
    srv = S.Server(0, hbm_budget_bytes=4 * RESERVE + (1 << 30), workers=4)
    rng = np.random.default_rng(21)
    specs = [("goldilocks" if i % 3 else "u32", int(rng.integers(10, 21)), bool(i % 2)) for i in range(60)]
    reqs = []
    for field, L, inv in specs:
        data, root, mod = make_request(srv, rng, field, L, inv)
        reqs.append((field, L, inv, data, data.array.copy(), root, mod))
    ids = []
    for field, L, inv, data, _, root, mod in reqs:  # bursty arrivals: overlap on purpose
        ids.append(srv.submit(field, L, data, inverse=inv, root=root, mod=mod))
        time.sleep(float(rng.exponential(2e-4)))
    lat = [srv.wait(i) for i in ids]
    assert all(v >= 0 for v in lat)
    st = srv.stats()
    assert st["completed"] == len(ids) and st["rounds"] >= 1
    ctx = S.Context(0)
    for k, (field, L, inv, data, x, root, mod) in enumerate(reqs):
        want = expected(field, L, inv, x, root, mod)
        assert np.array_equal(data.array, want), (k, field, L, inv)
        # the same transform alone, through the plain device path
        buf = ctx.upload(x)
        if field == "goldilocks":
            ctx.ntt_goldilocks(buf, L, 1, inverse=inv)
        else:
            ctx.ntt_u32(buf, L, 1, root, mod, inverse=inv)
        assert np.array_equal(buf.copy_to(np.empty_like(x)), data.array), k
        buf.free()
        data.free()
    ctx.close()
    srv.close()


def test_budget_admission_serializes_but_completes():
    # room for about two 2^20 Goldilocks requests at a time
    budget = RESERVE + 2 * (8 << 20) + 1024
    srv = S.Server(0, hbm_budget_bytes=budget, workers=1)
    rng = np.random.default_rng(5)
    items = []
    for i in range(8):
        data, _, _ = make_request(srv, rng, "goldilocks", 20, False)
        items.append((data, data.array.copy()))
    ids = [srv.submit("goldilocks", 20, d) for d, _ in items]
    for i in ids:
        srv.wait(i)
    st = srv.stats()
    assert st["peak_bytes"] <= budget - RESERVE
    for d, x in items:
        assert np.array_equal(d.array, O.stockham_gl(x, 20))
        d.free()
    srv.close()


def test_size_classes_share_the_budget():
    """Requests of many size classes through a small budget: idle cached
    buffers count against it and are released (other classes first) when a
    new class needs the room, so the bytes held never pass the budget and
    every request completes exactly."""
    room = 3 * (8 << 20)  # three 2^20 Goldilocks requests
    budget = RESERVE + room
    srv = S.Server(0, hbm_budget_bytes=budget, workers=1)
    rng = np.random.default_rng(31)
    sizes = [20, 19, 18, 17, 20, 16, 19, 15, 20, 18, 14, 20, 17, 19, 20, 13, 18, 20, 16, 20]
    items = []
    for wave in range(3):
        ids = []
        for L in sizes:
            data, _, _ = make_request(srv, rng, "goldilocks", L, bool(wave % 2))
            ids.append((L, data, data.array.copy(), srv.submit("goldilocks", L, data, inverse=bool(wave % 2))))
        for L, data, x, rid in ids:
            srv.wait(rid)
            items.append((L, bool(wave % 2), data, x))
        st = srv.stats()
        assert st["peak_bytes"] <= room, (wave, st)
        assert st["held_bytes"] <= room, (wave, st)
    for L, inv, data, x in items:
        assert np.array_equal(data.array, expected("goldilocks", L, inv, x, 0, 0)), (L, inv)
        data.free()
    srv.close()


def test_same_shape_requests_run_grouped():
    """Requests of one shape admitted in one round run as one grouped
    multi-pass run over pointer tables, in chunks that the worker's 64 MiB
    scratch holds (two-buffer schedule: Goldilocks 2^20 eight per chunk, so
    ten split 8 + 2; 2^21 four, so five split 4 + 1); a shape with one
    request runs alone.
    Every result must equal the oracle's.  A 2^22 request goes first so the
    burst behind it is admitted together."""
    srv = S.Server(0, hbm_budget_bytes=RESERVE + (2 << 30), workers=1)
    rng = np.random.default_rng(41)
    shapes = ([("goldilocks", 22, False)] + [("goldilocks", 18, False)] * 12 + [("goldilocks", 20, True)] * 10 +
              [("goldilocks", 21, False)] * 5 + [("u32", 17, False)] * 8 + [("u32", 19, True)] * 5 +
              [("goldilocks", 19, True)])
    reqs = []
    for field, L, inv in shapes:
        data, root, mod = make_request(srv, rng, field, L, inv)
        reqs.append((field, L, inv, data, data.array.copy(), root, mod))
    ids = [srv.submit(field, L, data, inverse=inv, root=root, mod=mod) for field, L, inv, data, _, root, mod in reqs]
    for i in ids:
        srv.wait(i)
    st = srv.stats()
    assert st["completed"] == len(ids)
    assert st["grouped"] >= 2, st
    for k, (field, L, inv, data, x, root, mod) in enumerate(reqs):
        assert np.array_equal(data.array, expected(field, L, inv, x, root, mod)), (k, field, L, inv)
        data.free()
    srv.close()


def test_concurrent_submitters():
    srv = S.Server(0, hbm_budget_bytes=2 * RESERVE + (1 << 30), workers=2)
    results, errors = [], []

    def client(seed):
        rng = np.random.default_rng(seed)
        try:
            for _ in range(10):
                L = int(rng.integers(8, 19))
                data, _, _ = make_request(srv, rng, "goldilocks", L, False)
                x = data.array.copy()
                rid = srv.submit("goldilocks", L, data)
                srv.wait(rid)
                results.append(np.array_equal(data.array, O.stockham_gl(x, L)))
                data.free()
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=client, args=(s,)) for s in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not errors, errors
    assert len(results) == 40 and all(results)
    srv.close()


def test_server_errors():
    with pytest.raises(ValueError):
        S.Server(0, hbm_budget_bytes=1 << 20, workers=1)  # budget below the worker reservation
    srv = S.Server(0, hbm_budget_bytes=RESERVE + (1 << 20), workers=1)
    data = srv.pinned(1 << 18, np.uint64)
    with pytest.raises(ValueError):
        srv.submit("goldilocks", 18, data)            # 2 MiB > 1 MiB budget
    with pytest.raises(ValueError):
        srv.submit("u32", 4, data, root=2, mod=17)    # bad root
    with pytest.raises(_lib.StnttError):
        srv.wait(987654321)                           # unknown id
    data.free()
    srv.close()


def test_server_close_frees_pinned_arrays():
    srv = S.Server(0, hbm_budget_bytes=RESERVE + (1 << 30), workers=1)
    data = srv.pinned(1 << 10, np.uint64)
    data.array[:] = np.arange(1 << 10, dtype=np.uint64)
    srv.wait(srv.submit("goldilocks", 10, data))
    srv.close()
    assert data.ptr is None
    data.free()  # no-op


def test_two_waiters_on_one_request():
    """Exactly one of two threads waiting on the same request collects it;
    the other is told it is gone (no use of a collected entry)."""
    srv = S.Server(0, hbm_budget_bytes=RESERVE + (1 << 30), workers=1)
    rng = np.random.default_rng(21)
    data, root, mod = make_request(srv, rng, "goldilocks", 20, False)
    x = data.array.copy()
    rid = srv.submit("goldilocks", 20, data)
    outcomes = []

    def waiter():
        try:
            srv.wait(rid)
            outcomes.append("ok")
        except _lib.StnttError as e:
            outcomes.append(e.code)

    threads = [threading.Thread(target=waiter) for _ in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=60)
    assert sorted(outcomes, key=str) == sorted(["ok", _lib.STNTT_ENOTFOUND], key=str)
    assert np.array_equal(data.array, expected("goldilocks", 20, False, x, root, mod))
    data.free()
    srv.close()
