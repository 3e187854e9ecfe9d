"""The batched and serving workloads at their full benchmark shapes, every
output checked.

* Batched mixed workload: 4096 independent transforms, log2 n uniform on
  10..18, field and direction one fair bit each (mixed_batch(seed=42)),
  captured into a HIP graph and replayed; each output equals the oracle.
* Serving stream: 200 Goldilocks requests with Poisson arrivals, log2 n
  uniform on 12..22, direction 50/50 (poisson_arrivals(seed=17)), under a
  64 GiB HBM budget with 4 workers, offered at twice the measured
  plain-sequential capacity; each result equals the oracle and the same
  transform run alone.

The reference's own check is GPU output == CPU output per transform
(src/tests/cuda/ntt/st_ntt_test.cu:37-44); the oracle plays the CPU side,
run on a thread pool (the C oracle releases the GIL).
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import stntt_amd as S
from oracle import oracle as O
from stntt_amd.workloads import mixed_batch, poisson_arrivals

pytestmark = pytest.mark.gpu
P = S.GL_P
POOL = 16


def oracle_of(field, inverse, L, x, root, mod):
    n = 1 << L
    if field == "goldilocks":
        return O.inverse_gl(x, L) if inverse else O.stockham_gl(x, L)
    return O.inverse_u32(x, n, root, mod) if inverse else O.stockham_u32(x, n, root, mod)


def test_mixed_batch_4096_graph_matches_oracle():
    specs = mixed_batch(seed=42, count=4096, lo=10, hi=18)
    rng = np.random.default_rng(4096)
    params = {}
    inputs, offsets, total = [], [], 0
    for sp in specs:
        n = 1 << sp.log2n
        if sp.field == "goldilocks":
            x, root, mod = rng.integers(0, P, size=n, dtype=np.uint64), 0, P
        else:
            if sp.log2n not in params:
                params[sp.log2n] = S.find_params(n, 11)
            root, mod = params[sp.log2n]
            x = rng.integers(0, mod, size=n, dtype=np.uint32)
        inputs.append((x, root, mod))
        offsets.append(total)
        total += x.nbytes
    print(f"4096 transforms, {total / 2**30:.2f} GiB", flush=True)
    host = np.empty(total, dtype=np.uint8)
    for (x, _, _), off in zip(inputs, offsets):
        host[off:off + x.nbytes] = x.view(np.uint8)
    ctx = S.Context(0)
    try:
        buf = ctx.upload(host)
        descs = [(sp.field, sp.inverse, sp.log2n, buf.ptr + off, root, mod)
                 for sp, (_, root, mod), off in zip(specs, inputs, offsets)]
        b = S.Batch(ctx, descs)
        b.capture()
        b.run_graph()
        ctx.sync()
        out = buf.copy_to(np.empty_like(host))
        b.close()
        buf.free()
    finally:
        ctx.close()
    print("batch done; checking against the oracle", flush=True)

    def check(i):
        sp, (x, root, mod), off = specs[i], inputs[i], offsets[i]
        got = out[off:off + x.nbytes].view(x.dtype)
        return np.array_equal(got, oracle_of(sp.field, sp.inverse, sp.log2n, x, root, mod))

    with ThreadPoolExecutor(POOL) as ex:
        ok = list(ex.map(check, range(len(specs))))
    bad = [i for i, v in enumerate(ok) if not v]
    assert not bad, f"{len(bad)} transforms differ, first {[(i, specs[i]) for i in bad[:5]]}"


def test_poisson_serving_trace_matches_oracle_and_alone():
    count = 200
    trace0 = poisson_arrivals(seed=17, count=count, rate_per_s=1.0, lo=12, hi=22)
    specs = [sp for _, sp in trace0]
    rng = np.random.default_rng(17)
    xs = [rng.integers(0, P, size=1 << sp.log2n, dtype=np.uint64) for sp in specs]

    # the same transforms alone, one after the other (upload, transform,
    # download): the results to match and the plain-sequential capacity
    ctx = S.Context(0)
    alone = []
    t0 = time.perf_counter()
    for sp, x in zip(specs, xs):
        buf = ctx.upload(x)
        ctx.ntt_goldilocks(buf, sp.log2n, 1, inverse=sp.inverse)
        alone.append(buf.copy_to(np.empty_like(x)))
        buf.free()
    seq_s = time.perf_counter() - t0
    ctx.close()
    capacity = count / seq_s
    rate = 2.0 * capacity
    trace = poisson_arrivals(seed=17, count=count, rate_per_s=rate, lo=12, hi=22)
    assert [sp for _, sp in trace] == specs  # the rate only scales the gaps
    print(f"sequential capacity {capacity:.0f}/s, offering {rate:.0f}/s", flush=True)

    srv = S.Server(0, hbm_budget_bytes=64 << 30, workers=4)
    try:
        datas = []
        for sp, x in zip(specs, xs):
            d = srv.pinned(x.size, np.uint64)
            d.array[:] = x
            datas.append(d)
        ids = []
        start = time.perf_counter()
        for (t, sp), d in zip(trace, datas):
            delay = start + t - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit("goldilocks", sp.log2n, d, inverse=sp.inverse))
        for i in ids:
            srv.wait(i)
        st = srv.stats()
        print(f"served: p50 {st['p50_ms']:.3f} ms, p99 {st['p99_ms']:.3f} ms, rounds {st['rounds']}", flush=True)
        assert st["completed"] == count
        assert st["peak_bytes"] <= 64 << 30
        got = [d.array.copy() for d in datas]
    finally:
        srv.close()

    for k, (sp, g, a) in enumerate(zip(specs, got, alone)):
        assert np.array_equal(g, a), f"request {k} {sp}: served result differs from the transform alone"

    def check(k):
        sp = specs[k]
        return np.array_equal(got[k], oracle_of("goldilocks", sp.inverse, sp.log2n, xs[k], 0, P))

    with ThreadPoolExecutor(POOL) as ex:
        ok = list(ex.map(check, range(count)))
    bad = [k for k, v in enumerate(ok) if not v]
    assert not bad, f"{len(bad)} served results differ from the oracle, first {[(k, specs[k]) for k in bad[:5]]}"
