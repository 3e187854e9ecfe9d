"""Benchmark: Goldilocks n = 2^24 forward NTT on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one forward transform of one resident 2^24-element uint64 buffer
(uniform in [0, P) from a fixed seed), in place.  value = whole-job
elements per second over exactly K steps bracketed by device
synchronisation.  Prints ONE JSON line.

roofline: the dominant kernel (the pass with the largest mean duration),
timed live with HIP events on the stream it runs on; achieved = its
algorithmic bytes per launch (each pass reads and writes the whole 2^24 x 8 B
buffer once: 2 n w = 256 MiB) / its mean duration, against the 8 TB/s HBM3E
peak.  traffic: HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE
(recorded in profiles/; null here when not supplied).

hbm_copy: the library's plain copy kernel moving the same 2^24 x 8 B buffer
over four buffers in turn (512 MiB, beyond the 256 MB MALL, as the passes'
ping-pong is): the HBM rate this footprint reaches in practice, with the
dominant pass and the whole transform as fractions of it.

Extra legs (--legs): n12 (Goldilocks 2^12, one transform), n20, gllarge,
u32large, orders (element orders, polynomial multiply), field64, batched
(4096 mixed transforms), serving (Poisson trace through the server).

cpu_baseline: the oracle's restatement of the reference's CPU Stockham
(src/ntt/ntt_cpu.cpp:46-90 with exact Goldilocks arithmetic), single
thread, on a bounded sample of the same workload; the host's core count is
reported beside it, and the batched and serving workloads also get an
all-cores CPU figure (the same restatement on a thread pool over a bounded
sample of their transforms).

Clock state: the GPU clock ramps up over the first ~50 ms of load and drops
when the host leaves it idle, so before the W warm-up steps the headline
workload runs for --settle-s seconds (default 0.2 s, untimed); the line
says so in "clock".
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

from stntt_amd import _lib

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
VALU_PEAK_G = 1024 * 2.4 / 4  # wave64 VALU instructions per ns: 256 CUs x 4 SIMDs, 2.4 GHz, 4 cycles each
METRIC = "Goldilocks 2^24 NTT GElem/s, %HBM peak; batched-vs-plain transforms/s"
P = 0xFFFFFFFF00000001


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # about 70 ms of warm-up: the GPU clock ramps up over the first ~50 ms
    # of load (a 2^24 Goldilocks transform takes ~0.26 ms cold, ~0.22 ms
    # after the ramp), then about 0.2 s of timed transforms
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--log2n", type=int, default=24)
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="FETCH_SIZE/WRITE_SIZE per dispatch measured by rocprofv3 (optional)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device copy of the same buffer (counter calibration)")
    ap.add_argument("--no-extra", action="store_true", help="headline workload only")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--legs", default="n12,n20,gllarge,u32large,orders,field64,batched,serving",
                    help="comma-separated extra legs to run after the headline")
    ap.add_argument("--no-batch-plain", action="store_true",
                    help="batched leg without the per-transform plain comparison")
    ap.add_argument("--serve-count", type=int, default=400)
    ap.add_argument("--settle-s", type=float, default=0.2,
                    help="seconds of untimed transforms before the warm-up steps (GPU clock ramp)")
    return ap.parse_args()


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_baseline(log2n: int, budget_s: float) -> dict:
    """Time the oracle on whole 2^log2n transforms until about budget_s."""
    from oracle import oracle as O
    n = 1 << log2n
    x = np.random.default_rng(7).integers(0, P, size=n, dtype=np.uint64)
    times = []
    t_end = time.perf_counter() + budget_s
    while time.perf_counter() < t_end or not times:
        t0 = time.perf_counter()
        O.stockham_gl(x, log2n)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"value": n / med / 1e9, "unit": "GElem/s", "cores": 1, "host_cores": host_cores(), "kind": "port",
            "sample": f"{len(times)} whole Goldilocks 2^{log2n} forward transforms, median "
                      f"{med * 1e3:.0f} ms each, single thread (oracle/ntt_oracle.c)"}


def cpu_small(log2n: int, budget_s: float = 0.2) -> dict:
    """The oracle on one Goldilocks 2^log2n transform, single thread: mean
    over at least 20 transforms and about budget_s."""
    from oracle import oracle as O
    x = np.random.default_rng(12).integers(0, P, size=1 << log2n, dtype=np.uint64)
    reps, t0 = 0, time.perf_counter()
    while reps < 20 or time.perf_counter() - t0 < budget_s:
        O.stockham_gl(x, log2n)
        reps += 1
    return {"forward_us": round((time.perf_counter() - t0) / reps * 1e6, 1), "cores": 1, "kind": "port",
            "sample": f"{reps} Goldilocks 2^{log2n} forward transforms, single thread (oracle/ntt_oracle.c)"}


def cpu_reference_u32(log2n: int, budget_s: float = 2.0) -> dict | None:
    """The reference's own CPU nttStockham (src/ntt/ntt_cpu.cpp:87-90,
    compiled in place into oracle/_ref by oracle/Makefile) on one 32-bit
    transform with findParams(n, 11), single thread, int32 data as it takes
    them; None where that build is absent."""
    from oracle import oracle as O
    if not O.ref_available():
        return None
    n = 1 << log2n
    root, mod = O.find_params(n, 11)
    base = np.random.default_rng(20).integers(0, 10, size=n).astype(np.int32)
    R = O.ref()
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < 3 or time.perf_counter() < t_end:
        x = base.copy()
        t0 = time.perf_counter()
        rc = R.ref_ntt_stockham(x, n, root, mod, 1, -1)
        times.append(time.perf_counter() - t0)
        if rc != 0:
            return None
    med = statistics.median(times)
    return {"value": n / med / 1e9, "unit": "GElem/s", "forward_ms": round(med * 1e3, 3), "cores": 1,
            "kind": "reference", "sample": f"{len(times)} 32-bit 2^{log2n} forward transforms (root {root}, "
                                         f"p {mod}), median, single thread (oracle/_ref: src/ntt/ntt_cpu.cpp)"}


def cpu_all_cores(specs, budget_s: float, label: str) -> dict:
    """The oracle on a thread pool of every host core (the C oracle releases
    the GIL) over the first transforms of a workload until about budget_s:
    transforms/s and elements/s of all cores."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    cores = host_cores()
    params = {sp.log2n: O.find_params(1 << sp.log2n, 11) for sp in specs if sp.field == "u32"}

    def one(i_sp):
        i, sp = i_sp
        n = 1 << sp.log2n
        rng = np.random.default_rng(1000 + i)  # one generator per task: threads share none
        if sp.field == "goldilocks":
            x = rng.integers(0, P, size=n, dtype=np.uint64)
            (O.inverse_gl if sp.inverse else O.stockham_gl)(x, sp.log2n)
        else:
            r, m = params[sp.log2n]
            x = rng.integers(0, m, size=n, dtype=np.uint32)
            (O.inverse_u32 if sp.inverse else O.stockham_u32)(x, n, r, m)
        return n

    done, elems = 0, 0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        while done < len(specs) and time.perf_counter() - t0 < budget_s:
            chunk = list(enumerate(specs))[done:done + 4 * cores]
            elems += sum(ex.map(one, chunk))
            done += len(chunk)
    dt = time.perf_counter() - t0
    return {"transforms_s": round(done / dt, 1), "gelem_s": round(elems / dt / 1e9, 4), "cores": cores,
            "kind": "port", "sample": f"first {done} transforms of the {label} ({elems} elements) in "
                                      f"{dt:.1f} s on {cores} threads (oracle/ntt_oracle.c)"}


def copy_rate(ctx, nbytes: int, reps: int = 200) -> float:
    """HBM rate of the library's plain copy kernel (one 8-byte word per
    lane), nbytes -> nbytes over four buffers walked in turn: 4 nbytes of
    footprint, more than the 256 MB MALL at nbytes = 128 MiB, so (like the
    ping-pong of a transform's passes) every pass streams from HBM.
    Returns TB/s of read + write bytes."""
    bufs = [ctx.alloc(nbytes) for _ in range(4)]
    try:
        t_end = time.perf_counter() + 0.05
        i = 0
        while time.perf_counter() < t_end:
            ctx.copy(bufs[(i + 1) & 3], bufs[i & 3], nbytes)
            i += 1
        ctx.sync()
        t0 = time.perf_counter()
        for i in range(reps):
            ctx.copy(bufs[(i + 1) & 3], bufs[i & 3], nbytes)
        ctx.sync()
        dt = (time.perf_counter() - t0) / reps
    finally:
        for b in bufs:
            b.free()
    return 2 * nbytes / dt / 1e12


def bench_n20(S, ctx, steps: int) -> dict:
    """2^20, batch 1, forward + inverse, both fields (device-resident)."""
    out = {}
    n = 1 << 20
    rng = np.random.default_rng(3)
    root, mod = S.find_params(n, 11)
    for field in ("u32", "goldilocks"):
        x = rng.integers(0, mod, size=n, dtype=np.uint32) if field == "u32" else \
            rng.integers(0, P, size=n, dtype=np.uint64)
        buf = ctx.upload(x)

        def step():
            if field == "u32":
                ctx.ntt_u32(buf, 20, 1, root, mod)
                ctx.ntt_u32(buf, 20, 1, root, mod, inverse=True)
            else:
                ctx.ntt_goldilocks(buf, 20, 1)
                ctx.ntt_goldilocks(buf, 20, 1, inverse=True)
        dt, _ = _time_steps(ctx, step, steps)  # every step is a round trip
        ok = np.array_equal(buf.copy_to(np.empty_like(x)), x)
        out[field] = {"us_per_fwd_inv": round(dt * 1e6, 2), "roundtrip_exact": bool(ok),
                      "gelem_s": round(2 * n / dt / 1e9, 3)}
        buf.free()
    return out


def bench_n12(S, ctx, steps: int) -> dict:
    """Goldilocks 2^12, one transform (the reference's own CPU-runnable
    size): device-resident forward and forward + inverse latency (the CPU
    leg times the oracle on the same transform: cpu_n2^12)."""
    n = 1 << 12
    x = np.random.default_rng(12).integers(0, P, size=n, dtype=np.uint64)
    buf = ctx.upload(x)
    fwd = lambda: ctx.ntt_goldilocks(buf, 12, 1)  # noqa: E731

    def pair():
        ctx.ntt_goldilocks(buf, 12, 1)
        ctx.ntt_goldilocks(buf, 12, 1, inverse=True)
    tf, nf = _time_steps(ctx, fwd, steps)
    for _ in range(nf):  # undo the forward transforms
        ctx.ntt_goldilocks(buf, 12, 1, inverse=True)
    tp, _ = _time_steps(ctx, pair, steps)
    ok = np.array_equal(buf.copy_to(np.empty_like(x)), x)
    buf.free()
    return {"forward_us": round(tf * 1e6, 2), "fwd_inv_us": round(tp * 1e6, 2), "roundtrip_exact": bool(ok)}


def _time_steps(ctx, step, steps: int, settle_s: float = 0.08) -> tuple[float, int]:
    """Seconds per step over `steps` back-to-back steps, after at least 3
    untimed steps and `settle_s` of the same work (the input upload before a
    leg leaves the GPU idle, and its clock ramps up again over ~50 ms of
    load).  Returns (seconds per step, steps run in total)."""
    calls = 0
    t_end = time.perf_counter() + settle_s
    while calls < 3 or time.perf_counter() < t_end:
        step()
        ctx.sync()
        calls += 1
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    return (time.perf_counter() - t0) / steps, calls + steps


def bench_orders_polymul(S, ctx, steps: int) -> dict:
    """Element orders at 2^24 (Goldilocks, device-resident) and polynomial
    multiply (pwcNtt shape: two forward transforms to bit-reversed order,
    pointwise product, inverse from bit-reversed order), the pointwise
    product fused into the inverse's first pass versus a separate kernel.
    Algorithmic bytes of one multiply: 2 n w per forward, 3 n w for the
    inverse (reads a and b, writes a) = 7 n w."""
    out = {}
    L = 24
    n = 1 << L
    rng = np.random.default_rng(5)
    x = rng.integers(0, P, size=n, dtype=np.uint64)
    buf = ctx.upload(x)
    orders = {}
    for name, kw in (("natural", {}), ("bitrev_out", {"bitrev_out": True}),
                     ("bitrev_in", {"bitrev_in": True})):
        orders[name] = round(_time_steps(ctx, lambda: ctx.ntt_goldilocks(buf, L, 1, **kw), steps)[0] * 1e3, 4)
    out["goldilocks_2^24_forward_ms"] = orders
    buf.free()

    for field, L, batches in (("goldilocks", 24, 1), ("goldilocks", 20, 1), ("u32", 20, 1), ("u32", 16, 64)):
        n = 1 << L
        if field == "u32":
            root, mod = S.find_params(n, 11)
            a = rng.integers(0, mod, size=batches * n, dtype=np.uint64).astype(np.uint32)
            b = rng.integers(0, mod, size=batches * n, dtype=np.uint64).astype(np.uint32)
        else:
            a = rng.integers(0, P, size=batches * n, dtype=np.uint64)
            b = rng.integers(0, P, size=batches * n, dtype=np.uint64)
        w = a.dtype.itemsize
        da, db, ra, rb = ctx.upload(a), ctx.upload(b), ctx.alloc(a.nbytes), ctx.alloc(b.nbytes)

        def reset():
            ctx.copy(ra, da, a.nbytes)
            ctx.copy(rb, db, b.nbytes)

        def fused():
            reset()
            if field == "u32":
                ctx.polymul_u32(ra, rb, L, batches, root, mod)
            else:
                ctx.polymul_goldilocks(ra, rb, L, batches)

        def unfused():
            reset()
            if field == "u32":
                ctx.ntt_u32(ra, L, batches, root, mod, bitrev_out=True)
                ctx.ntt_u32(rb, L, batches, root, mod, bitrev_out=True)
                ctx.pointwise_u32(ra, rb, batches * n, mod)
                ctx.ntt_u32(ra, L, batches, root, mod, inverse=True, bitrev_in=True)
            else:
                ctx.ntt_goldilocks(ra, L, batches, bitrev_out=True)
                ctx.ntt_goldilocks(rb, L, batches, bitrev_out=True)
                ctx.pointwise_goldilocks(ra, rb, batches * n)
                ctx.ntt_goldilocks(ra, L, batches, inverse=True, bitrev_in=True)

        def copies():
            reset()

        t_copy = _time_steps(ctx, copies, steps)[0]
        t_fused = _time_steps(ctx, fused, steps)[0] - t_copy
        got_fused = ra.copy_to(np.empty_like(a))
        t_unfused = _time_steps(ctx, unfused, steps)[0] - t_copy
        got_unfused = ra.copy_to(np.empty_like(a))
        alg = 7 * batches * n * w
        out[f"polymul_{field}_2^{L}_x{batches}"] = {
            "fused_ms": round(t_fused * 1e3, 4), "unfused_ms": round(t_unfused * 1e3, 4),
            "gelem_s": round(batches * n / t_fused / 1e9, 3),
            "hbm_frac": round(alg / t_fused / 1e9 / HBM_PEAK_GBS, 4),
            "fused_equals_unfused": bool(np.array_equal(got_fused, got_unfused)),
            "note": "times exclude the two input-restoring device copies timed separately"}
        for d in (da, db, ra, rb):
            d.free()
    return out


def bench_u32_large(S, ctx, steps: int) -> dict:
    """The reference's 32-bit field at 2^24 and 2^26 (findParams(n, 11)
    primes), forward and inverse, device-resident; round trip checked."""
    out = {}
    for L in (24, 26):
        n = 1 << L
        root, mod = S.find_params(n, 11)
        x = np.random.default_rng(L).integers(0, mod, size=n, dtype=np.uint32)
        buf = ctx.upload(x)
        fwd = lambda: ctx.ntt_u32(buf, L, 1, root, mod)  # noqa: E731
        inv = lambda: ctx.ntt_u32(buf, L, 1, root, mod, inverse=True)  # noqa: E731
        fw, nf = _time_steps(ctx, fwd, steps)
        iv, ni = _time_steps(ctx, inv, steps)
        # as many inverses as forwards: x is back
        for _ in range(nf - ni):
            inv()
        for _ in range(ni - nf):
            fwd()
        ok = np.array_equal(buf.copy_to(np.empty_like(x)), x)
        out[f"2^{L}"] = {"mod": mod, "forward_us": round(fw * 1e6, 2), "inverse_us": round(iv * 1e6, 2),
                         "gelem_s": round(n / fw / 1e9, 2), "roundtrip_exact": bool(ok)}
        buf.free()
    return out


def bench_goldilocks_large(S, ctx, steps: int) -> dict:
    """Goldilocks 2^24 inverse and 2^26 forward and inverse, device-resident:
    time per transform, GElem/s and the fraction of the 8 TB/s HBM roofline
    against the algorithmic 2 n w bytes; round trips checked."""
    out = {}
    for L in (24, 26):
        n = 1 << L
        x = np.random.default_rng(100 + L).integers(0, P, size=n, dtype=np.uint64)
        buf = ctx.upload(x)
        fwd = lambda: ctx.ntt_goldilocks(buf, L, 1)  # noqa: E731
        inv = lambda: ctx.ntt_goldilocks(buf, L, 1, inverse=True)  # noqa: E731
        fw, nf = _time_steps(ctx, fwd, steps)
        iv, ni = _time_steps(ctx, inv, steps)
        for _ in range(nf - ni):
            inv()
        for _ in range(ni - nf):
            fwd()
        ok = np.array_equal(buf.copy_to(np.empty_like(x)), x)
        alg = 2 * n * 8
        out[f"2^{L}"] = {"forward_us": round(fw * 1e6, 2), "inverse_us": round(iv * 1e6, 2),
                         "forward_gelem_s": round(n / fw / 1e9, 2), "inverse_gelem_s": round(n / iv / 1e9, 2),
                         "forward_hbm_frac": round(alg / fw / 1e9 / HBM_PEAK_GBS, 4),
                         "inverse_hbm_frac": round(alg / iv / 1e9 / HBM_PEAK_GBS, 4),
                         "roundtrip_exact": bool(ok)}
        buf.free()
    return out


def bench_field64(S, ctx, steps: int) -> dict:
    """The 64-bit prime field (Montgomery arithmetic) at 2^24 forward, for a
    prime above 2^63 and for P itself beside the special-form Goldilocks
    kernels (same words, so the ratio is the price of generality)."""
    out = {}
    L = 24
    n = 1 << L
    x = np.random.default_rng(9).integers(0, 2**64, size=n, dtype=np.uint64)
    buf = ctx.upload(x)
    for name, min_mod in (("p_above_2^63", 2**63 + 1), ("p_goldilocks", P)):
        root, mod = S.find_params(n, min_mod)
        t = _time_steps(ctx, lambda: ctx.ntt_u64(buf, L, 1, root, mod), steps)[0]
        out[name] = {"mod": mod, "forward_ms": round(t * 1e3, 4), "gelem_s": round(n / t / 1e9, 3)}
    t = _time_steps(ctx, lambda: ctx.ntt_goldilocks(buf, L, 1), steps)[0]
    out["goldilocks_kernels_forward_ms"] = round(t * 1e3, 4)
    buf.free()
    return out


def bench_batched(S, ctx, count: int, reps: int, plain: bool = True) -> dict:
    """4096 transforms, log2 n ~ U{10..18}, field and direction 50/50: the
    batch (one persistent launch for the LDS-resident sizes plus grouped
    tile passes for the rest, replayed as one HIP graph) versus
    per-transform plain launches."""
    from stntt_amd.workloads import mixed_batch
    specs = mixed_batch(seed=42, count=count, lo=10, hi=18)
    rng = np.random.default_rng(4)
    bufs, descs, elems, nbytes = [], [], 0, 0
    params = {}
    for sp in specs:
        n = 1 << sp.log2n
        if sp.field == "goldilocks":
            x, r, m = rng.integers(0, P, size=n, dtype=np.uint64), 0, P
        else:
            if sp.log2n not in params:
                params[sp.log2n] = S.find_params(n, 11)
            r, m = params[sp.log2n]
            x = rng.integers(0, m, size=n, dtype=np.uint32)
        b = ctx.upload(x)
        bufs.append(b)
        descs.append((sp.field, sp.inverse, sp.log2n, b, r, m))
        elems += n
        nbytes += 2 * x.nbytes
    batch = S.Batch(ctx, descs)
    batch.run()
    ctx.sync()
    batch.capture()
    t_batch = _time_steps(ctx, batch.run_graph, reps)[0]
    t_plain = float("nan")
    if plain:
        t_plain = _time_steps(ctx, batch.run_plain, max(1, reps // 4))[0]
    batch.close()
    for b in bufs:
        b.free()
    out = {"transforms": count, "elements": elems, "algorithmic_bytes": nbytes,
           "batched_ms": round(t_batch * 1e3, 3), "batched_transforms_s": round(count / t_batch, 1),
           "batched_hbm_frac": round(nbytes / t_batch / 1e9 / HBM_PEAK_GBS, 4)}
    if plain:
        out.update({"plain_ms": round(t_plain * 1e3, 3), "plain_transforms_s": round(count / t_plain, 1),
                    "speedup": round(t_plain / t_batch, 2)})
    return out


def bench_serving(S, count: int, workers: int) -> dict:
    """Poisson arrivals, Goldilocks, log2 n ~ U{12..22}, direction 50/50,
    64 GB admission budget: the continuous-batching server versus plain
    sequential launches (one request at a time, FIFO) on the same trace."""
    from stntt_amd.workloads import poisson_arrivals
    srv = S.Server(0, hbm_budget_bytes=64 << 30, workers=workers)
    rng = np.random.default_rng(5)
    trace0 = poisson_arrivals(seed=17, count=count, rate_per_s=1.0)
    bufs = []
    for _, sp in trace0:
        a = srv.pinned(1 << sp.log2n, np.uint64)
        a.array[:] = rng.integers(0, P, size=1 << sp.log2n, dtype=np.uint64)
        bufs.append(a)
    # plain sequential service time (H2D, transform, D2H, synchronous)
    ctx = S.Context(0)
    dev = ctx.alloc(8 << 22)

    def serve_plain(sp, a):
        nbytes = 8 << sp.log2n
        _lib.check(_lib.lib().stntt_memcpy_htod(ctx.handle, dev.ptr, a.ptr, nbytes, None))
        ctx.ntt_goldilocks(dev.ptr, sp.log2n, 1, inverse=sp.inverse)
        _lib.check(_lib.lib().stntt_memcpy_dtoh(ctx.handle, a.ptr, dev.ptr, nbytes, None))
    for (_, sp), a in zip(trace0[:20], bufs[:20]):
        serve_plain(sp, a)
    t0 = time.perf_counter()
    for (_, sp), a in zip(trace0, bufs):
        serve_plain(sp, a)
    seq_capacity = count / (time.perf_counter() - t0)

    def replay(rate, use_server):
        scale = 1.0 / rate  # trace0 has rate 1/s
        lat, ids = [], []
        start = time.perf_counter()
        for (t, sp), a in zip(trace0, bufs):
            due = start + t * scale
            while time.perf_counter() < due:
                pass
            if use_server:
                ids.append((srv.submit("goldilocks", sp.log2n, a, inverse=sp.inverse), due))
            else:
                serve_plain(sp, a)
                lat.append((time.perf_counter() - due) * 1e3)
        if use_server:
            for rid, _ in ids:
                srv.wait(rid)
        span = time.perf_counter() - start
        return lat, span

    out = {"requests": count, "workers": workers, "budget_bytes": 64 << 30,
           "plain_sequential_capacity_tps": round(seq_capacity, 1), "rates": []}
    replay(4.0 * seq_capacity, True)  # warm the server's device-buffer cache and plans
    for frac in (0.5, 0.8, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0):
        rate = frac * seq_capacity
        srv.reset_stats()
        _, span_srv = replay(rate, True)
        st = srv.stats()
        row = {"offered_x_sequential": frac, "offered_tps": round(rate, 1),
               "server": {"p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                          "transforms_s": round(count / span_srv, 1), "rounds": st["rounds"],
                          "batched_requests": st["batched"], "grouped_requests": st["grouped"],
                          "peak_bytes": st["peak_bytes"]}}
        if frac in (0.8, 2.0):
            lat_plain, span_plain = replay(rate, False)
            lp = np.array(lat_plain)
            row["plain_sequential"] = {"p50_ms": round(float(np.percentile(lp, 50)), 3),
                                       "p99_ms": round(float(np.percentile(lp, 99)), 3),
                                       "transforms_s": round(count / span_plain, 1)}
        out["rates"].append(row)
    # saturation: the most the server completed at any offered rate, and the
    # highest offered rate it kept up with (completed >= 95 % of offered)
    out["saturation_tps"] = max(r["server"]["transforms_s"] for r in out["rates"])
    # host<->device bytes per request (upload + download of n words of 8 B):
    # the rate the copies ran at, at saturation
    mean_bytes = float(np.mean([2 * (8 << sp.log2n) for _, sp in trace0]))
    out["pcie_bytes_per_request"] = round(mean_bytes)
    out["saturation_pcie_gbs"] = round(out["saturation_tps"] * mean_bytes / 1e9, 1)
    kept = [r["offered_tps"] for r in out["rates"] if r["server"]["transforms_s"] >= 0.95 * r["offered_tps"]]
    out["max_sustained_offered_tps"] = max(kept) if kept else None
    dev.free()
    ctx.close()
    for a in bufs:
        a.free()
    srv.close()
    return out


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the transform is not "
              "sharded across GPUs in this build", file=sys.stderr)
        return 2
    import stntt_amd as S

    L = args.log2n
    n = 1 << L
    ctx = S.Context(0)
    x = np.random.default_rng(42).integers(0, P, size=n, dtype=np.uint64)
    buf = ctx.upload(x)
    log(f"[bench] goldilocks 2^{L}: clock settle {args.settle_s} s, warmup {args.warmup}")
    t_end = time.perf_counter() + args.settle_s
    while time.perf_counter() < t_end:
        for _ in range(10):
            ctx.ntt_goldilocks(buf, L, 1)
        ctx.sync()
    for _ in range(args.warmup):
        ctx.ntt_goldilocks(buf, L, 1)
    ctx.sync()

    # timed region: exactly K steps between two device synchronisations
    log(f"[bench] timing {args.steps} steps")
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.ntt_goldilocks(buf, L, 1)
    ctx.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3) / 1e9

    if args.copy_calibration:
        dst = ctx.alloc(n * 8)
        for _ in range(3):
            ctx.copy(dst, buf, n * 8)
        ctx.sync()
        dst.free()

    # per-kernel durations with HIP events on the library's stream
    ctx.set_profiling(True)
    per_pass: list[list[float]] = []
    for _ in range(min(max(args.steps, 10), 200)):
        ctx.ntt_goldilocks(buf, L, 1)
        per_pass.append(ctx.profile_read())
    ctx.set_profiling(False)
    npass = len(per_pass[0])
    mean_pass = [statistics.mean(r[i] for r in per_pass) for i in range(npass)]
    dom = max(range(npass), key=lambda i: mean_pass[i])
    alg_bytes = 2 * n * 8
    achieved = alg_bytes / (mean_pass[dom] * 1e-3) / 1e9
    transform_ms = sum(mean_pass)
    traffic = None
    valu_instr = None
    if os.path.exists(args.traffic_json):
        try:
            prof = json.load(open(args.traffic_json)).get(f"goldilocks_2^{L}", {})
            if len(prof.get("pass_bytes", [])) == npass:
                traffic = prof["pass_bytes"][dom]
            if len(prof.get("pass_valu_instr", [])) == npass:
                valu_instr = prof["pass_valu_instr"][dom]
        except (OSError, ValueError, KeyError, TypeError):
            traffic = valu_instr = None

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GElem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: uniform uint64 in [0, P) from seed 42, resident in HBM",
        "config": {"workload": f"goldilocks_forward_n2^{L}_batch1", "field": "goldilocks", "log2n": L,
                   "batch": 1, "direction": "forward", "passes": npass},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"pass {dom} of {npass}", "kernel_ms": round(mean_pass[dom], 5)},
        # the transform as a whole: 2 n w bytes over the whole step
        "hbm_frac_transform": round(alg_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "hbm_frac_transform_kernels": round(alg_bytes / (transform_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "clock": f"{args.settle_s} s of untimed transforms, then {args.warmup} warm-up steps",
        # hardware queues HIP gives this process (GPU_MAX_HW_QUEUES, 4 unless set)
        "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES") or 4),
        # the bound that actually limits this kernel: VALU issue.  Wave64
        # instructions per launch (rocprofv3 SQ_INSTS_VALU, profiles/) over
        # the live mean duration, against 1024 SIMDs x 2.4 GHz / 4 cycles.
        "valu_roofline": None if valu_instr is None else {
            "bound": "valu", "achieved": round(valu_instr / (mean_pass[dom] * 1e-3) / 1e9, 1),
            "peak": VALU_PEAK_G, "unit": "G wave-instr/s",
            "frac": round(valu_instr / (mean_pass[dom] * 1e-3) / 1e9 / VALU_PEAK_G, 4),
            "instr_per_launch": valu_instr},
        "pass_ms": [round(v, 5) for v in mean_pass],
    }
    # what a plain copy of the same buffer reaches when, like the passes,
    # it cannot stay in the MALL: the HBM rate this access footprint allows
    copy_tbs = copy_rate(ctx, n * 8)
    out["hbm_copy"] = {"tbs": round(copy_tbs, 3), "bytes": 2 * n * 8, "footprint_bytes": 4 * n * 8,
                       "dominant_pass_frac_of_copy": round(achieved / 1e3 / copy_tbs, 4),
                       "transform_frac_of_copy": round(npass * alg_bytes / (transform_ms * 1e-3) / 1e12 / copy_tbs, 4)}
    legs = set() if args.no_extra else set(args.legs.split(","))
    if "n12" in legs:
        log("[bench] n=2^12 Goldilocks, one transform")
        out["n2^12_goldilocks"] = bench_n12(S, ctx, 200)
    if "n20" in legs:
        log("[bench] n=2^20 forward+inverse, both fields")
        out["n2^20_fwd_inv"] = bench_n20(S, ctx, 20)
    if "orders" in legs:
        log("[bench] element orders and polynomial multiply")
        out["orders_polymul"] = bench_orders_polymul(S, ctx, 20)
    if "gllarge" in legs:
        log("[bench] Goldilocks 2^24 inverse, 2^26 forward and inverse")
        out["goldilocks_large"] = bench_goldilocks_large(S, ctx, 20)
    if "u32large" in legs:
        log("[bench] 32-bit field at 2^24 and 2^26")
        out["u32_large"] = bench_u32_large(S, ctx, 20)
    if "field64" in legs:
        log("[bench] 64-bit prime field")
        out["field64_2^24"] = bench_field64(S, ctx, 20)
    if "batched" in legs:
        log(f"[bench] batched mixed workload ({args.batch_count} transforms)")
        out["batched_mixed"] = bench_batched(S, ctx, args.batch_count, 10, plain=not args.no_batch_plain)
    if "serving" in legs:
        log("[bench] serving stream")
        out["serving"] = bench_serving(S, args.serve_count, 4)
    if not args.no_cpu:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(L, args.cpu_seconds)
        if "n12" in legs:
            out["cpu_n2^12"] = cpu_small(12)
        if "n20" in legs:
            ref20 = cpu_reference_u32(20)
            if ref20 is not None:
                out["cpu_reference_u32_2^20"] = ref20
        if "batched" in legs or "serving" in legs:
            from stntt_amd.workloads import mixed_batch, poisson_arrivals
            log("[bench] cpu, all cores: batched and serving workloads")
            out["cpu_all_cores"] = {
                "batched_mixed": cpu_all_cores(mixed_batch(seed=42, count=args.batch_count, lo=10, hi=18),
                                               args.cpu_seconds / 2, "batched mixed workload"),
                "serving": cpu_all_cores([sp for _, sp in poisson_arrivals(seed=17, count=args.serve_count,
                                                                           rate_per_s=1.0)],
                                         args.cpu_seconds / 2, "serving trace")}
    buf.free()
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
