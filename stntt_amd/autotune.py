"""Autotuner: per (field, log2 n) radix splits, measured on the GPU.

Two choices are tuned and installed into a context through the C ABI:
  * the stages per tile pass of the multi-pass path (stntt_set_pass_split),
    timed on a single resident transform of that size;
  * the radix split of the batched kernel's sub-steps (stntt_set_batch_split),
    timed on a batch of same-shape transforms in one persistent launch.
The winners are written to tuning/gfx950.json next to this file, applied to
every new Context (Context(tuned=False) skips the file) and compiled into the
library by the next build (tuning/gen_tuned.py -> csrc/tuned_gfx950.hpp), so
contexts created through the C ABI start with them too.  The 64-bit prime field
("u64") has pass splits only (timed at a prime above 2^63; the Montgomery
kernels cost the same for every odd modulus).

    python -m stntt_amd.autotune            # on an MI355X: everything
    python -m stntt_amd.autotune --fields u32,u64 --parts pass
        # re-tune a subset; entries not re-tuned are kept from the table
"""
from __future__ import annotations

import itertools
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "gfx950.json")
GL_P = 0xFFFFFFFF00000001
FIELDS = ("u32", "goldilocks")  # pass and batch splits
PASS_FIELDS = FIELDS + ("u64",)  # pass splits only


def word_log2c(field: str) -> int:
    """log2 of the tile columns: 128-byte rows of 4- or 8-byte words."""
    return 5 if field == "u32" else 4


def single_pass_max_log(field: str) -> int:
    """Largest log2 n the library always runs in one pass (csrc/plan.hpp,
    narrow_max_logr): pass splits of these sizes are never used."""
    return 12 if field == "u32" else 10


def apply_table(ctx, path: str = TABLE) -> int:
    """Install a tuning table into ctx; returns the number of entries applied."""
    from .ntt import set_batch_split, set_pass_shape
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        table = json.load(f)
    k = 0
    for field in PASS_FIELDS:
        for L, logr in table.get("pass_split", {}).get(field, {}).items():
            logc = table.get("pass_logc", {}).get(field, {}).get(L) or [-1] * len(logr)
            set_pass_shape(ctx, field, int(L), logr, logc)
            k += 1
    for field in FIELDS:
        for L, logs in table.get("batch_split", {}).get(field, {}).items():
            set_batch_split(ctx, field, int(L), logs)
            k += 1
    return k


def table_pass_split(field: str, L: int, path: str = TABLE) -> tuple[list[int], list[int]] | None:
    """The tuned (stages per pass, tile columns log2 per pass or -1) of
    (field, 2^L) in the table, or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    v = table.get("pass_split", {}).get(field, {}).get(str(L))
    if not v:
        return None
    c = table.get("pass_logc", {}).get(field, {}).get(str(L)) or [-1] * len(v)
    return list(v), list(c)


BIG_LO, BIG_HI, BIG_TILE_LOG_WORDS = 10, 12, 14  # the large-radix kernel (csrc/ntt_big.hpp)


def big_candidates(L: int) -> list[tuple[list[int], list[int]]]:
    """Two large-radix passes (2^10..2^12 each) with every tile width the
    kernel takes (the full 2^14-word tile down to a quarter) that fits n/R >=
    C and, on the second pass, J >= C."""
    out = []
    for a in range(BIG_LO, BIG_HI + 1):
        b = L - a
        if not BIG_LO <= b <= BIG_HI:
            continue
        for ca in range(BIG_TILE_LOG_WORDS - a, BIG_TILE_LOG_WORDS - a - 3, -1):
            for cb in range(BIG_TILE_LOG_WORDS - b, BIG_TILE_LOG_WORDS - b - 3, -1):
                if L - a >= ca and L - b >= cb and a >= cb:
                    out.append(([a, b], [ca, cb]))
    return out


def pass_candidates(L: int, logc: int, lo: int = 4, hi: int = 9) -> list[list[int]]:
    out = []
    for np_ in (2, 3, 4):
        for parts in itertools.product(range(lo, hi + 1), repeat=np_):
            if sum(parts) != L:
                continue
            if any(L - p < logc for p in parts):
                continue
            if any(sum(parts[:i]) < logc for i in range(1, np_)):
                continue
            out.append(list(parts))
    # fewest passes first, then the most even splits
    out.sort(key=lambda p: (len(p), max(p) - min(p)))
    return out[:12]


def batch_candidates(L: int, streamed: bool) -> list[list[int]]:
    """Radix splits from {4, 8, 16} sub-steps (2^2..2^4), fewest sub-steps
    first, each multiset in descending and ascending order; streamed sizes
    need an even count (the ping-pong must end in the destination)."""
    nmin = (L + 3) // 4
    out = []
    for ns in range(nmin, nmin + 3):
        if streamed and ns % 2:
            continue
        for parts in itertools.combinations_with_replacement((4, 3, 2), ns):
            if sum(parts) != L:
                continue
            desc = sorted(parts, reverse=True)
            for order in (desc, desc[::-1]):
                if order not in out:
                    out.append(list(order))
    if not out:  # small sizes: fall back to radix-2 parts
        ns = L if not streamed or L % 2 == 0 else L + 1
        out.append([1] * ns if sum([1] * ns) == L else [1] * L)
    return out[:6]


def _time(fn, sync, reps: int) -> float:
    sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    sync()
    return (time.perf_counter() - t0) / reps


def measure(cands, install, fn, sync, reps: int, rounds: int = 3, warm_s: float = 0.3) -> dict:
    """Median time per candidate over `rounds` rounds in alternating order.

    The GPU clock ramps up over the first tenths of a second of load (and
    after every host-side pause, such as preparing the next input), so every
    candidate first runs once (plans, scratch), the workload then runs for
    `warm_s`, and no candidate is always timed first: timed in a fixed order
    right after an upload, the first candidate read up to 20 % slow."""
    for c in cands:
        install(c)
        fn()
    sync()
    t_end = time.perf_counter() + warm_s
    while time.perf_counter() < t_end:
        fn()
        sync()
    acc = {str(c): [] for c in cands}
    for r in range(rounds):
        for c in (cands if r % 2 == 0 else cands[::-1]):
            install(c)
            acc[str(c)].append(_time(fn, sync, reps))
    return {k: float(np.median(v)) for k, v in acc.items()}


def tune(logs_pass=range(8, 27), logs_batch=range(10, 19), verbose: bool = True, fields=PASS_FIELDS,
         parts=("pass", "batch")) -> dict:
    import stntt_amd as S
    from .ntt import set_batch_split, set_pass_shape
    ctx = S.Context(0, tuned=False)
    rng = np.random.default_rng(1)
    table = {"device": "gfx950", "pass_split": {f: {} for f in PASS_FIELDS}, "batch_split": {f: {} for f in FIELDS},
             "measured_us": {"pass": {f: {} for f in PASS_FIELDS}, "batch": {f: {} for f in FIELDS}}}
    if "pass" not in parts:
        logs_pass = ()
    if "batch" not in parts:
        logs_batch = ()
    for field in fields:
        logc = word_log2c(field)
        for L in logs_pass:
            cands = pass_candidates(L, logc)
            if not cands or L <= single_pass_max_log(field):  # the single-pass kernel only
                continue
            n = 1 << L
            if field == "u32":
                root, mod = S.find_params(n, 11)
                x = rng.integers(0, mod, size=n, dtype=np.uint32)
            elif field == "u64":
                root, mod = S.find_params(n, 2**63 + 1)
                x = rng.integers(0, mod, size=n, dtype=np.uint64)
            else:
                x = rng.integers(0, GL_P, size=n, dtype=np.uint64)
            buf = ctx.upload(x)
            if field == "u32":
                fn = lambda: ctx.ntt_u32(buf, L, 1, root, mod)  # noqa: E731
            elif field == "u64":
                fn = lambda: ctx.ntt_u64(buf, L, 1, root, mod)  # noqa: E731
            else:
                fn = lambda: ctx.ntt_goldilocks(buf, L, 1)  # noqa: E731
            # shapes: (stages per pass, tile columns per pass or -1)
            shapes = [(c, [-1] * len(c)) for c in cands] + (big_candidates(L) if field != "u64" else [])
            med = measure(shapes, lambda c: set_pass_shape(ctx, field, L, c[0], c[1]), fn, ctx.sync, 20)
            times = {k: round(v * 1e6, 2) for k, v in med.items()}
            best = min(shapes, key=lambda c: med[str(c)])
            best_t = med[str(best)]
            set_pass_shape(ctx, field, L, best[0], best[1])
            table["pass_split"][field][str(L)] = best[0]
            if any(x >= 0 for x in best[1]):
                table.setdefault("pass_logc", {}).setdefault(field, {})[str(L)] = best[1]
            table["measured_us"]["pass"][field][str(L)] = times
            if verbose:
                print(f"[tune] pass {field} 2^{L}: {best} {best_t * 1e6:.1f} us", flush=True)
            buf.free()
        for L in logs_batch if field in FIELDS else ():
            n = 1 << L
            count = max(64, (1 << 22) >> L)
            w = 8 if field == "goldilocks" else 4
            streamed = (w << L) > 32768
            if field == "u32":
                root, mod = S.find_params(n, 11)
            else:
                root, mod = 0, GL_P
            bufs = []
            for _ in range(count):
                x = rng.integers(0, mod if field == "u32" else GL_P, size=n,
                                 dtype=np.uint32 if field == "u32" else np.uint64)
                bufs.append(ctx.upload(x))
            # the split is fixed when a batch is built: one batch per candidate
            cands = batch_candidates(L, streamed)
            batches = {}
            for cand in cands:
                set_batch_split(ctx, field, L, cand)
                batches[str(cand)] = S.Batch(ctx, [(field, bool(i & 1), L, bb, root, mod)
                                                   for i, bb in enumerate(bufs)])
            cur = {}
            med = measure(cands, lambda c: cur.update(b=batches[str(c)]), lambda: cur["b"].run(), ctx.sync, 5)
            for b in batches.values():
                b.close()
            times = {k: round(v * 1e6, 2) for k, v in med.items()}
            best = min(cands, key=lambda c: med[str(c)])
            best_t = med[str(best)]
            set_batch_split(ctx, field, L, best)
            table["batch_split"][field][str(L)] = best
            table["measured_us"]["batch"][field][str(L)] = times
            if verbose:
                print(f"[tune] batch {field} 2^{L} x{count}: {best} {best_t * 1e6:.1f} us", flush=True)
            for bb in bufs:
                bb.free()
    ctx.close()
    return table


def merge(old: dict, new: dict) -> dict:
    """new's entries over old's, per (kind, field, log2 n)."""
    out = json.loads(json.dumps(old))
    out["device"] = new.get("device", out.get("device"))
    for kind in ("pass_split", "batch_split"):
        for field, entries in new.get(kind, {}).items():
            out.setdefault(kind, {}).setdefault(field, {}).update(entries)
            if kind == "pass_split":  # a re-tuned size drops its old tile widths
                for L in entries:
                    out.get("pass_logc", {}).get(field, {}).pop(L, None)
    for field, entries in new.get("pass_logc", {}).items():
        out.setdefault("pass_logc", {}).setdefault(field, {}).update(entries)
    for kind, per_field in new.get("measured_us", {}).items():
        for field, entries in per_field.items():
            out.setdefault("measured_us", {}).setdefault(kind, {}).setdefault(field, {}).update(entries)
    return out


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("out", nargs="?", default=TABLE)
    ap.add_argument("--fields", default=",".join(PASS_FIELDS),
                    help="comma list of u32, goldilocks, u64 (u64: pass splits only)")
    ap.add_argument("--parts", default="pass,batch", help="comma list of pass, batch")
    ap.add_argument("--logs", default="8-26", help="log2 n range of the pass splits, e.g. 19-26")
    args = ap.parse_args(argv)
    lo, hi = (int(v) for v in args.logs.split("-"))
    fields = tuple(f for f in args.fields.split(",") if f)
    parts = tuple(p for p in args.parts.split(",") if p)
    if any(f not in PASS_FIELDS for f in fields) or any(p not in ("pass", "batch") for p in parts):
        ap.error("unknown field or part")
    table = tune(logs_pass=range(lo, hi + 1), fields=fields, parts=parts)
    if os.path.exists(args.out):
        with open(args.out) as f:
            table = merge(json.load(f), table)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print("wrote", args.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
