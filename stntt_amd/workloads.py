"""Synthetic workloads of the benchmark configurations.

Seeded generators (numpy PCG64):
  * uniform elements: Goldilocks words in [0, P), 32-bit words in [0, p);
  * mixed batch: log2 n uniform on [lo, hi], field and direction one fair
    bit each (the batched configuration: 4096 transforms, log2 n in 10..18);
  * serving stream: Poisson arrivals, Goldilocks, log2 n uniform on 12..22.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

GL_P = 0xFFFFFFFF00000001


def uniform_gl(rng: np.random.Generator, count: int) -> np.ndarray:
    return rng.integers(0, GL_P, size=count, dtype=np.uint64)


def uniform_u32(rng: np.random.Generator, count: int, mod: int) -> np.ndarray:
    return rng.integers(0, mod, size=count, dtype=np.uint32)


@dataclass
class TransformSpec:
    field: str       # "u32" or "goldilocks"
    inverse: bool
    log2n: int


def mixed_batch(seed: int = 42, count: int = 4096, lo: int = 10, hi: int = 18) -> list[TransformSpec]:
    rng = np.random.default_rng(seed)
    logs = rng.integers(lo, hi + 1, size=count)
    fields = rng.integers(0, 2, size=count)
    dirs = rng.integers(0, 2, size=count)
    return [TransformSpec("goldilocks" if f else "u32", bool(d), int(L)) for L, f, d in zip(logs, fields, dirs)]


def poisson_arrivals(seed: int, count: int, rate_per_s: float, lo: int = 12, hi: int = 22):
    """(arrival_time_s, TransformSpec) pairs, Goldilocks, direction 50/50."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=count)
    times = np.cumsum(gaps)
    logs = rng.integers(lo, hi + 1, size=count)
    dirs = rng.integers(0, 2, size=count)
    return [(float(t), TransformSpec("goldilocks", bool(d), int(L))) for t, L, d in zip(times, logs, dirs)]
