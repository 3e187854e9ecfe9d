"""Deterministic input generator shared by the fixture script and the tests.

splitmix64 over the element index: value i of stream `seed` is
splitmix64(seed * 2^32 + i).  Vectorised with numpy uint64 arithmetic, so the
same sequence is reproduced on any host without storing it.
"""
from __future__ import annotations

import numpy as np

_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, count: int) -> np.ndarray:
    z = (np.arange(count, dtype=np.uint64) + np.uint64((seed << 32) & 0xFFFFFFFFFFFFFFFF))
    with np.errstate(over="ignore"):
        z = z * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def digits(seed: int, count: int) -> np.ndarray:
    """Values 0..9, the reference's own test distribution (src/util/rng.h:35)."""
    return (splitmix64(seed, count) % np.uint64(10)).astype(np.int32)
