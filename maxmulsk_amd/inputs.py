"""Seeded synthetic inputs (splitmix64 -> FP32), identical on every machine.

U[-1, 1) draws are k * 2^-23 - 1 with k the top 24 bits of the draw (exact in
FP32); log-uniform draws are +-2^e with e uniform on [lo, hi]. The test suite
checks this generator against the oracle's copy bit for bit.
"""
from __future__ import annotations

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, n: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = np.uint64(seed) * _GOLD + np.arange(n, dtype=np.uint64)
        z = x + _GOLD
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, shape: tuple[int, ...]) -> np.ndarray:
    n = int(np.prod(shape))
    k = (splitmix64(seed, n) >> np.uint64(40)).astype(np.float64)
    return (k * 2.0 ** -23 - 1.0).astype(np.float32).reshape(shape)


def log_uniform(seed: int, shape: tuple[int, ...], lo: float = -20.0, hi: float = 20.0) -> np.ndarray:
    n = int(np.prod(shape))
    z = splitmix64(seed, n)
    f = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    s = np.where((z & np.uint64(1)) == 1, -1.0, 1.0)
    return (s * np.exp2(lo + (hi - lo) * f)).astype(np.float32).reshape(shape)
