"""Counter-based seeded generator shared by the host, the device and the tests.

SplitMix64 (Steele, Lea and Flood 2014) evaluated at counter k of stream ``seed``:
z = seed + (k + 1) * 0x9E3779B97F4A7C15, then the SplitMix64 finaliser. A U[0,1) Float32 is
(z >> 40) * 2^-24. Counter-based draws make every trajectory's inputs reproducible
independently of batch order. (The reference draws from Julia's default RNG, which
cannot be reproduced outside Julia.)
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, k: np.ndarray) -> np.ndarray:
    k = np.asarray(k, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (k + np.uint64(1)) * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_f32(seed: int, k: np.ndarray) -> np.ndarray:
    """U[0,1) Float32 draws at counters k."""
    z = splitmix64(seed, k)
    return ((z >> np.uint64(40)).astype(np.float64) * (1.0 / 16777216.0)).astype(np.float32)


def uniform_table(seed: int, rows: int, n: int) -> np.ndarray:
    """(rows, n) table of U[0,1) Float32; entry (r, i) uses counter i*rows + r."""
    k = (np.arange(n, dtype=np.uint64)[None, :] * np.uint64(rows) + np.arange(rows, dtype=np.uint64)[:, None])
    return uniform_f32(seed, k)


def normal_table(seed: int, rows: int, n: int) -> np.ndarray:
    """(rows, n) standard normals by Box-Muller on pairs of uniforms (float64)."""
    u = uniform_table(seed, 2 * rows, n).astype(np.float64)
    u1 = np.clip(u[0::2], 2.0 ** -25, 1.0)
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
