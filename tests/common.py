"""Shared helpers for the tests: seeded elements and small definitional
evaluators written on the independent big-int model (tests/models.py)."""
from __future__ import annotations

import numpy as np

from models import gf_mul

MASK64 = (1 << 64) - 1


def splitmix(seed: int):
    """splitmix64 stream, as crates/gkr/src/lib.rs:986-999."""
    x = seed & MASK64
    while True:
        x = (x + 0x9E3779B97F4A7C15) & MASK64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def elements(seed: int, n: int) -> np.ndarray:
    """n elements F128::from([w0, w1]) from splitmix64(seed)."""
    s = splitmix(seed)
    out = np.empty((n, 2), dtype=np.uint64)
    for i in range(n):
        out[i, 0] = next(s)
        out[i, 1] = next(s)
    return out


def elements_fast(seed: int, n: int) -> np.ndarray:
    """Bulk uniform elements from numpy's PCG64 (large tables)."""
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2**64, size=(n, 2), dtype=np.uint64, endpoint=False)


def to_int(e) -> int:
    return int(e[0]) | (int(e[1]) << 64)


def ints(a: np.ndarray) -> list[int]:
    return [to_int(e) for e in np.asarray(a).reshape(-1, 2)]


def from_ints(xs) -> np.ndarray:
    return np.array([[x & MASK64, x >> 64] for x in xs], dtype=np.uint64).reshape(-1, 2)


def mle_eval_def(table: list[int], point: list[int]) -> int:
    """Sum_x table[x] * prod_i (x_i ? p_i : 1 + p_i) by definition."""
    n = len(point)
    acc = 0
    for x, v in enumerate(table):
        w = 1
        for i in range(n):
            w = gf_mul(w, point[i] if (x >> i) & 1 else point[i] ^ 1)
        acc ^= gf_mul(w, v)
    return acc


def horner(coeffs: list[int], x: int) -> int:
    v = 0
    for c in reversed(coeffs):
        v = gf_mul(v, x) ^ c
    return v
