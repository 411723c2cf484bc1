"""Shared helpers for the tests: seeded elements and small definitional
evaluators written on the independent big-int model (tests/models.py)."""
from __future__ import annotations

import struct

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


SHA256_K = [
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2]


def compress_py(state, block: bytes):
    """FIPS 180-4 section 6.2.2, in Python."""
    r = lambda x, k: ((x >> k) | (x << (32 - k))) & 0xffffffff  # noqa: E731
    w = list(struct.unpack(">16I", block))
    for i in range(16, 64):
        s0 = r(w[i - 15], 7) ^ r(w[i - 15], 18) ^ (w[i - 15] >> 3)
        s1 = r(w[i - 2], 17) ^ r(w[i - 2], 19) ^ (w[i - 2] >> 10)
        w.append((w[i - 16] + s0 + w[i - 7] + s1) & 0xffffffff)
    a, b, c, d, e, f, g, h = state
    for i in range(64):
        t1 = (h + (r(e, 6) ^ r(e, 11) ^ r(e, 25)) + ((e & f) ^ (~e & g)) + SHA256_K[i] + w[i]) & 0xffffffff
        t2 = ((r(a, 2) ^ r(a, 13) ^ r(a, 22)) + ((a & b) ^ (a & c) ^ (b & c))) & 0xffffffff
        a, b, c, d, e, f, g, h = (t1 + t2) & 0xffffffff, a, b, c, (d + t1) & 0xffffffff, e, f, g
    return [(x + y) & 0xffffffff for x, y in zip(state, (a, b, c, d, e, f, g, h))]
