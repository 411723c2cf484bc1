"""Pins the oracle's GF(2^128) arithmetic to answers it did not compute.

Mirrors crates/field/src/gf128.rs:453-462 (boundary products),
gf128/portable.rs:144-237 (definition and long-division cross-checks),
gf128/wide.rs:92-156 (deferred reduction) and tests/ghash_vectors.rs
(525 NIST GCM vectors)."""
import os
import random

import numpy as np
import pytest

from models import (aes128_encrypt, clmul_def, from_nist, gf_inv, gf_mul, parse_rsp,
                    reduce_longdiv, to_nist, words)
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
X, X63, X64, X127 = 2, 1 << 63, 1 << 64, 1 << 127


@pytest.fixture(params=["portable", "pclmul"])
def kernel(request):
    O.set_kernel(request.param)
    yield request.param
    O.set_kernel("portable")


def omul(a: int, b: int) -> int:
    return O.to_int(O.mul(a, b))


def test_boundary_products(kernel):
    # gf128.rs:453-462 / portable.rs:231-237: pin the polynomial and word order.
    assert omul(X, X63) == X64
    assert omul(X, X127) == 0x87
    assert omul(X64, X64) == 0x87
    assert O.to_int(O.lib().oracle_mul_x(O.f(X127))) == 0x87


def test_clmul64_matches_coefficient_definition(kernel):
    rng = random.Random(101)
    edges = [0, 1, 2, 2**64 - 1, 1 << 63, 0x87]
    cases = [(a, b) for a in edges for b in edges] + [
        (rng.getrandbits(64), rng.getrandbits(64)) for _ in range(512)]
    out = np.zeros(2, dtype=np.uint64)
    for a, b in cases:
        O.lib().oracle_clmul64(a, b, O.ptr(out))
        assert int(out[0]) | (int(out[1]) << 64) == clmul_def(a, b)


def test_reduce_matches_long_division():
    rng = random.Random(102)
    cases = [1 << i for i in range(256)] + [rng.getrandbits(256) for _ in range(512)]
    for v in cases:
        w = np.array([(v >> (64 * k)) & (2**64 - 1) for k in range(4)], dtype=np.uint64)
        assert O.to_int(O.lib().oracle_reduce(O.ptr(w))) == reduce_longdiv(v)


def test_mul_matches_bigint_model(kernel):
    rng = random.Random(103)
    edges = [0, 1, X, 0x87, X63, X64, X127, 2**128 - 1]
    cases = [(a, b) for a in edges for b in edges]
    cases += [(rng.getrandbits(128), rng.getrandbits(128)) for _ in range(1024)]
    for a, b in cases:
        assert omul(a, b) == gf_mul(a, b), (hex(a), hex(b))


def test_field_laws_square_mulx_inv(kernel):
    rng = random.Random(104)
    for _ in range(128):
        a, b, c = (rng.getrandbits(128) for _ in range(3))
        assert omul(a, b) == omul(b, a)
        assert omul(omul(a, b), c) == omul(a, omul(b, c))
        assert omul(a, b ^ c) == omul(a, b) ^ omul(a, c)
        assert O.to_int(O.lib().oracle_square(O.f(a))) == omul(a, a)
        assert O.to_int(O.lib().oracle_mul_x(O.f(a))) == omul(a, X)
    for _ in range(8):
        a = rng.getrandbits(128) | 1
        inv = O.to_int(O.lib().oracle_inv(O.f(a)))
        assert omul(a, inv) == 1
        assert inv == gf_inv(a)


def test_portable_and_pclmul_agree():
    rng = np.random.default_rng(7)
    a = rng.integers(0, 2**63, size=(4096, 2), dtype=np.uint64) * 2 + 1
    b = rng.integers(0, 2**63, size=(4096, 2), dtype=np.uint64) * 3
    O.set_kernel("portable")
    p = O.mul_batch(a, b)
    O.set_kernel("pclmul")
    q = O.mul_batch(a, b)
    O.set_kernel("portable")
    assert np.array_equal(p, q)


def test_deferred_reduction_is_linear():
    # wide.rs:92-156: accumulate unreduced, reduce once == reduce each term.
    rng = random.Random(203)
    a = np.array([words(rng.getrandbits(128)) for _ in range(64)], dtype=np.uint64)
    b = np.array([words(rng.getrandbits(128)) for _ in range(64)], dtype=np.uint64)
    eager = 0
    for x, y in zip(a, b):
        eager ^= gf_mul(O.to_int(x), O.to_int(y))
    assert O.inner_product(a, b) == eager


def test_nist_bit_order_shim_round_trips():
    for i in range(16):
        block = bytearray(16)
        block[i] = 0x80
        assert from_nist(bytes(block)) == 1 << (8 * i)
        assert to_nist(from_nist(bytes(block))) == bytes(block)


def test_aes128_fips197_example():
    # FIPS-197 Appendix C.1.
    key = bytes.fromhex("000102030405060708090a0b0c0d0e0f")
    pt = bytes.fromhex("00112233445566778899aabbccddeeff")
    assert aes128_encrypt(key, pt).hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"


def ghash(h: int, aad: bytes, ct: bytes, mul) -> bytes:
    x = 0
    for data in (aad, ct):
        for i in range(0, len(data), 16):
            blk = data[i: i + 16].ljust(16, b"\0")
            x = mul(x ^ from_nist(blk), h)
    lengths = (len(aad) * 8).to_bytes(8, "big") + (len(ct) * 8).to_bytes(8, "big")
    x = mul(x ^ from_nist(lengths), h)
    return to_nist(x)


def test_ghash_matches_nist_gcm_vectors():
    """All 525 NIST CAVS 14.0 GCM-128 96-bit-IV cases through the oracle's
    multiply (ghash_vectors.rs:129-161)."""
    with open(os.path.join(GOLDEN, "gcm_encrypt_128.rsp")) as fh:
        cases = parse_rsp(fh.read())
    assert len(cases) == 525
    O.set_kernel("pclmul")
    try:
        for i, c in enumerate(cases):
            assert len(c["IV"]) == 12
            h = from_nist(aes128_encrypt(c["Key"], bytes(16)))
            mask = aes128_encrypt(c["Key"], c["IV"] + b"\0\0\0\1")
            digest = ghash(h, c["AAD"], c["CT"], omul)
            tag = bytes(d ^ m for d, m in zip(digest, mask))[: c["tag_bytes"]]
            assert tag == c["Tag"], f"case {i}"
    finally:
        O.set_kernel("portable")


def test_ghash_vectors_also_pin_the_bigint_model():
    with open(os.path.join(GOLDEN, "gcm_encrypt_128.rsp")) as fh:
        cases = parse_rsp(fh.read())[:40]
    for c in cases:
        h = from_nist(aes128_encrypt(c["Key"], bytes(16)))
        mask = aes128_encrypt(c["Key"], c["IV"] + b"\0\0\0\1")
        digest = ghash(h, c["AAD"], c["CT"], gf_mul)
        assert bytes(d ^ m for d, m in zip(digest, mask))[: c["tag_bytes"]] == c["Tag"]


def test_sha256_matches_hashlib():
    import hashlib
    out = np.zeros(32, dtype=np.uint8)
    for msg in [b"", b"abc", b"a" * 55, b"a" * 56, b"a" * 64, bytes(range(256)) * 3]:
        O.lib().oracle_sha256(msg, len(msg), out.ctypes.data_as(O.C.c_void_p))
        assert out.tobytes() == hashlib.sha256(msg).digest()
