"""Keccak-f[1600] and the Keccak duplex transcript on the host: the oracle's
permutation (constants generated from FIPS 202's definitions) and the
library's (the code the device runs too) against hashlib's SHA3 and SHAKE,
and the duplex against SHAKE256, which it equals for a squeeze that follows
the start and messages only (rate 136, pad 0x1F)."""
import hashlib
import struct

import numpy as np
import pytest

import bitz_amd as G
from oracle import oracle as O

LENGTHS = [0, 1, 71, 72, 73, 135, 136, 137, 167, 168, 169, 200, 1000]
SPONGES = [("sha3_224", 144, 0x06, 28), ("sha3_256", 136, 0x06, 32), ("sha3_384", 104, 0x06, 48),
           ("sha3_512", 72, 0x06, 64), ("shake_128", 168, 0x1F, 300), ("shake_256", 136, 0x1F, 300)]


def expected(name, msg, out_len):
    h = getattr(hashlib, name)(msg)
    return h.digest(out_len) if name.startswith("shake") else h.digest()


@pytest.mark.parametrize("name,rate,pad,out_len", SPONGES)
def test_sponges_match_hashlib(name, rate, pad, out_len):
    rng = np.random.default_rng(rate)
    for n in LENGTHS:
        msg = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        want = expected(name, msg, out_len)
        assert O.keccak_hash(msg, rate, pad, out_len) == want, (name, n)
        assert G.keccak_hash(msg, rate, pad, out_len) == want, (name, n)


def framed(tag: bytes, data: bytes) -> bytes:
    return tag + struct.pack("<Q", len(data)) + data


def test_duplex_squeezes_after_messages_are_shake256():
    for domain in (b"", b"bitz", b"x" * 300):
        for msgs in ([], [(0, b"abc")], [(1, bytes(range(64))), (0, b"y" * 140)]):
            gt, ot = G.Transcript.keccak(domain), O.KeccakTranscript(domain)
            stream = framed(b"D", domain)
            for kind, m in msgs:
                gt.absorb(kind, m)
                O.absorb(ot, kind, m)
                stream += framed(b"P" if kind == 0 else b"N", m)
            want = int.from_bytes(hashlib.shake_256(stream + b"S").digest(16), "little")
            assert gt.squeeze() == want == O.squeeze(ot)


def test_duplex_matches_the_oracle_over_a_long_transcript():
    rng = np.random.default_rng(3)
    gt, ot = G.Transcript.keccak(b"long"), O.KeccakTranscript(b"long")
    for i in range(400):
        op = int(rng.integers(0, 3))
        if op == 2:
            assert gt.squeeze() == O.squeeze(ot), i
        else:
            m = rng.integers(0, 256, int(rng.integers(0, 300)), dtype=np.uint8).tobytes()
            gt.absorb(op, m)
            O.absorb(ot, op, m)
    assert gt.narg() == ot.narg()


@pytest.mark.gpu
def test_device_permutation_matches_hashlib():
    """The permutation's device build (one thread) against FIPS 202 answers."""
    rng = np.random.default_rng(17)
    for name, rate, pad, out_len in SPONGES:
        for n in (0, 1, rate - 1, rate, rate + 1, 3 * rate + 5):
            msg = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
            assert G.keccak_hash(msg, rate, pad, out_len, device=True) == expected(name, msg, out_len), (name, n)
