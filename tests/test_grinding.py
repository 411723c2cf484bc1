"""Grinding (crates/transcript/src/{prover.rs:29-41,pow.rs}) and
SecurityLevel::grinding_bits (crates/common/src/security.rs:25-32) on the
host: the oracle's BLAKE3 against the published known answers, the
reference's grinding_bits table, the library's proof-of-work search and
transcript grinding against the oracle's, and the reference's grinding
properties (zero bits leave the transcript unchanged; the seed binds the
label, the difficulty and the history)."""
import numpy as np
import pytest

import bitz_amd as G
from oracle import oracle as O


def test_blake3_known_answers():
    # BLAKE3's published test vectors (input bytes i % 251): lengths 0 and 1.
    assert O.blake3(b"").hex() == "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"
    assert O.blake3(bytes([0])).hex() == "2d3adedff11b61f14c886e35afa036736dcd87a74d27b5c1510225d0f592e213"


def test_grinding_bits_table():
    """security.rs:38-46."""
    for coefficient, expected in [(0, 0), (1, 0), (2, 1), (3, 2), (7, 3), (8, 3), (15, 4)]:
        for grinding_bits in (O.grinding_bits, G.grinding_bits):
            assert grinding_bits(128, coefficient) == expected
            assert grinding_bits(100, coefficient) == 0
    for grinding_bits in (O.grinding_bits, G.grinding_bits):
        assert grinding_bits(100, 1 << 28) == 0
        assert grinding_bits(100, (1 << 28) + 1) == 1


def test_pow_search_matches_the_oracle_and_the_definition():
    rng = np.random.default_rng(5)
    for i in range(60):
        seed = rng.integers(0, 256, 16, dtype=np.uint8).tobytes()
        bits = i % 11
        n = G.pow_find(seed, bits)
        assert n == O.pow_find(seed, bits)
        assert G.pow_valid(seed, n, bits) and O.pow_valid(seed, n, bits)
        assert not any(O.pow_valid(seed, k, bits) for k in range(n))  # the first valid nonce
        if bits:  # the definition: the hash's first four bytes, big-endian
            h = O.blake3(b"bitz-pcs-pow-v1" + seed + n.to_bytes(8, "little"))
            assert int.from_bytes(h[:4], "big") >> (32 - bits) == 0
    zero = bytes(16)
    assert G.pow_find(zero, 0) == 0 and G.pow_valid(zero, 0, 0) and not G.pow_valid(zero, 1, 0)
    assert not G.pow_valid(zero, 0, 33) and not O.pow_valid(zero, 0, 33)
    with pytest.raises(G.BitzError):
        G.pow_find(zero, 33)


def test_duplex_grinding_matches_the_oracle():
    for bits in (0, 1, 2, 8):
        gt, ot = G.Transcript.duplex(b"grind"), O.DuplexTranscript(b"grind")
        for t in (gt, ot):
            (t.prover_message if t is gt else (lambda d: O.absorb(ot, 1, d)))(b"first")
        gt.grind(b"test/cubic/v1", bits)
        O.grind(ot, b"test/cubic/v1", bits)
        assert gt.narg() == ot.narg()
        assert len(gt.narg()) == len(b"first") + (8 if bits else 0)
        assert gt.squeeze() == O.squeeze(ot)


def test_zero_bits_leave_the_transcript_unchanged_and_the_seed_binds_its_inputs():
    def after(label, bits, message):
        t = G.Transcript.duplex(b"grinding-test")
        t.public_message(message)
        t.grind(label, bits)
        return t.squeeze(), t.narg()

    plain = G.Transcript.duplex(b"grinding-test")
    plain.public_message(b"first")
    assert after(b"test/cubic/v1", 0, b"first") == (plain.squeeze(), b"")
    base = after(b"test/cubic/v1", 8, b"first")
    assert after(b"test/affine/v1", 8, b"first")[0] != base[0]
    assert after(b"test/cubic/v1", 9, b"first")[0] != base[0]
    assert after(b"test/cubic/v1", 8, b"second")[0] != base[0]
    with pytest.raises(G.BitzError):
        G.Transcript.duplex(b"x").grind(b"l", 33)


def test_a_grind_hook_takes_the_whole_step():
    events = []
    t = G.Transcript.callbacks(lambda k, d: events.append(("absorb", k, d)),
                               lambda: (events.append(("squeeze",)), bytes(16))[1],
                               grind=lambda label, bits: events.append(("grind", label, bits)))
    t.grind(b"gkr/cubic/v1", 2)
    t.grind(b"gkr/cubic/v1", 0)
    assert events == [("grind", b"gkr/cubic/v1", 2)]
    # without a hook, grinding runs over the callbacks: header, seed, nonce
    events.clear()
    t2 = G.Transcript.callbacks(lambda k, d: events.append(("absorb", k, d)),
                                lambda: (events.append(("squeeze",)), bytes(16))[1])
    t2.grind(b"gkr/cubic/v1", 2)
    nonce = O.pow_find(bytes(16), 2)
    assert events == [("absorb", 0, b"bitz-transcript-pow-v1"), ("absorb", 0, (12).to_bytes(8, "little")),
                      ("absorb", 0, b"gkr/cubic/v1"), ("absorb", 0, (2).to_bytes(4, "little")), ("squeeze",),
                      ("absorb", 1, nonce.to_bytes(8, "little"))]
