"""The proof container (crates/host/src/wire_proof.rs) and the transcript's
narg/hint streams (crates/transcript/src/proof.rs), on the host: the
reference's own wire-format cases (the grammar written out byte by byte,
every prefix and suffix refused, each header field's check and their
order, overflowing lengths, payload flips carried), the round trip over
random streams, and a GKR proof's narg in its container."""
import struct

import numpy as np
import pytest

import bitz_amd as G

HEADER_LEN = 32
OFF = {"magic": 0, "version": 8, "header_len": 10, "flags": 12, "narg": 16, "hint": 24}
ONE_RECORD = bytes([1] + [0] * 15)


def test_the_encoding_is_the_grammar_in_the_module_docs():
    """wire_proof.rs:199-223."""
    expect = (b"bitzPCS\0" + bytes([1, 0]) + bytes([0x20, 0]) + bytes(4) + bytes([4, 0, 0, 0, 0, 0, 0, 0]) +
              bytes([3, 0, 0, 0, 0, 0, 0, 0]) + b"NNNN" + b"HHH")
    assert G.encode_proof(b"NNNN", b"HHH") == expect
    assert G.decode_proof(expect) == (b"NNNN", b"HHH")
    assert len(G.encode_proof(ONE_RECORD)) == HEADER_LEN + 16
    assert G.decode_proof(G.encode_proof(ONE_RECORD)) == (ONE_RECORD, b"")
    assert G.encode_proof(b"", b"") == expect[:16] + bytes(16)
    assert G.decode_proof(G.encode_proof(b"")) == (b"", b"")


def variant(data):
    with pytest.raises(G.ProofDecodingError) as e:
        G.decode_proof(data)
    return e.value.variant, e.value.detail


def test_nothing_but_the_exact_encoding_is_accepted():
    """wire_proof.rs:225-263."""
    data = G.encode_proof(bytes([1, 2, 3, 4]), bytes([5, 6]))
    assert len(data) == HEADER_LEN + 6
    for n in range(len(data)):
        v, detail = variant(data[:n])
        if n < HEADER_LEN:
            assert v == "OutOfInput", n
        else:
            assert (v, detail) == ("LengthMismatch", len(data)), n
    for suffix in (b"\0", bytes(64), b"\xff" * 3):
        assert variant(data + suffix) == ("LengthMismatch", len(data))
    assert G.decode_proof(data) == (bytes([1, 2, 3, 4]), bytes([5, 6]))


def patched(offset, value: bytes):
    data = bytearray(G.encode_proof(ONE_RECORD))
    data[offset: offset + len(value)] = value
    return bytes(data)


def test_the_header_fields_are_checked_in_order():
    """wire_proof.rs:265-331."""
    data = bytearray(G.encode_proof(ONE_RECORD))
    data[OFF["magic"]] ^= 1
    assert variant(data)[0] == "BadMagic"
    data = bytearray(G.encode_proof(ONE_RECORD))
    data[OFF["magic"] + 7] = 0xFF
    data[OFF["version"]] = 9
    data[OFF["flags"]] = 1
    assert variant(data)[0] == "BadMagic"  # the order of the checks is part of the format
    for version in (0, 2, 0x0100, 0xFFFF):
        assert variant(patched(OFF["version"], struct.pack("<H", version))) == ("UnsupportedVersion", version)
    for header_len in (0, 31, 33, 40, 0xFFFF):
        assert variant(patched(OFF["header_len"], struct.pack("<H", header_len))) == ("BadHeaderLen", header_len)
    for bit in range(32):
        assert variant(patched(OFF["flags"], struct.pack("<I", 1 << bit))) == ("ReservedFlagsSet", 1 << bit)
    for name, truth in (("narg", 16), ("hint", 0)):
        for declared in (0, 15, 16, 17, 2**32 - 1):
            if declared != truth:
                assert variant(patched(OFF[name], struct.pack("<Q", declared)))[0] == "LengthMismatch"


def test_declared_lengths_that_overflow_are_refused():
    """wire_proof.rs:333-349."""
    assert variant(patched(OFF["narg"], struct.pack("<Q", 2**64 - 1)))[0] == "LengthOverflow"
    half = 2**63
    data = bytearray(G.encode_proof(ONE_RECORD))
    data[16:24] = struct.pack("<Q", half)
    data[24:32] = struct.pack("<Q", half)
    assert variant(bytes(data))[0] == "LengthOverflow"


def test_a_flip_in_the_payload_is_carried_not_caught():
    """wire_proof.rs:351-368."""
    honest = (bytes([1, 2, 3, 4]), bytes([5, 6]))
    data = G.encode_proof(*honest)
    for off in range(HEADER_LEN, len(data)):
        t = bytearray(data)
        t[off] ^= 0x80
        decoded = G.decode_proof(bytes(t))
        assert decoded != honest and G.encode_proof(*decoded) == bytes(t)


def test_round_trips_and_random_bytes_are_well_behaved():
    """host/src/lib.rs:46-118: an accepted container has one spelling."""
    rng = np.random.default_rng(9)
    for _ in range(300):
        narg = rng.integers(0, 256, int(rng.integers(0, 64)), dtype=np.uint8).tobytes()
        hints = rng.integers(0, 256, int(rng.integers(0, 64)), dtype=np.uint8).tobytes()
        data = G.encode_proof(narg, hints)
        assert G.decode_proof(data) == (narg, hints)
        junk = rng.integers(0, 256, int(rng.integers(0, 80)), dtype=np.uint8).tobytes()
        for candidate in (junk, data[: int(rng.integers(0, len(data) + 1))] + junk[:3]):
            try:
                got = G.decode_proof(candidate)
            except G.ProofDecodingError:
                continue
            assert G.encode_proof(*got) == candidate


def test_a_transcripts_proof_is_its_narg_and_hints_in_the_container():
    t = G.Transcript.duplex(b"proof")
    t.prover_message(b"\x01" * 32)
    c = t.squeeze()
    t.hint(b"\xaa" * 5)
    t.prover_message(b"\x02" * 16)
    assert t.hints() == b"\xaa" * 5
    assert G.decode_proof(t.proof()) == (t.narg(), b"\xaa" * 5)
    # hints never reach the sponge: the same messages without the hint squeeze the same
    u = G.Transcript.duplex(b"proof")
    u.prover_message(b"\x01" * 32)
    assert u.squeeze() == c
    u.prover_message(b"\x02" * 16)
    assert u.squeeze() == t.squeeze()
    cb = G.Transcript.callbacks(lambda k, d: None, lambda: bytes(16))
    with pytest.raises(G.BitzError):
        cb.hint(b"x")
