"""The SHA-256 R1CS producer (bitz_sha256_*), host only.

Pinned by what the reference's own tests hold:
  - compression_circuit: M is 20457 x 7145, A/B/C have 184 rows
    (tooling/cli/tests/sha256_spartan.rs:44-49; crates/circuit/src/
    sha256.rs tests: 7144 witnesses, 20456 bitz calls, 184 constraints);
  - the FIPS 180-4 "abc" block gives ABC_DIGEST (sha256.rs:33-37, 106-115);
  - the 2 KiB stream-order circuit's digest 10fc3c51... and its sizes
    (sha256.rs:715-775: 33 x 20456 + 1 rows of M, 16384 + 33 x 6376 + 1
    columns, 33 x 184 constraints);
and by what the mathematics fixes: every row holds over the integers for a
valid witness (A h * B h = C h, as check_witness, constraints.rs:335-352),
h = M (1 | f) over F2 (integer_witness, constraints.rs:306-333), digests
equal hashlib's or a plain restatement of FIPS 180-4 compression, and a
witness for another message breaks the statement's rows.  The exact column
order within these shapes follows the reference's call order (sha256.rs:
344-470) by construction of one description for both backends; no reference
output lists the matrices entry by entry, so beyond the pins above the order
is unpinned.
"""
import hashlib
import struct

import numpy as np
import pytest

import bitz_amd as G
import bitz_amd.sha256 as S
from common import compress_py

ABC_DIGEST = "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
DIGEST_2KB = "10fc3c51a152e90e5b90319b601d92ccf37290ef53c35ff92507687d8a911a08"
def unpack(words, n):
    return np.unpackbits(np.asarray(words, np.uint64).view(np.uint8), bitorder="little")[:n].astype(np.int64)


def check_witness(c: S.Sha256Circuit, w: S.Sha256Witness):
    """h = M (1 | f) over F2 and A h * B h = C h over the integers, from the
    exported matrices; returns the per-row integer values of C h."""
    ex = c.export()
    h, f = unpack(w.h, c.columns), unpack(w.f, c.witness_bits)
    one_f = np.concatenate([[1], f])
    ms, mp = (x.astype(np.int64) for x in ex["m"])
    rows_of = np.repeat(np.arange(c.columns), np.diff(ms))
    from_m = np.zeros(c.columns, np.int64)
    np.add.at(from_m, rows_of, one_f[mp])
    assert np.array_equal(from_m % 2, h), "h differs from M (1 | f)"
    assert h[0] == 1
    vals = []
    for k in "abc":
        st, co, cf, par = ex[k]
        assert np.array_equal(par[:, 0], (cf & 1).astype(np.uint64)) and not par[:, 1].any()
        r = np.zeros(c.rows, np.int64)
        np.add.at(r, np.repeat(np.arange(c.rows), np.diff(st.astype(np.int64))), h[co.astype(np.int64)] * cf)
        vals.append(r)
    return vals


def test_compression_matches_the_reference_shape_and_fips_abc():
    c = S.Sha256Circuit(S.COMPRESSION_CHAIN, 1)
    assert (c.rows, c.columns, c.witness_bits) == (184, 20457, 7144)
    assert (c.num_row_vars, c.num_column_vars) == (8, 15)  # sha256_spartan.rs:66-67
    abc = struct.pack(">16I", 0x61626380, *([0] * 14), 0x18)
    w = c.witness(abc)
    assert w.digest.hex() == ABC_DIGEST and w.satisfied
    a, b, cz = check_witness(c, w)
    assert np.array_equal(a * b, cz)
    assert not a.any() and not b.any() and not cz.any()  # every SHA-256 row is linear: C h = 0
    # f starts with the inputs: block words (bit j = word bit j), then the state
    f = unpack(w.f, 768)
    words = struct.unpack(">16I", abc) + S.FIPS_IV
    assert all(f[32 * i + j] == (words[i] >> j) & 1 for i in range(24) for j in range(32))


def test_block_aligned_2kb_matches_the_reference_golden_digest():
    msg = bytes((i // 8) % 256 for i in range(0, 2048 * 8, 8))
    c = S.Sha256Circuit(S.BLOCK_ALIGNED, 32)
    assert c.columns == 33 * 20456 + 1
    assert c.witness_bits + 1 == 2048 * 8 + 33 * 6376 + 1
    assert c.rows == 33 * 184
    w = c.witness(msg)
    assert w.digest.hex() == DIGEST_2KB == hashlib.sha256(msg).hexdigest()
    a, b, cz = check_witness(c, w)
    assert np.array_equal(a * b, cz) and w.satisfied
    # the message is the first 16384 Boolean witnesses, in stream order
    bits = np.unpackbits(np.frombuffer(msg, np.uint8))
    assert np.array_equal(unpack(w.f, 16384), bits.astype(np.int64))


@pytest.mark.parametrize("blocks", [0, 1, 3])
def test_block_aligned_digest_is_sha256_of_the_message(blocks):
    rng = np.random.default_rng(blocks)
    msg = rng.integers(0, 256, 64 * blocks, dtype=np.uint8).tobytes()
    c = S.Sha256Circuit(S.BLOCK_ALIGNED, blocks)
    w = c.witness(msg)
    assert w.digest.hex() == hashlib.sha256(msg).hexdigest() and w.satisfied
    assert c.rows == 184 * (blocks + 1)


@pytest.mark.parametrize("blocks", [1, 2, 5])
def test_compression_chain_with_a_chaining_value(blocks):
    rng = np.random.default_rng(40 + blocks)
    msg = rng.integers(0, 256, 64 * blocks, dtype=np.uint8).tobytes()
    iv = [int(x) for x in rng.integers(0, 2**32, 8, dtype=np.uint64)]
    c = S.Sha256Circuit(S.COMPRESSION_CHAIN, blocks)
    assert (c.rows, c.columns, c.witness_bits) == (184 * blocks, 20456 * blocks + 1, 6376 * blocks + 512 * blocks + 256)
    w = c.witness(msg, iv=iv)
    st = iv
    for b in range(blocks):
        st = compress_py(st, msg[64 * b: 64 * b + 64])
    assert w.digest == struct.pack(">8I", *st) and w.satisfied
    a, b_, cz = check_witness(c, w)
    assert np.array_equal(a * b_, cz)
    # the same matrices serve every message and chaining value
    w2 = c.witness(bytes(64 * blocks), iv=list(S.FIPS_IV))
    st = list(S.FIPS_IV)
    for b in range(blocks):
        st = compress_py(st, bytes(64))
    assert w2.satisfied and w2.digest == struct.pack(">8I", *st)


def test_chain_statement_rows():
    raw = b"the quick brown fox " * 7
    padded = raw + b"\x80" + b"\x00" * ((-(len(raw) + 9)) % 64) + struct.pack(">Q", 8 * len(raw))
    B = len(padded) // 64
    c = S.Sha256Circuit(S.CHAIN, B, statement=padded)
    # circuits.rs:167-209: 512 B input rows, 184 B SHA-256 rows, 256 output rows
    assert c.rows == 696 * B + 256
    assert c.columns == 1 + 512 * B + 20456 * B + 256
    assert c.witness_bits == 512 * B + 6376 * B
    w = c.witness(padded)
    assert w.digest.hex() == hashlib.sha256(raw).hexdigest() and w.satisfied
    a, b, cz = check_witness(c, w)
    assert np.array_equal(a * b, cz)
    ex = c.export()
    # constrain_bit (circuits.rs:212-217): A = [(0, 1)], B = [(k, 1)], C = [(0, 1)] iff the bit is 1
    sa, ca, fa = ex["a"][:3]
    sc, cc, fc = ex["c"][:3]
    first = struct.unpack(">I", padded[:4])[0]
    for i in range(32):
        assert list(ca[sa[i]:sa[i + 1]]) == [0] and list(fa[sa[i]:sa[i + 1]]) == [1]
        expect = [0] if (first >> i) & 1 else []
        assert list(cc[sc[i]:sc[i + 1]]) == expect
    # a witness for another message breaks the statement
    other = bytearray(padded)
    other[5] ^= 0x10
    assert not c.witness(bytes(other)).satisfied


def test_statement_digest_is_the_constraint_matrix_digest():
    c = S.Sha256Circuit(S.COMPRESSION_CHAIN, 1)
    ex = c.export()
    _, dg = c.prepare(matrices=False)
    csr = [(ex[k][0], ex[k][1], ex[k][3]) for k in "abc"]
    ref = G.constraint_matrix_digest(ex["m"], c.witness_bits + 1, c.rows, c.columns, csr)
    assert dg == ref and len(dg) == 32


def test_m_rows_follow_the_generator_conventions():
    c = S.Sha256Circuit(S.CHAIN, 1, statement=bytes(64))
    ms, mp = c.export()["m"]
    assert list(mp[ms[0]:ms[1]]) == [0]          # row 0: the constant one
    # the 512 input constraints record the inputs first (witness k -> column k + 1)
    assert all(list(mp[ms[1 + i]:ms[2 + i]]) == [i + 1] for i in range(512))
    # then compress records the FIPS IV state words: constants, [0] or []
    for i in range(256):
        bit = (S.FIPS_IV[i // 32] >> (i % 32)) & 1
        assert list(mp[ms[513 + i]:ms[514 + i]]) == ([0] if bit else [])


def test_errors():
    with pytest.raises(G.BitzError):
        S.Sha256Circuit(7, 1)
    with pytest.raises(G.BitzError):
        S.Sha256Circuit(S.COMPRESSION_CHAIN, 0)
    with pytest.raises(G.BitzError):
        S.Sha256Circuit(S.CHAIN, 1)  # the statement is missing
    c = S.Sha256Circuit(S.BLOCK_ALIGNED, 1)
    with pytest.raises(G.BitzError):
        c.witness(bytes(64), iv=list(S.FIPS_IV))
    with pytest.raises(G.BitzError):
        c.witness(bytes(63))
