"""GPU parity of prove_spartan_piop over SHA-256 R1CS instances from the
producer (bitz_sha256_*): matrices prepared on the device in the parity
projection, the assignment unpacked on the device from the witness's h,
Az/Bz/Cz on the device, then the PIOP (crates/spartan/src/piop.rs:67-112)
against the oracle's outer sumcheck, bind_and_batch and inner sumcheck
composed in the reference's transcript order, with Az/Bz/Cz recomputed on
the host from the exported matrices (0/1 coefficients times 0/1 entries:
an XOR of ANDs)."""
import hashlib
import struct

import numpy as np
import pytest

import bitz_amd as G
import bitz_amd.sha256 as S
from common import ints
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, scope="module")
def _fast_oracle():
    O.set_kernel("pclmul")
    yield
    O.set_kernel("portable")


def products_host(ex, h_bits, rows, padded):
    out = []
    for k in "abc":
        st, co, _, par = ex[k]
        term = (par[:, 0].astype(np.uint8) & h_bits[co.astype(np.int64)]).astype(np.uint8)
        acc = np.zeros(padded, np.uint8)
        np.bitwise_xor.at(acc, np.repeat(np.arange(rows), np.diff(st.astype(np.int64))), term)
        t = np.zeros((padded, 2), np.uint64)
        t[:, 0] = acc
        out.append(t)
    return out


def oracle_piop(label, digest, ex, rows, row_vars, col_vars, h_elems, az, bz, cz, keccak=False):
    ot = O.KeccakTranscript(label) if keccak else O.DuplexTranscript(label)
    O.absorb(ot, 0, digest)
    tau = np.array([[(x := O.squeeze(ot)) & (2**64 - 1), x >> 64] for _ in range(row_vars)], np.uint64)
    split = row_vars // 2
    lo = O.eq_table(tau[:split]) if split else np.array([[1, 0]], np.uint64)
    hi = O.eq_table(tau[split:]) if split < row_vars else np.array([[1, 0]], np.uint64)
    rc, outer = O.outer_sumcheck(ot, 0, lo, hi, az, bz, cz)
    assert rc == 0
    rho = O.squeeze(ot)
    a_c, b_c, c_c = ints(outer["claims"])
    claim = a_c ^ O.to_int(O.mul(rho, b_c)) ^ O.to_int(O.mul(O.to_int(O.mul(rho, rho)), c_c))
    mats = [O.CsrMatrix(ex[k][0], ex[k][1], ex[k][3]) for k in "abc"]
    d = O.bind_and_batch(mats, rows, row_vars, col_vars, outer["points"], rho)
    rc, inner = O.inner_sumcheck(ot, claim, d, h_elems)
    assert rc == 0
    return ot, outer, inner


def prove_and_check(c: S.Sha256Circuit, message: bytes, iv=None, label=b"sha256/gpu", keccak=False):
    """The PIOP of the statement on the device (duplex or Keccak transcript),
    each output against the oracle's composition in the reference's order."""
    w = c.witness(message, iv=iv)
    assert w.satisfied
    prepared, digest = c.prepare()
    rv, cv = c.num_row_vars, c.num_column_vars
    assignment = G.assignment_mle_words(w.h, c.columns, c.columns, device=True)
    products = prepared.r1cs_products(assignment)
    gt = G.Transcript.keccak(label) if keccak else G.Transcript.duplex(label)
    out = G.prove_spartan_piop(gt, prepared, digest, products, assignment)
    ex = c.export()
    h_bits = np.unpackbits(w.h.view(np.uint8), bitorder="little")[: 1 << cv]
    h_bits = np.concatenate([h_bits, np.zeros((1 << cv) - len(h_bits), np.uint8)])
    h_elems = np.zeros((1 << cv, 2), np.uint64)
    h_elems[:, 0] = h_bits
    assert np.array_equal(assignment.to_host(), h_elems)
    az, bz, cz = products_host(ex, h_bits, c.rows, 1 << rv)
    assert np.array_equal(products.az.to_host(), az)
    assert np.array_equal(products.bz.to_host(), bz)
    assert np.array_equal(products.cz.to_host(), cz)
    ot, outer, inner = oracle_piop(label, digest, ex, c.rows, rv, cv, h_elems, az, bz, cz, keccak)
    assert np.array_equal(out.outer_rounds, outer["rounds"])
    assert out.outer_claims == ints(outer["claims"])
    assert np.array_equal(out.inner_rounds, inner["rounds"])
    assert np.array_equal(out.inner_point, inner["points"])
    assert out.matrix_evaluation == ints(inner["evals"])[0]
    assert out.final_claim == inner["final_claim"]
    assert gt.narg() == ot.narg()
    return w, out


def test_compression_piop_matches_oracle():
    rng = np.random.default_rng(3)
    c = S.Sha256Circuit(S.COMPRESSION_CHAIN, 1)
    msg = rng.integers(0, 256, 64, dtype=np.uint8).tobytes()
    iv = [int(x) for x in rng.integers(0, 2**32, 8, dtype=np.uint64)]
    prove_and_check(c, msg, iv)


@pytest.mark.parametrize("blocks", [2, 5])
def test_chain_statement_piop_matches_oracle(blocks):
    rng = np.random.default_rng(10 + blocks)
    msg = rng.integers(0, 256, 64 * blocks, dtype=np.uint8).tobytes()
    c = S.Sha256Circuit(S.CHAIN, blocks, statement=msg)
    prove_and_check(c, msg, label=b"sha256-chain/%d" % blocks)


def test_block_aligned_piop_matches_oracle():
    msg = bytes((i // 8) % 256 for i in range(0, 2048 * 8, 8))[:128]
    c = S.Sha256Circuit(S.BLOCK_ALIGNED, 2)
    w, _ = prove_and_check(c, msg)
    assert w.digest == hashlib.sha256(msg).digest()


def test_sha256_chain_64_blocks_end_to_end():
    """The 64-block sha256-chain statement: digest of a padded message, every
    round polynomial and claim against the oracle."""
    rng = np.random.default_rng(64)
    raw = rng.integers(0, 256, 64 * 64 - 9 - 7, dtype=np.uint8).tobytes()
    padded = raw + b"\x80" + b"\x00" * ((-(len(raw) + 9)) % 64) + struct.pack(">Q", 8 * len(raw))
    assert len(padded) == 64 * 64
    c = S.Sha256Circuit(S.CHAIN, 64, statement=padded)
    assert (c.num_row_vars, c.num_column_vars) == (16, 21)
    w, out = prove_and_check(c, padded, label=b"sha256-chain/64")
    assert w.digest == hashlib.sha256(raw).digest()
