"""GPU parity of the Spartan PIOP pieces: Az/Bz/Cz (the reference test's
multiply + padded_mle, crates/spartan/src/piop.rs:352-372), the assignment
MLE (matrix.rs:392-420), and prove_spartan_piop (piop.rs:67-112) against the
oracle's outer sumcheck, bind_and_batch and inner sumcheck composed in the
reference's transcript order."""
import numpy as np
import pytest

import bitz_amd as G
from common import elements, elements_fast, ints, to_int
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, scope="module")
def _fast_oracle():
    O.set_kernel("pclmul")
    yield
    O.set_kernel("portable")


def random_csr(rng, rows, cols, per_row, seed):
    starts, cl = [0], []
    for _ in range(rows):
        c = np.unique(rng.integers(0, cols, size=per_row))
        cl += c.tolist()
        starts.append(len(cl))
    return np.array(starts, np.uint64), np.array(cl, np.uint32), elements_fast(seed, len(cl))


def multiply(csr, h, rows, padded):
    """The reference test's multiply + padded_mle, row by row through the oracle."""
    starts, cols, vals = csr
    out = np.zeros((padded, 2), np.uint64)
    for i in range(rows):
        k0, k1 = int(starts[i]), int(starts[i + 1])
        if k1 > k0:
            prods = O.mul_batch(vals[k0:k1], h[cols[k0:k1].astype(np.int64)])
            acc = 0
            for p in ints(prods):
                acc ^= p
            out[i] = [acc & (2**64 - 1), acc >> 64]
    return out


@pytest.mark.parametrize("rows,row_vars,col_vars,per_row", [(1, 0, 2, 2), (29, 5, 7, 4), (700, 10, 12, 9)])
def test_r1cs_products_match_multiply(rows, row_vars, col_vars, per_row):
    rng = np.random.default_rng(rows)
    csr = [random_csr(rng, rows, 1 << col_vars, per_row, 60 + m) for m in range(3)]
    h = elements_fast(70 + rows, 1 << col_vars)
    prepared = G.PreparedConstraintMatrices(rows, row_vars, col_vars, csr)
    out = prepared.r1cs_products(h)
    dev = prepared.r1cs_products(G.DeviceTable(h))
    for m, t in enumerate((out.az, out.bz, out.cz)):
        ref = multiply(csr[m], h, rows, 1 << row_vars)
        assert np.array_equal(t, ref)
        assert np.array_equal((dev.az, dev.bz, dev.cz)[m].to_host(), ref)


def test_assignment_mle_on_device():
    bits = np.array([1, 0, 1, 1, 0, 0, 0, 1, 1], np.uint8)
    out = G.assignment_mle(bits, 9)
    assert out.shape == (16, 2)
    assert [to_int(e) for e in out] == [int(b) for b in bits] + [0] * 7
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 2, size=5000).astype(np.uint8)
    bits[0] = 1
    out = G.assignment_mle(bits, 5000)
    assert out.shape == (8192, 2)
    assert np.array_equal(out[:5000, 0], bits.astype(np.uint64)) and not out[5000:].any() and not out[:, 1].any()


def prefix(cols, used):
    """The length of a nonzero prefix: all columns, or (used < 1) an odd
    count, so the prefix ends inside a pair."""
    return cols if used >= 1 else min(cols - 1, max(3, int(cols * used) | 1))


def satisfied_r1cs(rng, rows, row_vars, col_vars, per_row, used=1.0, h_used=1.0):
    """A, B random over the first `used` fraction of the columns; a Boolean h
    with h[0] = 1, zero past its first `h_used` fraction; C with one entry per
    row at the constant column, valued (A h)_i (B h)_i, so A h * B h = C h."""
    cols = 1 << col_vars
    span = prefix(cols, used)
    a = random_csr(rng, rows, span, per_row, 80)
    b = random_csr(rng, rows, span, per_row, 81)
    bits = rng.integers(0, 2, size=cols).astype(np.uint8)
    bits[prefix(cols, h_used):] = 0
    bits[0] = 1
    bits[prefix(cols, h_used) - 1] = 1
    h = np.zeros((cols, 2), np.uint64)
    h[:, 0] = bits
    ah, bh = multiply(a, h, rows, rows), multiply(b, h, rows, rows)
    cvals = O.mul_batch(ah, bh)
    c = (np.arange(rows + 1, dtype=np.uint64), np.zeros(rows, np.uint32), cvals)
    return [a, b, c], bits, h


@pytest.mark.parametrize("rows,row_vars,col_vars,per_row,dev,used,h_used",
                         [(13, 4, 6, 3, False, 1.0, 1.0), (300, 9, 11, 6, True, 1.0, 1.0),
                          (1000, 10, 13, 8, True, 1.0, 1.0), (700, 10, 12, 5, True, 0.6, 1.0),
                          (64, 6, 9, 4, False, 0.3, 0.5), (200, 8, 10, 4, True, 0.01, 0.004),
                          (500, 9, 13, 5, True, 0.2, 0.7)])
def test_spartan_piop_matches_oracle(rows, row_vars, col_vars, per_row, dev, used, h_used):
    """used, h_used < 1: the last columns of the matrices (so of D) or of h
    are zero, so the inner sumcheck's folds skip a zero tail that round 0
    measures (D longer than h, h longer than D)."""
    rng = np.random.default_rng(rows + col_vars)
    csr, bits, h = satisfied_r1cs(rng, rows, row_vars, col_vars, per_row, used, h_used)
    m_csr = random_csr(rng, 7, 1 << col_vars, 3, 90)[:2]
    digest = G.constraint_matrix_digest(m_csr, 1 << col_vars, rows, 1 << col_vars, csr)
    prepared = G.PreparedConstraintMatrices(rows, row_vars, col_vars, csr)
    assignment = G.assignment_mle(bits, 1 << col_vars)
    assert np.array_equal(assignment, h)
    products = prepared.r1cs_products(G.DeviceTable(assignment) if dev else assignment)
    gt = G.Transcript.duplex(b"spartan/gpu")
    out = G.prove_spartan_piop(gt, prepared, digest, products, G.DeviceTable(assignment) if dev else assignment)
    # The oracle in the reference's order (piop.rs:85-99).
    ot = O.DuplexTranscript(b"spartan/gpu")
    O.absorb(ot, 0, digest)
    tau = np.array([[(x := O.squeeze(ot)) & (2**64 - 1), x >> 64] for _ in range(row_vars)], np.uint64)
    split = row_vars // 2
    lo = O.eq_table(tau[:split]) if split else np.array([[1, 0]], np.uint64)
    hi = O.eq_table(tau[split:]) if split < row_vars else np.array([[1, 0]], np.uint64)
    az = multiply(csr[0], h, rows, 1 << row_vars)
    bz = multiply(csr[1], h, rows, 1 << row_vars)
    cz = multiply(csr[2], h, rows, 1 << row_vars)
    rc, outer = O.outer_sumcheck(ot, 0, lo, hi, az, bz, cz)
    assert rc == 0
    rho = O.squeeze(ot)
    a_c, b_c, c_c = ints(outer["claims"])
    claim = a_c ^ O.to_int(O.mul(rho, b_c)) ^ O.to_int(O.mul(O.to_int(O.mul(rho, rho)), c_c))
    mats = [O.CsrMatrix(*m) for m in csr]
    d = O.bind_and_batch(mats, rows, row_vars, col_vars, outer["points"], rho)
    rc, inner = O.inner_sumcheck(ot, claim, d, h)
    assert rc == 0
    assert np.array_equal(out.outer_rounds, outer["rounds"])
    assert out.outer_claims == ints(outer["claims"])
    assert np.array_equal(out.inner_rounds, inner["rounds"])
    assert np.array_equal(out.inner_point, inner["points"])
    assert out.matrix_evaluation == ints(inner["evals"])[0]
    assert out.final_claim == inner["final_claim"]
    assert gt.narg() == ot.narg()
    # The R1CS is satisfied (Az * Bz = Cz), so the outer proof of claim 0 verifies.
    vt = O.DuplexTranscript(b"spartan/gpu")
    O.absorb(vt, 0, digest)
    for _ in range(row_vars):
        O.squeeze(vt)
    assert O.outer_verify(vt, 0, tau, out.outer_rounds, np.array(
        [[c & (2**64 - 1), c >> 64] for c in out.outer_claims], np.uint64)) == 0
