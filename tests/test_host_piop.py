"""CPU checks of the PIOP's host logic: the constraint-matrix digest
(crates/spartan/src/matrix.rs:562-627) against an independent restatement
over hashlib, and the assignment checks of build_assignment_mle
(matrix.rs:392-404), which fail before any device work."""
import hashlib
import struct

import numpy as np
import pytest

import bitz_amd as G
from common import elements_fast


def u64(v: int) -> bytes:
    return struct.pack("<Q", v)


def digest_model(m_starts, m_pos, m_cols, rows, cols, csr) -> bytes:
    h = hashlib.sha256()
    h.update(b"bitz/spartan/constraint-matrices/v1")
    h.update(b"M")
    h.update(u64(len(m_starts) - 1))
    h.update(u64(m_cols))
    for i in range(len(m_starts) - 1):
        h.update(u64(int(m_starts[i + 1] - m_starts[i])))
        for k in range(int(m_starts[i]), int(m_starts[i + 1])):
            h.update(u64(int(m_pos[k])))
    for label, (starts, cs, vals) in zip((b"A", b"B", b"C"), csr):
        h.update(label)
        h.update(u64(rows))
        h.update(u64(cols))
        for i in range(rows):
            h.update(u64(int(starts[i + 1] - starts[i])))
            for k in range(int(starts[i]), int(starts[i + 1])):
                h.update(u64(int(cs[k])))
                h.update(u64(16))
                h.update(vals[k].astype("<u8").tobytes())  # F128::to_bytes: lo then hi, little-endian
    return h.digest()


def random_csr(rng, rows, cols, per_row, seed):
    starts, cl = [0], []
    for _ in range(rows):
        c = np.unique(rng.integers(0, cols, size=per_row))
        cl += c.tolist()
        starts.append(len(cl))
    return np.array(starts, np.uint64), np.array(cl, np.uint32), elements_fast(seed, len(cl))


@pytest.mark.parametrize("rows,cols,per_row", [(1, 1, 1), (37, 200, 5), (300, 1000, 9)])
def test_constraint_matrix_digest_matches_hashlib_model(rows, cols, per_row):
    rng = np.random.default_rng(rows + cols)
    csr = [random_csr(rng, rows, cols, per_row, 40 + m) for m in range(3)]
    m_starts, m_pos, _ = random_csr(rng, rows // 2 + 1, cols, per_row, 50)
    got = G.constraint_matrix_digest((m_starts, m_pos), cols, rows, cols, csr)
    assert got == digest_model(m_starts, m_pos, cols, rows, cols, csr)
    # any change of a coefficient, a position or a length changes it
    csr2 = [(s, c, v.copy()) for s, c, v in csr]
    csr2[1][2][0, 1] ^= 1
    assert G.constraint_matrix_digest((m_starts, m_pos), cols, rows, cols, csr2) != got
    assert G.constraint_matrix_digest((m_starts, m_pos), cols + 1, rows, cols, csr) != got


def test_assignment_checks_fail_before_device_work():
    with pytest.raises(G.BitzError) as e:
        G.assignment_mle(np.array([1, 0, 1], np.uint8), 4)
    assert e.value.variant == "InvalidAssignmentLength"
    with pytest.raises(G.BitzError) as e:
        G.assignment_mle(np.array([0, 1, 1], np.uint8), 3)
    assert e.value.variant == "InvalidAssignmentConstant"
