"""SHA-256 R1CS instances for the prover (bitz_sha256_* in the C ABI).

The reference builds its SHA-256 constraint systems with the circuit DSL
(crates/circuit/src/sha256.rs) run through ConstraintGenerator for M, A, B,
C (crates/circuit/src/constraints.rs:578-790) and Witgen for the witness
(crates/circuit/src/witgen.rs:425-500).  Here the library produces both on
the host (C++); this module exposes them as numpy arrays and as prepared
prover matrices.

Kinds:
  COMPRESSION_CHAIN  inputs = message words then the chaining value; one
                     block is the reference's compression_circuit.
  CHAIN              the CLI's sha256-chain statement: message and digest
                     bits constrained as rows (tooling/cli/src/circuits.rs).
  BLOCK_ALIGNED      sha256_block_aligned_circuit, the reference benchmark's
                     circuit (message stream + padding block).

Coefficients are integers; the prover works over GF(2^128) through the
parity projection Z -> F2 (a ring homomorphism, so A h * B h = C h over the
integers stays true in the field).  The reference draws a prime for its
own projection (tooling/cli/src/end_to_end.rs:401-431); it has no GF(2^128)
projection of these matrices.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import BitzError, PreparedConstraintMatrices, _check, init, lib

COMPRESSION_CHAIN, CHAIN, BLOCK_ALIGNED = 0, 1, 2
KIND_NAMES = {"sha256-compression-chain": COMPRESSION_CHAIN, "sha256-chain": CHAIN,
              "sha256-block-aligned": BLOCK_ALIGNED}
FIPS_IV = (0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19)


def _vars(n: int) -> int:
    return max(int(n) - 1, 0).bit_length()


@dataclass
class Sha256Witness:
    f: np.ndarray        # packed Boolean witness (inputs, then hint bits)
    h: np.ndarray        # packed assignment: 1, then every recorded bit
    digest: bytes        # final state, big-endian words
    satisfied: bool      # every row holds over the integers


class Sha256Circuit:
    """One SHA-256 constraint system; see the module notes for the kinds."""

    def __init__(self, kind: int, blocks: int, statement: bytes | None = None):
        self.kind, self.blocks = int(kind), int(blocks)
        buf = None
        if statement is not None:
            statement = bytes(statement)
            if len(statement) != 64 * self.blocks:
                raise BitzError(-7, "the statement is 64 bytes per block")
            buf = C.create_string_buffer(statement, len(statement))
        h = C.c_void_p()
        _check(lib().bitz_sha256_circuit(C.byref(h), self.kind, self.blocks, buf))
        self.h = h
        self.statement = statement
        rows, cols, wit, mpos = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_size_t()
        nnz = (C.c_size_t * 3)()
        _check(lib().bitz_sha256_shape(h, C.byref(rows), C.byref(cols), C.byref(wit), nnz, C.byref(mpos)))
        self.rows, self.columns, self.witness_bits = rows.value, cols.value, wit.value
        self.nnz = tuple(int(x) for x in nnz)
        self.m_positions = mpos.value
        self.num_row_vars, self.num_column_vars = _vars(self.rows), _vars(self.columns)

    def export(self) -> dict:
        """M, A, B, C as CSR arrays: {"m": (starts, positions), "a"/"b"/"c":
        (starts, columns, integer coefficients, parity values (n, 2) u64)}."""
        out = {}
        starts = [np.zeros(self.rows + 1, np.uint64) for _ in range(3)]
        cols = [np.zeros(max(n, 1), np.uint32) for n in self.nnz]
        coefs = [np.zeros(max(n, 1), np.int64) for n in self.nnz]
        par = [np.zeros((max(n, 1), 2), np.uint64) for n in self.nnz]
        m_starts = np.zeros(self.columns + 1, np.uint64)
        m_pos = np.zeros(max(self.m_positions, 1), np.uint32)
        arr = lambda xs: (C.c_void_p * 3)(*[x.ctypes.data for x in xs])  # noqa: E731
        _check(lib().bitz_sha256_export(self.h, arr(starts), arr(cols), arr(coefs), arr(par),
                                        m_starts.ctypes.data_as(C.c_void_p), m_pos.ctypes.data_as(C.c_void_p)))
        out["m"] = (m_starts, m_pos[: self.m_positions])
        for k, name in enumerate("abc"):
            n = self.nnz[k]
            out[name] = (starts[k], cols[k][:n], coefs[k][:n], par[k][:n])
        return out

    def m_transpose(self):
        """M in MaterializedMTranspose's layout (CSC, crates/circuit/src/
        matrix_transpose.rs:282-288): per column of M (the constant, then each
        witness bit), the assignment entries that read it, ascending.
        Returns (column_offsets, row_indices) as uint32 arrays."""
        ms, mp = self.export()["m"]
        rows_of = np.repeat(np.arange(self.columns, dtype=np.uint32), np.diff(ms.astype(np.int64)))
        order = np.argsort(mp, kind="stable")
        counts = np.bincount(mp, minlength=self.witness_bits + 1)
        offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
        return offsets, rows_of[order]

    def witness(self, message: bytes, iv=None) -> Sha256Witness:
        """The witness for `message` (64 bytes per block); iv (8 words) only
        for the compression chain (None = the FIPS IV)."""
        message = bytes(message)
        if len(message) != 64 * self.blocks:
            raise BitzError(-7, "the message is 64 bytes per block")
        f = np.zeros(max((self.witness_bits + 63) // 64, 1), np.uint64)
        h = np.zeros(max((self.columns + 63) // 64, 1), np.uint64)
        dg = (C.c_uint8 * 32)()
        sat = C.c_int32()
        ivp = (C.c_uint32 * 8)(*iv) if iv is not None else None
        buf = C.create_string_buffer(message, max(len(message), 1))
        _check(lib().bitz_sha256_witness(self.h, buf, ivp, f.ctypes.data_as(C.c_void_p),
                                         h.ctypes.data_as(C.c_void_p), dg, C.byref(sat)))
        return Sha256Witness(f, h, bytes(dg), bool(sat.value))

    def prepare(self, matrices: bool = True, digest: bool = True):
        """(PreparedConstraintMatrices of the parity projection or None, its
        statement digest (constraint_matrix_digest) or None)."""
        if matrices:
            init()
        h = C.c_void_p()
        dg = (C.c_uint8 * 32)()
        _check(lib().bitz_sha256_prepare(self.h, C.byref(h) if matrices else None, dg if digest else None))
        prepared = (PreparedConstraintMatrices._adopt(h, self.rows, self.num_row_vars, self.num_column_vars)
                    if matrices else None)
        return prepared, (bytes(dg) if digest else None)

    def __del__(self):
        try:
            lib().bitz_sha256_free(self.h)
        except Exception:
            pass
