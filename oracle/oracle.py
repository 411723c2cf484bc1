"""ctypes bindings for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product never imports it.

Field elements travel as numpy uint64 arrays of shape (n, 2) = [lo, hi],
the byte layout of the reference's `#[repr(C, align(16))] F128`
(crates/field/src/gf128.rs:60-68).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OF128(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64)]


ABSORB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(C.c_uint8), C.c_size_t)
SQUEEZE = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_uint8))


GRIND = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_uint8), C.c_size_t, C.c_uint32)


class Transcript(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("absorb", C.c_void_p), ("squeeze", C.c_void_p), ("grind", C.c_void_p)]


class Seeded(C.Structure):
    _fields_ = [("state", C.c_uint64), ("narg", C.POINTER(C.c_uint8)),
                ("narg_len", C.c_size_t), ("narg_cap", C.c_size_t),
                ("absorbed_bytes", C.c_uint64)]


class KeccakDuplex(C.Structure):
    _fields_ = [("a", C.c_uint64 * 25), ("pos", C.c_size_t), ("narg", C.POINTER(C.c_uint8)),
                ("narg_len", C.c_size_t), ("narg_cap", C.c_size_t)]


class Duplex(C.Structure):
    _fields_ = [("state", C.c_uint8 * 32), ("narg", C.POINTER(C.c_uint8)),
                ("narg_len", C.c_size_t), ("narg_cap", C.c_size_t)]


class CSR(C.Structure):
    _fields_ = [("row_starts", C.c_void_p), ("columns", C.c_void_p), ("values", C.c_void_p)]


def build() -> None:
    """Compile oracle/liboracle.so from the C restatement (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        L.oracle_mul.restype = OF128
        L.oracle_mul.argtypes = [OF128, OF128]
        L.oracle_square.restype = OF128
        L.oracle_square.argtypes = [OF128]
        L.oracle_mul_x.restype = OF128
        L.oracle_mul_x.argtypes = [OF128]
        L.oracle_inv.restype = OF128
        L.oracle_inv.argtypes = [OF128]
        L.oracle_reduce.restype = OF128
        L.oracle_reduce.argtypes = [P]
        L.oracle_clmul64.argtypes = [C.c_uint64, C.c_uint64, P]
        L.oracle_clmul128.argtypes = [OF128, OF128, P]
        L.oracle_mul_batch.argtypes = [P, P, P, C.c_size_t]
        L.oracle_set_kernel.argtypes = [C.c_int]
        L.oracle_get_kernel.restype = C.c_int
        L.oracle_seeded_init.argtypes = [P, C.c_uint64]
        L.oracle_seeded_free.argtypes = [P]
        L.oracle_seeded_transcript.restype = Transcript
        L.oracle_seeded_transcript.argtypes = [P]
        L.oracle_duplex_init.argtypes = [P, P, C.c_size_t]
        L.oracle_duplex_free.argtypes = [P]
        L.oracle_duplex_transcript.restype = Transcript
        L.oracle_duplex_transcript.argtypes = [P]
        L.oracle_sha256.argtypes = [P, C.c_size_t, P]
        L.oracle_eq_table.argtypes = [P, C.c_int, P]
        L.oracle_eq_eval.restype = OF128
        L.oracle_eq_eval.argtypes = [P, P, C.c_int]
        L.oracle_mle_fold.argtypes = [P, C.c_size_t, OF128, P]
        L.oracle_mle_evaluate.restype = OF128
        L.oracle_mle_evaluate.argtypes = [P, P, C.c_int]
        L.oracle_inner_product.restype = OF128
        L.oracle_inner_product.argtypes = [P, P, C.c_size_t]
        L.oracle_outer_sumcheck.restype = C.c_int
        L.oracle_outer_sumcheck.argtypes = [P, OF128, P, C.c_int, P, C.c_int, P, P, P,
                                            C.c_int, C.c_int, C.c_int, P, P, P, P]
        L.oracle_inner_sumcheck.restype = C.c_int
        L.oracle_inner_sumcheck.argtypes = [P, OF128, P, C.c_int, P, C.c_int, P, P, P, P]
        L.oracle_sumcheck_verify.restype = C.c_int
        L.oracle_sumcheck_verify.argtypes = [P, OF128, P, C.c_int, C.c_int, P, P]
        L.oracle_outer_verify.restype = C.c_int
        L.oracle_outer_verify.argtypes = [P, OF128, P, C.c_int, P, P, P]
        L.oracle_bind_and_batch.restype = C.c_int
        L.oracle_bind_and_batch.argtypes = [P, C.c_size_t, C.c_int, C.c_int, P, OF128, P]
        L.oracle_csr_multiply.argtypes = [P, C.c_size_t, P, P]
        L.oracle_evaluate_batched.restype = C.c_int
        L.oracle_evaluate_batched.argtypes = [P, C.c_size_t, C.c_int, C.c_int, P, OF128, P, P]
        L.oracle_mt_apply.argtypes = [P, C.c_size_t, P, P, P]
        L.oracle_mt_transpose_inner.restype = C.c_int
        L.oracle_mt_transpose_inner.argtypes = [P, C.c_size_t, P, P, C.c_size_t, P,
                                                C.c_size_t, P]
        L.oracle_gkr_fold_layers.restype = C.c_int
        L.oracle_gkr_fold_layers.argtypes = [P, C.c_int, C.c_int, P]
        L.oracle_gkr_eq_tables.argtypes = [P, C.c_int, P]
        L.oracle_gkr_prove_layer.restype = OF128
        L.oracle_gkr_prove_layer.argtypes = [P, P, C.c_int, P, P, P, C.c_uint32]
        L.oracle_gkr_prove.restype = OF128
        L.oracle_gkr_prove.argtypes = [P, P, C.c_int, P, C.c_int, P, C.c_uint32]
        L.oracle_transcript_grind.argtypes = [P, P, C.c_size_t, C.c_uint32]
        L.oracle_keccak_f1600.argtypes = [P]
        L.oracle_keccak_hash.argtypes = [P, C.c_size_t, C.c_int, C.c_uint8, P, C.c_size_t]
        L.oracle_keccak_duplex_init.argtypes = [P, P, C.c_size_t]
        L.oracle_keccak_duplex_free.argtypes = [P]
        L.oracle_keccak_duplex_transcript.restype = Transcript
        L.oracle_keccak_duplex_transcript.argtypes = [P]
        L.oracle_post_gkr_prove.restype = C.c_int
        L.oracle_post_gkr_prove.argtypes = [P, P, C.c_int, P, C.c_int, OF128, P, C.c_size_t, C.c_uint32, P, P]
        L.oracle_post_gkr_verify.restype = C.c_int
        L.oracle_post_gkr_verify.argtypes = [P, P, C.c_int, P, C.c_int, OF128, C.c_uint32, P, OF128, P]
        L.oracle_blake3_small.restype = C.c_int
        L.oracle_blake3_small.argtypes = [P, C.c_size_t, P]
        L.oracle_pow_valid.restype = C.c_int
        L.oracle_pow_valid.argtypes = [P, C.c_uint64, C.c_uint32]
        L.oracle_pow_find.restype = C.c_uint64
        L.oracle_pow_find.argtypes = [P, C.c_uint32]
        L.oracle_grinding_bits.restype = C.c_uint32
        L.oracle_grinding_bits.argtypes = [C.c_uint32, C.c_uint64]
        L.oracle_gkr_verify.restype = C.c_int
        L.oracle_gkr_verify.argtypes = [P, OF128, P, C.c_int, C.c_int, P, P, P]
        _lib = L
    return _lib


# ---- element helpers --------------------------------------------------------
def f(x) -> OF128:
    if isinstance(x, OF128):
        return x
    if isinstance(x, (tuple, list, np.ndarray)):
        return OF128(int(x[0]), int(x[1]))
    x = int(x)
    return OF128(x & (2**64 - 1), x >> 64)


def to_int(e) -> int:
    if isinstance(e, OF128):
        return e.lo | (e.hi << 64)
    return int(e[0]) | (int(e[1]) << 64)


def arr(n: int) -> np.ndarray:
    return np.zeros((n, 2), dtype=np.uint64)


def ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


def as_elems(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a.reshape(-1, 2)


def elem_array(e: OF128) -> np.ndarray:
    return np.array([[e.lo, e.hi]], dtype=np.uint64)


def set_kernel(kernel: str) -> None:
    lib().oracle_set_kernel({"portable": 0, "pclmul": 1}[kernel])


def mul(a, b) -> OF128:
    return lib().oracle_mul(f(a), f(b))


def mul_batch(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a, b = as_elems(a), as_elems(b)
    out = arr(len(a))
    lib().oracle_mul_batch(ptr(a), ptr(b), ptr(out), len(a))
    return out


# ---- transcripts ---------------------------------------------------------------
class SeededTranscript:
    """Challenges from splitmix64(seed) (crates/gkr/src/lib.rs:986-999)."""

    def __init__(self, seed: int):
        self._s = Seeded()
        lib().oracle_seeded_init(C.byref(self._s), seed)
        self.t = lib().oracle_seeded_transcript(C.byref(self._s))

    def __del__(self):
        try:
            lib().oracle_seeded_free(C.byref(self._s))
        except Exception:
            pass

    @property
    def ref(self):
        return C.byref(self.t)

    def narg(self) -> bytes:
        return C.string_at(self._s.narg, self._s.narg_len) if self._s.narg_len else b""


class DuplexTranscript:
    """SHA-256 ratchet: every challenge depends on every earlier message."""

    def __init__(self, domain: bytes):
        self._d = Duplex()
        buf = C.create_string_buffer(domain, len(domain))
        lib().oracle_duplex_init(C.byref(self._d), buf, len(domain))
        self.t = lib().oracle_duplex_transcript(C.byref(self._d))

    def __del__(self):
        try:
            lib().oracle_duplex_free(C.byref(self._d))
        except Exception:
            pass

    @property
    def ref(self):
        return C.byref(self.t)

    def narg(self) -> bytes:
        return C.string_at(self._d.narg, self._d.narg_len) if self._d.narg_len else b""


class KeccakTranscript:
    """The Keccak duplex (rate 136 bytes; framing in oracle/bitz_oracle.h)."""

    def __init__(self, domain: bytes):
        self._d = KeccakDuplex()
        buf = C.create_string_buffer(domain, max(len(domain), 1))
        lib().oracle_keccak_duplex_init(C.byref(self._d), buf, len(domain))
        self.t = lib().oracle_keccak_duplex_transcript(C.byref(self._d))

    def __del__(self):
        try:
            lib().oracle_keccak_duplex_free(C.byref(self._d))
        except Exception:
            pass

    @property
    def ref(self):
        return C.byref(self.t)

    def narg(self) -> bytes:
        return C.string_at(self._d.narg, self._d.narg_len) if self._d.narg_len else b""


def keccak_hash(msg: bytes, rate: int, pad: int, out_len: int) -> bytes:
    out = (C.c_uint8 * max(out_len, 1))()
    buf = C.create_string_buffer(msg, max(len(msg), 1))
    lib().oracle_keccak_hash(buf, len(msg), rate, pad, out, out_len)
    return bytes(out[:out_len])


def keccak_f1600(lanes) -> list:
    a = (C.c_uint64 * 25)(*lanes)
    lib().oracle_keccak_f1600(a)
    return list(a)


class ListTranscript:
    """Challenges from a fixed list (tests that need chosen challenges);
    messages are recorded.  With grind_hook, grinding is handed over whole
    and recorded as ("grind", label, bits) in `events` (beside ("absorb",
    kind, bytes) and ("squeeze",)); without it, grinding runs over absorb and
    squeeze (pow.rs framing) and its seed takes a challenge from the list."""

    def __init__(self, challenges, grind_hook: bool = False):
        self.challenges = [int(c) for c in challenges]
        self.messages = []
        self.events = []

        def _a(ctx, kind, ptr, n):
            data = C.string_at(ptr, n) if n else b""
            self.messages.append((kind, data))
            self.events.append(("absorb", kind, data))

        def _s(ctx, out):
            v = self.challenges.pop(0)
            self.events.append(("squeeze",))
            C.memmove(out, v.to_bytes(16, "little"), 16)

        def _g(ctx, label, n, bits):
            self.events.append(("grind", C.string_at(label, n), int(bits)))

        self._a, self._s, self._g = ABSORB(_a), SQUEEZE(_s), GRIND(_g)
        self.t = Transcript(None, C.cast(self._a, C.c_void_p), C.cast(self._s, C.c_void_p),
                            C.cast(self._g, C.c_void_p) if grind_hook else None)

    @property
    def ref(self):
        return C.byref(self.t)


def squeeze(tr) -> int:
    out = (C.c_uint8 * 16)()
    fn = SQUEEZE(tr.t.squeeze)
    fn(tr.t.ctx, out)
    return int.from_bytes(bytes(out), "little")


def absorb(tr, kind: int, data: bytes) -> None:
    buf = (C.c_uint8 * max(1, len(data))).from_buffer_copy(data or b"\0")
    ABSORB(tr.t.absorb)(tr.t.ctx, kind, buf, len(data))


# ---- poly ------------------------------------------------------------------------
def eq_table(r: np.ndarray) -> np.ndarray:
    r = as_elems(r)
    out = arr(1 << len(r))
    lib().oracle_eq_table(ptr(r), len(r), ptr(out))
    return out


def eq_eval(x: np.ndarray, y: np.ndarray) -> int:
    x, y = as_elems(x), as_elems(y)
    return to_int(lib().oracle_eq_eval(ptr(x), ptr(y), len(x)))


def mle_fold(t: np.ndarray, r) -> np.ndarray:
    t = as_elems(t)
    out = arr(len(t) // 2)
    lib().oracle_mle_fold(ptr(t), len(t), f(r), ptr(out))
    return out


def mle_evaluate(t: np.ndarray, point: np.ndarray) -> int:
    t, point = as_elems(t), as_elems(point)
    return to_int(lib().oracle_mle_evaluate(ptr(t), ptr(point), len(point)))


def inner_product(a: np.ndarray, b: np.ndarray) -> int:
    a, b = as_elems(a), as_elems(b)
    return to_int(lib().oracle_inner_product(ptr(a), ptr(b), len(a)))


# ---- spartan -----------------------------------------------------------------------
def outer_sumcheck(tr, claim, eq_low, eq_high, az, bz, cz):
    eq_low, eq_high = as_elems(eq_low), as_elems(eq_high)
    az, bz, cz = as_elems(az), as_elems(bz), as_elems(cz)
    lv, hv = int(len(eq_low)).bit_length() - 1, int(len(eq_high)).bit_length() - 1
    av, bv, cv = (int(len(x)).bit_length() - 1 for x in (az, bz, cz))
    n = av
    rounds = arr(4 * max(n, 1))
    points = arr(max(n, 1))
    final = arr(1)
    claims = arr(3)
    rc = lib().oracle_outer_sumcheck(tr.ref, f(claim), ptr(eq_low), lv, ptr(eq_high), hv,
                                     ptr(az), ptr(bz), ptr(cz), av, bv, cv, ptr(rounds),
                                     ptr(points), ptr(final), ptr(claims))
    if rc:
        return rc, None
    return 0, dict(rounds=rounds[: 4 * n].reshape(n, 4, 2), points=points[:n],
                   final_claim=to_int(final[0]), claims=claims)


def inner_sumcheck(tr, claim, matrix, witness):
    matrix, witness = as_elems(matrix), as_elems(witness)
    mv, wv = int(len(matrix)).bit_length() - 1, int(len(witness)).bit_length() - 1
    n = mv
    rounds = arr(3 * max(n, 1))
    points = arr(max(n, 1))
    final = arr(1)
    evals = arr(2)
    rc = lib().oracle_inner_sumcheck(tr.ref, f(claim), ptr(matrix), mv, ptr(witness), wv,
                                     ptr(rounds), ptr(points), ptr(final), ptr(evals))
    if rc:
        return rc, None
    return 0, dict(rounds=rounds[: 3 * n].reshape(n, 3, 2), points=points[:n],
                   final_claim=to_int(final[0]), evals=evals)


def sumcheck_verify(tr, claim, rounds: np.ndarray, coeffs: int):
    rounds = np.ascontiguousarray(rounds, dtype=np.uint64)
    n = rounds.shape[0] if rounds.size else 0
    points = arr(max(n, 1))
    final = arr(1)
    rc = lib().oracle_sumcheck_verify(tr.ref, f(claim), ptr(rounds.reshape(-1, 2) if rounds.size else arr(1)),
                                      n, coeffs, ptr(points), ptr(final))
    return rc, points[:n], to_int(final[0])


def outer_verify(tr, claim, tau, rounds, claims):
    tau = as_elems(tau)
    rounds = np.ascontiguousarray(rounds, dtype=np.uint64).reshape(-1, 2)
    claims = as_elems(claims)
    n = len(tau)
    points = arr(max(n, 1))
    return lib().oracle_outer_verify(tr.ref, f(claim), ptr(tau), n,
                                     ptr(rounds if len(rounds) else arr(1)), ptr(claims),
                                     ptr(points))


class CsrMatrix:
    """One sparse matrix in compressed-row form (crates/spartan/src/matrix.rs:97-125)."""

    def __init__(self, row_starts, columns, values):
        self.row_starts = np.ascontiguousarray(row_starts, dtype=np.uint64)
        self.columns = np.ascontiguousarray(columns, dtype=np.uint32)
        self.values = as_elems(values) if len(values) else arr(0)

    def c(self) -> CSR:
        return CSR(self.row_starts.ctypes.data, self.columns.ctypes.data if len(self.columns) else None,
                   self.values.ctypes.data if len(self.values) else None)


def bind_and_batch(mats, rows, row_vars, col_vars, row_point, rho):
    cs = (CSR * 3)(*[m.c() for m in mats])
    row_point = as_elems(row_point) if len(row_point) else arr(1)
    out = arr(1 << col_vars)
    rc = lib().oracle_bind_and_batch(cs, rows, row_vars, col_vars, ptr(row_point), f(rho), ptr(out))
    assert rc == 0
    return out


def csr_multiply(m: "CsrMatrix", rows: int, assignment, padded: int) -> np.ndarray:
    """The reference tests' multiply + padded_mle (piop.rs:352-372): A h per
    row, zero-padded to `padded` entries."""
    assignment = as_elems(assignment)
    out = arr(max(padded, rows, 1))
    cs = m.c()
    lib().oracle_csr_multiply(C.byref(cs), rows, ptr(assignment), ptr(out))
    return out[:padded]


def evaluate_batched(mats, rows, row_vars, col_vars, row_point, rho, col_point):
    cs = (CSR * 3)(*[m.c() for m in mats])
    row_point = as_elems(row_point) if len(row_point) else arr(1)
    col_point = as_elems(col_point) if len(col_point) else arr(1)
    out = arr(1)
    rc = lib().oracle_evaluate_batched(cs, rows, row_vars, col_vars, ptr(row_point), f(rho),
                                       ptr(col_point), ptr(out))
    assert rc == 0
    return to_int(out[0])


# ---- circuit ------------------------------------------------------------------------
def mt_apply(column_offsets, row_indices, challenges):
    column_offsets = np.ascontiguousarray(column_offsets, dtype=np.uint32)
    row_indices = np.ascontiguousarray(row_indices, dtype=np.uint32)
    challenges = as_elems(challenges)
    cols = len(column_offsets) - 1
    out = arr(cols)
    lib().oracle_mt_apply(ptr(column_offsets), cols, ptr(row_indices) if len(row_indices) else None,
                          ptr(challenges), ptr(out))
    return out


def mt_transpose_inner(column_offsets, row_indices, row_weights, col_weights):
    column_offsets = np.ascontiguousarray(column_offsets, dtype=np.uint32)
    row_indices = np.ascontiguousarray(row_indices, dtype=np.uint32)
    row_weights, col_weights = as_elems(row_weights), as_elems(col_weights)
    cols = len(column_offsets) - 1
    out = arr(cols)
    rc = lib().oracle_mt_transpose_inner(ptr(column_offsets), cols,
                                         ptr(row_indices) if len(row_indices) else None,
                                         ptr(row_weights), len(row_weights), ptr(col_weights),
                                         len(col_weights), ptr(out))
    assert rc == 0
    return out


# ---- gkr ------------------------------------------------------------------------------
def gkr_fold_layers(leaves, log_groups):
    leaves = as_elems(leaves)
    log_len = int(len(leaves)).bit_length() - 1
    total = sum(1 << j for j in range(log_groups, log_len + 1))
    store = arr(total)
    nl = lib().oracle_gkr_fold_layers(ptr(leaves), log_len, log_groups, ptr(store))
    layers, off = [], 0
    for j in range(nl):
        size = 1 << (log_len - j)
        layers.append(store[off: off + size])
        off += size
    return store, layers


def gkr_eq_tables(point):
    point = as_elems(point)
    n = len(point)
    out = arr((1 << n) - 1)
    lib().oracle_gkr_eq_tables(ptr(point), n, ptr(out))
    return out


def gkr_prove_layer(tr, point, children, grind_bits: int = 0):
    point, children = as_elems(point), as_elems(children)
    n = len(point)
    nxt = arr(n + 1)
    msgs = arr(2 * (n + 1))
    claim = lib().oracle_gkr_prove_layer(tr.ref, ptr(point) if n else ptr(arr(1)), n,
                                         ptr(children), ptr(nxt), ptr(msgs), grind_bits)
    return nxt, to_int(claim), msgs


def gkr_prove(tr, point, store, log_leaves, security_bits: int = 100):
    """gpgkr_prove; security_bits is the SecurityLevel (100 or 128)."""
    point = as_elems(point)
    leaf_point = arr(log_leaves)
    claim = lib().oracle_gkr_prove(tr.ref, ptr(point), len(point), ptr(store), log_leaves,
                                   ptr(leaf_point), security_bits)
    return leaf_point, to_int(claim)


def blake3(data: bytes) -> bytes:
    """BLAKE3 of at most 64 bytes (one block of the first chunk)."""
    out = (C.c_uint8 * 32)()
    buf = C.create_string_buffer(data, max(len(data), 1))
    assert lib().oracle_blake3_small(buf, len(data), out) == 0
    return bytes(out)


def pow_valid(seed: bytes, nonce: int, bits: int) -> bool:
    return bool(lib().oracle_pow_valid(C.create_string_buffer(seed, 16), nonce, bits))


def pow_find(seed: bytes, bits: int) -> int:
    return int(lib().oracle_pow_find(C.create_string_buffer(seed, 16), bits))


def grinding_bits(security_bits: int, coefficient: int) -> int:
    return int(lib().oracle_grinding_bits(security_bits, coefficient))


def grind(tr, label: bytes, bits: int) -> None:
    buf = C.create_string_buffer(label, max(len(label), 1))
    lib().oracle_transcript_grind(tr.ref, buf, len(label), bits)


def gkr_verify(tr, claim, point, layers, messages):
    point = as_elems(point)
    messages = as_elems(messages)
    total = len(point) + layers
    lp = arr(max(total, 1))
    lc = arr(1)
    rc = lib().oracle_gkr_verify(tr.ref, f(claim), ptr(point), len(point), layers, ptr(messages),
                                 ptr(lp), ptr(lc))
    return rc, lp[:total], to_int(lc[0])


# ---- post_gkr ------------------------------------------------------------------------
def post_gkr_prove(tr, rows, columns, target, packed, security_bits: int = 100):
    """crates/post_gkr prove_factors: returns (rc, point, evaluation)."""
    rows, columns, packed = as_elems(rows), as_elems(columns), as_elems(packed)
    t, s = len(rows).bit_length() - 1, len(columns).bit_length() - 1
    point = arr(t + s)
    ev = arr(1)
    rc = lib().oracle_post_gkr_prove(tr.ref, ptr(rows), t, ptr(columns), s, f(target), ptr(packed), len(packed),
                                     security_bits, ptr(point), ptr(ev))
    return rc, point, to_int(ev[0])


def post_gkr_verify(tr, rows, columns, target, messages, evaluation, security_bits: int = 100):
    """crates/post_gkr verify_factors over the prover's messages: (rc, point)."""
    rows, columns, messages = as_elems(rows), as_elems(columns), as_elems(messages)
    t, s = len(rows).bit_length() - 1, len(columns).bit_length() - 1
    point = arr(t + s)
    rc = lib().oracle_post_gkr_verify(tr.ref, ptr(rows), t, ptr(columns), s, f(target), security_bits,
                                      ptr(messages), f(evaluation), ptr(point))
    return rc, point
