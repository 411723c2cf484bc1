"""bitz_amd — MI355X (gfx950) GF(2^128) sumcheck engine for BitZ's provers.

Python mirror of the reference's prover interface for the sumcheck path,
over the C ABI in include/bitz_hip.h (libbitz_hip.so, built in-tree by
bitz_amd/build.py).  Names, argument meaning and error behaviour follow the
reference crates:

  poly:     eq_table, make_equality_factors, mle_fold, mle_evaluate
            (crates/poly/src/{eq.rs,mle.rs})
  spartan:  prove_outer_sumcheck, prove_inner_sumcheck,
            PreparedConstraintMatrices.bind_and_batch
            (crates/spartan/src/{sumcheck.rs,matrix.rs})
  circuit:  MaterializedMTranspose.apply / transpose_inner
            (crates/circuit/src/matrix_transpose.rs)
  gkr:      grand_product, gpgkr_prove (crates/gkr/src/lib.rs)

Field elements are numpy uint64 arrays of shape (n, 2) = [lo, hi], the byte
layout of the reference's F128.  Every computation runs in the HIP library;
there is no CPU fallback, and importing this package raises if the library
is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbitz_hip.so")

DEVICE = 1  # BITZ_DEVICE


class BitzError(RuntimeError):
    """A failure reported by the C ABI (status code + message)."""

    NAMES = {
        -1: "InvalidProductDimensions",
        -2: "InvalidEqualityDimensions",
        -3: "InvalidRoundClaim",
        -4: "InvalidTerminalClaim",
        -5: "InvalidRowPointLength",
        -6: "TooManyChallenges",
        -7: "InvalidArgument",
        -8: "InvalidState",
        -9: "InvalidPointLength",
        -10: "InvalidAssignmentDimensions",
        -11: "InvalidAssignmentLength",
        -12: "InvalidAssignmentConstant",
        -13: "WitnessLengthMismatch",
        -14: "ClaimDoesNotHold",
        -20: "OutOfInput",
        -21: "BadMagic",
        -22: "UnsupportedVersion",
        -23: "BadHeaderLen",
        -24: "ReservedFlagsSet",
        -25: "LengthOverflow",
        -26: "LengthMismatch",
        -100: "DeviceError",
        -101: "AllocationError",
    }

    def __init__(self, code: int, message: str):
        self.code = code
        self.variant = self.NAMES.get(code, f"Error{code}")
        super().__init__(f"{self.variant} ({code}): {message}")


class SumcheckError(BitzError):
    """crates/spartan/src/sumcheck.rs:39-47 SumcheckError."""


class ProofDecodingError(BitzError):
    """crates/host/src/wire_proof.rs:43-60 ProofDecodingError; `detail` is the
    variant's value (the version, header length, flags or declared length)."""

    def __init__(self, code: int, message: str, detail: int = 0):
        super().__init__(code, message)
        self.detail = detail


class F128(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64)]


_lib = None


def lib():
    """Loads libbitz_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python bitz_amd/build.py` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        P, U32, SZ, I32 = C.c_void_p, C.c_uint32, C.c_size_t, C.c_int32
        sig = {
            "bitz_init": (I32, [I32]),
            "bitz_last_error": (C.c_char_p, []),
            "bitz_device_alloc": (I32, [C.POINTER(C.c_void_p), SZ]),
            "bitz_device_free": (I32, [P]),
            "bitz_memcpy_h2d": (I32, [P, P, SZ]),
            "bitz_memcpy_d2h": (I32, [P, P, SZ]),
            "bitz_memcpy_d2d": (I32, [P, P, SZ]),
            "bitz_synchronize": (I32, []),
            "bitz_release_cache": (I32, []),
            "bitz_profile_enable": (I32, [I32]),
            "bitz_profile_only": (I32, [C.c_char_p]),
            "bitz_prove_outer_batch": (I32, [P, U32, U32]),
            "bitz_outer_batch_config": (I32, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
            "bitz_outer_batch_release": (None, []),
            "bitz_outer_batch_set_config": (I32, [U32, U32]),
            "bitz_profile_read": (I32, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_double)]),
            "bitz_transcript_callbacks": (P, [P, P, P]),
            "bitz_transcript_seeded": (P, [C.c_uint64]),
            "bitz_transcript_duplex": (P, [P, SZ]),
            "bitz_transcript_free": (None, [P]),
            "bitz_transcript_absorb": (I32, [P, I32, P, SZ]),
            "bitz_transcript_squeeze": (I32, [P, P]),
            "bitz_transcript_narg": (SZ, [P, P, SZ]),
            "bitz_mul_batch": (I32, [P, P, P, SZ, U32]),
            "bitz_gkr_prove_with_challenges": (I32, [P, P, U32, P, SZ, P, P, P, P]),
            "bitz_prove_inner_sumcheck_with_challenges": (I32, [P, F128, P, U32, P, U32, U32, P, P, P, P]),
            "bitz_prove_inner_sumcheck_with_challenges_noalloc": (I32, [P, F128, P, U32, P, U32, P, P, P, P, P]),
            "bitz_r1cs_products": (I32, [P, P, P, P, P, U32]),
            "bitz_assignment_mle": (I32, [P, SZ, SZ, P, U32]),
            "bitz_assignment_pack": (I32, [P, SZ, SZ, P, U32]),
            "bitz_r1cs_products_packed": (I32, [P, P, P, P, P, U32]),
            "bitz_prove_spartan_piop_packed": (I32, [P, P, P, P, P, P, U32, P, U32, U32, P, P, P, P, P, P]),
            "bitz_prove_inner_sumcheck_packed": (I32, [P, F128, P, U32, P, U32, U32, P, P, P, P]),
            "bitz_inner_begin_packed": (I32, [C.POINTER(C.c_void_p), P, U32, P, U32, U32]),
            "bitz_piop_engine_submit_packed": (I32, [P, C.c_uint64, P, P, P, P, P, P, U32, P, U32]),
            "bitz_piop_engine_submit_assignment": (I32, [P, C.c_uint64, P, P, P, P, SZ, SZ]),
            "bitz_server_create": (I32, [C.POINTER(C.c_void_p), U32, U32, U32, C.c_uint64]),
            "bitz_server_add_circuit": (I32, [P, P, P, P, C.c_uint64, C.POINTER(U32)]),
            "bitz_server_serve": (I32, [P, P, SZ]),
            "bitz_server_free": (None, [P]),
            "bitz_constraint_matrix_digest": (I32, [SZ, SZ, P, P, SZ, SZ, P, P, P, P]),
            "bitz_prove_spartan_piop": (I32, [P, P, P, P, P, P, U32, P, U32, U32, P, P, P, P, P, P]),
            "bitz_eq_table": (I32, [P, U32, P, U32]),
            "bitz_mle_fold": (I32, [P, C.POINTER(U32), P, U32, U32]),
            "bitz_mle_evaluate": (I32, [P, U32, P, U32, P, U32]),
            "bitz_prove_outer_sumcheck": (I32, [P, F128, P, U32, P, U32, P, U32, P, U32, P, U32, U32,
                                                P, P, P, P]),
            "bitz_prove_outer_sumcheck_with_challenges": (I32, [P, F128, P, U32, P, U32, P, U32, P, U32, P, U32,
                                                                U32, P, P, P, P]),
            "bitz_outer_claim": (I32, [P, U32, P, U32, P, P, P, U32, U32, P]),
            "bitz_outer_begin": (I32, [C.POINTER(C.c_void_p), P, U32, P, U32, P, U32, P, U32, P, U32, U32]),
            "bitz_outer_eq_mode": (I32, [P]),
            "bitz_outer_round0": (I32, [P, P]),
            "bitz_outer_fold_next": (I32, [P, F128, P]),
            "bitz_outer_finish": (I32, [P, P]),
            "bitz_outer_tables": (I32, [P, P, P, P, C.POINTER(SZ)]),
            "bitz_outer_free": (None, [P]),
            "bitz_prove_inner_sumcheck": (I32, [P, F128, P, U32, P, U32, U32, P, P, P, P]),
            "bitz_inner_begin": (I32, [C.POINTER(C.c_void_p), P, U32, P, U32, U32]),
            "bitz_inner_round0": (I32, [P, P]),
            "bitz_inner_fold_next": (I32, [P, F128, P]),
            "bitz_inner_finish": (I32, [P, P]),
            "bitz_inner_free": (None, [P]),
            "bitz_matrices_prepare": (I32, [C.POINTER(C.c_void_p), SZ, U32, U32, P, P, P]),
            "bitz_bind_and_batch": (I32, [P, P, U32, F128, P, U32]),
            "bitz_matrices_free": (None, [P]),
            "bitz_matrices_device_bytes": (I32, [P, C.POINTER(C.c_uint64)]),
            "bitz_piop_engine_device_bytes": (I32, [P, C.c_uint64, C.POINTER(C.c_uint64)]),
            "bitz_server_resident_bytes": (I32, [P, C.POINTER(C.c_uint64)]),
            "bitz_mt_prepare": (I32, [C.POINTER(C.c_void_p), SZ, SZ, P, P]),
            "bitz_mt_apply": (I32, [P, P, SZ, P, U32]),
            "bitz_mt_transpose_inner": (I32, [P, P, SZ, P, SZ, P, U32]),
            "bitz_mt_free": (None, [P]),
            "bitz_gkr_build": (I32, [C.POINTER(C.c_void_p), P, U32, U32, P, U32]),
            "bitz_gkr_build_bittable": (I32, [C.POINTER(C.c_void_p), P, U32, P, U32, P, U32]),
            "bitz_gkr_prove": (I32, [P, P, P, U32, U32, P, P]),
            "bitz_transcript_set_grind": (I32, [P, P]),
            "bitz_transcript_keccak": (P, [P, SZ]),
            "bitz_set_device_fiat_shamir": (I32, [I32]),
            "bitz_set_engine_positions": (I32, [I32]),
            "bitz_keccak_hash": (I32, [P, SZ, I32, C.c_uint8, P, SZ, U32]),
            "bitz_post_gkr_prove": (I32, [P, P, U32, P, U32, F128, P, SZ, U32, U32, P, P]),
            "bitz_transcript_hint": (I32, [P, P, SZ]),
            "bitz_transcript_hints": (SZ, [P, P, SZ]),
            "bitz_transcript_proof": (I32, [P, P, SZ, C.POINTER(SZ)]),
            "bitz_proof_encode": (I32, [P, SZ, P, SZ, P, SZ, C.POINTER(SZ)]),
            "bitz_proof_decode": (I32, [P, SZ, C.POINTER(SZ), C.POINTER(SZ), C.POINTER(SZ), C.POINTER(SZ),
                                        C.POINTER(C.c_uint64)]),
            "bitz_transcript_grind": (I32, [P, P, SZ, U32]),
            "bitz_pow_find": (I32, [P, U32, C.POINTER(C.c_uint64)]),
            "bitz_pow_valid": (I32, [P, C.c_uint64, U32]),
            "bitz_grinding_bits": (U32, [U32, C.c_uint64]),
            "bitz_gkr_layer": (I32, [P, U32, P, SZ, C.POINTER(SZ)]),
            "bitz_gkr_free": (None, [P]),
            "bitz_piop_engine_create": (I32, [C.POINTER(C.c_void_p)]),
            "bitz_piop_engine_submit": (I32, [P, C.c_uint64, P, P, P, P, P, P, U32, P, U32]),
            "bitz_piop_engine_step": (I32, [P, C.POINTER(SZ)]),
            "bitz_piop_engine_launch": (I32, [P, C.POINTER(SZ)]),
            "bitz_piop_engine_run": (I32, [P]),
            "bitz_piop_engine_wait": (I32, [P, U32, C.POINTER(SZ)]),
            "bitz_piop_engine_finish": (I32, [P, C.POINTER(SZ)]),
            "bitz_piop_engine_done": (I32, [P, P, SZ, C.POINTER(SZ)]),
            "bitz_piop_engine_take": (I32, [P, C.c_uint64, P, P, P, P, P, P]),
            "bitz_piop_engine_free": (None, [P]),
            "bitz_sha256_circuit": (I32, [C.POINTER(C.c_void_p), U32, SZ, P]),
            "bitz_sha256_shape": (I32, [P, C.POINTER(SZ), C.POINTER(SZ), C.POINTER(SZ), P, C.POINTER(SZ)]),
            "bitz_sha256_export": (I32, [P, P, P, P, P, P, P]),
            "bitz_sha256_witness": (I32, [P, P, P, P, P, P, C.POINTER(I32)]),
            "bitz_sha256_prepare": (I32, [P, C.POINTER(C.c_void_p), P]),
            "bitz_sha256_free": (None, [P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, cls=BitzError) -> None:
    if rc != 0:
        raise cls(rc, (lib().bitz_last_error() or b"").decode())


_initialized = False


def init(device: int = 0) -> None:
    global _initialized
    if not _initialized:
        _check(lib().bitz_init(device))
        _initialized = True


# ---------------------------------------------------------------- elements --
def elems(a) -> np.ndarray:
    """(n, 2) uint64 contiguous view/copy of a table of field elements."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a.reshape(-1, 2)


def _p(a):
    """An address for a c_void_p parameter (plain ints convert without the
    cost of a ctypes cast; None is NULL)."""
    if isinstance(a, (DeviceTable, PackedAssignment)):
        return a.ptr
    return a.ctypes.data if a is not None and a.size else None


def f128(x) -> F128:
    if isinstance(x, F128):
        return x
    if isinstance(x, (tuple, list, np.ndarray)):
        return F128(int(x[0]), int(x[1]))
    x = int(x)
    return F128(x & (2**64 - 1), x >> 64)


def to_int(e) -> int:
    if isinstance(e, F128):
        return e.lo | (e.hi << 64)
    return int(e[0]) | (int(e[1]) << 64)


def _nvars(n: int) -> int:
    if n <= 0 or n & (n - 1):
        raise BitzError(-7, f"table length {n} is not a power of two")
    return n.bit_length() - 1


def _challenges(challenges, n: int) -> np.ndarray:
    """The first n challenges as a contiguous (n, 2) array (one zero row when
    n = 0, so the C entry point gets a valid pointer).  The C entry points
    read exactly num_vars challenges; fewer is an error here, before any
    call, so a short array never reads past its end."""
    ch = elems(challenges) if len(challenges) else np.zeros((0, 2), np.uint64)
    if len(ch) < n:
        raise BitzError(-7, f"{len(ch)} challenges for {n} rounds")
    return np.ascontiguousarray(ch[:n]) if n else np.zeros((1, 2), np.uint64)


class DeviceTable:
    """A table of field elements resident in HBM (bitz_device_alloc)."""

    def __init__(self, host: np.ndarray | None = None, length: int | None = None):
        init()
        if host is not None:
            host = elems(host)
            length = len(host)
        self.len = int(length)
        ptr = C.c_void_p()
        _check(lib().bitz_device_alloc(C.byref(ptr), max(self.len, 1) * 16))
        self.ptr = ptr.value
        if host is not None and self.len:
            _check(lib().bitz_memcpy_h2d(C.c_void_p(self.ptr), _p(host), self.len * 16))

    def to_host(self) -> np.ndarray:
        out = np.zeros((self.len, 2), dtype=np.uint64)
        if self.len:
            _check(lib().bitz_memcpy_d2h(_p(out), C.c_void_p(self.ptr), self.len * 16))
        return out

    def __len__(self):
        return self.len

    def copy_from(self, other: "DeviceTable", offset: int = 0) -> None:
        """Device-to-device copy of other's elements into this table, from
        entry `offset` on."""
        n = min(self.len - offset, other.len)
        if n > 0:
            _check(lib().bitz_memcpy_d2d(C.c_void_p(self.ptr + 16 * offset), C.c_void_p(other.ptr), n * 16))

    def free(self):
        if self.ptr:
            lib().bitz_device_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PackedAssignment:
    """A Boolean assignment kept as its packed bits (bitz_assignment_pack):
    2^num_vars bits, bit j = words[j // 64] >> (j % 64) & 1, zero past the
    assignment; on the host (numpy words) or in HBM (device=True).  What the
    *_packed entry points take in place of build_assignment_mle's table
    (crates/spartan/src/matrix.rs:392-420); the checks are that function's."""

    def __init__(self, words, bit_len: int, columns: int, device: bool = False):
        words = np.ascontiguousarray(words, dtype=np.uint64)
        if not len(words):
            words = np.zeros(1, np.uint64)
        if bit_len > 64 * len(words):
            raise BitzError(-7, "fewer words than bits")
        padded = 1 << max(int(columns - 1).bit_length(), 0) if columns > 1 else 1
        self.num_vars = padded.bit_length() - 1
        self.len = padded
        self.nwords = (padded + 63) // 64
        self.device = device
        self.ptr = 0
        if device:
            init()
            ptr = C.c_void_p()
            _check(lib().bitz_device_alloc(C.byref(ptr), self.nwords * 8))
            self.ptr = ptr.value
            try:
                _check(lib().bitz_assignment_pack(_p(words), bit_len, columns, C.c_void_p(self.ptr), DEVICE))
            except BaseException:
                self.free()
                raise
        else:
            self.words = np.zeros(self.nwords, np.uint64)
            _check(lib().bitz_assignment_pack(_p(words), bit_len, columns, _p(self.words), 0))
            self.ptr = self.words.ctypes.data

    @classmethod
    def from_bits(cls, bits, columns: int, device: bool = False) -> "PackedAssignment":
        """From an array of 0/1 values (bit j of the assignment at index j)."""
        bits = np.ascontiguousarray(bits, dtype=np.uint8)
        b = np.packbits(bits, bitorder="little")
        b = np.concatenate([b, np.zeros((-len(b)) % 8, np.uint8)])
        return cls(b.view(np.uint64) if len(b) else np.zeros(1, np.uint64), len(bits), columns, device)

    def to_table(self) -> np.ndarray:
        """The {0, 1} elements build_assignment_mle would give (host)."""
        words = self.words if not self.device else self._download()
        bits = np.unpackbits(words.view(np.uint8), bitorder="little")[: self.len]
        out = np.zeros((self.len, 2), np.uint64)
        out[:, 0] = bits
        return out

    def _download(self) -> np.ndarray:
        out = np.zeros(self.nwords, np.uint64)
        _check(lib().bitz_memcpy_d2h(_p(out), C.c_void_p(self.ptr), self.nwords * 8))
        return out

    def free(self):
        if self.device and self.ptr:
            lib().bitz_device_free(C.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _table(t):
    """(pointer-able, length, flags) for a host array or DeviceTable."""
    if isinstance(t, DeviceTable):
        return t, t.len, DEVICE
    t = elems(t)
    return t, len(t), 0


def profile_enable(on: bool = True) -> None:
    """Time the sumcheck kernels with HIP events on their own streams."""
    init()
    _check(lib().bitz_profile_enable(1 if on else 0))


def profile_only(kernel: str | None) -> None:
    """Time only launches named `kernel` (None: all)."""
    init()
    _check(lib().bitz_profile_only(kernel.encode() if kernel else None))


def profile_read(kernel: str) -> dict:
    ms, n, by = C.c_double(), C.c_uint64(), C.c_double()
    _check(lib().bitz_profile_read(kernel.encode(), C.byref(ms), C.byref(n), C.byref(by)))
    return {"ms": ms.value, "launches": n.value, "bytes": by.value}


def synchronize() -> None:
    _check(lib().bitz_synchronize())


# ------------------------------------------------------------- transcripts --
ABSORB_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.POINTER(C.c_uint8), C.c_size_t)
SQUEEZE_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_uint8))
GRIND_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_uint8), C.c_size_t, C.c_uint32)


class Transcript:
    """Fiat-Shamir transcript behind the C ABI's bitz_transcript."""

    def __init__(self, handle):
        if not handle:
            raise BitzError(-7, "transcript creation failed")
        self.h = C.c_void_p(handle)

    # Transcripts are host objects: creating one does not touch the device.
    @classmethod
    def seeded(cls, seed: int) -> "Transcript":
        return cls(lib().bitz_transcript_seeded(seed & (2**64 - 1)))

    @classmethod
    def duplex(cls, domain: bytes) -> "Transcript":
        buf = C.create_string_buffer(domain, len(domain))
        return cls(lib().bitz_transcript_duplex(buf, len(domain)))

    @classmethod
    def keccak(cls, domain: bytes) -> "Transcript":
        """The Keccak duplex (bitz_transcript_keccak); its challenges can be
        drawn on the device."""
        buf = C.create_string_buffer(domain, max(len(domain), 1))
        return cls(lib().bitz_transcript_keccak(buf, len(domain)))

    @classmethod
    def callbacks(cls, absorb, squeeze, grind=None) -> "Transcript":
        """absorb(kind: int, data: bytes) -> None; squeeze() -> 16 bytes;
        grind(label: bytes, bits: int) -> None, the caller's own grinding
        (ProverState::grind), or None to frame grinding over absorb/squeeze."""

        def _a(ctx, kind, ptr, n):
            absorb(kind, C.string_at(ptr, n) if n else b"")

        def _s(ctx, out):
            data = squeeze()
            C.memmove(out, data, 16)

        a, s = ABSORB_FN(_a), SQUEEZE_FN(_s)
        t = cls(lib().bitz_transcript_callbacks(None, C.cast(a, C.c_void_p), C.cast(s, C.c_void_p)))
        t._keep = (a, s)
        if grind is not None:
            def _g(ctx, label, n, bits):
                grind(C.string_at(label, n) if n else b"", int(bits))

            g = GRIND_FN(_g)
            _check(lib().bitz_transcript_set_grind(t.h, C.cast(g, C.c_void_p)))
            t._keep = (a, s, g)
        return t

    def hint(self, data: bytes) -> None:
        """ProverState::hint: bytes beside the narg string, never absorbed."""
        buf = C.create_string_buffer(bytes(data), max(len(data), 1))
        _check(lib().bitz_transcript_hint(self.h, buf, len(data)))

    def hints(self) -> bytes:
        n = lib().bitz_transcript_hints(self.h, None, 0)
        buf = (C.c_uint8 * max(n, 1))()
        lib().bitz_transcript_hints(self.h, buf, n)
        return bytes(buf[:n])

    def proof(self) -> bytes:
        """ProverState::finish, then the wire container (wire_proof::encode)."""
        n = C.c_size_t()
        _check(lib().bitz_transcript_proof(self.h, None, 0, C.byref(n)))
        buf = (C.c_uint8 * max(n.value, 1))()
        _check(lib().bitz_transcript_proof(self.h, buf, n.value, C.byref(n)))
        return bytes(buf[: n.value])

    def grind(self, label: bytes, bits: int) -> None:
        """ProverState::grind (crates/transcript/src/prover.rs:29-41)."""
        buf = C.create_string_buffer(label, max(len(label), 1))
        _check(lib().bitz_transcript_grind(self.h, buf, len(label), bits))

    def absorb(self, kind: int, data: bytes) -> None:
        buf = C.create_string_buffer(data, max(len(data), 1))
        _check(lib().bitz_transcript_absorb(self.h, kind, buf, len(data)))

    def public_message(self, data) -> None:
        self.absorb(0, bytes(elems(data)) if not isinstance(data, bytes) else data)

    def prover_message(self, data) -> None:
        self.absorb(1, bytes(elems(data)) if not isinstance(data, bytes) else data)

    def squeeze(self) -> int:
        out = F128()
        _check(lib().bitz_transcript_squeeze(self.h, C.byref(out)))
        return to_int(out)

    def narg(self) -> bytes:
        n = lib().bitz_transcript_narg(self.h, None, 0)
        buf = (C.c_uint8 * max(n, 1))()
        lib().bitz_transcript_narg(self.h, buf, n)
        return bytes(buf[:n])

    def __del__(self):
        try:
            lib().bitz_transcript_free(self.h)
        except Exception:
            pass


# ------------------------------------------------------------------- poly ---
MUL_LANE, MUL_LANE2, MUL_CONST, MUL_ROWS, MUL_MFMA = 0x100, 0x200, 0x400, 0x800, 0x1000  # BITZ_MUL_*


def mul_batch(a, b, mode: int = 0) -> np.ndarray:
    """a[i] * b[i] on the device; mode selects the product primitive
    (MUL_CONST: a[i] * b[0] through b[0]'s constant table; MUL_MFMA: the
    same product on the int8 matrix cores)."""
    init()
    a, b = elems(a), elems(b)
    out = np.zeros_like(a)
    _check(lib().bitz_mul_batch(_p(a), _p(b), _p(out), len(a), mode))
    return out


def mul_batch_device(a: "DeviceTable", b: "DeviceTable", out: "DeviceTable", mode: int = 0) -> None:
    """mul_batch over tables resident on the device (out[i] = a[i] * b[i], or
    a[i] * b[0] for MUL_CONST and MUL_MFMA); with profiling on, the launch is
    timed as mul_comb, mul_lane, mul_lane2, mul_const, mul_rows or mul_mfma."""
    init()
    n = min(a.len, out.len)
    if mode not in (MUL_CONST, MUL_MFMA) and b.len < n:
        raise BitzError(-7, "b is shorter than a")
    _check(lib().bitz_mul_batch(C.c_void_p(a.ptr), C.c_void_p(b.ptr), C.c_void_p(out.ptr), n, mode | DEVICE))


def eq_table(r) -> np.ndarray:
    """poly::eq_table (crates/poly/src/eq.rs:63-99)."""
    init()
    r = elems(r)
    out = np.zeros((1 << len(r), 2), dtype=np.uint64)
    _check(lib().bitz_eq_table(_p(r), len(r), _p(out), 0))
    return out


def make_equality_factors(point):
    """poly::make_equality_factors (eq.rs:107-115): split at len/2."""
    point = elems(point)
    split = len(point) // 2
    return eq_table(point[:split]), eq_table(point[split:])


def mle_fold(evaluations, r) -> np.ndarray:
    """DenseMultilinearExtension::fold (mle.rs:136-156), returning the folded table."""
    init()
    ev = elems(evaluations).copy()
    r = elems(r)
    nv = C.c_uint32(_nvars(len(ev)))
    _check(lib().bitz_mle_fold(_p(ev), C.byref(nv), _p(r), len(r), 0))
    return ev[: 1 << nv.value]


def mle_evaluate(evaluations, point) -> int:
    """DenseMultilinearExtension::evaluate (mle.rs:158-172)."""
    init()
    ev, point = elems(evaluations), elems(point)
    out = F128()
    _check(lib().bitz_mle_evaluate(_p(ev), _nvars(len(ev)), _p(point), len(point), C.byref(out), 0))
    return to_int(out)


def mle_evaluate_device(table: "DeviceTable", point) -> int:
    """mle_evaluate over a table resident on the device (read in place); with
    profiling on, its first fold pass is timed as mle_fold_pass."""
    init()
    point = elems(point) if len(point) else np.zeros((0, 2), np.uint64)
    if table.len != 1 << len(point):
        raise BitzError(-7, "point width differs from num_vars")
    dpt, dout = DeviceTable(point), DeviceTable(length=1)  # BITZ_DEVICE: every element array on the device
    _check(lib().bitz_mle_evaluate(C.c_void_p(table.ptr), len(point), C.c_void_p(dpt.ptr), len(point),
                                   C.c_void_p(dout.ptr), DEVICE))
    return to_int(dout.to_host()[0])


# ---------------------------------------------------------------- spartan ---
@dataclass
class R1csProductMles:
    """sumcheck.rs:122-128: dense Boolean-row MLEs of Az, Bz, Cz."""
    az: object
    bz: object
    cz: object


@dataclass
class OuterSumcheckOutput:
    """sumcheck.rs:143-185: round polynomials [c0,c1,c2,c3], r_x, final claim,
    and the terminal claims [Az(r_x), Bz(r_x), Cz(r_x)]."""
    round_polynomials: np.ndarray  # (n, 4, 2)
    eval_points: np.ndarray        # (n, 2)
    final_claim: int
    az_mle_claim: int
    bz_mle_claim: int
    cz_mle_claim: int


@dataclass
class InnerSumcheckOutput:
    """sumcheck.rs:199-212."""
    round_polynomials: np.ndarray  # (n, 3, 2)
    eval_points: np.ndarray
    final_claim: int
    batched_matrix_evaluation: int
    witness_evaluation: int


def _factor(f, flags):
    """An eq factor on the side the products are on (DeviceTables pass through)."""
    if isinstance(f, DeviceTable):
        if not flags & DEVICE:
            raise BitzError(-7, "device eq factors need device products")
        return f
    f = elems(f)
    return DeviceTable(f) if flags & DEVICE else f


def prove_outer_sumcheck(transcript: Transcript, initial_claim, eq_factors, products: R1csProductMles):
    """prove_outer_sumcheck (crates/spartan/src/sumcheck.rs:252-384) on the GPU."""
    init()
    (az, na, fa), (bz, nb, fb), (cz, nc, fc) = (_table(t) for t in (products.az, products.bz, products.cz))
    flags = fa
    if not (fa == fb == fc):
        raise BitzError(-7, "products must all be host arrays or all DeviceTables")
    eq_low, eq_high = (_factor(f, flags) for f in eq_factors)
    n = _nvars(na)
    rounds = np.zeros((max(n, 1) * 4, 2), dtype=np.uint64)
    points = np.zeros((max(n, 1), 2), dtype=np.uint64)
    final = F128()
    claims = np.zeros((3, 2), dtype=np.uint64)
    rc = lib().bitz_prove_outer_sumcheck(
        transcript.h, f128(initial_claim), _p(eq_low), _nvars(len(eq_low)), _p(eq_high), _nvars(len(eq_high)),
        _p(az), n, _p(bz), _nvars(nb), _p(cz), _nvars(nc), flags, _p(rounds), _p(points), C.byref(final),
        _p(claims))
    _check(rc, SumcheckError)
    return OuterSumcheckOutput(rounds[: 4 * n].reshape(n, 4, 2), points[:n], to_int(final),
                               to_int(claims[0]), to_int(claims[1]), to_int(claims[2]))


def prove_outer_sumcheck_with_challenges(challenges, initial_claim, eq_factors, products: R1csProductMles):
    """prove_outer_sumcheck against challenges fixed in advance: every round
    is queued back to back, the round sums are read once."""
    (az, na, fa), (bz, nb, fb), (cz, nc, fc) = (_table(t) for t in (products.az, products.bz, products.cz))
    if not (fa == fb == fc):
        raise BitzError(-7, "products must all be host arrays or all DeviceTables")
    n = _nvars(na)
    ch = _challenges(challenges, n)
    init()
    eq_low, eq_high = (_factor(f, fa) for f in eq_factors)
    m = max(n, 1)
    out = np.zeros((5 * m + 3, 2), dtype=np.uint64)  # rounds | points | claims, one allocation
    rounds, points, claims = out[: 4 * m], out[4 * m: 5 * m], out[5 * m:]
    final = F128()
    base = out.ctypes.data
    rc = lib().bitz_prove_outer_sumcheck_with_challenges(
        _p(ch), f128(initial_claim), _p(eq_low), _nvars(len(eq_low)), _p(eq_high), _nvars(len(eq_high)),
        _p(az), n, _p(bz), _nvars(nb), _p(cz), _nvars(nc), fa, base, base + 64 * m, C.byref(final),
        base + 80 * m)
    _check(rc, SumcheckError)
    cl = claims.tolist()
    return OuterSumcheckOutput(rounds[: 4 * n].reshape(n, 4, 2), points[:n], to_int(final),
                               cl[0][0] | (cl[0][1] << 64), cl[1][0] | (cl[1][1] << 64), cl[2][0] | (cl[2][1] << 64))


class _OuterInstance(C.Structure):
    """bitz_outer_instance (include/bitz_hip.h)."""
    _fields_ = [("challenges", C.c_void_p), ("initial_claim", F128),
                ("eq_low", C.c_void_p), ("low_vars", C.c_uint32),
                ("eq_high", C.c_void_p), ("high_vars", C.c_uint32),
                ("az", C.c_void_p), ("bz", C.c_void_p), ("cz", C.c_void_p), ("num_vars", C.c_uint32),
                ("round_polys", C.c_void_p), ("eval_points", C.c_void_p),
                ("final_claim", C.c_void_p), ("claims", C.c_void_p)]


BATCH_EAGER = 2
BATCH_NO_TUNE = 4


def _instance_dtype() -> np.dtype:
    """A numpy view of bitz_outer_instance with the C struct's offsets."""
    fields = {}
    for name, ctype in _OuterInstance._fields_:
        off = getattr(_OuterInstance, name).offset
        fields[name] = (np.dtype((np.uint64, 2)) if ctype is F128 else
                        np.dtype(np.uint32) if ctype is C.c_uint32 else np.dtype(np.uint64), off)
    return np.dtype({"names": list(fields), "formats": [f for f, _ in fields.values()],
                     "offsets": [o for _, o in fields.values()], "itemsize": C.sizeof(_OuterInstance)})


class OuterBatch:
    """A batch of outer sumchecks marshalled once for bitz_prove_outer_batch
    (the instance records, challenges and output arrays), so that repeated
    calls cost the C call alone: prove() runs the batch and fills `rounds`,
    `points`, `finals` and `claims` in place; outputs() wraps them per
    instance.  The tables are borrowed and must outlive the object."""

    def __init__(self, instances, graph: bool = True, tune: bool = True):
        init()
        self.keep = list(instances)
        self.rec, self.arrays = _outer_batch_records(self.keep)
        self.m = len(self.keep)
        self.flags = DEVICE | (0 if graph else BATCH_EAGER) | (0 if tune else BATCH_NO_TUNE)

    def prove(self) -> None:
        _check(lib().bitz_prove_outer_batch(C.c_void_p(self.rec.ctypes.data), self.m, self.flags), SumcheckError)

    def outputs(self) -> list:
        ns, off, _, rounds, points, finals, claims = self.arrays
        m = self.m
        fin = [x[0] | (x[1] << 64) for x in finals[:m].tolist()]
        cl = [x[0] | (x[1] << 64) for x in claims[: 3 * m].tolist()]
        return [OuterSumcheckOutput(rounds[4 * off[i]: 4 * off[i + 1]].reshape(ns[i], 4, 2),
                                    points[off[i]: off[i + 1]].copy(), fin[i], cl[3 * i], cl[3 * i + 1],
                                    cl[3 * i + 2]) for i in range(m)]


def _outer_batch_records(instances):
    ns, chs, tabs = [], [], []
    for ch, claim, eq, prods in instances:
        az, bz, cz = prods.az, prods.bz, prods.cz
        if not all(isinstance(t, DeviceTable) for t in (az, bz, cz)):
            raise BitzError(-7, "batched proofs take DeviceTables")
        lo, hi = (_factor(f, DEVICE) for f in eq)
        n = _nvars(az.len)
        _nvars(bz.len), _nvars(cz.len)
        ns.append(n)
        chs.append(_challenges(ch, n)[:n])
        tabs.append((claim, lo, hi, az, bz, cz))
    m = len(tabs)
    off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    total = int(off[-1])
    allch = np.ascontiguousarray(np.concatenate(chs) if total else np.zeros((1, 2), np.uint64), dtype=np.uint64)
    rounds = np.zeros((max(4 * total, 1), 2), dtype=np.uint64)
    points = np.zeros((max(total, 1), 2), dtype=np.uint64)
    finals = np.zeros((max(m, 1), 2), dtype=np.uint64)
    claims = np.zeros((max(m, 1) * 3, 2), dtype=np.uint64)
    rec = np.zeros(max(m, 1), dtype=_instance_dtype())
    if m:
        rec["challenges"][:m] = allch.ctypes.data + 16 * off[:-1]
        rec["initial_claim"][:m] = [[c & (2**64 - 1), (c >> 64) & (2**64 - 1)] for c, *_ in tabs]
        rec["eq_low"][:m] = [t[1].ptr for t in tabs]
        rec["low_vars"][:m] = [_nvars(t[1].len) for t in tabs]
        rec["eq_high"][:m] = [t[2].ptr for t in tabs]
        rec["high_vars"][:m] = [_nvars(t[2].len) for t in tabs]
        rec["az"][:m] = [t[3].ptr for t in tabs]
        rec["bz"][:m] = [t[4].ptr for t in tabs]
        rec["cz"][:m] = [t[5].ptr for t in tabs]
        rec["num_vars"][:m] = ns
        rec["round_polys"][:m] = rounds.ctypes.data + 64 * off[:-1]
        rec["eval_points"][:m] = points.ctypes.data + 16 * off[:-1]
        rec["final_claim"][:m] = finals.ctypes.data + 16 * np.arange(m)
        rec["claims"][:m] = claims.ctypes.data + 48 * np.arange(m)
    # the factor tables _factor made and the arrays the records point into stay alive with the records
    return rec, (ns, off, (allch, tabs), rounds, points, finals, claims)


def prove_outer_batch(instances, graph: bool = True, tune: bool = True) -> list:
    """Many independent outer sumchecks with challenges fixed in advance, in
    one batched call (bitz_prove_outer_batch).  instances: iterable of
    (challenges, initial_claim, (eq_low, eq_high), R1csProductMles) with
    DeviceTables for the products and factors.  Returns one
    OuterSumcheckOutput per instance, equal to
    prove_outer_sumcheck_with_challenges on the same inputs.  graph=False
    launches the kernels directly instead of replaying the captured graph;
    tune=False keeps the default tile shapes."""
    b = OuterBatch(instances, graph, tune)
    b.prove()
    return b.outputs()


def outer_batch_config() -> tuple:
    """(thread_bits, tail_log) of the last batch: the tile shapes the
    autotuner chose (or the defaults)."""
    tb, tl = C.c_uint32(), C.c_uint32()
    _check(lib().bitz_outer_batch_config(C.byref(tb), C.byref(tl)))
    return tb.value, tl.value


def outer_batch_set_config(thread_bits: int = 0, tail_log: int = 0) -> None:
    """Fixes the tile shapes of later batches; thread_bits=0 returns to the
    autotuner."""
    _check(lib().bitz_outer_batch_set_config(thread_bits, tail_log))


def outer_batch_release() -> None:
    """Frees the cached batch plan (device buffers, captured graph)."""
    lib().bitz_outer_batch_release()


def outer_claim(eq_factors, products: R1csProductMles) -> int:
    """sum_x eq(x) (Az Bz + Cz)(x) on the device (sumcheck.rs:367-370)."""
    init()
    (az, na, fa), (bz, _, _), (cz, _, _) = (_table(t) for t in (products.az, products.bz, products.cz))
    eq_low, eq_high = (_factor(f, fa) for f in eq_factors)
    out = F128()
    _check(lib().bitz_outer_claim(_p(eq_low), _nvars(len(eq_low)), _p(eq_high), _nvars(len(eq_high)), _p(az),
                                  _p(bz), _p(cz), _nvars(na), fa, C.byref(out)), SumcheckError)
    return to_int(out)


class OuterSumcheck:
    """Per-round handle (bitz_outer_*): the caller keeps the transcript."""

    def __init__(self, eq_factors, products: R1csProductMles):
        init()
        (az, na, fa), (bz, nb, _), (cz, nc, _) = (_table(t) for t in (products.az, products.bz, products.cz))
        eq_low, eq_high = (_factor(f, fa) for f in eq_factors)
        self._keep = (eq_low, eq_high, az, bz, cz)
        h = C.c_void_p()
        _check(lib().bitz_outer_begin(C.byref(h), _p(eq_low), _nvars(len(eq_low)), _p(eq_high),
                                      _nvars(len(eq_high)), _p(az), _nvars(na), _p(bz), _nvars(nb), _p(cz),
                                      _nvars(nc), fa), SumcheckError)
        self.h = h

    @property
    def eq_mode(self) -> bool:
        return bool(lib().bitz_outer_eq_mode(self.h))

    def round0(self):
        out = np.zeros((3, 2), dtype=np.uint64)
        _check(lib().bitz_outer_round0(self.h, _p(out)))
        return out

    def fold_next(self, r):
        out = np.zeros((3, 2), dtype=np.uint64)
        _check(lib().bitz_outer_fold_next(self.h, f128(r), _p(out)))
        return out

    def finish(self):
        out = np.zeros((3, 2), dtype=np.uint64)
        _check(lib().bitz_outer_finish(self.h, _p(out)))
        return out

    def tables(self):
        n = C.c_size_t()
        _check(lib().bitz_outer_tables(self.h, None, None, None, C.byref(n)))
        a, b, c = (np.zeros((n.value, 2), dtype=np.uint64) for _ in range(3))
        _check(lib().bitz_outer_tables(self.h, _p(a), _p(b), _p(c), C.byref(n)))
        return a, b, c

    def __del__(self):
        try:
            lib().bitz_outer_free(self.h)
        except Exception:
            pass


def prove_inner_sumcheck(transcript: Transcript, initial_claim, batched_matrix, witness):
    """prove_inner_sumcheck (crates/spartan/src/sumcheck.rs:442-552) on the GPU."""
    init()
    (d, nd, fd), (w, nw, fw) = _table(batched_matrix), _table(witness)
    if fd != fw:
        raise BitzError(-7, "matrix and witness must both be host arrays or both DeviceTables")
    n = _nvars(nd)
    rounds = np.zeros((max(n, 1) * 3, 2), dtype=np.uint64)
    points = np.zeros((max(n, 1), 2), dtype=np.uint64)
    final = F128()
    evals = np.zeros((2, 2), dtype=np.uint64)
    rc = lib().bitz_prove_inner_sumcheck(transcript.h, f128(initial_claim), _p(d), n, _p(w), _nvars(nw), fd,
                                         _p(rounds), _p(points), C.byref(final), _p(evals))
    _check(rc, SumcheckError)
    return InnerSumcheckOutput(rounds[: 3 * n].reshape(n, 3, 2), points[:n], to_int(final),
                               to_int(evals[0]), to_int(evals[1]))


def prove_inner_sumcheck_packed(transcript: Transcript, initial_claim, batched_matrix,
                                assignment: PackedAssignment):
    """prove_inner_sumcheck against a Boolean assignment given as packed bits
    (bitz_prove_inner_sumcheck_packed): the same outputs as
    prove_inner_sumcheck over assignment.to_table()."""
    init()
    d, nd, fd = _table(batched_matrix)
    if bool(fd & DEVICE) != assignment.device:
        raise BitzError(-7, "matrix and assignment must both be on the host or both on the device")
    n = _nvars(nd)
    rounds = np.zeros((max(n, 1) * 3, 2), dtype=np.uint64)
    points = np.zeros((max(n, 1), 2), dtype=np.uint64)
    final = F128()
    evals = np.zeros((2, 2), dtype=np.uint64)
    rc = lib().bitz_prove_inner_sumcheck_packed(transcript.h, f128(initial_claim), _p(d), n, assignment.ptr,
                                                assignment.num_vars, fd, _p(rounds), _p(points), C.byref(final),
                                                _p(evals))
    _check(rc, SumcheckError)
    return InnerSumcheckOutput(rounds[: 3 * n].reshape(n, 3, 2), points[:n], to_int(final),
                               to_int(evals[0]), to_int(evals[1]))


def prove_inner_sumcheck_with_challenges(challenges, initial_claim, batched_matrix, witness):
    """prove_inner_sumcheck with the challenges known in advance
    (bitz_prove_inner_sumcheck_with_challenges): rounds queued back to back."""
    (d, nd, fd), (w, nw, fw) = _table(batched_matrix), _table(witness)
    if fd != fw:
        raise BitzError(-7, "matrix and witness must both be host arrays or both DeviceTables")
    n = _nvars(nd)
    ch = _challenges(challenges, n)
    init()
    rounds = np.zeros((max(n, 1) * 3, 2), dtype=np.uint64)
    points = np.zeros((max(n, 1), 2), dtype=np.uint64)
    final = F128()
    evals = np.zeros((2, 2), dtype=np.uint64)
    rc = lib().bitz_prove_inner_sumcheck_with_challenges(_p(ch), f128(initial_claim), _p(d), n, _p(w), _nvars(nw),
                                                         fd, _p(rounds), _p(points), C.byref(final), _p(evals))
    _check(rc, SumcheckError)
    return InnerSumcheckOutput(rounds[: 3 * n].reshape(n, 3, 2), points[:n], to_int(final),
                               to_int(evals[0]), to_int(evals[1]))


NOALLOC_MAX_VARS = 20


def prove_inner_sumcheck_with_challenges_noalloc(challenges, initial_claim, batched_matrix: DeviceTable,
                                                 witness: DeviceTable):
    """prove_inner_sumcheck_with_challenges through the allocation-free entry
    point (bitz_prove_inner_sumcheck_with_challenges_noalloc): device tables
    of at most 2^NOALLOC_MAX_VARS entries, scratch in the library's static
    device array (concurrent calls on one GPU race)."""
    if not (isinstance(batched_matrix, DeviceTable) and isinstance(witness, DeviceTable)):
        raise BitzError(-7, "the allocation-free prover takes DeviceTables")
    n = _nvars(batched_matrix.len)
    nw = _nvars(witness.len)
    ch = _challenges(challenges, n)
    init()
    m = max(n, 1)
    out = np.zeros((5 * m + 2, 2), dtype=np.uint64)  # rounds | points | evals
    base = out.ctypes.data
    final = F128()
    rc = lib().bitz_prove_inner_sumcheck_with_challenges_noalloc(
        _p(ch), f128(initial_claim), batched_matrix.ptr, n, witness.ptr, nw, None, base, base + 48 * m,
        C.byref(final), base + 80 * m)
    _check(rc, SumcheckError)
    ev = out[5 * m:]
    return InnerSumcheckOutput(out[: 3 * n].reshape(n, 3, 2), out[3 * m: 3 * m + n], to_int(final),
                               to_int(ev[0]), to_int(ev[1]))


class PreparedConstraintMatrices:
    """crates/spartan/src/matrix.rs:54-64: A, B, C prepared once (CSR in,
    column-major on device) for repeated bind_and_batch calls."""

    def __init__(self, rows: int, num_row_vars: int, num_column_vars: int, csr):
        """csr: three (row_starts[rows+1], columns, values) triples for A, B, C."""
        init()
        self.rows, self.num_row_vars, self.num_column_vars = rows, num_row_vars, num_column_vars
        starts = [np.ascontiguousarray(m[0], dtype=np.uint64) for m in csr]
        cols = [np.ascontiguousarray(m[1], dtype=np.uint32) for m in csr]
        vals = [elems(m[2]) if len(m[2]) else np.zeros((1, 2), np.uint64) for m in csr]
        self._keep = (starts, cols, vals)
        arr_s = (C.c_void_p * 3)(*[s.ctypes.data for s in starts])
        arr_c = (C.c_void_p * 3)(*[c.ctypes.data if len(c) else None for c in cols])
        arr_v = (C.c_void_p * 3)(*[v.ctypes.data for v in vals])
        h = C.c_void_p()
        _check(lib().bitz_matrices_prepare(C.byref(h), rows, num_row_vars, num_column_vars, arr_s, arr_c, arr_v))
        self.h = h

    @classmethod
    def _adopt(cls, h, rows: int, num_row_vars: int, num_column_vars: int) -> "PreparedConstraintMatrices":
        """Wraps matrices the library prepared itself (bitz_sha256_prepare)."""
        self = cls.__new__(cls)
        self.rows, self.num_row_vars, self.num_column_vars = rows, num_row_vars, num_column_vars
        self._keep = ()
        self.h = h
        return self

    def device_bytes(self) -> int:
        """Device bytes the prepared matrices hold."""
        b = C.c_uint64()
        _check(lib().bitz_matrices_device_bytes(self.h, C.byref(b)))
        return b.value

    def bind_and_batch(self, row_point, rho, device: bool = False, out: "DeviceTable | None" = None):
        """matrix.rs:429-468: D(j) = sum_i eq(i, r_x)(A + rho B + rho^2 C)[i, j];
        a host array, or (device=True) a DeviceTable (out, when given, is
        reused: 2^num_column_vars entries)."""
        if device:
            if not isinstance(row_point, DeviceTable):
                row_point = DeviceTable(elems(row_point) if len(row_point) else np.zeros((1, 2), np.uint64))
            if out is None:
                out = DeviceTable(length=1 << self.num_column_vars)
            elif out.len < 1 << self.num_column_vars:
                raise BitzError(-7, "output table too short")
            n = self.num_row_vars
            _check(lib().bitz_bind_and_batch(self.h, C.c_void_p(row_point.ptr), n, f128(rho), C.c_void_p(out.ptr),
                                             DEVICE))
            return out
        row_point = elems(row_point) if len(row_point) else np.zeros((0, 2), np.uint64)
        out = np.zeros((1 << self.num_column_vars, 2), dtype=np.uint64)
        _check(lib().bitz_bind_and_batch(self.h, _p(row_point), len(row_point), f128(rho), _p(out), 0))
        return out

    def r1cs_products(self, assignment):
        """Az, Bz, Cz = A h, B h, C h padded to 2^num_row_vars (the reference
        test's multiply + padded_mle, crates/spartan/src/piop.rs:352-372):
        host arrays for a host assignment, DeviceTables for a DeviceTable."""
        if isinstance(assignment, PackedAssignment):
            h, n, fl = assignment, assignment.len, DEVICE if assignment.device else 0
        else:
            h, n, fl = _table(assignment)
        if n != 1 << self.num_column_vars:
            raise BitzError(-7, "the assignment must span the padded columns")
        rows = 1 << self.num_row_vars
        if fl & DEVICE:
            outs = [DeviceTable(length=rows) for _ in range(3)]
            ptrs = [C.c_void_p(o.ptr) for o in outs]
        else:
            outs = [np.zeros((rows, 2), dtype=np.uint64) for _ in range(3)]
            ptrs = [_p(o) for o in outs]
        fn = lib().bitz_r1cs_products_packed if isinstance(h, PackedAssignment) else lib().bitz_r1cs_products
        _check(fn(self.h, _p(h), *ptrs, fl))
        return R1csProductMles(*outs)

    def __del__(self):
        try:
            lib().bitz_matrices_free(self.h)
        except Exception:
            pass


def assignment_mle(bits, columns: int, device: bool = False):
    """build_assignment_mle (crates/spartan/src/matrix.rs:392-420): a
    Boolean assignment (array of 0/1) as {0, 1} elements padded to a power of
    two, on the host or (device=True) as a DeviceTable.  The length and
    constant-bit checks run before any device work."""
    bits = np.ascontiguousarray(bits, dtype=np.uint8)
    words = np.packbits(bits, bitorder="little")
    words = np.concatenate([words, np.zeros((-len(words)) % 8, np.uint8)]).view(np.uint64) if len(words) else \
        np.zeros(1, np.uint64)
    padded = 1 << max(int(columns - 1).bit_length(), 0) if columns > 1 else 1
    if device and len(bits) == columns and columns and bits[0] & 1:
        out = DeviceTable(length=padded)
        _check(lib().bitz_assignment_mle(_p(words), len(bits), columns, C.c_void_p(out.ptr), DEVICE))
        return out
    out = np.zeros((padded, 2), dtype=np.uint64)
    _check(lib().bitz_assignment_mle(_p(words), len(bits), columns, _p(out), 0))
    return out


def assignment_mle_words(words, bit_len: int, columns: int, device: bool = False):
    """assignment_mle over an already packed assignment (bit i = words[i // 64]
    >> (i % 64) & 1, as PackedWitness::bit), e.g. a SHA-256 witness's h."""
    words = np.ascontiguousarray(words, dtype=np.uint64)
    if not len(words):
        words = np.zeros(1, np.uint64)
    padded = 1 << max(int(columns - 1).bit_length(), 0) if columns > 1 else 1
    if device:
        out = DeviceTable(length=padded)
        _check(lib().bitz_assignment_mle(_p(words), bit_len, columns, C.c_void_p(out.ptr), DEVICE))
        return out
    out = np.zeros((padded, 2), dtype=np.uint64)
    _check(lib().bitz_assignment_mle(_p(words), bit_len, columns, _p(out), 0))
    return out


def constraint_matrix_digest(m_csr, m_columns: int, rows: int, columns: int, csr) -> bytes:
    """constraint_matrix_digest (crates/spartan/src/matrix.rs:562-627): m_csr =
    (row_starts, positions) of M; csr = three (row_starts, columns, values).
    Host computation (no device)."""
    m_starts = np.ascontiguousarray(m_csr[0], dtype=np.uint64)
    m_pos = np.ascontiguousarray(m_csr[1], dtype=np.uint32) if len(m_csr[1]) else np.zeros(1, np.uint32)
    starts = [np.ascontiguousarray(m[0], dtype=np.uint64) for m in csr]
    cols = [np.ascontiguousarray(m[1], dtype=np.uint32) if len(m[1]) else np.zeros(1, np.uint32) for m in csr]
    vals = [elems(m[2]) if len(m[2]) else np.zeros((1, 2), np.uint64) for m in csr]
    arr_s = (C.c_void_p * 3)(*[x.ctypes.data for x in starts])
    arr_c = (C.c_void_p * 3)(*[x.ctypes.data for x in cols])
    arr_v = (C.c_void_p * 3)(*[x.ctypes.data for x in vals])
    out = (C.c_uint8 * 32)()
    _check(lib().bitz_constraint_matrix_digest(len(m_starts) - 1, m_columns, _p(m_starts), _p(m_pos), rows, columns,
                                               arr_s, arr_c, arr_v, out))
    return bytes(out)


@dataclass
class SpartanPiopOutput:
    """SpartanPiopProof (crates/spartan/src/piop.rs:18-22) and its terminal
    ScaledMleEvaluationClaim (piop.rs:101-105)."""
    outer_rounds: np.ndarray     # (row_vars, 4, 2)
    outer_claims: list           # [Az, Bz, Cz](r_x)
    inner_rounds: np.ndarray     # (col_vars, 3, 2)
    inner_point: np.ndarray      # r_y
    matrix_evaluation: int       # D(r_y)
    final_claim: int


def prove_spartan_piop(transcript: Transcript, matrices: "PreparedConstraintMatrices", digest: bytes,
                       products: R1csProductMles, assignment) -> SpartanPiopOutput:
    """prove_spartan_piop (crates/spartan/src/piop.rs:67-112) with the engine's
    outer sumcheck, bind_and_batch and inner sumcheck."""
    init()
    (az, na, fa), (bz, _, _), (cz, _, _) = (_table(t) for t in (products.az, products.bz, products.cz))
    packed = isinstance(assignment, PackedAssignment)
    if packed:
        hv, nh, fh = assignment, assignment.len, DEVICE if assignment.device else 0
    else:
        hv, nh, fh = _table(assignment)
    if fh != fa:
        raise BitzError(-7, "products and assignment must be on the same side")
    rv, cv = _nvars(na), _nvars(nh)
    fn = lib().bitz_prove_spartan_piop_packed if packed else lib().bitz_prove_spartan_piop
    outer = np.zeros((max(rv, 1) * 4, 2), dtype=np.uint64)
    claims = np.zeros((3, 2), dtype=np.uint64)
    inner = np.zeros((max(cv, 1) * 3, 2), dtype=np.uint64)
    point = np.zeros((max(cv, 1), 2), dtype=np.uint64)
    meval, final = F128(), F128()
    dg = (C.c_uint8 * 32).from_buffer_copy(digest)
    _check(fn(transcript.h, matrices.h, dg, _p(az), _p(bz), _p(cz), rv, _p(hv), cv, fa, _p(outer), _p(claims),
              _p(inner), _p(point), C.byref(meval), C.byref(final)), SumcheckError)
    return SpartanPiopOutput(outer[: 4 * rv].reshape(rv, 4, 2), [to_int(c) for c in claims],
                             inner[: 3 * cv].reshape(cv, 3, 2), point[:cv], to_int(meval), to_int(final))


class PiopEngine:
    """Many prove_spartan_piop proofs with their rounds batched
    (bitz_piop_engine_*): submit proofs (device products and assignment),
    step until they finish, take each one's SpartanPiopOutput.  The inputs of
    a proof are kept alive here until it is taken."""

    def __init__(self, background: bool = False):
        """background: a thread of the engine steps while proofs are in
        flight (then use wait/done/take, not step/launch/finish)."""
        init()
        h = C.c_void_p()
        _check(lib().bitz_piop_engine_create(C.byref(h)))
        self.h = h
        self._held: dict[int, tuple] = {}
        self.launched = False
        if background:
            _check(lib().bitz_piop_engine_run(h))

    def wait(self, timeout_s: float) -> int:
        """Blocks until a finished proof waits to be taken (or the timeout);
        returns how many wait."""
        n = C.c_size_t()
        _check(lib().bitz_piop_engine_wait(self.h, int(timeout_s * 1e6), C.byref(n)))
        return n.value

    def submit(self, tag: int, transcript: Transcript, matrices: "PreparedConstraintMatrices", digest: bytes,
               products: R1csProductMles, assignment) -> None:
        """assignment: a DeviceTable, or a PackedAssignment on the device."""
        (az, na, fa), (bz, _, _), (cz, _, _) = (_table(t) for t in (products.az, products.bz, products.cz))
        packed = isinstance(assignment, PackedAssignment)
        if packed:
            hv, nh, fh = assignment, assignment.len, DEVICE if assignment.device else 0
        else:
            hv, nh, fh = _table(assignment)
        if not (fa & DEVICE and fh & DEVICE):
            raise BitzError(-7, "the engine takes device tables")
        fn = lib().bitz_piop_engine_submit_packed if packed else lib().bitz_piop_engine_submit
        dg = (C.c_uint8 * 32).from_buffer_copy(digest)
        # Held before the native submit: a background engine may finish the
        # proof (and another thread take it) before submit returns.  A
        # rejected submit (a tag in use, say) restores what was held before.
        before = self._held.get(tag)
        self._held[tag] = (transcript, matrices, products, assignment, _nvars(na), _nvars(nh))
        try:
            _check(fn(self.h, tag, transcript.h, matrices.h, dg, _p(az), _p(bz), _p(cz), _nvars(na), _p(hv),
                      _nvars(nh)), SumcheckError)
        except BaseException:
            if before is None:
                self._held.pop(tag, None)
            else:
                self._held[tag] = before
            raise

    def submit_assignment(self, tag: int, transcript: Transcript, matrices: "PreparedConstraintMatrices",
                          digest: bytes, words, bit_len: int, columns: int) -> None:
        """A proof from the packed assignment on the host (PackedWitness
        words): the engine uploads it and computes Az, Bz, Cz itself
        (bitz_piop_engine_submit_assignment)."""
        words = np.ascontiguousarray(words, dtype=np.uint64)
        if not len(words):
            words = np.zeros(1, np.uint64)
        if bit_len > 64 * len(words):
            raise BitzError(-7, "fewer words than bits")
        dg = (C.c_uint8 * 32).from_buffer_copy(digest)
        before = self._held.get(tag)
        self._held[tag] = (transcript, matrices, None, None, matrices.num_row_vars, matrices.num_column_vars)
        try:
            _check(lib().bitz_piop_engine_submit_assignment(self.h, tag, transcript.h, matrices.h, dg, _p(words),
                                                            bit_len, columns), SumcheckError)
        except BaseException:
            if before is None:
                self._held.pop(tag, None)
            else:
                self._held[tag] = before
            raise

    def step(self) -> int:
        n = C.c_size_t()
        _check(lib().bitz_piop_engine_step(self.h, C.byref(n)))
        return n.value

    def launch(self) -> int:
        """Enqueues one step without waiting; returns the proofs in it."""
        n = C.c_size_t()
        _check(lib().bitz_piop_engine_launch(self.h, C.byref(n)))
        self.launched = n.value > 0
        return n.value

    def finish(self) -> int:
        """Completes the launched step; returns the proofs still in flight."""
        n = C.c_size_t()
        self.launched = False
        _check(lib().bitz_piop_engine_finish(self.h, C.byref(n)))
        return n.value

    @property
    def in_flight(self) -> int:
        return len(self._held) - len(self.done())

    def done(self) -> list[int]:
        n = C.c_size_t()
        _check(lib().bitz_piop_engine_done(self.h, None, 0, C.byref(n)))
        tags = np.zeros(max(n.value, 1), dtype=np.uint64)
        _check(lib().bitz_piop_engine_done(self.h, _p(tags), len(tags), C.byref(n)))
        return [int(t) for t in tags[: n.value]]

    def device_bytes(self, tag: int) -> int:
        """Device bytes the proof holds from its submission until taken."""
        b = C.c_uint64()
        _check(lib().bitz_piop_engine_device_bytes(self.h, tag, C.byref(b)))
        return b.value

    def take(self, tag: int) -> SpartanPiopOutput:
        rv, cv = self._held[tag][4], self._held[tag][5]
        outer = np.zeros((rv * 4, 2), dtype=np.uint64)
        claims = np.zeros((3, 2), dtype=np.uint64)
        inner = np.zeros((cv * 3, 2), dtype=np.uint64)
        point = np.zeros((cv, 2), dtype=np.uint64)
        meval, final = F128(), F128()
        rc = lib().bitz_piop_engine_take(self.h, tag, _p(outer), _p(claims), _p(inner), _p(point), C.byref(meval),
                                         C.byref(final))
        if rc != -8:  # not still in flight: the engine has released the proof and its inputs
            self._held.pop(tag, None)
        _check(rc, SumcheckError)
        return SpartanPiopOutput(outer.reshape(rv, 4, 2), [to_int(c) for c in claims], inner.reshape(cv, 3, 2),
                                 point, to_int(meval), to_int(final))

    def __del__(self):
        try:
            lib().bitz_piop_engine_free(self.h)
        except Exception:
            pass


class MaterializedMTranspose:
    """crates/circuit/src/matrix_transpose.rs:282-441 (CSC of M)."""

    def __init__(self, row_count: int, column_offsets, row_indices):
        init()
        self.offsets = np.ascontiguousarray(column_offsets, dtype=np.uint32)
        self.rows = np.ascontiguousarray(row_indices, dtype=np.uint32)
        self.row_count = row_count
        self.columns = len(self.offsets) - 1
        h = C.c_void_p()
        _check(lib().bitz_mt_prepare(C.byref(h), row_count, self.columns, _p(self.offsets),
                                     _p(self.rows) if len(self.rows) else None))
        self.h = h

    def apply(self, challenges) -> np.ndarray:
        ch = elems(challenges) if len(challenges) else np.zeros((0, 2), np.uint64)
        out = np.zeros((self.columns, 2), dtype=np.uint64)
        _check(lib().bitz_mt_apply(self.h, _p(ch), len(ch), _p(out), 0))
        return out

    def transpose_inner(self, row_weights, column_weights) -> np.ndarray:
        rw, cw = elems(row_weights), elems(column_weights)
        out = np.zeros((self.columns, 2), dtype=np.uint64)
        _check(lib().bitz_mt_transpose_inner(self.h, _p(rw), len(rw), _p(cw), len(cw), _p(out), 0))
        return out

    def apply_device(self, challenges: "DeviceTable", out: "DeviceTable") -> None:
        """apply with the challenges and the output resident on the device."""
        if challenges.len != self.row_count or out.len < self.columns:
            raise BitzError(-7, "one challenge per row of M and one output per column")
        _check(lib().bitz_mt_apply(self.h, C.c_void_p(challenges.ptr), challenges.len, C.c_void_p(out.ptr), DEVICE))

    def transpose_inner_device(self, row_weights: "DeviceTable", column_weights: "DeviceTable",
                               out: "DeviceTable") -> None:
        """transpose_inner with the weights and the output on the device."""
        if out.len < self.columns:
            raise BitzError(-7, "one output per column of M")
        _check(lib().bitz_mt_transpose_inner(self.h, C.c_void_p(row_weights.ptr), row_weights.len,
                                             C.c_void_p(column_weights.ptr), column_weights.len,
                                             C.c_void_p(out.ptr), DEVICE))

    def __del__(self):
        try:
            lib().bitz_mt_free(self.h)
        except Exception:
            pass


# -------------------------------------------------------------------- gkr ---
class LayerWitnesses:
    """The product tree's layers, resident on device (crates/gkr/src/lib.rs:896-904)."""

    def __init__(self, handle, log_leaves, log_groups):
        self.h, self.log_leaves, self.log_groups = handle, log_leaves, log_groups

    def layer(self, j: int) -> np.ndarray:
        n = C.c_size_t()
        _check(lib().bitz_gkr_layer(self.h, j, None, 0, C.byref(n)))
        out = np.zeros((n.value, 2), dtype=np.uint64)
        _check(lib().bitz_gkr_layer(self.h, j, _p(out), n.value, C.byref(n)))
        return out

    def __del__(self):
        try:
            lib().bitz_gkr_free(self.h)
        except Exception:
            pass


def grand_product(leaves, log_groups: int):
    """fold_layers (lib.rs:874-894) over dense leaves: returns (output, LayerWitnesses)."""
    init()
    lv, n, flags = _table(leaves)
    log_leaves = _nvars(n)
    out = np.zeros((1 << log_groups, 2), dtype=np.uint64)
    h = C.c_void_p()
    _check(lib().bitz_gkr_build(C.byref(h), _p(lv), log_leaves, log_groups, _p(out), flags))
    return out, LayerWitnesses(h, log_leaves, log_groups)


def grand_product_bittable(row_images, packed, log_rows: int, log_columns: int):
    """grand_product over a BitTable (crates/gkr/src/lib.rs:785-802): tree c
    multiplies row_images[b] over the rows b whose bit (column c, row b) is set,
    bit (c, b) being bit (c << log_rows) | b of `packed` (u64 words).  Returns
    (output, LayerWitnesses) like grand_product."""
    init()
    images = elems(row_images)
    words = np.ascontiguousarray(packed, dtype=np.uint64)
    if len(images) != 1 << log_rows:
        raise BitzError(-7, "one image per row of the table")
    if len(words) * 64 != 1 << (log_rows + log_columns):
        raise BitzError(-7, "the table has 2^(log_rows + log_columns) bits")
    out = np.zeros((1 << log_columns, 2), dtype=np.uint64)
    h = C.c_void_p()
    _check(lib().bitz_gkr_build_bittable(C.byref(h), _p(images), log_rows, _p(words), log_columns, _p(out), 0))
    return out, LayerWitnesses(h, log_rows + log_columns, log_columns)


def gpgkr_prove_with_challenges(point, challenges, witnesses: LayerWitnesses):
    """gpgkr_prove with the challenges known in advance
    (bitz_gkr_prove_with_challenges): per layer from the top, its round
    challenges in squeeze order, then t.  Returns (messages (rounds, 2, 2),
    layer_evals (layers, 2, 2), leaf point, claim)."""
    point = elems(point) if len(point) else np.zeros((0, 2), np.uint64)
    ch = elems(challenges)
    n0, nl = witnesses.log_groups, witnesses.log_leaves
    rounds = sum(range(n0, nl))
    layers = nl - n0
    msgs = np.zeros((max(rounds, 1) * 2, 2), dtype=np.uint64)
    evals = np.zeros((max(layers, 1) * 2, 2), dtype=np.uint64)
    leaf_point = np.zeros((nl, 2), dtype=np.uint64)
    claim = F128()
    _check(lib().bitz_gkr_prove_with_challenges(witnesses.h, _p(point), len(point), _p(ch), len(ch), _p(msgs),
                                                _p(evals), _p(leaf_point), C.byref(claim)))
    return (msgs[: 2 * rounds].reshape(rounds, 2, 2), evals[: 2 * layers].reshape(layers, 2, 2), leaf_point,
            to_int(claim))


def set_engine_positions(positions: int) -> None:
    """Rounds a PIOP engine step may advance each device-transcript proof by
    (1 to 16), or 0 (the default: 6 while an engine holds fewer than 8
    proofs, 4 from 8 on); the proofs do not depend on it."""
    _check(lib().bitz_set_engine_positions(int(positions)))


def set_device_fiat_shamir(on: bool) -> None:
    """Whether provers draw a Keccak transcript's challenges on the device
    (the default) or through the host transcript, round by round."""
    _check(lib().bitz_set_device_fiat_shamir(1 if on else 0))


def keccak_hash(msg: bytes, rate: int, pad: int, out_len: int, device: bool = False) -> bytes:
    """A one-message Keccak sponge (rate in bytes; pad 0x06 SHA3, 0x1F SHAKE),
    on the host or by one device thread."""
    if device:
        init()
    out = (C.c_uint8 * max(out_len, 1))()
    buf = C.create_string_buffer(bytes(msg), max(len(msg), 1))
    _check(lib().bitz_keccak_hash(buf, len(msg), rate, pad, out, out_len, DEVICE if device else 0))
    return bytes(out[:out_len])


def encode_proof(narg: bytes, hints: bytes = b"") -> bytes:
    """crates/host/src/wire_proof.rs encode: the 32-byte header, narg, hints."""
    narg, hints = bytes(narg), bytes(hints)
    n = C.c_size_t()
    a = C.create_string_buffer(narg, max(len(narg), 1))
    b = C.create_string_buffer(hints, max(len(hints), 1))
    _check(lib().bitz_proof_encode(a, len(narg), b, len(hints), None, 0, C.byref(n)))
    out = (C.c_uint8 * n.value)()
    _check(lib().bitz_proof_encode(a, len(narg), b, len(hints), out, n.value, C.byref(n)))
    return bytes(out)


def decode_proof(data: bytes) -> tuple[bytes, bytes]:
    """wire_proof.rs decode: (narg, hints) of the exact encoding; anything
    else raises ProofDecodingError with the reference's variant."""
    data = bytes(data)
    buf = C.create_string_buffer(data, max(len(data), 1))
    no, nl, ho, hl, detail = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_uint64()
    rc = lib().bitz_proof_decode(buf, len(data), C.byref(no), C.byref(nl), C.byref(ho), C.byref(hl), C.byref(detail))
    if rc:
        raise ProofDecodingError(rc, (lib().bitz_last_error() or b"").decode(), detail.value)
    return data[no.value: no.value + nl.value], data[ho.value: ho.value + hl.value]


BITS100, BITS128 = 100, 128  # SecurityLevel (crates/common/src/security.rs:7-10)


def grinding_bits(security: int, coefficient: int) -> int:
    """SecurityLevel::grinding_bits (crates/common/src/security.rs:25-32)."""
    return int(lib().bitz_grinding_bits(security, coefficient))


def pow_find(seed: bytes, bits: int) -> int:
    """crates/transcript/src/pow.rs:21-33: the first valid nonce."""
    nonce = C.c_uint64()
    _check(lib().bitz_pow_find(C.create_string_buffer(bytes(seed), 16), bits, C.byref(nonce)))
    return nonce.value


def pow_valid(seed: bytes, nonce: int, bits: int) -> bool:
    """crates/transcript/src/pow.rs:38-52."""
    return bool(lib().bitz_pow_valid(C.create_string_buffer(bytes(seed), 16), nonce, bits))


def gpgkr_prove(transcript: Transcript, point, witnesses: LayerWitnesses, security: int = BITS100):
    """gpgkr_prove (lib.rs:24-79) over dense layers: returns (leaf point,
    claim).  security: BITS100 (no grinding) or BITS128 (2 bits of grinding
    before every cubic round's challenge, through the transcript's grind)."""
    point = elems(point) if len(point) else np.zeros((0, 2), np.uint64)
    leaf_point = np.zeros((witnesses.log_leaves, 2), dtype=np.uint64)
    claim = F128()
    _check(lib().bitz_gkr_prove(witnesses.h, transcript.h, _p(point), len(point), security, _p(leaf_point),
                                C.byref(claim)))
    return leaf_point, to_int(claim)


# --------------------------------------------------------------- post_gkr --
def post_gkr_prove(transcript: Transcript, rows, columns, target, packed, security: int = BITS100):
    """post_gkr::prove (crates/post_gkr/src/lib.rs:85-148) on the GPU: the
    claim <rows (x) columns, f> = target over the bits f (2^t x 2^s, t >= 7,
    `packed` 128 bits per element column after column, or DeviceTables)
    reduced to MLE[f](point) = evaluation.  Returns (point (t + s, 2),
    evaluation)."""
    init()
    (r, nr, fr), (c, nc, fc), (p, npk, fp) = _table(rows), _table(columns), _table(packed)
    if not (fr == fc == fp):
        raise BitzError(-7, "rows, columns and packed must all be host arrays or all DeviceTables")
    t, s = _nvars(nr), _nvars(nc)
    point = np.zeros((t + s, 2), dtype=np.uint64)
    ev = F128()
    _check(lib().bitz_post_gkr_prove(transcript.h, _p(r), t, _p(c), s, f128(target), _p(p), npk, security, fr,
                                     _p(point), C.byref(ev)))
    return point, to_int(ev)
