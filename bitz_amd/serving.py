"""Serving: a proving loop over SHA-256 proving requests of uneven size with
admission by an HBM budget and concurrent requests overlapping on the GPU.

A request is the sumcheck stage of a proof that a chain of B SHA-256
compressions maps its chaining value through its message blocks: the
witness (host, C++: bitz_sha256_witness), the assignment (kept as its packed
bits), Az/Bz/Cz and prove_spartan_piop (outer sumcheck, bind_and_batch, inner sumcheck;
crates/spartan/src/piop.rs:67-112), every table on the device.  The circuit
is the compression chain of the SHA-256 producer: one block is the
reference's compression_circuit (crates/circuit/src/sha256.rs:476-490), B
blocks chain it, and the message and chaining value are witness inputs, so
one prepared circuit (matrices in the parity projection, statement digest)
per block count serves every request of that size, as the reference's
benchmark prepares one instance per size (tooling/cli/benches/
sha256_spartan.rs:45-90).

Scheduling: requests wait in arrival order; one is admitted when a slot is
free and its HBM footprint fits the budget beside the requests in flight.
Two ways to run the admitted ones:
  - serve_batched (continuous batching, native: bitz_server_serve): preparer
    threads compute each admitted request's witness, then it joins a PIOP
    engine (which uploads the packed assignment and computes Az/Bz/Cz on a
    stream of the proof's own) that advances every proof in flight by several
    rounds per step (its Keccak transcript's challenges drawn on the device),
    the rounds of one kind across all proofs in one launch; a request joins
    at the next step and leaves when done.
  - serve: every admitted request runs on its own host thread with its own
    HIP streams, so the requests' kernels overlap on the device.
Each result depends only on its request (its circuit, its message and
chaining value, its transcript), not on the mix or the mode.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from collections import deque
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass, field

import numpy as np

from . import (PackedAssignment, PreparedConstraintMatrices, SpartanPiopOutput, SumcheckError, Transcript, _check,
               init, lib, prove_spartan_piop, to_int)
from . import sha256 as sha


class _ServeRequest(C.Structure):
    """bitz_serve_request (include/bitz_hip.h)."""
    _fields_ = [("rid", C.c_uint64), ("circuit", C.c_uint32), ("engine", C.c_uint32), ("message", C.c_void_p),
                ("iv", C.c_uint32 * 8), ("arrival_s", C.c_double), ("admitted_s", C.c_double),
                ("done_s", C.c_double), ("digest", C.c_uint8 * 32), ("status", C.c_int32),
                ("outer_rounds", C.c_void_p), ("outer_claims", C.c_void_p), ("inner_rounds", C.c_void_p),
                ("inner_point", C.c_void_p), ("matrix_eval", C.c_void_p), ("final_claim", C.c_void_p)]


_SERVE_DTYPE = np.dtype(_ServeRequest)  # the same layout as a numpy record


class NativeServer:
    """The library's proving server (bitz_server_*): admission, preparer
    threads, PIOP engines and harvesters all native, so serving holds no
    Python lock while it runs."""

    def __init__(self, slots: int, engines: int, preparers: int, hbm_budget: int):
        init()
        h = C.c_void_p()
        _check(lib().bitz_server_create(C.byref(h), slots, engines, preparers, hbm_budget))
        self.h = h
        self.ids: dict[int, int] = {}
        self.config = (slots, engines, preparers, hbm_budget)

    def resident_bytes(self) -> int:
        """Device bytes held for the server's life (circuits' matrices and the
        engines' bind buffers); requests share the rest of the budget."""
        b = C.c_uint64()
        _check(lib().bitz_server_resident_bytes(self.h, C.byref(b)))
        return b.value

    def add(self, c: "Circuit") -> int:
        if c.blocks not in self.ids:
            cid = C.c_uint32()
            dg = (C.c_uint8 * 32).from_buffer_copy(c.digest)
            _check(lib().bitz_server_add_circuit(self.h, c.circuit.h, c.matrices.h, dg, c.footprint(),
                                                 C.byref(cid)))
            self.ids[c.blocks] = cid.value
        return self.ids[c.blocks]

    def __del__(self):
        try:
            lib().bitz_server_free(self.h)
        except Exception:
            pass


@dataclass
class Circuit:
    blocks: int
    circuit: sha.Sha256Circuit
    matrices: PreparedConstraintMatrices
    digest: bytes

    @property
    def rows(self) -> int:
        return self.circuit.rows

    @property
    def columns(self) -> int:
        return self.circuit.columns

    @property
    def row_vars(self) -> int:
        return self.circuit.num_row_vars

    @property
    def col_vars(self) -> int:
        return self.circuit.num_column_vars

    @property
    def nnz(self) -> int:
        return sum(self.circuit.nnz)

    def footprint(self) -> int:
        """Device bytes a request holds at its peak: the packed assignment
        (2^col_vars bits), D (2^col_vars elements) with the inner sumcheck's
        scratch (1.5 x 2^col_vars), the three products (3 x 2^row_vars) with
        the outer scratch (2.25 x), and 256 KiB of small tables (transcript
        state and outputs, eq factors, challenge and H(i) tables).  A PIOP
        engine's measured bytes per request (PiopEngine.device_bytes) stay
        within it for 1 to 256 blocks: 91 % of it at 1 block, 99.9 % at 256."""
        return (16 * int((1 << self.col_vars) * 2.5 + (1 << self.row_vars) * 5.25) + (1 << self.col_vars) // 8 +
                (256 << 10))


def make_circuit(blocks: int) -> Circuit:
    c = sha.Sha256Circuit(sha.COMPRESSION_CHAIN, blocks)
    matrices, digest = c.prepare()
    return Circuit(blocks, c, matrices, digest)


@dataclass
class Request:
    rid: int
    arrival: float  # seconds from the start of serving
    blocks: int
    message: bytes  # 64 bytes per block
    iv: tuple       # the chaining value, 8 words
    admitted: float = 0.0
    done: float = 0.0
    result: SpartanPiopOutput | None = field(default=None, repr=False)
    digest: bytes = b""  # the chain's output state

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def random_input(rng, blocks: int) -> tuple[bytes, tuple]:
    """Message bytes and chaining value, uniform from the generator."""
    return (rng.integers(0, 256, 64 * blocks, dtype=np.uint8).tobytes(),
            tuple(int(x) for x in rng.integers(0, 2**32, 8, dtype=np.uint64)))


class ProvingServer:
    """Admits requests by HBM budget and slots; see the module notes."""

    def __init__(self, hbm_budget: int = 32 << 30, slots: int = 8):
        init()
        self.hbm_budget, self.slots = hbm_budget, slots
        self.circuits: dict[int, Circuit] = {}
        self._lock = threading.Lock()
        self.peak_bytes = 0      # largest footprint in flight during the last serve()
        self.resident_bytes = 0  # serve_batched: circuits' matrices and bind buffers, held throughout
        self.peak_in_flight = 0  # most requests in flight during the last serve()
        self._native: NativeServer | None = None  # serve_batched's server, kept across calls

    @staticmethod
    def transcript(rid: int) -> Transcript:
        """A request's transcript: the Keccak duplex, so the engine draws its
        challenges on the device (several rounds per step, no host round trip
        per round)."""
        return Transcript.keccak(b"bitz/serve/%d" % rid)

    def circuit(self, blocks: int) -> Circuit:
        with self._lock:
            if blocks not in self.circuits:
                self.circuits[blocks] = make_circuit(blocks)
            return self.circuits[blocks]

    def prove(self, rid: int, blocks: int, message: bytes, iv) -> tuple[SpartanPiopOutput, bytes]:
        """One request: the witness on the host, then assignment, products and
        PIOP with every table on the device.  Returns (proof, output state)."""
        c = self.circuit(blocks)
        w = c.circuit.witness(message, iv=iv)
        h = PackedAssignment(w.h, c.columns, c.columns, device=True)
        products = c.matrices.r1cs_products(h)
        return prove_spartan_piop(self.transcript(rid), c.matrices, c.digest, products, h), w.digest

    def prepare(self, r: "Request"):
        """A request's device inputs: witness (host), the assignment's packed
        bits, Az/Bz/Cz."""
        c = self.circuit(r.blocks)
        w = c.circuit.witness(r.message, iv=r.iv)
        h = PackedAssignment(w.h, c.columns, c.columns, device=True)
        return c, h, c.matrices.r1cs_products(h), w.digest

    def serve_batched(self, requests: list[Request], preparers: int = 4, engines: int | None = None) -> list[Request]:
        """Continuous batching on the library's native server
        (bitz_server_serve): requests are admitted in arrival order by slots
        and HBM budget; preparer threads compute each one's witness, then it
        joins the least loaded of `engines` PIOP engines, which upload its
        packed assignment, compute Az/Bz/Cz and advance every proof they hold
        by up to 6 rounds per step (challenges drawn on the device from the
        Keccak transcript), each round kind in one launch across the proofs;
        harvester threads take finished proofs and free their slots.  engines
        defaults to one per 8 slots."""
        engines = engines or max(1, self.slots // 8)
        config = (self.slots, engines, preparers, self.hbm_budget)
        if self._native is None or self._native.config != config:
            self._native = None
            self._native = NativeServer(*config)
        ns = self._native
        m = len(requests)
        # The request records as one numpy array in bitz_serve_request's layout, filled a column at a time;
        # the messages back to back in one buffer; one output array for every request: rows [outer rounds
        # (4 rv) | claims (3) | inner rounds (3 cv) | inner point (cv) | matrix eval | final claim] each.
        reqs = np.zeros(max(m, 1), dtype=_SERVE_DTYPE)
        circ = [self.circuit(r.blocks) for r in requests]
        dims = np.array([(c.row_vars, c.col_vars) for c in circ], dtype=np.int64).reshape(-1, 2)
        rv_a, cv_a = dims[:, 0], dims[:, 1]
        starts = np.concatenate([[0], np.cumsum(4 * rv_a + 4 * cv_a + 5)]).astype(np.int64)
        out = np.zeros((max(int(starts[-1]), 1), 2), np.uint64)
        msgs = np.frombuffer(b"".join(r.message for r in requests) or b"\0", dtype=np.uint8)
        moff = np.concatenate([[0], np.cumsum([len(r.message) for r in requests])]).astype(np.int64)
        if m:
            ids = {b: ns.add(c) for b, c in zip((r.blocks for r in requests), circ)}
            b = out.ctypes.data + 16 * starts[:-1]
            reqs["rid"][:m] = [r.rid for r in requests]
            reqs["circuit"][:m] = [ids[r.blocks] for r in requests]
            reqs["message"][:m] = msgs.ctypes.data + moff[:-1]
            reqs["iv"][:m] = np.array([r.iv for r in requests], dtype=np.uint64) & 0xffffffff
            reqs["arrival_s"][:m] = [r.arrival for r in requests]
            reqs["outer_rounds"][:m] = b
            reqs["outer_claims"][:m] = b + 64 * rv_a
            reqs["inner_rounds"][:m] = b + 16 * (4 * rv_a + 3)
            reqs["inner_point"][:m] = b + 16 * (4 * rv_a + 3 + 3 * cv_a)
            reqs["matrix_eval"][:m] = b + 16 * (4 * rv_a + 3 + 4 * cv_a)
            reqs["final_claim"][:m] = b + 16 * (4 * rv_a + 4 + 4 * cv_a)
        self.peak_bytes = self.peak_in_flight = 0
        t0 = time.perf_counter()
        _check(lib().bitz_server_serve(ns.h, C.c_void_p(reqs.ctypes.data), m), SumcheckError)
        self.native_seconds = time.perf_counter() - t0  # the native server's part of the call
        if m:
            # the five scalar outputs of every request (claims, matrix eval, final claim) as ints in one pass
            sc = starts[:-1] + 4 * rv_a
            rows = np.concatenate([np.stack([sc, sc + 1, sc + 2], axis=1).ravel(),
                                   np.stack([sc + 3 + 4 * cv_a, sc + 4 + 4 * cv_a], axis=1).ravel()])
            vals = [lo | (hi << 64) for lo, hi in out[rows].tolist()]
            adm, done = reqs["admitted_s"][:m].tolist(), reqs["done_s"][:m].tolist()
            digs = reqs["digest"][:m].tobytes()
            for i, r in enumerate(requests):
                rv, cv, o = int(rv_a[i]), int(cv_a[i]), int(sc[i])
                r.admitted, r.done = adm[i], done[i]
                r.digest = digs[32 * i:32 * i + 32]
                r.result = SpartanPiopOutput(out[o - 4 * rv:o].reshape(rv, 4, 2), vals[3 * i:3 * i + 3],
                                             out[o + 3:o + 3 + 3 * cv].reshape(cv, 3, 2),
                                             out[o + 3 + 3 * cv:o + 3 + 4 * cv], vals[3 * m + 2 * i],
                                             vals[3 * m + 2 * i + 1])
        # the peaks of the admission that just ran (replayed from the times),
        # the resident circuits' bytes included
        self.resident_bytes = ns.resident_bytes()
        fp = {b: self.circuits[b].footprint() for b in {r.blocks for r in requests}}
        ev = sorted([(r.admitted, 1, fp[r.blocks]) for r in requests] +
                    [(r.done, -1, -fp[r.blocks]) for r in requests], key=lambda x: (x[0], x[1]))
        n, b = 0, self.resident_bytes
        self.peak_bytes = b
        for _, dn, db in ev:
            n += dn
            b += db
            self.peak_in_flight, self.peak_bytes = max(self.peak_in_flight, n), max(self.peak_bytes, b)
        return requests

    def serve(self, requests: list[Request]) -> list[Request]:
        """Runs requests at their arrival times (seconds from now); returns
        them with admission and completion times and results."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        for r in pending:
            self.circuit(r.blocks)  # circuits are prepared ahead, like a deployed service
        in_flight: dict = {}
        used = 0
        self.peak_bytes = self.peak_in_flight = 0
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=self.slots) as pool:
            while pending or in_flight:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival <= now and len(in_flight) < self.slots:
                    fp = self.circuits[pending[0].blocks].footprint()
                    if in_flight and used + fp > self.hbm_budget:
                        break
                    r = pending.popleft()
                    r.admitted = now
                    used += fp
                    in_flight[pool.submit(self.prove, r.rid, r.blocks, r.message, r.iv)] = (r, fp)
                    self.peak_bytes = max(self.peak_bytes, used)
                    self.peak_in_flight = max(self.peak_in_flight, len(in_flight))
                timeout = None
                if pending and len(in_flight) < self.slots:
                    timeout = max(0.0, pending[0].arrival - (time.perf_counter() - t0))
                if not in_flight:
                    time.sleep(timeout or 0)
                    continue
                done, _ = wait(list(in_flight), timeout=timeout, return_when=FIRST_COMPLETED)
                for f in done:
                    r, fp = in_flight.pop(f)
                    r.result, r.digest = f.result()
                    r.done = time.perf_counter() - t0
                    used -= fp
        return requests


def poisson_requests(rate: float, count: int, seed: int, max_log_blocks: int = 8) -> list[Request]:
    """count requests with exponential inter-arrival times at `rate` per
    second, blocks = 2^U with U uniform in {0 .. max_log_blocks}."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        blocks = 1 << int(rng.integers(0, max_log_blocks + 1))
        out.append(Request(i, t, blocks, *random_input(rng, blocks)))
    return out
