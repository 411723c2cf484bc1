"""Serving: a proving loop over Spartan PIOP requests of uneven size with
admission by an HBM budget and concurrent requests overlapping on the GPU.

A request is the sumcheck stage of a SHA-256-chain proof of B blocks:
assignment MLE, Az/Bz/Cz, and prove_spartan_piop (outer sumcheck,
bind_and_batch, inner sumcheck; crates/spartan/src/piop.rs:67-112).  The
reference has no SHA-256 R1CS producer over GF(2^128) (its product builder
works over a prime field, crates/spartan/src/matrix.rs:336-370), so each
block count gets a synthetic circuit of the chain's shape: 696 B + 256
constraint rows and 20457 B assignment columns (tooling/cli/src/circuits.rs:
167-209, tooling/cli/tests/sha256_spartan.rs:44-49), eight nonzeros per row
in each of A, B and C.  Circuits are prepared once per block count; a
request brings its own assignment bits.

Scheduling: requests wait in arrival order; one is admitted when a slot is
free and its HBM footprint fits the budget beside the requests in flight.
Every admitted request runs on its own host thread with its own HIP streams,
so the requests' kernels overlap on the device; a request that arrives while
others run starts as soon as a slot and memory free up.  Each result depends
only on its request (its circuit, its bits, its transcript), not on the mix.
"""
from __future__ import annotations

import threading
import time
from collections import deque
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass, field

import numpy as np

from . import (PreparedConstraintMatrices, SpartanPiopOutput, Transcript, assignment_mle,
               constraint_matrix_digest, init, prove_spartan_piop)

ROWS_PER_BLOCK, ROWS_FIXED = 696, 256
COLUMNS_PER_BLOCK = 20457
NNZ_PER_ROW = 8


def chain_shape(blocks: int) -> tuple[int, int, int, int]:
    """(rows, columns, row_vars, col_vars) of a SHA-256 chain of `blocks` blocks."""
    rows = ROWS_PER_BLOCK * blocks + ROWS_FIXED
    cols = COLUMNS_PER_BLOCK * blocks
    return rows, cols, max(rows - 1, 1).bit_length(), max(cols - 1, 1).bit_length()


def _random_csr(rng, rows: int, cols: int, k: int):
    c = np.sort(rng.integers(0, cols, size=(rows, k), dtype=np.int64), axis=1)
    keep = np.ones_like(c, dtype=bool)
    keep[:, 1:] = c[:, 1:] != c[:, :-1]
    counts = keep.sum(axis=1)
    starts = np.zeros(rows + 1, np.uint64)
    np.cumsum(counts, out=starts[1:])
    columns = c[keep].astype(np.uint32)
    values = rng.integers(0, 2**64, size=(len(columns), 2), dtype=np.uint64)
    return starts, columns, values


@dataclass
class Circuit:
    blocks: int
    rows: int
    columns: int
    row_vars: int
    col_vars: int
    matrices: PreparedConstraintMatrices
    digest: bytes
    nnz: int = 0

    def footprint(self) -> int:
        """Device bytes a request holds at its peak: the assignment and D
        (2^col_vars each) with the inner sumcheck's scratch (1.5 x 2^col_vars),
        the three products (3 x 2^row_vars) with the outer scratch (2.25 x)."""
        return 16 * int((1 << self.col_vars) * 3.5 + (1 << self.row_vars) * 5.25)


def make_circuit(blocks: int, seed: int = 0) -> Circuit:
    rows, cols, rv, cv = chain_shape(blocks)
    rng = np.random.default_rng(seed * 1000 + blocks)
    csr = [_random_csr(rng, rows, cols, NNZ_PER_ROW) for _ in range(3)]
    m = csr[0][:2]  # the positions of M (a stand-in with A's pattern)
    digest = constraint_matrix_digest(m, cols, rows, cols, csr)
    return Circuit(blocks, rows, cols, rv, cv, PreparedConstraintMatrices(rows, rv, cv, csr), digest,
                   sum(len(m[1]) for m in csr))


@dataclass
class Request:
    rid: int
    arrival: float  # seconds from the start of serving
    blocks: int
    bits: np.ndarray
    admitted: float = 0.0
    done: float = 0.0
    result: SpartanPiopOutput | None = field(default=None, repr=False)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def request_bits(rng, columns: int) -> np.ndarray:
    bits = rng.integers(0, 2, size=columns, dtype=np.uint8)
    bits[0] = 1  # the R1CS constant one
    return bits


class ProvingServer:
    """Admits requests by HBM budget and slots; see the module notes."""

    def __init__(self, hbm_budget: int = 32 << 30, slots: int = 8, seed: int = 0):
        init()
        self.hbm_budget, self.slots, self.seed = hbm_budget, slots, seed
        self.circuits: dict[int, Circuit] = {}
        self._lock = threading.Lock()
        self.peak_bytes = 0      # largest footprint in flight during the last serve()
        self.peak_in_flight = 0  # most requests in flight during the last serve()

    def circuit(self, blocks: int) -> Circuit:
        with self._lock:
            if blocks not in self.circuits:
                self.circuits[blocks] = make_circuit(blocks, self.seed)
            return self.circuits[blocks]

    def prove(self, rid: int, blocks: int, bits: np.ndarray) -> SpartanPiopOutput:
        """One request, every table on the device: assignment, products, PIOP."""
        c = self.circuit(blocks)
        h = assignment_mle(bits, c.columns, device=True)
        products = c.matrices.r1cs_products(h)
        t = Transcript.duplex(b"bitz/serve/%d" % rid)
        return prove_spartan_piop(t, c.matrices, c.digest, products, h)

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
                    in_flight[pool.submit(self.prove, r.rid, r.blocks, r.bits)] = (r, fp)
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
                    r.result = f.result()
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
        out.append(Request(i, t, blocks, request_bits(rng, chain_shape(blocks)[1])))
    return out
