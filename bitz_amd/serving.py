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
  - serve_batched (continuous batching): a few threads prepare each request
    (witness, assignment, products), then it joins a PiopEngine that advances
    every proof in flight by one round per step, the rounds of one kind
    across all proofs in one launch; a request joins at the next step and
    leaves when done.
  - serve: every admitted request runs on its own host thread with its own
    HIP streams, so the requests' kernels overlap on the device.
Each result depends only on its request (its circuit, its message and
chaining value, its transcript), not on the mix or the mode.
"""
from __future__ import annotations

import threading
import time
from collections import deque
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass, field

import numpy as np

from . import (PackedAssignment, PiopEngine, PreparedConstraintMatrices, SpartanPiopOutput, Transcript, init,
               prove_spartan_piop)
from . import sha256 as sha


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
        the outer scratch (2.25 x)."""
        return 16 * int((1 << self.col_vars) * 2.5 + (1 << self.row_vars) * 5.25) + (1 << self.col_vars) // 8


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
        self.peak_in_flight = 0  # most requests in flight during the last serve()
        self._engines: list[PiopEngine] = []  # kept across serve_batched calls

    def engines(self, count: int) -> list[PiopEngine]:
        """count background PIOP engines, created on first use and kept (each
        holds a stream, a thread and its staging buffers)."""
        while len(self._engines) < count:
            self._engines.append(PiopEngine(background=True))
        return self._engines[:count]

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
        t = Transcript.duplex(b"bitz/serve/%d" % rid)
        return prove_spartan_piop(t, c.matrices, c.digest, products, h), w.digest

    def prepare(self, r: "Request"):
        """A request's device inputs: witness (host), the assignment's packed
        bits, Az/Bz/Cz."""
        c = self.circuit(r.blocks)
        w = c.circuit.witness(r.message, iv=r.iv)
        h = PackedAssignment(w.h, c.columns, c.columns, device=True)
        return c, h, c.matrices.r1cs_products(h), w.digest

    def serve_batched(self, requests: list[Request], preparers: int = 4, engines: int | None = None) -> list[Request]:
        """Continuous batching: admitted requests are prepared (witness,
        assignment, products) by a few threads, then join a PiopEngine (the
        least loaded of `engines`, each with its own stream and thread) that
        advances every proof it holds by one round per step, each round kind
        in one launch across the proofs.  A request joins at its engine's next
        step and leaves when its proof is done; admission is by slots and HBM
        budget as in serve().  engines defaults to one per 8 slots."""
        engines = engines or max(1, self.slots // 8)
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        for r in pending:
            self.circuit(r.blocks)
        pool_engines = self.engines(engines)
        load = [0] * engines
        by_tag: dict[int, tuple] = {}
        used = 0
        self.peak_bytes = self.peak_in_flight = 0
        t0 = time.perf_counter()
        lock = threading.Lock()

        def admit_one(r: Request, fp: int) -> None:
            c, h, products, digest = self.prepare(r)
            r.digest = digest
            with lock:
                k = min(range(engines), key=lambda i: load[i])
                load[k] += 1
                by_tag[r.rid] = (r, fp, h, products, k)
            pool_engines[k].submit(r.rid, Transcript.duplex(b"bitz/serve/%d" % r.rid), c.matrices, c.digest,
                                   products, h)

        in_flight = 0
        with ThreadPoolExecutor(max_workers=preparers) as pool:
            futures = []
            while pending or in_flight:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival <= now and in_flight < self.slots:
                    fp = self.circuits[pending[0].blocks].footprint()
                    if in_flight and used + fp > self.hbm_budget:
                        break
                    r = pending.popleft()
                    r.admitted = now
                    used += fp
                    in_flight += 1
                    futures.append(pool.submit(admit_one, r, fp))
                    self.peak_bytes = max(self.peak_bytes, used)
                    self.peak_in_flight = max(self.peak_in_flight, in_flight)
                for f in [f for f in futures if f.done()]:
                    f.result()  # a failed preparation raises here
                    futures.remove(f)
                timeout = 0.0005
                if pending and in_flight < self.slots:
                    timeout = min(timeout, max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                if in_flight and (pool_engines[0].wait(timeout / engines) if engines == 1 else
                                  sum(e.wait(timeout / engines) for e in pool_engines)):
                    for k, e in enumerate(pool_engines):
                        for tag in e.done():
                            with lock:
                                r, fp, _, _, _ = by_tag.pop(tag)
                                load[k] -= 1
                            r.result = e.take(tag)
                            r.done = time.perf_counter() - t0
                            used -= fp
                            in_flight -= 1
                elif not in_flight and pending:
                    time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
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
