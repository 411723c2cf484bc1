"""Serving: every result is the same alone and in the mix, admission keeps
the footprint in flight within the HBM budget, each request proves its own
SHA-256 chain (output state = the compressions of its message from its
chaining value), and a served proof is the PIOP the oracle composes."""
import struct

import numpy as np
import pytest

import bitz_amd as G
from bitz_amd import serving as S
from common import compress_py
from test_gpu_sha256 import prove_and_check

pytestmark = pytest.mark.gpu


def same(a, b):
    return (np.array_equal(a.outer_rounds, b.outer_rounds) and a.outer_claims == b.outer_claims and
            np.array_equal(a.inner_rounds, b.inner_rounds) and np.array_equal(a.inner_point, b.inner_point) and
            a.matrix_evaluation == b.matrix_evaluation and a.final_claim == b.final_claim)


def chain(iv, message):
    st = list(iv)
    for b in range(len(message) // 64):
        st = compress_py(st, message[64 * b: 64 * b + 64])
    return struct.pack(">8I", *st)


def test_results_are_the_same_alone_and_in_the_mix():
    srv = S.ProvingServer(hbm_budget=1 << 30, slots=4)
    reqs = S.poisson_requests(rate=5000.0, count=12, seed=3, max_log_blocks=3)
    alone = [srv.prove(r.rid, r.blocks, r.message, r.iv) for r in reqs]
    served = srv.serve(reqs)
    assert srv.peak_in_flight > 1  # the requests did overlap
    for (proof, digest), r in zip(alone, served):
        assert r.result is not None and same(proof, r.result)
        assert r.digest == digest == chain(r.iv, r.message)
        assert r.done >= r.admitted >= r.arrival - 1e-6


def test_batched_serving_gives_the_same_proofs():
    srv = S.ProvingServer(hbm_budget=1 << 30, slots=6)
    reqs = S.poisson_requests(rate=5000.0, count=12, seed=6, max_log_blocks=4)
    alone = [srv.prove(r.rid, r.blocks, r.message, r.iv) for r in reqs]
    served = srv.serve_batched(reqs)
    assert srv.peak_in_flight > 1
    for (proof, digest), r in zip(alone, served):
        assert r.result is not None and same(proof, r.result)
        assert r.digest == digest == chain(r.iv, r.message)
        assert r.done >= r.admitted >= r.arrival - 1e-6


def test_batched_admission_keeps_the_budget():
    """The native server's requests share what the resident circuits (their
    matrices and the engines' bind buffers) leave of the budget: with room
    for two of the largest requests beside them, the resident bytes plus the
    footprints in flight stay within the budget."""
    srv = S.ProvingServer(hbm_budget=1 << 34, slots=8)
    reqs = S.poisson_requests(rate=5000.0, count=10, seed=7, max_log_blocks=3)
    fps = [srv.circuit(r.blocks).footprint() for r in reqs]
    srv.serve_batched(reqs)
    resident = srv.resident_bytes
    assert resident > 0
    srv.hbm_budget = resident + 2 * max(fps)
    srv.serve_batched(reqs)
    assert srv.resident_bytes == resident
    assert srv.peak_bytes <= srv.hbm_budget
    assert all(r.result is not None for r in reqs)


@pytest.mark.parametrize("blocks", [1, 8, 64, 256])
def test_footprint_covers_a_requests_device_bytes(blocks):
    """What one request holds on the device in a PIOP engine, measured (every
    buffer as allocated, from its submission until it is taken), is within
    the footprint admission budgets for it."""
    c = S.make_circuit(blocks)
    msg, iv = S.random_input(np.random.default_rng(blocks), blocks)
    w = c.circuit.witness(msg, iv=list(iv))
    eng = G.PiopEngine()
    eng.submit_assignment(1, G.Transcript.keccak(b"fp/%d" % blocks), c.matrices, c.digest, w.h, c.columns,
                          c.columns)
    used = eng.device_bytes(1)
    steps = 0
    while eng.step():
        steps += 1
        assert steps < 100
    eng.take(1)
    print(f"blocks {blocks}: {used} device bytes, footprint {c.footprint()} ({used / c.footprint():.3f})")
    assert 0 < used <= c.footprint()


def test_admission_keeps_the_budget():
    srv = S.ProvingServer(hbm_budget=1 << 30, slots=8)
    reqs = S.poisson_requests(rate=5000.0, count=10, seed=4, max_log_blocks=3)
    fps = [srv.circuit(r.blocks).footprint() for r in reqs]
    srv.hbm_budget = 2 * max(fps)  # at most two of the largest at once
    srv.serve(reqs)
    assert srv.peak_bytes <= srv.hbm_budget
    assert all(r.result is not None for r in reqs)


def test_a_served_proof_matches_the_oracle():
    srv = S.ProvingServer(hbm_budget=1 << 30, slots=2)
    reqs = [r for r in S.poisson_requests(rate=5000.0, count=8, seed=5, max_log_blocks=2) if r.blocks == 2][:1]
    assert reqs
    r = srv.serve(reqs)[0]
    _, ref = prove_and_check(srv.circuit(2).circuit, r.message, list(r.iv), label=b"bitz/serve/%d" % r.rid,
                             keccak=True)
    assert same(r.result, ref)
