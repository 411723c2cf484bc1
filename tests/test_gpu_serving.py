"""Serving: every result is the same alone and in the mix, admission keeps
the footprint in flight within the HBM budget, and the served proofs are
the PIOP the oracle composes (checked for one request)."""
import numpy as np
import pytest

import bitz_amd as G
from bitz_amd import serving as S

pytestmark = pytest.mark.gpu


def same(a, b):
    return (np.array_equal(a.outer_rounds, b.outer_rounds) and a.outer_claims == b.outer_claims and
            np.array_equal(a.inner_rounds, b.inner_rounds) and np.array_equal(a.inner_point, b.inner_point) and
            a.matrix_evaluation == b.matrix_evaluation and a.final_claim == b.final_claim)


def test_results_are_the_same_alone_and_in_the_mix():
    srv = S.ProvingServer(hbm_budget=1 << 30, slots=4, seed=1)
    reqs = S.poisson_requests(rate=5000.0, count=12, seed=3, max_log_blocks=3)
    alone = [srv.prove(r.rid, r.blocks, r.bits) for r in reqs]
    served = srv.serve(reqs)
    assert srv.peak_in_flight > 1  # the requests did overlap
    for a, r in zip(alone, served):
        assert r.result is not None and same(a, r.result)
        assert r.done >= r.admitted >= r.arrival - 1e-6


def test_admission_keeps_the_budget():
    srv = S.ProvingServer(hbm_budget=1 << 30, slots=8, seed=1)
    reqs = S.poisson_requests(rate=5000.0, count=10, seed=4, max_log_blocks=3)
    fps = [srv.circuit(r.blocks).footprint() for r in reqs]
    srv.hbm_budget = 2 * max(fps)  # at most two of the largest at once
    srv.serve(reqs)
    assert srv.peak_bytes <= srv.hbm_budget
    assert all(r.result is not None for r in reqs)
