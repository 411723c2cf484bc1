"""The allocation-free queued inner prover
(bitz_prove_inner_sumcheck_with_challenges_noalloc) against the allocating
one (bitz_prove_inner_sumcheck_with_challenges) on the same inputs, bit for
bit: two calls in a row, the second smaller (the static scratch keeps the
first call's contents, which must not leak into the second), a zero
challenge, num_vars 0 and 1, and the size limit."""
import numpy as np
import pytest

import bitz_amd as G
from common import elements, elements_fast, ints
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def both(n, seed, zero_at=None):
    d, w = elements_fast(seed, 1 << n), elements_fast(seed + 1, 1 << n)
    claim = O.inner_product(d, w)
    ch = elements(seed + 2, n) if n else np.zeros((0, 2), np.uint64)
    if zero_at is not None:
        ch[zero_at] = 0
    dd, dw = G.DeviceTable(d), G.DeviceTable(w)
    ref = G.prove_inner_sumcheck_with_challenges(ch, claim, dd, dw)
    out = G.prove_inner_sumcheck_with_challenges_noalloc(ch, claim, dd, dw)
    return ref, out, (d, w, ch, claim)


def assert_same(ref, out):
    assert np.array_equal(out.round_polynomials, ref.round_polynomials)
    assert np.array_equal(out.eval_points, ref.eval_points)
    assert out.final_claim == ref.final_claim
    assert (out.batched_matrix_evaluation, out.witness_evaluation) == (ref.batched_matrix_evaluation,
                                                                       ref.witness_evaluation)


def test_two_calls_in_a_row_the_second_smaller():
    for n, seed in ((20, 40), (12, 50)):
        print(f"noalloc inner n={n}", flush=True)
        ref, out, _ = both(n, seed)
        assert_same(ref, out)


def test_matches_the_oracle_with_a_zero_challenge():
    ref, out, (d, w, ch, claim) = both(14, 60, zero_at=3)
    assert_same(ref, out)
    rc, o = O.inner_sumcheck(O.ListTranscript(ints(ch)), claim, d, w)
    assert rc == 0
    assert np.array_equal(out.round_polynomials, o["rounds"]) and out.final_claim == o["final_claim"]


@pytest.mark.parametrize("n", [0, 1, 2, 11])
def test_small_tables(n):
    ref, out, _ = both(n, 70 + n)
    assert_same(ref, out)


def test_rejects_tables_beyond_the_static_scratch():
    t = G.DeviceTable(length=1 << 21)
    with pytest.raises(G.BitzError) as e:
        G.prove_inner_sumcheck_with_challenges_noalloc(elements(1, 21), 0, t, t)
    assert e.value.variant == "InvalidArgument" and "maximum of 20" in str(e.value)
    with pytest.raises(G.BitzError):
        G.prove_inner_sumcheck_with_challenges_noalloc(elements(1, 3), 0, elements(2, 8), elements(3, 8))
