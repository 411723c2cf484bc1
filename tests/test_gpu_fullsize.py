"""GPU parity at the workloads' full sizes: the HIP path against the oracle
(PCLMUL products, pinned to the bit-serial kernel by test_oracle_field) on
the same seeded inputs, bit for bit.

- the isolated outer sumcheck at 2^22 and 2^24 elements per table, queued
  (challenges fixed in advance) and transcript-driven (duplex transcript, a
  host round trip per round): every round polynomial, point and claim;
- the grand product over 2^22 leaves, full depth (2 trees) and the
  end-to-end shape (2^15 trees, 7 layers): every layer message, the leaf
  point and claim, queued, transcript-driven, and with a Keccak transcript
  whose challenges the device draws;
- the batch of 256 outer sumchecks with num_vars uniform in [12, 22]
  (seed 2024, the batched workload's instances), each instance against the
  oracle;
- a served 256-block compression-chain request in a mix, against the
  oracle's composition of the PIOP.
"""
import numpy as np
import pytest

import bitz_amd as G
from common import elements, elements_fast, ints
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, scope="module")
def _fast_oracle():
    O.set_kernel("pclmul")
    yield
    O.set_kernel("portable")


def say(msg: str) -> None:
    print(msg, flush=True)


def true_claim(eq, a, b, c):
    return O.inner_product(eq, O.mul_batch(a, b)) ^ O.inner_product(eq, c)


def assert_outer_equal(ref, out):
    assert np.array_equal(out.round_polynomials, ref["rounds"])
    assert np.array_equal(out.eval_points, ref["points"])
    assert out.final_claim == ref["final_claim"]
    assert [out.az_mle_claim, out.bz_mle_claim, out.cz_mle_claim] == ints(ref["claims"])


@pytest.mark.parametrize("n", [22, 24])
def test_outer_full_size_queued_and_transcript(n):
    say(f"outer full size n={n}: inputs")
    a, b, c = (elements_fast(2200 + 3 * n + k, 1 << n) for k in range(3))
    tau = elements(2300 + n, n)
    lo, hi = O.eq_table(tau[: n // 2]), O.eq_table(tau[n // 2:])
    claim = true_claim(O.eq_table(tau), a, b, c)
    ch = elements(2400 + n, n)
    ch_ints = ints(ch)
    rc, ref = O.outer_sumcheck(O.ListTranscript(ch_ints), claim, lo, hi, a, b, c)
    assert rc == 0
    dev = G.R1csProductMles(G.DeviceTable(a), G.DeviceTable(b), G.DeviceTable(c))
    eq = (G.DeviceTable(lo), G.DeviceTable(hi))
    assert G.outer_claim(eq, dev) == claim
    say(f"outer full size n={n}: queued")
    assert_outer_equal(ref, G.prove_outer_sumcheck_with_challenges(ch, claim, eq, dev))
    say(f"outer full size n={n}: transcript-driven")
    ot = O.DuplexTranscript(b"outer/full/%d" % n)
    rc, tref = O.outer_sumcheck(ot, claim, lo, hi, a, b, c)
    assert rc == 0
    gt = G.Transcript.duplex(b"outer/full/%d" % n)
    assert_outer_equal(tref, G.prove_outer_sumcheck(gt, claim, eq, dev))
    assert O.squeeze(ot) == gt.squeeze()


def gkr_case(log_leaves, log_groups, seed):
    leaves = elements_fast(seed, 1 << log_leaves)
    zero = (leaves[:, 0] == 0) & (leaves[:, 1] == 0)
    leaves[zero, 0] = 1  # nonzero leaves
    return leaves


@pytest.mark.parametrize("log_groups", [1, 15])
def test_gkr_full_size_queued_and_transcript(log_groups):
    L = 22
    say(f"gkr full size 2^{L} leaves, 2^{log_groups} trees")
    leaves = gkr_case(L, log_groups, 2500 + log_groups)
    store, layers = O.gkr_fold_layers(leaves, log_groups)
    out, wit = G.grand_product(G.DeviceTable(leaves), log_groups)
    assert np.array_equal(out, layers[-1])
    for j in (0, 1, len(layers) // 2, len(layers) - 1):
        assert np.array_equal(wit.layer(j), layers[j]), j
    point = elements(2600 + log_groups, log_groups)
    count = sum(n + 1 for n in range(log_groups, L))
    ch = elements(2700 + log_groups, count)
    ot = O.ListTranscript(ints(ch))
    ref_point, ref_claim = O.gkr_prove(ot, point, store, L)
    msgs, evals, gp, gc = G.gpgkr_prove_with_challenges(point, ch, wit)
    assert np.array_equal(gp, ref_point) and gc == ref_claim
    sent = [np.frombuffer(b, dtype=np.uint64).reshape(2, 2) for kind, b in ot.messages if kind == 1]
    expect_msgs, expect_evals, k = [], [], 0
    for n in range(log_groups, L):
        expect_msgs += sent[k: k + n]
        expect_evals.append(sent[k + n])
        k += n + 1
    assert k == len(sent)
    assert np.array_equal(msgs, np.array(expect_msgs).reshape(msgs.shape))
    assert np.array_equal(evals, np.array(expect_evals).reshape(evals.shape))
    say("gkr full size: transcript-driven")
    _, wit = G.grand_product(G.DeviceTable(leaves), log_groups)
    dt = O.DuplexTranscript(b"gkr/full")
    ref_point, ref_claim = O.gkr_prove(dt, point, store, L)
    gt = G.Transcript.duplex(b"gkr/full")
    gp, gc = G.gpgkr_prove(gt, point, wit)
    assert np.array_equal(gp, ref_point) and gc == ref_claim
    assert gt.narg() == dt.narg()
    say("gkr full size: Keccak transcript, challenges drawn on the device")
    _, wit = G.grand_product(G.DeviceTable(leaves), log_groups)
    kt = O.KeccakTranscript(b"gkr/full/keccak")
    ref_point, ref_claim = O.gkr_prove(kt, point, store, L)
    gt = G.Transcript.keccak(b"gkr/full/keccak")
    gp, gc = G.gpgkr_prove(gt, point, wit)
    assert np.array_equal(gp, ref_point) and gc == ref_claim
    assert gt.narg() == kt.narg()
    assert gt.squeeze() == O.squeeze(kt)


def test_batch_of_256_instances_at_full_size():
    """The batched workload's 256 instances (num_vars uniform in [12, 22],
    seed 2024): one batched call, each instance against the oracle on its own
    challenges and its true claim, which the device computes too."""
    import bench
    host_pools, spec = bench.batch_workload(2024, 256)
    say(f"batch of {len(spec)}: {sum(1 << n for n, *_ in spec)} elements per table")
    pools = [G.DeviceTable(p) for p in host_pools]

    class View(G.DeviceTable):  # a window of a device pool (owns nothing)
        def __init__(self, pool, off, n):
            self.ptr, self.len, self.pool = pool.ptr + 16 * off, n, pool

        def free(self):
            pass

    insts, cases = [], []
    for n, off, tau, ch in spec:
        a, b, c = (p[off: off + (1 << n)] for p in host_pools)
        lo, hi = O.eq_table(tau[: n // 2]), O.eq_table(tau[n // 2:])
        claim = true_claim(O.eq_table(tau), a, b, c)
        prods = G.R1csProductMles(*(View(p, off, 1 << n) for p in pools))
        eq = (G.DeviceTable(lo), G.DeviceTable(hi))
        assert G.outer_claim(eq, prods) == claim
        insts.append((ch, claim, eq, prods))
        cases.append((ch, claim, lo, hi, a, b, c))
    say("batch of 256: proving")
    outs = G.prove_outer_batch(insts)
    say("batch of 256: oracle per instance")
    for i, ((ch, claim, lo, hi, a, b, c), out) in enumerate(zip(cases, outs)):
        rc, ref = O.outer_sumcheck(O.ListTranscript(ints(ch)), claim, lo, hi, a, b, c)
        assert rc == 0
        assert_outer_equal(ref, out)
    G.outer_batch_release()


def test_served_256_block_request_matches_the_oracle():
    """A 256-block compression-chain request served with continuous batching
    beside smaller requests: the proof equals the oracle's composition of the
    PIOP (and the smaller ones their proofs alone)."""
    from bitz_amd import serving as S
    from test_gpu_serving import chain, same
    from test_gpu_sha256 import prove_and_check
    srv = S.ProvingServer(hbm_budget=8 << 30, slots=6)
    rng = np.random.default_rng(256)
    reqs = [S.Request(i, 0.0002 * i, b, *S.random_input(rng, b)) for i, b in enumerate([4, 256, 1, 16, 2])]
    say("serve: 256-block request in a mix of 5")
    served = srv.serve_batched(reqs)
    assert srv.peak_in_flight > 1
    for r in served:
        assert r.result is not None and r.digest == chain(r.iv, r.message)
    big = served[1]
    assert big.blocks == 256
    say("serve: oracle composition of the 256-block proof")
    _, ref = prove_and_check(srv.circuit(256).circuit, big.message, list(big.iv), label=b"bitz/serve/%d" % big.rid,
                             keccak=True)
    assert same(big.result, ref)
    for r in served:
        if r.blocks != 256:
            alone, _ = srv.prove(r.rid, r.blocks, r.message, r.iv)
            assert same(alone, r.result)
