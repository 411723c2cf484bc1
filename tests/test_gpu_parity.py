"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs, bit for bit.  Every test here needs an MI355X."""
import numpy as np
import pytest

import bitz_amd as G
from common import elements, elements_fast, from_ints, ints, to_int
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, scope="module")
def _fast_oracle():
    O.set_kernel("pclmul")  # same products as the bit-serial kernel, pinned by test_oracle_field
    yield
    O.set_kernel("portable")


EDGES = [0, 1, 2, 0x87, 1 << 63, 1 << 64, 1 << 127, 2**128 - 1, (1 << 128) - 2]


def test_mul_batch_matches_oracle():
    a = [x for x in EDGES for _ in EDGES] + ints(elements(1, 4096))
    b = [y for _ in EDGES for y in EDGES] + ints(elements(2, 4096))
    A, B = from_ints(a), from_ints(b)
    assert np.array_equal(G.mul_batch(A, B), O.mul_batch(A, B))
    A, B = elements_fast(3, 1 << 20), elements_fast(4, 1 << 20)
    assert np.array_equal(G.mul_batch(A, B), O.mul_batch(A, B))


@pytest.mark.parametrize("mode", [G.MUL_LANE, G.MUL_LANE2, G.MUL_ROWS])
def test_lane_table_products_match_oracle(mode):
    """The per-lane LDS table product (3-bit windows, a's top two bits added
    separately), its two-operand form and the four-per-wave row product on
    every pair of edge values (zero,
    one, the reduction constant, the top bits) and on random elements, odd
    length included."""
    edges = EDGES + [1 << 126, 3 << 126, (1 << 126) - 1]
    a = [x for x in edges for _ in edges] + ints(elements(11, 4097))
    b = [y for _ in edges for y in edges] + ints(elements(12, 4097))
    A, B = from_ints(a), from_ints(b)
    assert np.array_equal(G.mul_batch(A, B, mode), O.mul_batch(A, B))


@pytest.mark.parametrize("mode", [G.MUL_CONST, G.MUL_MFMA])
@pytest.mark.parametrize("c", EDGES + [0xDEADBEEF << 70 | 12345])
def test_constant_table_product_matches_oracle(c, mode):
    """a[i] * c through c's 4-bit constant table (the perm-addressed lookups of
    the fold and weight products), and as c's F2 matrix on the int8 matrix
    cores; 3009 elements leave a partial group of 16."""
    a = from_ints(EDGES + ints(elements(13, 3000)))
    b = from_ints([c] * len(a))
    assert np.array_equal(G.mul_batch(a, b, mode), O.mul_batch(a, b))


@pytest.mark.parametrize("k", [0, 1, 5, 12, 13, 20])
def test_eq_table_matches_oracle(k):
    r = elements(10 + k, k) if k else np.zeros((0, 2), np.uint64)
    assert np.array_equal(G.eq_table(r), O.eq_table(r))


def test_mle_fold_and_evaluate_match_oracle():
    t = elements_fast(5, 1 << 14)
    pt = elements(6, 14)
    cur = t
    for i in range(5):
        cur = O.mle_fold(cur, pt[i])
    assert np.array_equal(G.mle_fold(t, pt[:5]), cur)
    assert G.mle_evaluate(t, pt) == O.mle_evaluate(t, pt)
    with pytest.raises(G.BitzError) as e:
        G.mle_fold(t[:4], pt[:3])
    assert e.value.variant == "TooManyChallenges"


@pytest.mark.parametrize("n", [0, 1, 2, 3, 13, 21])
def test_mle_evaluate_on_device_tables_matches_oracle(n):
    """Two variables a pass through the challenges' constant tables, one for
    an odd last variable, the device table read in place (sizes 1 .. 2^21)."""
    t = elements_fast(40 + n, 1 << n)
    pt = elements_fast(60 + n, n)
    d = G.DeviceTable(t)
    assert G.mle_evaluate_device(d, pt) == O.mle_evaluate(t, pt)
    assert np.array_equal(d.to_host(), t)  # the input is not written
    if n >= 3:
        ref = t
        for i in range(3):
            ref = O.mle_fold(ref, pt[i])
        assert np.array_equal(G.mle_fold(t, pt[:3]), ref)


def outer_case(n, split, seed, satisfied=False):
    a = elements_fast(seed, 1 << n)
    b = elements_fast(seed + 1, 1 << n)
    c = O.mul_batch(a, b) if satisfied else elements_fast(seed + 2, 1 << n)
    tau = elements(seed + 3, n) if n else np.zeros((0, 2), np.uint64)
    lo = O.eq_table(tau[:split]) if split else from_ints([1])
    hi = O.eq_table(tau[split:]) if split < n else from_ints([1])
    return a, b, c, tau, lo, hi


def run_outer_both(n, split, seed, claim=0, satisfied=False, domain=b"outer/gpu"):
    a, b, c, tau, lo, hi = outer_case(n, split, seed, satisfied)
    ot = O.DuplexTranscript(domain)
    rc, ref = O.outer_sumcheck(ot, claim, lo, hi, a, b, c)
    assert rc == 0
    gt = G.Transcript.duplex(domain)
    out = G.prove_outer_sumcheck(gt, claim, (lo, hi), G.R1csProductMles(a, b, c))
    return ref, out, (a, b, c, tau, lo, hi), (ot, gt)


def assert_outer_equal(ref, out):
    assert np.array_equal(out.round_polynomials, ref["rounds"])
    assert np.array_equal(out.eval_points, ref["points"])
    assert out.final_claim == ref["final_claim"]
    assert [out.az_mle_claim, out.bz_mle_claim, out.cz_mle_claim] == ints(ref["claims"])


@pytest.mark.parametrize("n", [0, 1, 2, 3, 6, 11, 16])
def test_outer_sumcheck_matches_oracle(n):
    ref, out, _, tr = run_outer_both(n, n // 2, 100 + n, claim=0x1234 + n)
    assert_outer_equal(ref, out)
    # the transcripts saw the same bytes: their next challenges agree
    assert O.squeeze(tr[0]) == tr[1].squeeze()


@pytest.mark.parametrize("split", range(6))
def test_outer_every_split(split):
    ref, out, *_ = run_outer_both(5, split, 7, satisfied=True)
    assert_outer_equal(ref, out)


def test_outer_handle_tables_match_oracle_folds():
    n = 9
    a, b, c, tau, lo, hi = outer_case(n, 4, 55)
    h = G.OuterSumcheck((lo, hi), G.R1csProductMles(a, b, c))
    h.round0()
    rs = elements(56, n)
    ta, tb, tc = a, b, c
    for k in range(n):
        h.fold_next(rs[k])
        ta, tb, tc = O.mle_fold(ta, rs[k]), O.mle_fold(tb, rs[k]), O.mle_fold(tc, rs[k])
        ga, gb, gc = h.tables()
        assert np.array_equal(ga, ta) and np.array_equal(gb, tb) and np.array_equal(gc, tc), k
    assert np.array_equal(h.finish(), np.vstack([ta, tb, tc]))


def test_outer_dimension_errors_leave_transcript_untouched():
    a, b, c, tau, lo, hi = outer_case(3, 1, 9)
    t = G.Transcript.duplex(b"dims")
    with pytest.raises(G.SumcheckError) as e:
        G.prove_outer_sumcheck(t, 0, (lo, hi), G.R1csProductMles(a, b[:4], c))
    assert e.value.variant == "InvalidProductDimensions"
    with pytest.raises(G.SumcheckError) as e:
        G.prove_outer_sumcheck(t, 0, (lo, lo), G.R1csProductMles(a, b, c))
    assert e.value.variant == "InvalidEqualityDimensions"
    assert t.squeeze() == G.Transcript.duplex(b"dims").squeeze()


def true_claim(a, b, c, tau):
    eq = O.eq_table(tau)
    return O.inner_product(eq, O.mul_batch(a, b)) ^ O.inner_product(eq, c)


def test_outer_n20_matches_oracle_and_verifies():
    a, b, c, tau, lo, hi = outer_case(20, 10, 2024)
    claim = true_claim(a, b, c, tau)
    ref, out, *_ = run_outer_both(20, 10, 2024, claim=claim)
    assert_outer_equal(ref, out)
    vt = O.DuplexTranscript(b"outer/gpu")
    claims = np.array([[out.az_mle_claim & (2**64 - 1), out.az_mle_claim >> 64],
                       [out.bz_mle_claim & (2**64 - 1), out.bz_mle_claim >> 64],
                       [out.cz_mle_claim & (2**64 - 1), out.cz_mle_claim >> 64]], dtype=np.uint64)
    assert O.outer_verify(vt, claim, tau, out.round_polynomials, claims) == 0


@pytest.mark.parametrize("n", [0, 1, 3, 12, 13, 18])
def test_inner_sumcheck_matches_oracle(n):
    d, w = elements_fast(300 + n, 1 << n), elements_fast(400 + n, 1 << n)
    claim = O.inner_product(d, w)
    ot = O.DuplexTranscript(b"inner/gpu")
    rc, ref = O.inner_sumcheck(ot, claim, d, w)
    assert rc == 0
    gt = G.Transcript.duplex(b"inner/gpu")
    out = G.prove_inner_sumcheck(gt, claim, d, w)
    assert np.array_equal(out.round_polynomials, ref["rounds"])
    assert np.array_equal(out.eval_points, ref["points"])
    assert out.final_claim == ref["final_claim"]
    assert [out.batched_matrix_evaluation, out.witness_evaluation] == ints(ref["evals"])
    assert O.squeeze(ot) == gt.squeeze()


@pytest.mark.parametrize("n", [0, 1, 2, 9, 10, 11, 12, 16, 20])
def test_inner_sumcheck_with_challenges_matches_oracle(n):
    """The queued inner prover (grid-wide rounds, then the LDS tail) against
    the oracle on the same challenge list, host and device tables."""
    d, w = elements_fast(500 + n, 1 << n), elements_fast(600 + n, 1 << n)
    claim = O.inner_product(d, w)
    ch = elements(700 + n, n) if n else np.zeros((0, 2), np.uint64)
    if n > 4:
        ch[2] = 0  # a zero challenge folds to the low half
    rc, ref = O.inner_sumcheck(O.ListTranscript(ints(ch)), claim, d, w)
    assert rc == 0
    for dev in (False, True):
        args = (G.DeviceTable(d), G.DeviceTable(w)) if dev else (d, w)
        out = G.prove_inner_sumcheck_with_challenges(ch, claim, *args)
        assert np.array_equal(out.round_polynomials, ref["rounds"])
        assert np.array_equal(out.eval_points, ref["points"])
        assert out.final_claim == ref["final_claim"]
        assert [out.batched_matrix_evaluation, out.witness_evaluation] == ints(ref["evals"])


def test_inner_dimension_error():
    with pytest.raises(G.SumcheckError) as e:
        G.prove_inner_sumcheck(G.Transcript.seeded(1), 0, from_ints([1, 2]), from_ints([1, 2, 3, 4]))
    assert e.value.variant == "InvalidProductDimensions"


def random_csr(rng, rows, cols, per_row, seed):
    starts, colsl = [0], []
    for _ in range(rows):
        cs = np.unique(rng.integers(0, cols, size=per_row))
        colsl += cs.tolist()
        starts.append(len(colsl))
    return np.array(starts, np.uint64), np.array(colsl, np.uint32), elements_fast(seed, len(colsl))


@pytest.mark.parametrize("rows,row_vars,col_vars,per_row", [(13, 4, 5, 6), (700, 10, 17, 24), (1, 0, 3, 2)])
def test_bind_and_batch_matches_oracle(rows, row_vars, col_vars, per_row):
    rng = np.random.default_rng(rows)
    csr = [random_csr(rng, rows, 1 << col_vars, per_row, 500 + m) for m in range(3)]
    rx = elements(7, row_vars) if row_vars else np.zeros((0, 2), np.uint64)
    rho = 0xDEADBEEF0123456789ABCDEF55AA55AA
    prepared = G.PreparedConstraintMatrices(rows, row_vars, col_vars, csr)
    d = prepared.bind_and_batch(rx, rho)
    mats = [O.CsrMatrix(*m) for m in csr]
    assert np.array_equal(d, O.bind_and_batch(mats, rows, row_vars, col_vars, rx, rho))
    with pytest.raises(G.BitzError) as e:
        prepared.bind_and_batch(elements(8, row_vars + 1), rho)
    assert e.value.variant == "InvalidRowPointLength"


@pytest.mark.parametrize("rows,row_vars,col_vars", [(5000, 13, 10), (300, 9, 8)])
def test_bind_and_products_with_long_columns_and_zero_values(rows, row_vars, col_vars):
    """Every row holds columns 0 and 5 (long columns, summed in runs by
    workgroups) and a third of the values are zero (entries the device
    drops); D and A h, B h, C h against the oracle and the definition."""
    rng = np.random.default_rng(rows + 1)
    csr = []
    for m in range(3):
        starts, cl = [0], []
        for _ in range(rows):
            cs = np.unique(np.concatenate([[0, 5], rng.integers(0, 1 << col_vars, size=4)]))
            cl += cs.tolist()
            starts.append(len(cl))
        vals = elements_fast(900 + m, len(cl))
        vals[rng.integers(0, 3, size=len(cl)) == 0] = 0
        csr.append((np.array(starts, np.uint64), np.array(cl, np.uint32), vals))
    rx = elements(17, row_vars)
    rho = 0x0123456789ABCDEFFEDCBA9876543210
    prepared = G.PreparedConstraintMatrices(rows, row_vars, col_vars, csr)
    mats = [O.CsrMatrix(*m) for m in csr]
    assert np.array_equal(prepared.bind_and_batch(rx, rho), O.bind_and_batch(mats, rows, row_vars, col_vars, rx, rho))
    h = elements_fast(31, 1 << col_vars)
    out = prepared.r1cs_products(h)
    for m, t in enumerate((out.az, out.bz, out.cz)):
        st, co, va = csr[m]
        ref = np.zeros((1 << row_vars, 2), np.uint64)
        terms = O.mul_batch(va, h[co.astype(np.int64)])
        np.bitwise_xor.at(ref, np.repeat(np.arange(rows), np.diff(st.astype(np.int64))), terms)
        assert np.array_equal(t, ref)


def test_mt_apply_and_transpose_inner_match_oracle():
    rng = np.random.default_rng(3)
    rows, cols = 5000, 3000
    offsets, idx = [0], []
    for _ in range(cols):
        idx += sorted(set(rng.integers(0, rows, size=int(rng.integers(0, 9))).tolist()))
        offsets.append(len(idx))
    r = elements_fast(9, rows)
    mt = G.MaterializedMTranspose(rows, offsets, idx)
    assert np.array_equal(mt.apply(r), O.mt_apply(offsets, idx, r))
    rw, cw = elements_fast(10, 64), elements_fast(11, 128)
    assert np.array_equal(mt.transpose_inner(rw, cw), O.mt_transpose_inner(offsets, idx, rw, cw))


def test_mt_long_columns_match_oracle():
    """Columns past the one-thread limit (64 entries) go to a workgroup each:
    lengths at the limit, one over it, the 2 KiB SHA-256 circuit's constant
    column (300) and a column longer than a workgroup's stride, beside short
    and empty ones."""
    rng = np.random.default_rng(5)
    rows = 20000
    lens = [64, 65, 300, 1000, 0, 3, 257, 1, 9, 4096] + [int(x) for x in rng.integers(0, 10, size=500)]
    offsets, idx = [0], []
    for n in lens:
        idx += sorted(rng.choice(rows, size=n, replace=False).tolist())
        offsets.append(len(idx))
    mt = G.MaterializedMTranspose(rows, offsets, idx)
    r = elements_fast(21, rows)
    assert np.array_equal(mt.apply(r), O.mt_apply(offsets, idx, r))
    rw, cw = elements_fast(22, 128), elements_fast(23, 256)
    assert np.array_equal(mt.transpose_inner(rw, cw), O.mt_transpose_inner(offsets, idx, rw, cw))


def test_mt_on_the_sha256_2kib_circuit_matches_oracle():
    """M^T of the SHA-256 2 KiB block-aligned circuit (the reference's matrix
    benchmark, crates/circuit/benches/sha256_matrix_products.rs): 675049 rows,
    226793 columns, its constant column 300 entries long."""
    from bitz_amd import sha256 as S
    c = S.Sha256Circuit(S.BLOCK_ALIGNED, 32)
    offsets, idx = c.m_transpose()
    assert c.columns == 675049 and len(offsets) - 1 == 226793 and offsets[1] - offsets[0] == 300
    mt = G.MaterializedMTranspose(c.columns, offsets, idx)
    r = elements_fast(24, c.columns)
    assert np.array_equal(mt.apply(r), O.mt_apply(offsets, idx, r))
    rw, cw = elements_fast(25, 1 << 10), elements_fast(26, 1 << 10)
    assert np.array_equal(mt.transpose_inner(rw, cw), O.mt_transpose_inner(offsets, idx, rw, cw))


@pytest.mark.parametrize("log_leaves,log_groups,zero", [(6, 2, None), (10, 3, 0), (14, 7, 6), (16, 1, None)])
def test_gkr_matches_oracle(log_leaves, log_groups, zero):
    leaves = elements_fast(70 + log_leaves, 1 << log_leaves)
    store, layers = O.gkr_fold_layers(leaves, log_groups)
    out, wit = G.grand_product(leaves, log_groups)
    assert np.array_equal(out, layers[-1])
    for j in (0, 1, len(layers) - 1):
        assert np.array_equal(wit.layer(j), layers[j])
    point = elements(80, log_groups)
    if zero is not None:
        point[zero] = 0
    ot = O.DuplexTranscript(b"gkr/gpu")
    ref_point, ref_claim = O.gkr_prove(ot, point, store, log_leaves)
    gt = G.Transcript.duplex(b"gkr/gpu")
    gp, gc = G.gpgkr_prove(gt, point, wit)
    assert np.array_equal(gp, ref_point) and gc == ref_claim
    assert gt.narg() == ot.narg()
    assert gc == O.mle_evaluate(leaves, gp)


@pytest.mark.parametrize("log_leaves,log_groups", [(12, 0), (13, 0), (16, 1), (20, 3), (9, 9)])
def test_gkr_layers_from_device_leaves(log_leaves, log_groups):
    """Every layer of the product tree, with the leaves already in device
    memory (the first layer also copies them into the store) and across the
    switch from grid-wide layers to the one-workgroup small tree."""
    leaves = elements_fast(60 + log_leaves, 1 << log_leaves)
    _, layers = O.gkr_fold_layers(leaves, log_groups)
    out, wit = G.grand_product(G.DeviceTable(leaves), log_groups)
    assert np.array_equal(out, layers[-1])
    for j in range(len(layers)):
        assert np.array_equal(wit.layer(j), layers[j]), j


@pytest.mark.parametrize("log_leaves,log_groups,zero", [(6, 2, None), (12, 1, 4), (15, 7, None), (18, 1, None)])
def test_gkr_with_challenges_matches_oracle(log_leaves, log_groups, zero):
    """The queued GKR prover against the oracle driven by the same challenge
    list: every prover_message, every [L(r), R(r)], the leaf point and claim."""
    leaves = elements_fast(90 + log_leaves, 1 << log_leaves)
    store, layers = O.gkr_fold_layers(leaves, log_groups)
    point = elements(91, log_groups)
    count = sum(n + 1 for n in range(log_groups, log_leaves))
    ch = elements(92 + log_leaves, count)
    if zero is not None:
        point[0] = 0
        ch[zero] = 0  # a zero challenge makes a later round send its at-one sum
    ot = O.ListTranscript(ints(ch))
    ref_point, ref_claim = O.gkr_prove(ot, point, store, log_leaves)
    _, wit = G.grand_product(leaves, log_groups)
    msgs, evals, gp, gc = G.gpgkr_prove_with_challenges(point, ch, wit)
    assert np.array_equal(gp, ref_point) and gc == ref_claim
    sent = [np.frombuffer(b, dtype=np.uint64).reshape(2, 2) for kind, b in ot.messages if kind == 1]
    expect_msgs, expect_evals, k = [], [], 0
    for n in range(log_groups, log_leaves):
        expect_msgs += sent[k: k + n]
        expect_evals.append(sent[k + n])
        k += n + 1
    assert k == len(sent)
    assert np.array_equal(msgs, np.array(expect_msgs).reshape(msgs.shape))
    assert np.array_equal(evals, np.array(expect_evals).reshape(evals.shape))
    assert gc == O.mle_evaluate(leaves, gp)


def bittable_instance(t, s, bits, seed):
    """The reference's test instance (crates/gkr/src/lib.rs:1002-1014): packed
    words (u128, as lo/hi u64 pairs) then the row images, from splitmix(seed)."""
    n_words = (1 << (t + s)) // 128
    els = elements(seed, n_words + (1 << t))
    packed = els[:n_words].copy()
    if bits == "clear":
        packed[:] = 0
    elif bits == "set":
        packed[:] = np.uint64(2**64 - 1)
    return packed.reshape(-1), els[n_words:]


def bittable_leaves(packed, images, t, s):
    """leaves() of the reference's tests (lib.rs:1023-1037): leaf (b << s) | c
    is images[b] where bit (c, b) is set, one elsewhere."""
    bits = np.unpackbits(packed.view(np.uint8), bitorder="little")
    leaves = np.zeros((1 << (t + s), 2), np.uint64)
    leaves[:, 0] = 1
    for b in range(1 << t):
        cols = np.arange(1 << s)
        on = bits[(cols << t) | b].astype(bool)
        leaves[(b << s) | cols[on]] = images[b]
    return leaves


@pytest.mark.parametrize("t,s,bits", [(7, 4, "drawn"), (9, 6, "drawn"), (8, 5, "clear"), (7, 5, "set"),
                                      (11, 6, "drawn")])
def test_gkr_bittable_matches_the_dense_tree(t, s, bits):
    packed, images = bittable_instance(t, s, bits, 4 + t)
    leaves = bittable_leaves(packed, images, t, s)
    store, layers = O.gkr_fold_layers(leaves, s)
    out, wit = G.grand_product_bittable(images, packed, t, s)
    assert np.array_equal(out, layers[-1])
    for j in (0, 1, 2, 3, len(layers) - 1):
        assert np.array_equal(wit.layer(j), layers[j])
    point = elements(60 + t, s)
    ot = O.DuplexTranscript(b"gkr/bittable")
    ref_point, ref_claim = O.gkr_prove(ot, point, store, t + s)
    gt = G.Transcript.duplex(b"gkr/bittable")
    gp, gc = G.gpgkr_prove(gt, point, wit)
    assert np.array_equal(gp, ref_point) and gc == ref_claim and gt.narg() == ot.narg()


def test_gkr_bittable_needs_sixteen_trees():
    packed, images = bittable_instance(7, 3, "drawn", 1)
    with pytest.raises(G.BitzError):
        G.grand_product_bittable(images, packed, 7, 3)


def test_gkr_rejects_a_wrong_point_length():
    leaves = elements_fast(1, 1 << 6)
    _, wit = G.grand_product(leaves, 2)
    with pytest.raises(G.BitzError) as e:
        G.gpgkr_prove(G.Transcript.seeded(1), elements(2, 3), wit)
    assert e.value.variant == "InvalidPointLength"


# ---- the two outer-sumcheck modes ---------------------------------------------
def outer_with(a, b, c, lo, hi, claim, challenges):
    ot = O.ListTranscript(challenges)
    rc, ref = O.outer_sumcheck(ot, claim, lo, hi, a, b, c)
    assert rc == 0
    it = iter(int(x) for x in challenges)
    gt = G.Transcript.callbacks(lambda kind, data: None, lambda: next(it).to_bytes(16, "little"))
    out = G.prove_outer_sumcheck(gt, claim, (lo, hi), G.R1csProductMles(a, b, c))
    assert_outer_equal(ref, out)
    return out


def mode_of(lo, hi, a, b, c):
    return G.OuterSumcheck((lo, hi), G.R1csProductMles(a, b, c)).eq_mode


def test_outer_eq_mode_is_used_for_eq_factors():
    a, b, c, tau, lo, hi = outer_case(12, 6, 91)
    assert mode_of(lo, hi, a, b, c)
    ch = ints(elements(92, 12))
    outer_with(a, b, c, lo, hi, 5, ch)


@pytest.mark.parametrize("zero_at", [0, 1, 5, 11])
def test_outer_eq_mode_with_zero_tau(zero_at):
    n = 12
    a, b, c = elements_fast(93, 1 << n), elements_fast(94, 1 << n), elements_fast(95, 1 << n)
    tau = elements(96, n)
    tau[zero_at] = 0
    lo, hi = O.eq_table(tau[:6]), O.eq_table(tau[6:])
    assert mode_of(lo, hi, a, b, c)
    outer_with(a, b, c, lo, hi, 0, ints(elements(97, n)))


def test_outer_generic_mode_for_non_eq_factors_and_tau_one():
    n = 10
    a, b, c = elements_fast(98, 1 << n), elements_fast(99, 1 << n), elements_fast(100, 1 << n)
    lo, hi = elements_fast(101, 1 << 5), elements_fast(102, 1 << 5)  # arbitrary tables
    assert not mode_of(lo, hi, a, b, c)
    outer_with(a, b, c, lo, hi, 3, ints(elements(103, n)))
    tau = elements(104, n)
    tau[2] = (1, 0)  # the element 1: eq_table(tau) has T[0] = 0, tau cannot be recovered
    lo, hi = O.eq_table(tau[:5]), O.eq_table(tau[5:])
    assert not mode_of(lo, hi, a, b, c)
    outer_with(a, b, c, lo, hi, 0, ints(elements(105, n)))


def test_outer_eq_mode_when_eq_bound_vanishes():
    """r_k = 1 + tau_k makes eq(tau, r) zero from round k on."""
    n = 12
    a, b, c, tau, lo, hi = outer_case(n, 6, 106)
    ch = ints(elements(107, n))
    ch[3] = to_int(tau[3]) ^ 1
    out = outer_with(a, b, c, lo, hi, 9, ch)
    assert all(to_int(x) == 0 for x in out.round_polynomials[5][[0, 2, 3]])


@pytest.mark.parametrize("n,split,eq", [(1, 0, True), (9, 4, True), (14, 7, True), (10, 5, False)])
def test_outer_with_challenges_matches_oracle(n, split, eq):
    a, b, c, tau, lo, hi = outer_case(n, split, 120 + n)
    if not eq:
        lo, hi = elements_fast(121, len(lo)), elements_fast(122, len(hi))
    ch = elements(123, n)
    ot = O.ListTranscript(ints(ch))
    rc, ref = O.outer_sumcheck(ot, 77, lo, hi, a, b, c)
    assert rc == 0
    out = G.prove_outer_sumcheck_with_challenges(ch, 77, (lo, hi), G.R1csProductMles(a, b, c))
    assert_outer_equal(ref, out)
    dev = G.R1csProductMles(G.DeviceTable(a), G.DeviceTable(b), G.DeviceTable(c))
    assert_outer_equal(ref, G.prove_outer_sumcheck_with_challenges(ch, 77, (lo, hi), dev))


# ---- pipelined prover: the optimistic pass and its second pass ----------------
def pipelined(ch, claim, lo, hi, a, b, c):
    """The pipelined prover and how many times its tail kernel ran (1 = the
    optimistic pass was accepted, 2 = a second pass was needed)."""
    G.profile_enable(True)
    try:
        out = G.prove_outer_sumcheck_with_challenges(ch, claim, (lo, hi), G.R1csProductMles(a, b, c))
        passes = G.profile_read("outer_tail")["launches"]
    finally:
        G.profile_enable(False)
    return out, passes


def oracle_with(ch, claim, lo, hi, a, b, c):
    rc, ref = O.outer_sumcheck(O.ListTranscript(ints(ch)), claim, lo, hi, a, b, c)
    assert rc == 0
    return ref


@pytest.mark.parametrize("n,split", [(1, 0), (1, 1), (2, 1), (6, 0), (6, 6), (10, 3), (11, 5), (13, 0),
                                     (13, 6), (13, 13), (16, 9), (20, 10)])
def test_pipelined_true_claim_takes_one_pass(n, split):
    a, b, c, tau, lo, hi = outer_case(n, split, 300 + 7 * n + split)
    claim = true_claim(a, b, c, tau)
    ch = elements(301 + n, n)
    out, passes = pipelined(ch, claim, lo, hi, a, b, c)
    assert_outer_equal(oracle_with(ch, claim, lo, hi, a, b, c), out)
    assert passes == 1


@pytest.mark.parametrize("n", [3, 12, 15])
def test_pipelined_wrong_claim_is_caught(n):
    """Round 0 recovers q(1) from the claim; a wrong claim fails the final
    check and the second pass computes q(1) from the tables."""
    a, b, c, tau, lo, hi = outer_case(n, n // 2, 400 + n)
    claim = true_claim(a, b, c, tau) ^ 0x5A5A
    ch = elements(401 + n, n)
    out, passes = pipelined(ch, claim, lo, hi, a, b, c)
    assert_outer_equal(oracle_with(ch, claim, lo, hi, a, b, c), out)
    assert passes == 2


@pytest.mark.parametrize("zero_at", [0, 2, 11, 13])
def test_pipelined_zero_tau(zero_at):
    n = 14
    a, b, c = elements_fast(410, 1 << n), elements_fast(411, 1 << n), elements_fast(412, 1 << n)
    tau = elements(413, n)
    tau[zero_at] = 0
    lo, hi = O.eq_table(tau[:7]), O.eq_table(tau[7:])
    claim = true_claim(a, b, c, tau)
    ch = elements(414, n)
    out, passes = pipelined(ch, claim, lo, hi, a, b, c)
    assert_outer_equal(oracle_with(ch, claim, lo, hi, a, b, c), out)
    assert passes == 2


@pytest.mark.parametrize("zero_r", [0, 5, 13])
def test_pipelined_zero_challenge_computes_q1_from_the_start(zero_r):
    n = 14
    a, b, c, tau, lo, hi = outer_case(n, 7, 420)
    claim = true_claim(a, b, c, tau) ^ 1  # wrong, and undetectable through r = 0
    ch = elements(421, n)
    ch[zero_r] = 0
    out, passes = pipelined(ch, claim, lo, hi, a, b, c)
    assert_outer_equal(oracle_with(ch, claim, lo, hi, a, b, c), out)
    assert passes == 1


def test_pipelined_non_eq_factors_use_the_general_kernels():
    n = 12
    a, b, c = elements_fast(430, 1 << n), elements_fast(431, 1 << n), elements_fast(432, 1 << n)
    tau = elements(433, n)
    lo, hi = O.eq_table(tau[:6]), O.eq_table(tau[6:])
    bad = [lo.copy(), hi.copy()]
    bad[0][5, 0] ^= 1                                  # one entry off
    scaled = O.mul_batch(hi, np.tile(from_ints([3]), (len(hi), 1)))  # c * eq: products agree, sum != 1
    for lo2, hi2 in [(bad[0], hi), (lo, scaled)]:
        ch = elements(434, n)
        out, passes = pipelined(ch, 0, lo2, hi2, a, b, c)
        assert_outer_equal(oracle_with(ch, 0, lo2, hi2, a, b, c), out)
        assert passes == 1
        assert not mode_of(lo2, hi2, a, b, c)


@pytest.mark.parametrize("split", [0, 3, 9, 13])
def test_per_round_eq_mode_every_split(split):
    n = 13
    a, b, c, tau, lo, hi = outer_case(n, split, 440 + split)
    assert mode_of(lo, hi, a, b, c)
    outer_with(a, b, c, lo, hi, 11, ints(elements(441, n)))


# ---- batched prover -----------------------------------------------------------
def test_outer_batch_matches_oracle_per_instance():
    """Uneven instances in one batch, each equal to the oracle on its own
    challenges: true and wrong claims, a zero challenge, a zero tau, factors
    that are not eq tables, every split."""
    cases = []
    specs = [(1, 0, True), (2, 1, True), (5, 2, False), (10, 5, True), (11, 0, True), (12, 6, True),
             (13, 13, False), (14, 7, True), (16, 9, True), (20, 10, True), (21, 10, True), (12, 6, "zero_r"),
             (12, 6, "zero_tau"), (12, 6, "not_eq")]
    for j, (n, split, kind) in enumerate(specs):
        a, b, c, tau, lo, hi = outer_case(n, split, 500 + 11 * j)
        ch = elements(501 + j, n)
        claim = true_claim(a, b, c, tau)
        if kind is False:
            claim ^= 0xABC
        elif kind == "zero_r":
            ch[3] = 0
            claim ^= 7
        elif kind == "zero_tau":
            tau[8] = 0
            lo, hi = O.eq_table(tau[:split]), O.eq_table(tau[split:])
            claim = true_claim(a, b, c, tau)
        elif kind == "not_eq":
            hi = hi.copy()
            hi[3, 1] ^= 1
        cases.append((ch, claim, lo, hi, a, b, c))
    insts = [(ch, claim, (G.DeviceTable(lo), G.DeviceTable(hi)),
              G.R1csProductMles(G.DeviceTable(a), G.DeviceTable(b), G.DeviceTable(c)))
             for ch, claim, lo, hi, a, b, c in cases]
    outs = G.prove_outer_batch(insts)
    for (ch, claim, lo, hi, a, b, c), out in zip(cases, outs):
        assert_outer_equal(oracle_with(ch, claim, lo, hi, a, b, c), out)


def test_outer_batch_of_many_small_instances():
    rng = np.random.default_rng(7)
    cases = []
    for j in range(48):
        n = int(rng.integers(1, 15))
        a, b, c, tau, lo, hi = outer_case(n, n // 2, 900 + 13 * j)
        cases.append((elements(901 + j, n), true_claim(a, b, c, tau), lo, hi, a, b, c))
    insts = [(ch, claim, (G.DeviceTable(lo), G.DeviceTable(hi)),
              G.R1csProductMles(G.DeviceTable(a), G.DeviceTable(b), G.DeviceTable(c)))
             for ch, claim, lo, hi, a, b, c in cases]
    outs = G.prove_outer_batch(insts)
    for (ch, claim, lo, hi, a, b, c), out in zip(cases, outs):
        assert_outer_equal(oracle_with(ch, claim, lo, hi, a, b, c), out)


def test_outer_batch_marshalled_once_proves_again():
    """An OuterBatch's records reused: every call reproves the batch and
    fills the same outputs (checked against the oracle after each)."""
    cases = batch_cases([(5, 2), (11, 5), (13, 13)], 1500)
    b = G.OuterBatch(batch_insts(cases))
    for _ in range(2):
        b.prove()
        for (ch, claim, lo, hi, x, y, z), out in zip(cases, b.outputs()):
            assert_outer_equal(oracle_with(ch, claim, lo, hi, x, y, z), out)


def batch_cases(specs, seed0):
    cases = []
    for j, (n, split) in enumerate(specs):
        a, b, c, tau, lo, hi = outer_case(n, split, seed0 + 17 * j)
        cases.append((elements(seed0 + 1 + j, n), true_claim(a, b, c, tau), lo, hi, a, b, c))
    return cases


def batch_insts(cases):
    return [(ch, claim, (G.DeviceTable(lo), G.DeviceTable(hi)),
             G.R1csProductMles(G.DeviceTable(a), G.DeviceTable(b), G.DeviceTable(c)))
            for ch, claim, lo, hi, a, b, c in cases]


@pytest.mark.parametrize("thread_bits,tail_log", [(17, 10), (14, 10), (12, 9), (8, 4)])
def test_outer_batch_every_tile_shape(thread_bits, tail_log):
    """Fixed tile shapes (multi-iteration grid-wide rounds with H(i) tables
    from thread_bits below num_vars - 1, short tails) against the oracle."""
    specs = [(3, 1), (9, 4), (12, 6), (15, 7), (16, 16), (18, 9)]
    cases = batch_cases(specs, 700 + thread_bits)
    G.outer_batch_set_config(thread_bits, tail_log)
    try:
        for graph in (True, False):
            outs = G.prove_outer_batch(batch_insts(cases), graph=graph)
            assert G.outer_batch_config() == (thread_bits, tail_log)
            for (ch, claim, lo, hi, a, b, c), out in zip(cases, outs):
                assert_outer_equal(oracle_with(ch, claim, lo, hi, a, b, c), out)
    finally:
        G.outer_batch_set_config(0, 0)


def test_outer_batch_graph_replays_track_new_data():
    """The captured graph is replayed for batches of the same shapes with new
    tables, challenges and claims, with eager batches and other device work
    between the replays; every replay equals the eager result and the oracle."""
    specs = [(4, 2), (11, 5), (13, 6), (16, 8)]
    eager_ref = None
    for rep in range(3):
        cases = batch_cases(specs, 800 + 100 * rep)
        outs = G.prove_outer_batch(batch_insts(cases), graph=True)
        eager = G.prove_outer_batch(batch_insts(cases), graph=False, tune=False)
        G.mul_batch(elements_fast(rep, 4096), elements_fast(rep + 1, 4096))
        for (ch, claim, lo, hi, a, b, c), out, e in zip(cases, outs, eager):
            ref = oracle_with(ch, claim, lo, hi, a, b, c)
            assert_outer_equal(ref, out)
            assert_outer_equal(ref, e)
        eager_ref = outs
    assert eager_ref is not None
    G.outer_batch_release()


def test_outer_batch_rejects_a_short_challenge_array():
    """Each instance needs num_vars challenges; one short array fails the
    whole call before anything is concatenated or launched."""
    cases = batch_cases([(5, 2), (7, 3)], 990)
    insts = batch_insts(cases)
    ch, claim, eq, prods = insts[1]
    insts[1] = (ch[:-1], claim, eq, prods)
    with pytest.raises(G.BitzError) as e:
        G.prove_outer_batch(insts)
    assert "challenges" in str(e.value)
    outs = G.prove_outer_batch(batch_insts(cases))
    for (ch, claim, lo, hi, a, b, c), out in zip(cases, outs):
        assert_outer_equal(oracle_with(ch, claim, lo, hi, a, b, c), out)


# ---- GKR at Bits128: grinding before every cubic round's challenge ---------------
@pytest.mark.parametrize("log_leaves,log_groups", [(8, 2), (13, 3)])
def test_gkr_bits128_grinds_like_the_oracle(log_leaves, log_groups):
    """Every nonce in the narg string and every challenge after it equal the
    oracle's (crates/gkr/src/lib.rs:144: grinding_bits(3) = 2 at Bits128);
    Bits100 adds none."""
    leaves = elements_fast(140 + log_leaves, 1 << log_leaves)
    store, _ = O.gkr_fold_layers(leaves, log_groups)
    point = elements(141, log_groups)
    ot = O.DuplexTranscript(b"gkr/bits128")
    ref_point, ref_claim = O.gkr_prove(ot, point, store, log_leaves, 128)
    _, wit = G.grand_product(leaves, log_groups)
    gt = G.Transcript.duplex(b"gkr/bits128")
    gp, gc = G.gpgkr_prove(gt, point, wit, G.BITS128)
    assert np.array_equal(gp, ref_point) and gc == ref_claim and gt.narg() == ot.narg()
    assert gc == O.mle_evaluate(leaves, gp)
    # the serialized proof: the narg string (nonces included) in the wire container
    assert G.decode_proof(gt.proof()) == (ot.narg(), b"")
    assert gt.proof() == G.encode_proof(ot.narg())
    rounds = sum(range(log_groups, log_leaves))
    layers = log_leaves - log_groups
    assert len(gt.narg()) == 32 * rounds + 32 * layers + 8 * rounds  # messages, [L, R], one nonce per round
    _, wit = G.grand_product(leaves, log_groups)
    g100 = G.Transcript.duplex(b"gkr/bits128")
    G.gpgkr_prove(g100, point, wit, G.BITS100)
    assert len(g100.narg()) == 32 * rounds + 32 * layers


def test_gkr_bits128_hands_grinding_to_the_callers_transcript():
    """With a grind hook the prover's calls, in order, are the oracle's:
    each round's message, grind("gkr/cubic/v1", 2), its challenge; no grind
    before t."""
    log_leaves, log_groups = 7, 2
    leaves = elements_fast(150, 1 << log_leaves)
    store, _ = O.gkr_fold_layers(leaves, log_groups)
    point = elements(151, log_groups)
    count = sum(n + 1 for n in range(log_groups, log_leaves))
    ch = ints(elements(152, count))
    ot = O.ListTranscript(ch, grind_hook=True)
    ref_point, ref_claim = O.gkr_prove(ot, point, store, log_leaves, 128)
    events, it = [], iter(ch)
    gt = G.Transcript.callbacks(lambda kind, data: events.append(("absorb", kind, data)),
                                lambda: (events.append(("squeeze",)), next(it).to_bytes(16, "little"))[1],
                                grind=lambda label, bits: events.append(("grind", label, bits)))
    _, wit = G.grand_product(leaves, log_groups)
    gp, gc = G.gpgkr_prove(gt, point, wit, G.BITS128)
    assert np.array_equal(gp, ref_point) and gc == ref_claim
    assert events == ot.events
    assert sum(e[0] == "grind" for e in events) == sum(range(log_groups, log_leaves))
    assert all(e == ("grind", b"gkr/cubic/v1", 2) for e in events if e[0] == "grind")


def test_gkr_rejects_an_unknown_security_level():
    leaves = elements_fast(1, 1 << 6)
    _, wit = G.grand_product(leaves, 2)
    with pytest.raises(G.BitzError) as e:
        G.gpgkr_prove(G.Transcript.seeded(1), elements(2, 2), wit, 120)
    assert e.value.variant == "InvalidArgument"
