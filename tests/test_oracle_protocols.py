"""Pins the oracle's eq/fold/sumcheck/matrix/GKR restatements by definition
(brute force on tiny inputs, closed invariants), not by retyping them.

References: crates/poly/src/eq.rs:126-356, mle.rs:465-551,
crates/spartan/src/sumcheck.rs:1203-1607, matrix.rs:858-875,
crates/circuit/src/matrix_transpose.rs, crates/gkr/src/lib.rs:935-1275."""
import random

import numpy as np
import pytest

from common import elements, from_ints, horner, ints, mle_eval_def, to_int
from models import gf_inv, gf_mul
from oracle import oracle as O


def eq_def(tau, p):
    """prod_i (tau_i p_i + (1 + tau_i)(1 + p_i)) by the eq.rs:8-10 definition."""
    acc = 1
    for t, x in zip(tau, p):
        acc = gf_mul(acc, gf_mul(t, x) ^ gf_mul(t ^ 1, x ^ 1))
    return acc


# ---------------------------------------------------------------- poly ----
def test_eq_table_two_variables_by_hand():
    # eq.rs:171-196
    r0, r1 = 2, 7
    t = ints(O.eq_table(from_ints([r0, r1])))
    assert t == [gf_mul(1 ^ r0, 1 ^ r1), gf_mul(r0, 1 ^ r1), gf_mul(1 ^ r0, r1), gf_mul(r0, r1)]


@pytest.mark.parametrize("k", [0, 1, 3, 6])
def test_eq_table_matches_definition_and_sums_to_one(k):
    r = ints(elements(11 + k, k))
    t = ints(O.eq_table(from_ints(r))) if k else ints(O.eq_table(np.zeros((0, 2), np.uint64)))
    assert len(t) == 1 << k
    total = 0
    for j, v in enumerate(t):
        assert v == eq_def(r, [(j >> i) & 1 for i in range(k)])
        total ^= v
    assert total == 1


def test_boolean_eq_table_is_one_hot():
    sel = 0b1010
    t = ints(O.eq_table(from_ints([(sel >> i) & 1 for i in range(4)])))
    assert t == [1 if j == sel else 0 for j in range(16)]


def test_eq_eval_matches_definition():
    x, y = ints(elements(3, 9)), ints(elements(4, 9))
    assert O.eq_eval(from_ints(x), from_ints(y)) == eq_def(x, y)
    assert O.eq_eval(from_ints(x), from_ints(y)) == O.eq_eval(from_ints(y), from_ints(x))


def test_fold_selects_children_and_binds_little_endian():
    ev = from_ints(list(range(8)))
    assert ints(O.mle_fold(ev, 0)) == [0, 2, 4, 6]
    assert ints(O.mle_fold(ev, 1)) == [1, 3, 5, 7]
    tab = ints(elements(5, 16))
    pt = ints(elements(6, 4))
    cur = from_ints(tab)
    for r in pt:
        cur = O.mle_fold(cur, r)
    assert to_int(cur[0]) == mle_eval_def(tab, pt)
    assert O.mle_evaluate(from_ints(tab), from_ints(pt)) == mle_eval_def(tab, pt)


# -------------------------------------------------------------- spartan ----
def outer_inputs(n, split, seed=0x5A17A11C, satisfied=True):
    a = ints(elements(seed, 1 << n))
    b = ints(elements(seed + 1, 1 << n))
    c = [gf_mul(x, y) for x, y in zip(a, b)] if satisfied else ints(elements(seed + 2, 1 << n))
    tau = [2 * i + 2 for i in range(n)]  # sumcheck.rs:1186-1188
    return a, b, c, tau, O.eq_table(from_ints(tau[:split]) if split else np.zeros((0, 2), np.uint64)), \
        O.eq_table(from_ints(tau[split:]) if split < n else np.zeros((0, 2), np.uint64))


def claim_def(a, b, c, tau):
    """sum_x eq(tau, x) (A(x) B(x) + C(x)) over the hypercube."""
    n = len(tau)
    total = 0
    for x in range(1 << n):
        total ^= gf_mul(eq_def(tau, [(x >> i) & 1 for i in range(n)]), gf_mul(a[x], b[x]) ^ c[x])
    return total


def outer_round_def(a, b, c, tau, rs, X):
    """g_k(X) = sum_y eq(tau, (r, X, y)) (A B + C)(r, X, y), the definition."""
    n, k = len(tau), len(rs)
    total = 0
    for y in range(1 << (n - k - 1)):
        p = list(rs) + [X] + [(y >> i) & 1 for i in range(n - k - 1)]
        va, vb, vc = mle_eval_def(a, p), mle_eval_def(b, p), mle_eval_def(c, p)
        total ^= gf_mul(eq_def(tau, p), gf_mul(va, vb) ^ vc)
    return total


@pytest.mark.parametrize("n", [1, 2, 3])
@pytest.mark.parametrize("satisfied", [True, False])
def test_outer_rounds_match_definition(n, satisfied):
    a, b, c, tau, lo, hi = outer_inputs(n, n // 2, satisfied=satisfied)
    claim = claim_def(a, b, c, tau)
    if satisfied:
        assert claim == 0
    tr = O.SeededTranscript(77)
    rc, out = O.outer_sumcheck(tr, claim, lo, hi, from_ints(a), from_ints(b), from_ints(c))
    assert rc == 0
    rs = ints(out["points"])
    for k in range(n):
        coeffs = ints(out["rounds"][k])
        for X in (0, 1, 2, 3):  # a cubic is fixed by four evaluations
            assert horner(coeffs, X) == outer_round_def(a, b, c, tau, rs[:k], X), (k, X)
    fa, fb, fc = ints(out["claims"])
    assert fa == mle_eval_def(a, rs) and fb == mle_eval_def(b, rs) and fc == mle_eval_def(c, rs)
    assert out["final_claim"] == gf_mul(eq_def(tau, rs), gf_mul(fa, fb) ^ fc)


def test_outer_every_split_gives_the_same_proof_and_verifies():
    n = 5
    proofs = []
    for split in range(n + 1):
        a, b, c, tau, lo, hi = outer_inputs(n, split, satisfied=False)
        claim = claim_def(a, b, c, tau)
        tr = O.DuplexTranscript(b"outer/split")
        rc, out = O.outer_sumcheck(tr, claim, lo, hi, from_ints(a), from_ints(b), from_ints(c))
        assert rc == 0
        proofs.append(out["rounds"].copy())
        vt = O.DuplexTranscript(b"outer/split")
        assert O.outer_verify(vt, claim, from_ints(tau), out["rounds"], out["claims"]) == 0
        bad = out["claims"].copy()
        bad[0, 0] ^= 1
        vt = O.DuplexTranscript(b"outer/split")
        assert O.outer_verify(vt, claim, from_ints(tau), out["rounds"], bad) == O_ERR_TERMINAL
    for p in proofs[1:]:
        assert np.array_equal(p, proofs[0])


O_ERR_TERMINAL = -4


@pytest.mark.parametrize("n", [0, 1, 6, 10])
def test_outer_verifier_accepts_and_terminal_identity_holds(n):
    a, b, c, tau, lo, hi = outer_inputs(n, n // 2)
    tr = O.DuplexTranscript(b"outer")
    rc, out = O.outer_sumcheck(tr, 0, lo, hi, from_ints(a), from_ints(b), from_ints(c))
    assert rc == 0 and out["rounds"].shape[0] == n
    vt = O.DuplexTranscript(b"outer")
    assert O.outer_verify(vt, 0, from_ints(tau) if n else np.zeros((1, 2), np.uint64)[:0],
                          out["rounds"], out["claims"]) == 0
    rs = out["points"]
    if n:
        assert O.mle_evaluate(from_ints(a), rs) == to_int(out["claims"][0])
        assert O.mle_evaluate(from_ints(c), rs) == to_int(out["claims"][2])


def test_outer_rejects_dimensions():
    a = from_ints([1])
    b = from_ints([2, 3])
    one = from_ints([1])
    rc, _ = O.outer_sumcheck(O.SeededTranscript(1), 0, one, one, a, b, a)
    assert rc == -1
    a2 = from_ints([1, 2])
    rc, _ = O.outer_sumcheck(O.SeededTranscript(1), 0, one, one, a2, a2, a2)
    assert rc == -2


def inner_round_def(d, w, rs, X):
    n, k = (len(d).bit_length() - 1), len(rs)
    total = 0
    for y in range(1 << (n - k - 1)):
        p = list(rs) + [X] + [(y >> i) & 1 for i in range(n - k - 1)]
        total ^= gf_mul(mle_eval_def(d, p), mle_eval_def(w, p))
    return total


@pytest.mark.parametrize("n", [1, 2, 4])
def test_inner_rounds_match_definition(n):
    d, w = ints(elements(21, 1 << n)), ints(elements(22, 1 << n))
    claim = 0
    for x in range(1 << n):
        claim ^= gf_mul(d[x], w[x])
    rc, out = O.inner_sumcheck(O.SeededTranscript(5), claim, from_ints(d), from_ints(w))
    assert rc == 0
    rs = ints(out["points"])
    for k in range(n):
        coeffs = ints(out["rounds"][k])
        for X in (0, 1, 2):
            assert horner(coeffs, X) == inner_round_def(d, w, rs[:k], X)
    assert ints(out["evals"]) == [mle_eval_def(d, rs), mle_eval_def(w, rs)]
    assert out["final_claim"] == gf_mul(mle_eval_def(d, rs), mle_eval_def(w, rs))


@pytest.mark.parametrize("n", [0, 1, 3, 12, 13])
def test_inner_verifies(n):
    d, w = elements(31 + n, 1 << n), elements(41 + n, 1 << n)
    claim = O.inner_product(d, w)
    rc, out = O.inner_sumcheck(O.DuplexTranscript(b"inner"), claim, d, w)
    assert rc == 0
    rc, pts, final = O.sumcheck_verify(O.DuplexTranscript(b"inner"), claim, out["rounds"], 3)
    assert rc == 0 and final == out["final_claim"]
    assert final == O.to_int(O.mul(out["evals"][0], out["evals"][1]))
    if n:
        bad = out["rounds"].copy()
        bad[0, 1, 0] ^= 1
        rc, _, _ = O.sumcheck_verify(O.DuplexTranscript(b"inner"), claim, bad, 3)
        assert rc == -3


def test_inner_zero_vars_and_dimension_error():
    rc, out = O.inner_sumcheck(O.SeededTranscript(1), 35, from_ints([5]), from_ints([7]))
    assert rc == 0 and ints(out["evals"]) == [5, 7] and out["final_claim"] == 35
    rc, _ = O.inner_sumcheck(O.SeededTranscript(1), 0, from_ints([1, 2]), from_ints([1, 2, 3, 4]))
    assert rc == -1


# --------------------------------------------------------------- matrix ----
def random_csr(rng, rows, cols, per_row, seed):
    starts, colsl, vals = [0], [], []
    for _ in range(rows):
        cs = sorted(set(rng.randrange(cols) for _ in range(per_row)))
        colsl += cs
        starts.append(len(colsl))
    vals = elements(seed, len(colsl))
    return O.CsrMatrix(starts, colsl, vals)


def test_bind_and_batch_matches_dense_brute_force_and_direct_evaluation():
    rng = random.Random(11)
    rows, row_vars, col_vars = 13, 4, 5
    mats = [random_csr(rng, rows, 1 << col_vars, 6, 100 + m) for m in range(3)]
    rx = elements(7, row_vars)
    ry = elements(8, col_vars)
    rho = 0x1234567890ABCDEF1122334455667788
    d = O.bind_and_batch(mats, rows, row_vars, col_vars, rx, rho)
    # dense brute force on the big-int model
    rxi, ryi = ints(rx), ints(ry)
    scales = [1, rho, gf_mul(rho, rho)]
    dense = [0] * (1 << col_vars)
    for m, M in enumerate(mats):
        for i in range(rows):
            wi = eq_def(rxi, [(i >> b) & 1 for b in range(row_vars)])
            for k in range(int(M.row_starts[i]), int(M.row_starts[i + 1])):
                dense[int(M.columns[k])] ^= gf_mul(gf_mul(wi, scales[m]), to_int(M.values[k]))
    assert ints(d) == dense
    direct = O.evaluate_batched(mats, rows, row_vars, col_vars, rx, rho, ry)
    assert direct == mle_eval_def(dense, ryi)
    assert O.mle_evaluate(d, ry) == direct


def test_csr_multiply_matches_the_dense_product():
    """A h per row (the reference tests' multiply), zero-padded: against the
    dense matrix-vector product on the big-int model."""
    rng = random.Random(12)
    rows, cols = 11, 16
    M = random_csr(rng, rows, cols, 5, 200)
    h = elements(13, cols)
    out = O.csr_multiply(M, rows, h, 16)
    dense = [0] * 16
    hi = ints(h)
    for i in range(rows):
        for k in range(int(M.row_starts[i]), int(M.row_starts[i + 1])):
            dense[i] ^= gf_mul(to_int(M.values[k]), hi[int(M.columns[k])])
    assert ints(out) == dense


def test_mt_apply_and_transpose_inner_match_dense():
    rng = random.Random(3)
    rows, cols = 40, 17
    offsets, idx = [0], []
    for _ in range(cols):
        idx += sorted(set(rng.randrange(rows) for _ in range(rng.randrange(0, 7))))
        offsets.append(len(idx))
    r = elements(9, rows)
    out = ints(O.mt_apply(offsets, idx, r))
    ri = ints(r)
    for c in range(cols):
        acc = 0
        for k in range(offsets[c], offsets[c + 1]):
            acc ^= ri[idx[k]]
        assert out[c] == acc
    rw, cw = elements(10, 8), elements(12, 8)  # 8 x 8 >= 40 rows
    out = ints(O.mt_transpose_inner(offsets, idx, rw, cw))
    rwi, cwi = ints(rw), ints(cw)
    for c in range(cols):
        acc = 0
        for k in range(offsets[c], offsets[c + 1]):
            acc ^= gf_mul(rwi[idx[k] & 7], cwi[idx[k] >> 3])
        assert out[c] == acc


# ------------------------------------------------------------------ gkr ----
@pytest.mark.parametrize("log_len,log_groups", [(2, 1), (5, 0), (6, 2), (8, 3)])
def test_fold_layers_products(log_len, log_groups):
    leaves = elements(50 + log_len, 1 << log_len)
    _, layers = O.gkr_fold_layers(leaves, log_groups)
    li = ints(leaves)
    groups = 1 << log_groups
    for j, v in enumerate(ints(layers[-1])):
        p = 1
        for x in li[j::groups]:
            p = gf_mul(p, x)
        assert v == p
    # every layer keeps each group's product (lib.rs:935-963)
    for layer in layers:
        vals = ints(layer)
        for j in range(groups):
            p = 1
            for x in vals[j::groups]:
                p = gf_mul(p, x)
            q = 1
            for x in li[j::groups]:
                q = gf_mul(q, x)
            assert p == q


def test_fold_layers_all_zero_regression():
    # crates/gkr/proptest-regressions/lib.txt:7-10
    _, layers = O.gkr_fold_layers(np.zeros((4, 2), np.uint64), 1)
    assert ints(layers[-1]) == [0, 0]


def test_gkr_eq_tables_are_prefix_eq_tables():
    pt = elements(60, 5)
    st = ints(O.gkr_eq_tables(pt))
    for i in range(5):
        tab = ints(O.eq_table(pt[:i])) if i else [1]
        assert st[(1 << i) - 1:(1 << (i + 1)) - 1] == tab


@pytest.mark.parametrize("log_leaves,log_groups,zero", [(6, 2, None), (9, 3, None), (7, 3, 0), (7, 3, 2)])
def test_gkr_round_trip(log_leaves, log_groups, zero):
    leaves = elements(70 + log_leaves, 1 << log_leaves)
    store, layers = O.gkr_fold_layers(leaves, log_groups)
    point = elements(80, log_groups)
    if zero is not None:
        point[zero] = 0  # sends_one branch (lib.rs:1103-1143)
    claim = O.mle_evaluate(layers[-1], point)
    tr = O.DuplexTranscript(b"gkr")
    leaf_point, leaf_claim = O.gkr_prove(tr, point, store, log_leaves)
    assert leaf_claim == O.mle_evaluate(leaves, leaf_point)
    msgs = np.frombuffer(tr.narg(), dtype=np.uint64).reshape(-1, 2)
    layers_n = log_leaves - log_groups
    assert len(msgs) == sum(n + 1 for n in range(log_groups, log_leaves)) * 2
    rc, vp, vc = O.gkr_verify(O.DuplexTranscript(b"gkr"), claim, point, layers_n, msgs)
    assert rc == 0 and vc == leaf_claim and np.array_equal(vp, leaf_point)
    rc, _, _ = O.gkr_verify(O.DuplexTranscript(b"gkr"), claim ^ 1, point, layers_n, msgs)
    assert rc != 0
