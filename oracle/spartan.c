/* Spartan outer/inner sumchecks and matrix binding, restating
 * crates/spartan/src/sumcheck.rs and crates/spartan/src/matrix.rs.
 * Test infrastructure: the checker for the GPU path and the CPU baseline. */
#include <stdlib.h>
#include <string.h>

#include "bitz_oracle.h"

static inline of128 fadd(of128 a, of128 b) {
    of128 c = {a.lo ^ b.lo, a.hi ^ b.hi};
    return c;
}
static const of128 ZERO = {0, 0};

static void absorb_elems(oracle_transcript* t, int kind, const of128* e, size_t n) {
    /* Canonical encoding: 16 bytes little-endian, lo first (gf128.rs:103-109). */
    t->absorb(t->ctx, kind, (const uint8_t*)e, n * sizeof(of128));
}

static of128 squeeze(oracle_transcript* t) {
    uint8_t b[16];
    t->squeeze(t->ctx, b);
    of128 r;
    memcpy(&r.lo, b, 8);
    memcpy(&r.hi, b + 8, 8);
    return r;
}

/* sumcheck.rs:784-794: Horner from the top coefficient. */
static of128 eval_poly(const of128* c, int n, of128 x) {
    of128 v = ZERO;
    for (int i = n - 1; i >= 0; --i) v = fadd(oracle_mul(v, x), c[i]);
    return v;
}

/* sumcheck.rs:760-781 + 803-829: c1 = claim - c0 - (c0 + c2 + ...), absorb
 * [c0..cD] as a public message, squeeze r, claim <- g(r). */
static of128 recover_and_sample(oracle_transcript* t, of128* claim,
                                const of128* without_linear, int degree,
                                of128* round_out, of128* point_out) {
    of128 c[4];
    c[0] = without_linear[0];
    c[1] = ZERO;
    for (int i = 2; i <= degree; ++i) c[i] = without_linear[i - 1];
    of128 at_one = ZERO;
    for (int i = 0; i <= degree; ++i) at_one = fadd(at_one, c[i]);
    c[1] = fadd(fadd(*claim, c[0]), at_one);
    absorb_elems(t, 0, c, (size_t)degree + 1);
    of128 r = squeeze(t);
    *claim = eval_poly(c, degree + 1, r);
    memcpy(round_out, c, sizeof(of128) * (size_t)(degree + 1));
    *point_out = r;
    return r;
}

/* sumcheck.rs:835-871 EqualityPairs::pair. */
static void eq_pair(const of128* low, int low_bits, const of128* high, size_t p, of128 out[2]) {
    if (low_bits == 0) {
        out[0] = oracle_mul(low[0], high[2 * p]);
        out[1] = oracle_mul(low[0], high[2 * p + 1]);
        return;
    }
    size_t low_mask = ((size_t)1 << low_bits) - 1;
    size_t lp = p & (low_mask >> 1);
    of128 hw = high[p >> (low_bits - 1)];
    out[0] = oracle_mul(low[2 * lp], hw);
    out[1] = oracle_mul(low[2 * lp + 1], hw);
}

/* sumcheck.rs:724-744 cubic_contribution. */
static void cubic_contribution(const of128 eq[2], const of128 a[2], const of128 b[2],
                               const of128 c[2], of128 acc[3]) {
    of128 ed = fadd(eq[1], eq[0]), ad = fadd(a[1], a[0]), bd = fadd(b[1], b[0]),
          cd = fadd(c[1], c[0]);
    of128 r0 = fadd(oracle_mul(a[0], b[0]), c[0]);
    of128 rl = fadd(fadd(oracle_mul(a[0], bd), oracle_mul(ad, b[0])), cd);
    of128 rq = oracle_mul(ad, bd);
    acc[0] = fadd(acc[0], oracle_mul(eq[0], r0));
    acc[1] = fadd(acc[1], fadd(oracle_mul(eq[0], rq), oracle_mul(ed, rl)));
    acc[2] = fadd(acc[2], oracle_mul(ed, rq));
}

static inline of128 interp(of128 z, of128 o, of128 r) {
    return fadd(z, oracle_mul(r, fadd(o, z)));
}

static int ilog2z(size_t x) {
    int k = 0;
    while (((size_t)1 << k) < x) ++k;
    return k;
}

/* sumcheck.rs:874-889 compute_coefficients_without_linear. */
static void coeffs_from_tables(const of128* az, const of128* bz, const of128* cz, size_t len,
                               const of128* low, size_t low_len, const of128* high,
                               of128 out[3]) {
    out[0] = out[1] = out[2] = ZERO;
    int low_bits = ilog2z(low_len);
    for (size_t p = 0; p < len / 2; ++p) {
        of128 e[2];
        eq_pair(low, low_bits, high, p, e);
        cubic_contribution(e, az + 2 * p, bz + 2 * p, cz + 2 * p, out);
    }
}

/* sumcheck.rs:932-946 fold_table. */
static void fold_table(const of128* in, of128* out, size_t out_len, of128 r) {
    for (size_t i = 0; i < out_len; ++i) out[i] = interp(in[2 * i], in[2 * i + 1], r);
}

/* sumcheck.rs:963-1025 fold_products_and_compute_next: fold A, B, C at r and
 * accumulate the next [c0, c2, c3] from the folded pairs. */
static void fold_and_next(const of128* a, const of128* b, const of128* c, of128* ao, of128* bo,
                          of128* co, size_t out_len, of128 r, const of128* low, size_t low_len,
                          const of128* high, of128 out[3]) {
    fold_table(a, ao, out_len, r);
    fold_table(b, bo, out_len, r);
    fold_table(c, co, out_len, r);
    coeffs_from_tables(ao, bo, co, out_len, low, low_len, high, out);
}

int oracle_outer_sumcheck(oracle_transcript* t, of128 initial_claim, const of128* eq_low_in,
                          int low_vars, const of128* eq_high_in, int high_vars, const of128* az_in,
                          const of128* bz_in, const of128* cz_in, int az_vars, int bz_vars,
                          int cz_vars, of128* round_polys, of128* eval_points,
                          of128* final_claim, of128 claims[3]) {
    /* sumcheck.rs:258-269: dimension checks before touching the transcript. */
    if (bz_vars != az_vars || cz_vars != az_vars) return ORACLE_ERR_PRODUCT_DIMS;
    if (low_vars < 0 || high_vars < 0 || low_vars + high_vars != az_vars)
        return ORACLE_ERR_EQUALITY_DIMS;
    int n = az_vars;
    size_t len = (size_t)1 << n;
    size_t low_len = (size_t)1 << low_vars, high_len = (size_t)1 << high_vars;

    of128* A = malloc(len * sizeof(of128));
    of128* B = malloc(len * sizeof(of128));
    of128* C = malloc(len * sizeof(of128));
    of128* As = malloc((len / 2 + 1) * sizeof(of128));
    of128* Bs = malloc((len / 2 + 1) * sizeof(of128));
    of128* Cs = malloc((len / 2 + 1) * sizeof(of128));
    of128* L = malloc(low_len * sizeof(of128));
    of128* H = malloc(high_len * sizeof(of128));
    of128* Ls = malloc((low_len / 2 + 1) * sizeof(of128));
    of128* Hs = malloc((high_len / 2 + 1) * sizeof(of128));
    memcpy(A, az_in, len * sizeof(of128));
    memcpy(B, bz_in, len * sizeof(of128));
    memcpy(C, cz_in, len * sizeof(of128));
    memcpy(L, eq_low_in, low_len * sizeof(of128));
    memcpy(H, eq_high_in, high_len * sizeof(of128));

    of128 claim = initial_claim;
    int round = 0;
    of128 co[3];
    coeffs_from_tables(A, B, C, len, L, low_len, H, co);

    /* Phase 1: bind the low equality variables (sumcheck.rs:294-323). */
    while (low_len > 1) {
        of128 r = recover_and_sample(t, &claim, co, 3, round_polys + 4 * round,
                                     eval_points + round);
        ++round;
        size_t next_len = len / 2, next_low = low_len / 2;
        fold_table(L, Ls, next_low, r);
        if (next_len > 1) {
            fold_and_next(A, B, C, As, Bs, Cs, next_len, r, Ls, next_low, H, co);
        } else {
            fold_table(A, As, 1, r);
            fold_table(B, Bs, 1, r);
            fold_table(C, Cs, 1, r);
        }
        of128* tmp;
        tmp = A; A = As; As = tmp;
        tmp = B; B = Bs; Bs = tmp;
        tmp = C; C = Cs; Cs = tmp;
        tmp = L; L = Ls; Ls = tmp;
        len = next_len;
        low_len = next_low;
    }
    /* Phase 2: bind the high variables (sumcheck.rs:328-362). */
    while (high_len > 1) {
        of128 r = recover_and_sample(t, &claim, co, 3, round_polys + 4 * round,
                                     eval_points + round);
        ++round;
        size_t next_high = high_len / 2;
        if (next_high == 1) {
            fold_table(A, As, 1, r);
            fold_table(B, Bs, 1, r);
            fold_table(C, Cs, 1, r);
            fold_table(H, Hs, 1, r);
        } else {
            fold_table(H, Hs, next_high, r);
            fold_and_next(A, B, C, As, Bs, Cs, next_high, r, L, low_len, Hs, co);
        }
        of128* tmp;
        tmp = A; A = As; As = tmp;
        tmp = B; B = Bs; Bs = tmp;
        tmp = C; C = Cs; Cs = tmp;
        tmp = H; H = Hs; Hs = tmp;
        len = next_high;
        high_len = next_high;
    }
    claims[0] = A[0];
    claims[1] = B[0];
    claims[2] = C[0];
    absorb_elems(t, 0, claims, 3); /* sumcheck.rs:372 */
    *final_claim = claim;
    free(A); free(B); free(C); free(As); free(Bs); free(Cs);
    free(L); free(H); free(Ls); free(Hs);
    return ORACLE_OK;
}

/* sumcheck.rs:442-552 prove_inner_sumcheck. */
int oracle_inner_sumcheck(oracle_transcript* t, of128 initial_claim, const of128* matrix,
                          int matrix_vars, const of128* witness, int witness_vars,
                          of128* round_polys, of128* eval_points, of128* final_claim,
                          of128 evals[2]) {
    if (witness_vars != matrix_vars) return ORACLE_ERR_PRODUCT_DIMS;
    int n = matrix_vars;
    size_t len = (size_t)1 << n;
    of128 claim = initial_claim;
    if (n == 0) {
        evals[0] = matrix[0];
        evals[1] = witness[0];
        *final_claim = claim;
        return ORACLE_OK;
    }
    of128* D = malloc(len * sizeof(of128));
    of128* W = malloc(len * sizeof(of128));
    of128* Ds = malloc(len / 2 * sizeof(of128));
    of128* Ws = malloc(len / 2 * sizeof(of128));
    memcpy(D, matrix, len * sizeof(of128));
    memcpy(W, witness, len * sizeof(of128));
    /* sumcheck.rs:555-566 per pair: [m0 w0, (m1 - m0)(w1 - w0)]. */
    of128 co[2] = {ZERO, ZERO};
    for (size_t p = 0; p < len / 2; ++p) {
        co[0] = fadd(co[0], oracle_mul(D[2 * p], W[2 * p]));
        co[1] = fadd(co[1], oracle_mul(fadd(D[2 * p + 1], D[2 * p]),
                                       fadd(W[2 * p + 1], W[2 * p])));
    }
    for (int round = 0; round < n; ++round) {
        of128 r = recover_and_sample(t, &claim, co, 2, round_polys + 3 * round,
                                     eval_points + round);
        size_t next = len / 2;
        fold_table(D, Ds, next, r);
        fold_table(W, Ws, next, r);
        co[0] = co[1] = ZERO;
        for (size_t p = 0; p + 1 < next; p += 2) {
            co[0] = fadd(co[0], oracle_mul(Ds[p], Ws[p]));
            co[1] = fadd(co[1], oracle_mul(fadd(Ds[p + 1], Ds[p]), fadd(Ws[p + 1], Ws[p])));
        }
        of128* tmp;
        tmp = D; D = Ds; Ds = tmp;
        tmp = W; W = Ws; Ws = tmp;
        len = next;
    }
    evals[0] = D[0];
    evals[1] = W[0];
    *final_claim = claim;
    free(D); free(W); free(Ds); free(Ws);
    return ORACLE_OK;
}

/* sumcheck.rs:63-111 SumcheckProof::verify. */
int oracle_sumcheck_verify(oracle_transcript* t, of128 initial_claim, const of128* round_polys,
                           int rounds, int coeffs, of128* eval_points, of128* final_claim) {
    if (coeffs <= 0) return ORACLE_ERR_ARG;
    of128 claim = initial_claim;
    for (int round = 0; round < rounds; ++round) {
        const of128* c = round_polys + (size_t)coeffs * round;
        absorb_elems(t, 0, c, (size_t)coeffs);
        of128 at_one = ZERO;
        for (int i = 0; i < coeffs; ++i) at_one = fadd(at_one, c[i]);
        of128 s = fadd(c[0], at_one);
        if (s.lo != claim.lo || s.hi != claim.hi) return ORACLE_ERR_ROUND_CLAIM;
        of128 r = squeeze(t);
        claim = eval_poly(c, coeffs, r);
        eval_points[round] = r;
    }
    *final_claim = claim;
    return ORACLE_OK;
}

/* sumcheck.rs:218-243 OuterSumcheckProof::verify. */
int oracle_outer_verify(oracle_transcript* t, of128 initial_claim, const of128* tau, int n,
                        const of128* round_polys, const of128 claims[3], of128* eval_points) {
    of128 final_claim;
    int rc = oracle_sumcheck_verify(t, initial_claim, round_polys, n, 4, eval_points, &final_claim);
    if (rc) return rc;
    absorb_elems(t, 0, claims, 3);
    of128 expected = oracle_mul(oracle_eq_eval(tau, eval_points, n),
                                fadd(oracle_mul(claims[0], claims[1]), claims[2]));
    if (expected.lo != final_claim.lo || expected.hi != final_claim.hi) return ORACLE_ERR_TERMINAL;
    return ORACLE_OK;
}

/* matrix.rs:429-468 bind_and_batch: D(j) = sum_i eq(i, r_x) (A + rho B + rho^2 C)[i, j].
 * The reference accumulates per 2^16-column chunk; exact arithmetic makes
 * the nonzero order irrelevant, so this walks the rows directly. */
int oracle_bind_and_batch(const oracle_csr mats[3], size_t rows, int row_vars, int col_vars,
                          const of128* row_point, of128 rho, of128* out) {
    (void)rows;
    size_t nrows = (size_t)1 << row_vars;
    of128* w = malloc(nrows * sizeof(of128));
    oracle_eq_table(row_point, row_vars, w);
    of128 scales[3] = {{1, 0}, rho, oracle_mul(rho, rho)};
    size_t ncols = (size_t)1 << col_vars;
    memset(out, 0, ncols * sizeof(of128));
    for (int m = 0; m < 3; ++m) {
        for (size_t i = 0; i < rows; ++i) {
            of128 rs = oracle_mul(w[i], scales[m]);
            for (uint64_t k = mats[m].row_starts[i]; k < mats[m].row_starts[i + 1]; ++k) {
                uint32_t col = mats[m].columns[k];
                out[col] = fadd(out[col], oracle_mul(rs, mats[m].values[k]));
            }
        }
    }
    free(w);
    return ORACLE_OK;
}

/* piop.rs:352-364 multiply: one product per entry, summed per row. */
void oracle_csr_multiply(const oracle_csr* m, size_t rows, const of128* assignment, of128* out) {
    for (size_t i = 0; i < rows; ++i) {
        of128 sum = ZERO;
        for (uint64_t k = m->row_starts[i]; k < m->row_starts[i + 1]; ++k)
            sum = fadd(sum, oracle_mul(m->values[k], assignment[m->columns[k]]));
        out[i] = sum;
    }
}

/* matrix.rs:477-529 evaluate_batched, independent of bind_and_batch:
 * sum over nonzeros of scale * eq(i, r_x) * eq(j, r_y) * value. */
int oracle_evaluate_batched(const oracle_csr mats[3], size_t rows, int row_vars, int col_vars,
                            const of128* row_point, of128 rho, const of128* col_point,
                            of128* result) {
    size_t nrows = (size_t)1 << row_vars, ncols = (size_t)1 << col_vars;
    of128* rw = malloc(nrows * sizeof(of128));
    of128* cw = malloc(ncols * sizeof(of128));
    oracle_eq_table(row_point, row_vars, rw);
    oracle_eq_table(col_point, col_vars, cw);
    of128 scales[3] = {{1, 0}, rho, oracle_mul(rho, rho)};
    of128 acc = ZERO;
    for (int m = 0; m < 3; ++m) {
        of128 msum = ZERO;
        for (size_t i = 0; i < rows; ++i) {
            of128 span = ZERO;
            for (uint64_t k = mats[m].row_starts[i]; k < mats[m].row_starts[i + 1]; ++k)
                span = fadd(span, oracle_mul(cw[mats[m].columns[k]], mats[m].values[k]));
            msum = fadd(msum, oracle_mul(rw[i], span));
        }
        acc = fadd(acc, oracle_mul(scales[m], msum));
    }
    *result = acc;
    free(rw);
    free(cw);
    return ORACLE_OK;
}
