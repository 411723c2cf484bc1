/* The post-GKR inner-product sumcheck, restated (test infrastructure).
 *
 * crates/post_gkr/src/lib.rs:85-292 (prove_factors, Factors, bind_first,
 * advance) and crates/post_gkr/src/sumcheck.rs:63-215 (the degree-two
 * rounds).  The claim is <rows (x) columns, f> = target over the 2^t x 2^s
 * committed bits f (bit (c << t) | b weighs rows[b] columns[c]), packed 128
 * to an element, column after column (t >= 7). */
#include <stdlib.h>
#include <string.h>

#include "bitz_oracle.h"

static inline of128 fadd(of128 a, of128 b) {
    of128 r = {a.lo ^ b.lo, a.hi ^ b.hi};
    return r;
}
static inline int bit_of(of128 e, int k) { return (int)(((k < 64 ? e.lo >> k : e.hi >> (k - 64))) & 1); }
static const of128 OZERO = {0, 0};
static const of128 OONE = {1, 0};
static const uint8_t LABEL[] = "bitz/pcs/sumcheck/v1"; /* lib.rs:71 */

static of128 squeeze_elem(oracle_transcript* t) {
    uint8_t b[16];
    t->squeeze(t->ctx, b);
    of128 r;
    memcpy(&r.lo, b, 8);
    memcpy(&r.hi, b + 8, 8);
    return r;
}

/* lib.rs:286-289 advance: h' = a0 + r (h + a2 + r a2). */
static of128 advance(of128 h, of128 a0, of128 a2, of128 r) {
    return fadd(a0, oracle_mul(r, fadd(fadd(h, a2), oracle_mul(r, a2))));
}

/* sumcheck.rs:203-207 folded: entry j = t[2j] + r (t[2j] + t[2j+1]). */
static void fold(of128* t, size_t len, of128 r) {
    for (size_t j = 0; j < len / 2; ++j) t[j] = fadd(t[2 * j], oracle_mul(r, fadd(t[2 * j], t[2 * j + 1])));
}

int oracle_post_gkr_prove(oracle_transcript* tr, const of128* rows, int t, const of128* columns, int s,
                          of128 target, const of128* packed, size_t packed_len, uint32_t security_bits,
                          of128* point, of128* evaluation) {
    if (t < 7) return ORACLE_ERR_ARG;
    const size_t per_column = (size_t)1 << (t - 7), ncols = (size_t)1 << s;
    if (packed_len != per_column * ncols) return ORACLE_POST_GKR_WITNESS_LENGTH; /* lib.rs:133-135 */
    /* lib.rs:189-202 weighted_sum and lib.rs:208-232 first_message, element by element */
    uint64_t sum[4] = {0}, a0w[4] = {0}, a2w[4] = {0};
    for (size_t e = 0; e < packed_len; ++e) {
        const of128* r = rows + ((e % per_column) << 7);
        const of128 col = columns[e / per_column];
        of128 all = OZERO, a0 = OZERO, a2 = OZERO;
        for (int v = 0; v < 128; ++v)
            if (bit_of(packed[e], v)) all = fadd(all, r[v]);
        for (int v = 0; v < 128; v += 2) {
            const int f0 = bit_of(packed[e], v), f1 = bit_of(packed[e], v + 1);
            if (f0) a0 = fadd(a0, r[v]);
            if (f0 != f1) a2 = fadd(a2, fadd(r[v], r[v + 1]));
        }
        uint64_t w[4];
        oracle_clmul128(col, all, w);
        for (int i = 0; i < 4; ++i) sum[i] ^= w[i];
        oracle_clmul128(col, a0, w);
        for (int i = 0; i < 4; ++i) a0w[i] ^= w[i];
        oracle_clmul128(col, a2, w);
        for (int i = 0; i < 4; ++i) a2w[i] ^= w[i];
    }
    const of128 ws = oracle_reduce(sum);
    if (ws.lo != target.lo || ws.hi != target.hi) return ORACLE_POST_GKR_CLAIM; /* lib.rs:136-138 */
    of128 msg[2] = {oracle_reduce(a0w), oracle_reduce(a2w)};
    const uint32_t gbits = oracle_grinding_bits(security_bits, 2);
    tr->absorb(tr->ctx, 1, (const uint8_t*)msg, sizeof msg); /* lib.rs:141 prover_message */
    oracle_transcript_grind(tr, LABEL, sizeof LABEL - 1, gbits);
    of128 r = squeeze_elem(tr);
    point[0] = r;
    of128 h = advance(target, msg[0], msg[1], r);
    /* lib.rs:236-244 fold: the folded rows tensored with the columns;
     * lib.rs:265-284 bind_first: each bit pair to 0, 1 + r, r or 1. */
    const size_t half_rows = (size_t)1 << (t - 1), n = half_rows * ncols;
    of128* W = malloc(n * sizeof(of128));
    of128* V = malloc(n * sizeof(of128));
    of128* rf = malloc(half_rows * sizeof(of128));
    for (size_t j = 0; j < half_rows; ++j) rf[j] = fadd(rows[2 * j], oracle_mul(r, fadd(rows[2 * j], rows[2 * j + 1])));
    for (size_t c = 0; c < ncols; ++c)
        for (size_t j = 0; j < half_rows; ++j) W[c * half_rows + j] = oracle_mul(columns[c], rf[j]);
    const of128 vals[4] = {OZERO, fadd(r, OONE), r, OONE};
    for (size_t e = 0; e < packed_len; ++e)
        for (int p = 0; p < 64; ++p) V[e * 64 + p] = vals[bit_of(packed[e], 2 * p) | (bit_of(packed[e], 2 * p + 1) << 1)];
    /* sumcheck.rs:122-139: the remaining rounds */
    size_t len = n;
    int k = 1;
    while (len > 1) {
        uint64_t x0[4] = {0}, x2[4] = {0}, w[4];
        for (size_t j = 0; j < len; j += 2) {
            oracle_clmul128(W[j], V[j], w);
            for (int i = 0; i < 4; ++i) x0[i] ^= w[i];
            oracle_clmul128(fadd(W[j], W[j + 1]), fadd(V[j], V[j + 1]), w);
            for (int i = 0; i < 4; ++i) x2[i] ^= w[i];
        }
        of128 m[2] = {oracle_reduce(x0), oracle_reduce(x2)};
        tr->absorb(tr->ctx, 1, (const uint8_t*)m, sizeof m);
        oracle_transcript_grind(tr, LABEL, sizeof LABEL - 1, gbits);
        r = squeeze_elem(tr);
        point[k++] = r;
        h = advance(h, m[0], m[1], r);
        fold(W, len, r);
        fold(V, len, r);
        len /= 2;
    }
    /* sumcheck.rs:172-177: the closing evaluation as a prover message */
    *evaluation = V[0];
    tr->absorb(tr->ctx, 1, (const uint8_t*)&V[0], 16);
    const of128 expect = oracle_mul(W[0], V[0]);
    free(W);
    free(V);
    free(rf);
    return (expect.lo == h.lo && expect.hi == h.hi) ? ORACLE_OK : ORACLE_POST_GKR_CLAIM;
}

/* lib.rs:104-118 verify_factors with sumcheck.rs:99-198: replay the rounds
 * from the prover's messages, then MLE[rows](rho_b) MLE[columns](rho_c) v = h. */
int oracle_post_gkr_verify(oracle_transcript* tr, const of128* rows, int t, const of128* columns, int s,
                           of128 target, uint32_t security_bits, const of128* messages, of128 evaluation,
                           of128* point) {
    const int rounds = t + s;
    const uint32_t gbits = oracle_grinding_bits(security_bits, 2);
    of128 h = target;
    for (int k = 0; k < rounds; ++k) {
        const of128 m[2] = {messages[2 * k], messages[2 * k + 1]};
        tr->absorb(tr->ctx, 1, (const uint8_t*)m, sizeof m);
        oracle_transcript_grind(tr, LABEL, sizeof LABEL - 1, gbits);
        point[k] = squeeze_elem(tr);
        h = advance(h, m[0], m[1], point[k]);
    }
    tr->absorb(tr->ctx, 1, (const uint8_t*)&evaluation, 16);
    const of128 w = oracle_mul(oracle_mle_evaluate(rows, point, t), oracle_mle_evaluate(columns, point + t, s));
    const of128 got = oracle_mul(w, evaluation);
    return (got.lo == h.lo && got.hi == h.hi) ? ORACLE_OK : ORACLE_POST_GKR_MISMATCH;
}
