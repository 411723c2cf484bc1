/* M^T gathers (crates/circuit/src/matrix_transpose.rs) and the dense
 * grand-product GKR (crates/gkr/src/lib.rs).  Test infrastructure. */
#include <stdlib.h>
#include <string.h>

#include "bitz_oracle.h"

static inline of128 fadd(of128 a, of128 b) {
    of128 c = {a.lo ^ b.lo, a.hi ^ b.hi};
    return c;
}
static const of128 ZERO = {0, 0};
static const of128 ONE = {1, 0};

/* matrix_transpose.rs:327-375: out[col] = XOR of challenges[row] over the
 * column's rows. */
void oracle_mt_apply(const uint32_t* column_offsets, size_t columns, const uint32_t* row_indices,
                     const of128* challenges, of128* out) {
    for (size_t c = 0; c < columns; ++c) {
        of128 acc = ZERO;
        for (uint32_t k = column_offsets[c]; k < column_offsets[c + 1]; ++k)
            acc = fadd(acc, challenges[row_indices[k]]);
        out[c] = acc;
    }
}

/* matrix_transpose.rs:377-441: out[c] = sum row_w[idx & mask] col_w[idx >> log_rows],
 * one unreduced product per nonzero and one reduction per column. */
int oracle_mt_transpose_inner(const uint32_t* column_offsets, size_t columns,
                              const uint32_t* row_indices, const of128* row_weights,
                              size_t n_row_weights, const of128* col_weights,
                              size_t n_col_weights, of128* out) {
    if (n_row_weights == 0 || (n_row_weights & (n_row_weights - 1))) return ORACLE_ERR_ARG;
    (void)n_col_weights;
    int log_rows = 0;
    while (((size_t)1 << log_rows) < n_row_weights) ++log_rows;
    size_t mask = n_row_weights - 1;
    for (size_t c = 0; c < columns; ++c) {
        uint64_t acc[4] = {0, 0, 0, 0};
        for (uint32_t k = column_offsets[c]; k < column_offsets[c + 1]; ++k) {
            size_t idx = row_indices[k];
            uint64_t w[4];
            oracle_clmul128(row_weights[idx & mask], col_weights[idx >> log_rows], w);
            acc[0] ^= w[0]; acc[1] ^= w[1]; acc[2] ^= w[2]; acc[3] ^= w[3];
        }
        out[c] = oracle_reduce(acc);
    }
    return ORACLE_OK;
}

/* lib.rs:874-894 fold_layers: node k of the next layer multiplies node k of
 * the left half and node k of the right half, until `groups` remain. */
int oracle_gkr_fold_layers(const of128* leaves, int log_len, int log_groups, of128* store) {
    size_t len = (size_t)1 << log_len;
    memcpy(store, leaves, len * sizeof(of128));
    of128* cur = store;
    int layers = 1;
    while (log_len > log_groups) {
        size_t half = len / 2;
        of128* next = cur + len;
        for (size_t k = 0; k < half; ++k) next[k] = oracle_mul(cur[k], cur[k + half]);
        cur = next;
        len = half;
        --log_len;
        ++layers;
    }
    return layers;
}

/* lib.rs:663-693 EqTables::new: table i = eq(z_{<i}, .) at [2^i - 1, 2^(i+1) - 1),
 * each coordinate doubling the previous table as its new top bit. */
void oracle_gkr_eq_tables(const of128* point, int n, of128* storage) {
    storage[0] = ONE;
    size_t prev_off = 0, prev_len = 1;
    for (int i = 0; i + 1 < n; ++i) {
        of128 z = point[i];
        size_t off = prev_off + prev_len;
        for (size_t k = 0; k < prev_len; ++k) {
            of128 e = storage[prev_off + k];
            of128 tmp = oracle_mul(z, e);
            storage[off + k] = fadd(e, tmp);
            storage[off + prev_len + k] = tmp;
        }
        prev_off = off;
        prev_len *= 2;
    }
}

static of128 squeeze(oracle_transcript* t) {
    uint8_t b[16];
    t->squeeze(t->ctx, b);
    of128 r;
    memcpy(&r.lo, b, 8);
    memcpy(&r.hi, b + 8, 8);
    return r;
}

/* lib.rs:630-632: eq(r, z) = 1 + r + z in characteristic 2. */
static inline of128 eq_factor(of128 r, of128 z) {
    of128 o = {1 ^ r.lo ^ z.lo, r.hi ^ z.hi};
    return o;
}

static inline int is_zero(of128 a) { return a.lo == 0 && a.hi == 0; }

/* lib.rs:88-159 prove_layer / prove_rounds, dense. */
static const uint8_t CUBIC_GRINDING_LABEL[] = "gkr/cubic/v1"; /* lib.rs:10 */

of128 oracle_gkr_prove_layer(oracle_transcript* t, const of128* point, int n,
                             const of128* children, of128* next_point, of128* messages, uint32_t grind_bits) {
    size_t mid = (size_t)1 << n;
    of128* left = malloc(mid * sizeof(of128));
    of128* right = malloc(mid * sizeof(of128));
    memcpy(left, children, mid * sizeof(of128));
    memcpy(right, children + mid, mid * sizeof(of128));
    size_t eq_size = (n > 0) ? (((size_t)1 << n) - 1) : 1;
    of128* eqs = malloc(eq_size * sizeof(of128));
    if (n > 0) oracle_gkr_eq_tables(point, n, eqs);
    of128 eq_bound = ONE;
    int msg = 0;
    for (int i = n - 1; i >= 0; --i) {
        of128 z = point[i];
        size_t half = (size_t)1 << i;
        const of128* eq = eqs + (half - 1);
        int at_one = is_zero(z); /* sends_one, lib.rs:622-624 */
        uint64_t sb[4] = {0, 0, 0, 0}, si[4] = {0, 0, 0, 0};
        for (size_t x = 0; x < half; ++x) {
            of128 l0 = left[x], l1 = left[x + half], r0 = right[x], r1 = right[x + half];
            of128 lb = at_one ? l1 : l0, rb = at_one ? r1 : r0;
            uint64_t w[4];
            /* lib.rs:164-219 round_sums: e joins the left factor. */
            oracle_clmul128(oracle_mul(eq[x], lb), rb, w);
            sb[0] ^= w[0]; sb[1] ^= w[1]; sb[2] ^= w[2]; sb[3] ^= w[3];
            oracle_clmul128(oracle_mul(eq[x], fadd(l1, l0)), fadd(r1, r0), w);
            si[0] ^= w[0]; si[1] ^= w[1]; si[2] ^= w[2]; si[3] ^= w[3];
        }
        of128 m[2] = {oracle_mul(eq_bound, oracle_reduce(sb)), oracle_mul(eq_bound, oracle_reduce(si))};
        t->absorb(t->ctx, 1, (const uint8_t*)m, sizeof m); /* prover_message, lib.rs:142 */
        if (messages) {
            messages[2 * msg] = m[0];
            messages[2 * msg + 1] = m[1];
        }
        ++msg;
        /* lib.rs:144: the grinding boundary before the round's challenge. */
        oracle_transcript_grind(t, CUBIC_GRINDING_LABEL, sizeof CUBIC_GRINDING_LABEL - 1, grind_bits);
        of128 r = squeeze(t);
        /* lib.rs:223-243 bind: l0 += r (l1 - l0), r0 += r (r1 - r0). */
        for (size_t x = 0; x < half; ++x) {
            left[x] = fadd(left[x], oracle_mul(r, fadd(left[x + half], left[x])));
            right[x] = fadd(right[x], oracle_mul(r, fadd(right[x + half], right[x])));
        }
        next_point[i] = r;
        eq_bound = oracle_mul(eq_bound, eq_factor(r, z));
    }
    of128 fin[2] = {left[0], right[0]};
    t->absorb(t->ctx, 1, (const uint8_t*)fin, sizeof fin); /* lib.rs:154 */
    if (messages) {
        messages[2 * msg] = fin[0];
        messages[2 * msg + 1] = fin[1];
    }
    of128 tt = squeeze(t);
    next_point[n] = tt;
    of128 claim = fadd(fin[0], oracle_mul(tt, fadd(fin[1], fin[0])));
    free(left);
    free(right);
    free(eqs);
    return claim;
}

/* gpgkr_prove (lib.rs:24-79) over dense layers: store holds every layer
 * from the leaves (2^log_leaves) up to the output (2^n_out). */
of128 oracle_gkr_prove(oracle_transcript* t, const of128* point, int n_out, const of128* store,
                       int log_leaves, of128* leaf_point, uint32_t security_bits) {
    const uint32_t grind_bits = oracle_grinding_bits(security_bits, 3);
    int layers = log_leaves - n_out; /* sumchecks: one per layer below the output */
    of128* pt = malloc((size_t)(log_leaves + 1) * sizeof(of128));
    of128* nxt = malloc((size_t)(log_leaves + 1) * sizeof(of128));
    memcpy(pt, point, (size_t)n_out * sizeof(of128));
    /* offsets of each layer in store */
    size_t* off = malloc((size_t)(layers + 1) * sizeof(size_t));
    off[0] = 0;
    for (int j = 0; j < layers; ++j) off[j + 1] = off[j] + ((size_t)1 << (log_leaves - j));
    of128 claim = ZERO;
    int n = n_out;
    for (int j = layers - 1; j >= 0; --j) {
        /* children = layer j (2^(n+1) nodes), parent point has n coordinates */
        claim = oracle_gkr_prove_layer(t, pt, n, store + off[j], nxt, NULL, grind_bits);
        memcpy(pt, nxt, (size_t)(n + 1) * sizeof(of128));
        ++n;
    }
    memcpy(leaf_point, pt, (size_t)log_leaves * sizeof(of128));
    free(pt);
    free(nxt);
    free(off);
    return claim;
}

/* lib.rs:717-772 gpgkr_verify / verify_layer. `messages` are the prover's
 * narg records in order: per layer, one [sum_b, sum_inf] per round and a
 * final [L(r), R(r)]. */
int oracle_gkr_verify(oracle_transcript* t, of128 claim, const of128* point, int n_out,
                      int layers, const of128* messages, of128* leaf_point, of128* leaf_claim) {
    int total = n_out + layers;
    of128* pt = malloc((size_t)(total + 1) * sizeof(of128));
    of128* nxt = malloc((size_t)(total + 1) * sizeof(of128));
    memcpy(pt, point, (size_t)n_out * sizeof(of128));
    size_t mi = 0;
    int n = n_out;
    int ok = 0;
    for (int layer = 0; layer < layers; ++layer) {
        of128 eq_bound = ONE;
        for (int i = n - 1; i >= 0; --i) {
            of128 z = pt[i];
            of128 m[2] = {messages[2 * mi], messages[2 * mi + 1]};
            ++mi;
            t->absorb(t->ctx, 1, (const uint8_t*)m, sizeof m);
            of128 sum0, sum1;
            if (is_zero(z)) {
                sum0 = claim;
                sum1 = m[0];
            } else {
                sum0 = m[0];
                /* (claim - (1 - z) sum_b) / z */
                of128 one_minus_z = {1 ^ z.lo, z.hi};
                sum1 = oracle_mul(fadd(claim, oracle_mul(one_minus_z, m[0])), oracle_inv(z));
            }
            of128 r = squeeze(t);
            nxt[i] = r;
            of128 eq_r = eq_factor(r, z);
            /* quadratic_at (lib.rs:614-616): at0 + r((at1 - at0) + (r - 1) at_inf) */
            of128 r_minus_one = {r.lo ^ 1, r.hi};
            of128 q = fadd(sum0, oracle_mul(r, fadd(fadd(sum1, sum0), oracle_mul(r_minus_one, m[1]))));
            claim = oracle_mul(eq_r, q);
            eq_bound = oracle_mul(eq_bound, eq_r);
        }
        of128 m[2] = {messages[2 * mi], messages[2 * mi + 1]};
        ++mi;
        t->absorb(t->ctx, 1, (const uint8_t*)m, sizeof m);
        of128 lhs = oracle_mul(oracle_mul(eq_bound, m[0]), m[1]);
        if (lhs.lo != claim.lo || lhs.hi != claim.hi) {
            ok = ORACLE_ERR_TERMINAL;
            break;
        }
        of128 tt = squeeze(t);
        nxt[n] = tt;
        claim = fadd(m[0], oracle_mul(tt, fadd(m[1], m[0])));
        memcpy(pt, nxt, (size_t)(n + 1) * sizeof(of128));
        ++n;
    }
    if (ok == 0) {
        memcpy(leaf_point, pt, (size_t)total * sizeof(of128));
        *leaf_claim = claim;
    }
    free(pt);
    free(nxt);
    return ok;
}
