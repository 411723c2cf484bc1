/* Multilinear helpers, restating crates/poly/src/{eq.rs,mle.rs,f128.rs}.
 * Test infrastructure. */
#include <stdlib.h>
#include <string.h>

#include "bitz_oracle.h"

static inline of128 fadd(of128 a, of128 b) {
    of128 c = {a.lo ^ b.lo, a.hi ^ b.hi};
    return c;
}

/* eq.rs:63-99: table[j] = prod_i (bit_i(j) ? r_i : 1 - r_i), built by
 * in-place level doubling: the lower half holds the parents, the upper half
 * receives their one-children (one multiply per entry). */
void oracle_eq_table(const of128* r, int k, of128* table) {
    size_t n = (size_t)1 << k;
    memset(table, 0, n * sizeof(of128));
    table[0].lo = 1;
    for (int i = 0; i < k; ++i) {
        size_t half = (size_t)1 << i;
        for (size_t j = 0; j < half; ++j) {
            of128 parent = table[j];
            of128 one = oracle_mul(parent, r[i]);
            table[j] = fadd(parent, one); /* parent - parent*r */
            table[j + half] = one;
        }
    }
}

/* eq.rs:22-41: prod_i ((1 - x_i) + y_i (2 x_i - 1)); in characteristic 2
 * 2 x_i - 1 = 1, so each factor is 1 + x_i + y_i. */
of128 oracle_eq_eval(const of128* x, const of128* y, int n) {
    of128 acc = {1, 0};
    for (int i = 0; i < n; ++i) {
        of128 f = {1 ^ x[i].lo ^ y[i].lo, x[i].hi ^ y[i].hi};
        acc = i == 0 ? f : oracle_mul(acc, f);
    }
    return acc;
}

/* mle.rs:240-258: f'(b) = f(0,b) + r (f(1,b) - f(0,b)) on adjacent pairs. */
void oracle_mle_fold(const of128* in, size_t len, of128 r, of128* out) {
    for (size_t i = 0; i < len / 2; ++i) {
        of128 a = in[2 * i], b = in[2 * i + 1];
        out[i] = fadd(a, oracle_mul(r, fadd(a, b)));
    }
}

/* mle.rs:174-238 evaluate_exact: split on the last variable, recurse on
 * both halves, interpolate at the last coordinate. */
static of128 eval_rec(const of128* e, const of128* r, int n) {
    if (n == 0) return e[0];
    size_t half = (size_t)1 << (n - 1);
    of128 zero = eval_rec(e, r, n - 1);
    of128 one = eval_rec(e + half, r, n - 1);
    return fadd(zero, oracle_mul(r[n - 1], fadd(zero, one)));
}

of128 oracle_mle_evaluate(const of128* evals, const of128* point, int n) {
    return eval_rec(evals, point, n);
}

/* f128.rs:22-28: sum of products, reduced once. */
of128 oracle_inner_product(const of128* a, const of128* b, size_t n) {
    uint64_t acc[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < n; ++i) {
        uint64_t w[4];
        oracle_clmul128(a[i], b[i], w);
        acc[0] ^= w[0];
        acc[1] ^= w[1];
        acc[2] ^= w[2];
        acc[3] ^= w[3];
    }
    return oracle_reduce(acc);
}
