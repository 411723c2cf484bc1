/* GF(2^128) = F2[X]/(X^128 + X^7 + X^2 + X + 1), the GHASH field in the
 * non-reflected basis (crates/field/src/gf128.rs:1-8, REDUCTION = 0x87 at :52).
 * Test infrastructure: the oracle the GPU path is checked against. */
#include <wmmintrin.h>
#include <emmintrin.h>

#include "bitz_oracle.h"

#define REDUCTION 0x87ULL

static int g_kernel = 0;

void oracle_set_kernel(int kernel) { g_kernel = kernel; }
int oracle_get_kernel(void) { return g_kernel; }

/* portable.rs:11-24: 64-step shift/XOR carry-less product. */
void oracle_clmul64(uint64_t a, uint64_t b, uint64_t out[2]) {
    uint64_t lo = 0, hi = 0;
    for (int i = 0; i < 64; ++i) {
        if ((b >> i) & 1) {
            lo ^= a << i;
            if (i > 0) hi ^= a >> (64 - i);
        }
    }
    out[0] = lo;
    out[1] = hi;
}

__attribute__((target("pclmul,sse2")))
static void clmul64_pclmul(uint64_t a, uint64_t b, uint64_t out[2]) {
    __m128i x = _mm_set_epi64x(0, (long long)a);
    __m128i y = _mm_set_epi64x(0, (long long)b);
    __m128i z = _mm_clmulepi64_si128(x, y, 0x00);
    out[0] = (uint64_t)_mm_cvtsi128_si64(z);
    out[1] = (uint64_t)_mm_cvtsi128_si64(_mm_unpackhi_epi64(z, z));
}

static inline void clmul64(uint64_t a, uint64_t b, uint64_t out[2]) {
    if (g_kernel == 1)
        clmul64_pclmul(a, b, out);
    else
        oracle_clmul64(a, b, out);
}

/* portable.rs:28-34: schoolbook 128x128, four words ascending. */
void oracle_clmul128(of128 a, of128 b, uint64_t out[4]) {
    uint64_t l[2], h[2], m[2], n[2];
    clmul64(a.lo, b.lo, l);
    clmul64(a.hi, b.hi, h);
    clmul64(a.lo, b.hi, m);
    clmul64(a.hi, b.lo, n);
    out[0] = l[0];
    out[1] = l[1] ^ m[0] ^ n[0];
    out[2] = h[0] ^ m[1] ^ n[1];
    out[3] = h[1];
}

/* portable.rs:37-43: w * (X^7 + X^2 + X + 1) as (low, high). */
static inline void times_g(uint64_t w, uint64_t* lo, uint64_t* hi) {
    *lo = w ^ (w << 1) ^ (w << 2) ^ (w << 7);
    *hi = (w >> 63) ^ (w >> 62) ^ (w >> 57);
}

/* portable.rs:50-57: fold the two high words down with X^128 = g, then the
 * <= 7 spilled bits once more. */
of128 oracle_reduce(const uint64_t r[4]) {
    uint64_t a0, a1, b0, b1;
    times_g(r[2], &a0, &a1);
    times_g(r[3], &b0, &b1);
    of128 out;
    out.lo = r[0] ^ a0 ^ b1 ^ (b1 << 1) ^ (b1 << 2) ^ (b1 << 7);
    out.hi = r[1] ^ a1 ^ b0;
    return out;
}

of128 oracle_mul(of128 a, of128 b) {
    uint64_t w[4];
    oracle_clmul128(a, b, w);
    return oracle_reduce(w);
}

/* portable.rs:67-71: the cross terms cancel in characteristic 2. */
of128 oracle_square(of128 a) {
    uint64_t l[2], h[2], w[4];
    clmul64(a.lo, a.lo, l);
    clmul64(a.hi, a.hi, h);
    w[0] = l[0];
    w[1] = l[1];
    w[2] = h[0];
    w[3] = h[1];
    return oracle_reduce(w);
}

/* portable.rs:114-120. */
of128 oracle_mul_x(of128 a) {
    uint64_t carry = (uint64_t)0 - (a.hi >> 63);
    of128 out;
    out.lo = (a.lo << 1) ^ (carry & REDUCTION);
    out.hi = (a.hi << 1) | (a.lo >> 63);
    return out;
}

/* a^(2^128 - 2) = a^-1 for a != 0 (pow.rs:87-103 computes the same power by
 * Itoh-Tsujii; the value is fixed by the exponent, not the addition chain).
 * 2^128 - 2 = sum_{i=1}^{127} 2^i, so the result is prod_{i=1}^{127} a^(2^i). */
of128 oracle_inv(of128 a) {
    of128 acc = {1, 0};
    of128 p = a;
    for (int i = 1; i < 128; ++i) {
        p = oracle_square(p);
        acc = oracle_mul(acc, p);
    }
    return acc;
}

void oracle_mul_batch(const of128* a, const of128* b, of128* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = oracle_mul(a[i], b[i]);
}
