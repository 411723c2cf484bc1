/* Challenge sources for the oracle (test infrastructure).
 *
 * The reference keeps Fiat-Shamir in spongefish (crates/transcript/src/
 * {domain.rs,prover.rs,challenge.rs}); that crate is absent here, so the
 * oracle exposes transcripts as callbacks and offers two stand-ins:
 *  - seeded: challenges from the splitmix64 stream the reference's GKR tests
 *    use (crates/gkr/src/lib.rs:986-999), F128::from([w0, w1]);
 *  - duplex: a SHA-256 ratchet over every absorbed message, so challenges
 *    depend on every earlier message. */
#include <stdlib.h>
#include <string.h>

#include "bitz_oracle.h"

/* ---- FIPS 180-4 SHA-256 ---------------------------------------------------- */
static const uint32_t K256[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

static inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void sha256_block(uint32_t h[8], const uint8_t* p) {
    uint32_t w[64];
    for (int i = 0; i < 16; ++i)
        w[i] = ((uint32_t)p[4 * i] << 24) | ((uint32_t)p[4 * i + 1] << 16) |
               ((uint32_t)p[4 * i + 2] << 8) | (uint32_t)p[4 * i + 3];
    for (int i = 16; i < 64; ++i) {
        uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
        uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int i = 0; i < 64; ++i) {
        uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
        uint32_t ch = (e & f) ^ (~e & g);
        uint32_t t1 = hh + S1 + ch + K256[i] + w[i];
        uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
        uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
        uint32_t t2 = S0 + mj;
        hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

void oracle_sha256(const uint8_t* msg, size_t len, uint8_t out[32]) {
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                     0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    size_t full = len / 64;
    for (size_t i = 0; i < full; ++i) sha256_block(h, msg + 64 * i);
    uint8_t tail[128];
    size_t rem = len - 64 * full;
    memset(tail, 0, sizeof tail);
    memcpy(tail, msg + 64 * full, rem);
    tail[rem] = 0x80;
    size_t tail_len = (rem + 9 <= 64) ? 64 : 128;
    uint64_t bits = (uint64_t)len * 8;
    for (int i = 0; i < 8; ++i) tail[tail_len - 1 - i] = (uint8_t)(bits >> (8 * i));
    sha256_block(h, tail);
    if (tail_len == 128) sha256_block(h, tail + 64);
    for (int i = 0; i < 8; ++i) {
        out[4 * i] = (uint8_t)(h[i] >> 24);
        out[4 * i + 1] = (uint8_t)(h[i] >> 16);
        out[4 * i + 2] = (uint8_t)(h[i] >> 8);
        out[4 * i + 3] = (uint8_t)h[i];
    }
}

static void narg_append(uint8_t** buf, size_t* len, size_t* cap, const uint8_t* bytes, size_t n) {
    if (*len + n > *cap) {
        size_t c = *cap ? *cap : 256;
        while (c < *len + n) c *= 2;
        *buf = (uint8_t*)realloc(*buf, c);
        *cap = c;
    }
    memcpy(*buf + *len, bytes, n);
    *len += n;
}

/* ---- seeded ------------------------------------------------------------------ */
uint64_t oracle_splitmix_next(uint64_t* state) {
    *state += 0x9E3779B97F4A7C15ULL;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void seeded_absorb(void* ctx, int kind, const uint8_t* bytes, size_t len) {
    oracle_seeded* s = (oracle_seeded*)ctx;
    s->absorbed_bytes += len;
    if (kind == 1) narg_append(&s->narg, &s->narg_len, &s->narg_cap, bytes, len);
}

static void seeded_squeeze(void* ctx, uint8_t out[16]) {
    oracle_seeded* s = (oracle_seeded*)ctx;
    uint64_t w0 = oracle_splitmix_next(&s->state);
    uint64_t w1 = oracle_splitmix_next(&s->state);
    memcpy(out, &w0, 8);
    memcpy(out + 8, &w1, 8);
}

void oracle_seeded_init(oracle_seeded* s, uint64_t seed) {
    memset(s, 0, sizeof *s);
    s->state = seed;
}

void oracle_seeded_free(oracle_seeded* s) {
    free(s->narg);
    s->narg = NULL;
}

oracle_transcript oracle_seeded_transcript(oracle_seeded* s) {
    oracle_transcript t = {s, seeded_absorb, seeded_squeeze, NULL};
    return t;
}

/* ---- SHA-256 duplex ----------------------------------------------------------
 * absorb(kind, m): state = SHA256(state || "A" || kind || le64(len) || m)
 * squeeze:         out = SHA256(state || "S")[0..16];
 *                  state = SHA256(state || "R") */
static void duplex_absorb(void* ctx, int kind, const uint8_t* bytes, size_t len) {
    oracle_duplex* d = (oracle_duplex*)ctx;
    size_t n = 32 + 2 + 8 + len;
    uint8_t* buf = (uint8_t*)malloc(n);
    memcpy(buf, d->state, 32);
    buf[32] = 'A';
    buf[33] = (uint8_t)kind;
    uint64_t l = len;
    memcpy(buf + 34, &l, 8);
    if (len) memcpy(buf + 42, bytes, len);
    oracle_sha256(buf, n, d->state);
    free(buf);
    if (kind == 1) narg_append(&d->narg, &d->narg_len, &d->narg_cap, bytes, len);
}

static void duplex_squeeze(void* ctx, uint8_t out[16]) {
    oracle_duplex* d = (oracle_duplex*)ctx;
    uint8_t buf[33], h[32];
    memcpy(buf, d->state, 32);
    buf[32] = 'S';
    oracle_sha256(buf, 33, h);
    memcpy(out, h, 16);
    buf[32] = 'R';
    oracle_sha256(buf, 33, d->state);
}

void oracle_duplex_init(oracle_duplex* d, const uint8_t* domain, size_t len) {
    memset(d, 0, sizeof *d);
    oracle_sha256(domain, len, d->state);
}

void oracle_duplex_free(oracle_duplex* d) {
    free(d->narg);
    d->narg = NULL;
}

oracle_transcript oracle_duplex_transcript(oracle_duplex* d) {
    oracle_transcript t = {d, duplex_absorb, duplex_squeeze, NULL};
    return t;
}

/* ---- BLAKE3 (one chunk, one block) and proof-of-work grinding -------------
 * The reference grinds with crates/transcript/src/pow.rs:21-52: nonce = the
 * first n in 0, 1, ... with BLAKE3(POW_HASH_TAG || seed || le64(n)) whose
 * first four bytes, read big-endian, have at least `bits` leading zeros.
 * BLAKE3 (the published specification): an input of at most 64 bytes is one
 * block of the first chunk, compressed once from the IV with counter 0,
 * block length = input length, flags CHUNK_START | CHUNK_END | ROOT. */
static const uint32_t B3_IV[8] = {0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A,
                                  0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19};
static const uint8_t B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
static void b3_g(uint32_t* s, int a, int b, int c, int d, uint32_t x, uint32_t y) {
    s[a] = s[a] + s[b] + x;
    s[d] = rotr(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + y;
    s[d] = rotr(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 7);
}
int oracle_blake3_small(const uint8_t* in, size_t len, uint8_t out[32]) {
    if (len > 64) return ORACLE_ERR_ARG;
    uint8_t block[64];
    memset(block, 0, sizeof block);
    if (len) memcpy(block, in, len);
    uint32_t m[16], s[16];
    for (int i = 0; i < 16; ++i)
        m[i] = (uint32_t)block[4 * i] | ((uint32_t)block[4 * i + 1] << 8) | ((uint32_t)block[4 * i + 2] << 16) |
               ((uint32_t)block[4 * i + 3] << 24);
    for (int i = 0; i < 8; ++i) s[i] = B3_IV[i];
    for (int i = 0; i < 4; ++i) s[8 + i] = B3_IV[i];
    s[12] = 0;
    s[13] = 0;
    s[14] = (uint32_t)len;
    s[15] = 1u | 2u | 8u;
    for (int round = 0; round < 7; ++round) {
        b3_g(s, 0, 4, 8, 12, m[0], m[1]);
        b3_g(s, 1, 5, 9, 13, m[2], m[3]);
        b3_g(s, 2, 6, 10, 14, m[4], m[5]);
        b3_g(s, 3, 7, 11, 15, m[6], m[7]);
        b3_g(s, 0, 5, 10, 15, m[8], m[9]);
        b3_g(s, 1, 6, 11, 12, m[10], m[11]);
        b3_g(s, 2, 7, 8, 13, m[12], m[13]);
        b3_g(s, 3, 4, 9, 14, m[14], m[15]);
        uint32_t p[16];
        for (int i = 0; i < 16; ++i) p[i] = m[B3_PERM[i]];
        memcpy(m, p, sizeof m);
    }
    for (int i = 0; i < 8; ++i) {
        uint32_t w = s[i] ^ s[i + 8];
        out[4 * i] = (uint8_t)w;
        out[4 * i + 1] = (uint8_t)(w >> 8);
        out[4 * i + 2] = (uint8_t)(w >> 16);
        out[4 * i + 3] = (uint8_t)(w >> 24);
    }
    return ORACLE_OK;
}

static const char POW_HASH_TAG[] = "bitz-pcs-pow-v1";              /* pow.rs:5 */
static const char POW_TRANSCRIPT_TAG[] = "bitz-transcript-pow-v1"; /* pow.rs:6 */

/* pow.rs:38-52 valid. */
int oracle_pow_valid(const uint8_t seed[16], uint64_t nonce, uint32_t bits) {
    if (bits > 32) return 0;
    if (bits == 0) return nonce == 0;
    uint8_t msg[15 + 16 + 8], h[32];
    memcpy(msg, POW_HASH_TAG, 15);
    memcpy(msg + 15, seed, 16);
    for (int i = 0; i < 8; ++i) msg[31 + i] = (uint8_t)(nonce >> (8 * i));
    oracle_blake3_small(msg, sizeof msg, h);
    uint32_t first = ((uint32_t)h[0] << 24) | ((uint32_t)h[1] << 16) | ((uint32_t)h[2] << 8) | h[3];
    uint32_t lz = 0;
    while (lz < 32 && !(first & (0x80000000u >> lz))) ++lz;
    return lz >= bits;
}

/* pow.rs:21-33 find. */
uint64_t oracle_pow_find(const uint8_t seed[16], uint32_t bits) {
    uint64_t nonce = 0;
    while (!oracle_pow_valid(seed, nonce, bits)) ++nonce;
    return nonce;
}

/* crates/common/src/security.rs:25-32 grinding_bits: ceil(log2(c)) + target - 128, at least 0. */
uint32_t oracle_grinding_bits(uint32_t security_bits, uint64_t coefficient) {
    if (coefficient == 0) return 0;
    uint32_t log_c = 0;
    while (((uint64_t)1 << log_c) < coefficient) ++log_c;
    return security_bits + log_c > 128 ? security_bits + log_c - 128 : 0;
}

/* ProverState::grind (crates/transcript/src/prover.rs:29-41) over a
 * transcript's own calls: the header of pow.rs:11-16 as public messages
 * (the tag, le64(label length), the label, le32(bits)), a squeezed seed, the
 * nonce as an 8-byte prover message.  A transcript with a grind hook hands
 * the whole step to it instead.  Zero bits leave the transcript untouched. */
void oracle_transcript_grind(oracle_transcript* t, const uint8_t* label, size_t len, uint32_t bits) {
    if (bits == 0) return;
    if (t->grind) {
        t->grind(t->ctx, label, len, bits);
        return;
    }
    uint8_t le[8];
    t->absorb(t->ctx, 0, (const uint8_t*)POW_TRANSCRIPT_TAG, sizeof POW_TRANSCRIPT_TAG - 1);
    for (int i = 0; i < 8; ++i) le[i] = (uint8_t)((uint64_t)len >> (8 * i));
    t->absorb(t->ctx, 0, le, 8);
    t->absorb(t->ctx, 0, label, len);
    for (int i = 0; i < 4; ++i) le[i] = (uint8_t)(bits >> (8 * i));
    t->absorb(t->ctx, 0, le, 4);
    uint8_t seed[16];
    t->squeeze(t->ctx, seed);
    uint64_t nonce = oracle_pow_find(seed, bits);
    for (int i = 0; i < 8; ++i) le[i] = (uint8_t)(nonce >> (8 * i));
    t->absorb(t->ctx, 1, le, 8);
}

/* ---- Keccak-f[1600] and the Keccak duplex ---------------------------------
 * FIPS 202 section 3: the round constants from the LFSR rc(t) (x^8 + x^6 +
 * x^5 + x^4 + 1), RC_i bit 2^j - 1 = rc(j + 7 i); the rho offsets from the
 * walk (x, y) = (1, 0), offset (t + 1)(t + 2) / 2, (x, y) <- (y, 2x + 3y). */
static uint64_t K_RC[24];
static int K_RHO[25];
static int k_ready = 0;
static int k_rc_bit(int t) {
    if (t % 255 == 0) return 1;
    unsigned r = 1; /* R[i] at bit i; R = 10000000 */
    for (int i = 1; i <= t % 255; ++i) {
        r <<= 1; /* R = 0 || R; then R[0], R[4], R[5], R[6] ^= R[8]; Trunc8 */
        if (r & 0x100) r ^= 0x171;
    }
    return r & 1;
}
static void k_init(void) {
    if (k_ready) return;
    for (int i = 0; i < 24; ++i) {
        uint64_t v = 0;
        for (int j = 0; j <= 6; ++j)
            if (k_rc_bit(j + 7 * i)) v |= (uint64_t)1 << ((1 << j) - 1);
        K_RC[i] = v;
    }
    K_RHO[0] = 0;
    int x = 1, y = 0;
    for (int t = 0; t < 24; ++t) {
        K_RHO[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
        int nx = y, ny = (2 * x + 3 * y) % 5;
        x = nx;
        y = ny;
    }
    k_ready = 1;
}
static uint64_t k_rot(uint64_t v, int n) { return n ? (v << n) | (v >> (64 - n)) : v; }
void oracle_keccak_f1600(uint64_t a[25]) {
    k_init();
    for (int round = 0; round < 24; ++round) {
        uint64_t c[5], d[5], b[25];
        for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
        for (int x = 0; x < 5; ++x) d[x] = c[(x + 4) % 5] ^ k_rot(c[(x + 1) % 5], 1);
        for (int i = 0; i < 25; ++i) a[i] ^= d[i % 5];
        for (int x = 0; x < 5; ++x)
            for (int y = 0; y < 5; ++y) b[y + 5 * ((2 * x + 3 * y) % 5)] = k_rot(a[x + 5 * y], K_RHO[x + 5 * y]);
        for (int i = 0; i < 25; ++i) {
            int x = i % 5, y = i / 5;
            a[i] = b[i] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
        }
        a[0] ^= K_RC[round];
    }
}
/* A one-message sponge (pad 0x06: SHA3, 0x1F: SHAKE), for the known answers. */
void oracle_keccak_hash(const uint8_t* msg, size_t n, int rate, uint8_t pad, uint8_t* out, size_t out_len) {
    uint64_t a[25];
    memset(a, 0, sizeof a);
    uint8_t* s = (uint8_t*)a; /* little-endian lanes */
    size_t pos = 0;
    for (size_t i = 0; i < n; ++i) {
        s[pos] ^= msg[i];
        if (++pos == (size_t)rate) {
            oracle_keccak_f1600(a);
            pos = 0;
        }
    }
    s[pos] ^= pad;
    s[rate - 1] ^= 0x80;
    oracle_keccak_f1600(a);
    size_t o = 0;
    while (o < out_len) {
        for (int i = 0; i < rate && o < out_len; ++i, ++o) out[o] = s[i];
        if (o < out_len) oracle_keccak_f1600(a);
    }
}
/* The Keccak duplex transcript (rate 136): start 'D' || le64(len) || domain;
 * messages ('P' | 'N') || le64(len) || m; squeeze 'S', pad 0x1F .. 0x80,
 * permute, the first 16 bytes. */
static void kd_bytes(oracle_keccak_duplex* d, const uint8_t* p, size_t n) {
    uint8_t* s = (uint8_t*)d->a;
    for (size_t i = 0; i < n; ++i) {
        s[d->pos] ^= p[i];
        if (++d->pos == 136) {
            oracle_keccak_f1600(d->a);
            d->pos = 0;
        }
    }
}
static void kd_framed(oracle_keccak_duplex* d, uint8_t tag, const uint8_t* p, size_t n) {
    uint8_t head[9];
    head[0] = tag;
    for (int i = 0; i < 8; ++i) head[1 + i] = (uint8_t)((uint64_t)n >> (8 * i));
    kd_bytes(d, head, 9);
    kd_bytes(d, p, n);
}
static void kd_absorb(void* ctx, int kind, const uint8_t* bytes, size_t len) {
    oracle_keccak_duplex* d = (oracle_keccak_duplex*)ctx;
    kd_framed(d, kind == 0 ? 'P' : 'N', bytes, len);
    if (kind == 1) narg_append(&d->narg, &d->narg_len, &d->narg_cap, bytes, len);
}
static void kd_squeeze(void* ctx, uint8_t out[16]) {
    oracle_keccak_duplex* d = (oracle_keccak_duplex*)ctx;
    const uint8_t tag = 'S';
    kd_bytes(d, &tag, 1);
    uint8_t* s = (uint8_t*)d->a;
    s[d->pos] ^= 0x1F;
    s[135] ^= 0x80;
    oracle_keccak_f1600(d->a);
    d->pos = 0;
    memcpy(out, s, 16);
}
void oracle_keccak_duplex_init(oracle_keccak_duplex* d, const uint8_t* domain, size_t len) {
    memset(d, 0, sizeof *d);
    kd_framed(d, 'D', domain, len);
}
void oracle_keccak_duplex_free(oracle_keccak_duplex* d) {
    free(d->narg);
    d->narg = NULL;
}
oracle_transcript oracle_keccak_duplex_transcript(oracle_keccak_duplex* d) {
    oracle_transcript t = {d, kd_absorb, kd_squeeze, NULL};
    return t;
}
