// Host-side transcripts behind the C ABI's bitz_transcript.
//
// The reference's Fiat–Shamir sponge is spongefish (crates/transcript), kept
// on the caller's side of the boundary: callers pass absorb/squeeze
// callbacks.  Two built-in sources serve benchmarks and tests:
//  - seeded: splitmix64(seed) challenges, F128::from([w0, w1]), the stream
//    the reference's GKR tests draw test data from (crates/gkr/src/lib.rs:986-999);
//  - duplex: a SHA-256 ratchet binding every challenge to every earlier message.
#include <cpuid.h>
#include <immintrin.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "transcript.hpp"

namespace bitz {

namespace {
const uint32_t K256[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

void compress(uint32_t h[8], const uint8_t* p) {
    uint32_t w[64];
    for (int i = 0; i < 16; ++i)
        w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 | (uint32_t)p[4 * i + 2] << 8 |
               (uint32_t)p[4 * i + 3];
    for (int i = 16; i < 64; ++i) {
        const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
        const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t v[8];
    memcpy(v, h, sizeof v);
    for (int i = 0; i < 64; ++i) {
        const uint32_t t1 = v[7] + (rotr(v[4], 6) ^ rotr(v[4], 11) ^ rotr(v[4], 25)) +
                            ((v[4] & v[5]) ^ (~v[4] & v[6])) + K256[i] + w[i];
        const uint32_t t2 = (rotr(v[0], 2) ^ rotr(v[0], 13) ^ rotr(v[0], 22)) +
                            ((v[0] & v[1]) ^ (v[0] & v[2]) ^ (v[1] & v[2]));
        memmove(v + 1, v, 7 * sizeof(uint32_t));
        v[4] += t1;
        v[0] = t1 + t2;
    }
    for (int i = 0; i < 8; ++i) h[i] += v[i];
}

// The same compression with the x86 SHA extensions (SHA256RNDS2 runs two
// rounds on the state split as ABEF / CDGH; SHA256MSG1/MSG2 extend the
// message schedule four words at a time), as Intel documents them.
__attribute__((target("sha,sse4.1,ssse3"))) void compress_shani(uint32_t h[8], const uint8_t* p, size_t blocks) {
    const __m128i bswap = _mm_set_epi64x(0x0c0d0e0f08090a0bLL, 0x0405060700010203LL);
    __m128i tmp = _mm_shuffle_epi32(_mm_loadu_si128(reinterpret_cast<const __m128i*>(h)), 0xB1);
    __m128i cdgh = _mm_shuffle_epi32(_mm_loadu_si128(reinterpret_cast<const __m128i*>(h + 4)), 0x1B);
    __m128i abef = _mm_alignr_epi8(tmp, cdgh, 8);
    cdgh = _mm_blend_epi16(cdgh, tmp, 0xF0);
    for (; blocks; --blocks, p += 64) {
        const __m128i abef0 = abef, cdgh0 = cdgh;
        __m128i w[4];
        for (int q = 0; q < 16; ++q) {
            if (q < 4) {
                w[q] = _mm_shuffle_epi8(_mm_loadu_si128(reinterpret_cast<const __m128i*>(p + 16 * q)), bswap);
            } else {
                __m128i x = _mm_sha256msg1_epu32(w[q & 3], w[(q + 1) & 3]);
                x = _mm_add_epi32(x, _mm_alignr_epi8(w[(q + 3) & 3], w[(q + 2) & 3], 4));
                w[q & 3] = _mm_sha256msg2_epu32(x, w[(q + 3) & 3]);
            }
            __m128i m = _mm_add_epi32(w[q & 3], _mm_loadu_si128(reinterpret_cast<const __m128i*>(K256 + 4 * q)));
            cdgh = _mm_sha256rnds2_epu32(cdgh, abef, m);
            abef = _mm_sha256rnds2_epu32(abef, cdgh, _mm_shuffle_epi32(m, 0x0E));
        }
        abef = _mm_add_epi32(abef, abef0);
        cdgh = _mm_add_epi32(cdgh, cdgh0);
    }
    tmp = _mm_shuffle_epi32(abef, 0x1B);
    cdgh = _mm_shuffle_epi32(cdgh, 0xB1);
    _mm_storeu_si128(reinterpret_cast<__m128i*>(h), _mm_blend_epi16(tmp, cdgh, 0xF0));
    _mm_storeu_si128(reinterpret_cast<__m128i*>(h + 4), _mm_alignr_epi8(cdgh, tmp, 8));
}

bool cpu_has_sha() {
    unsigned a, b, c, d;
    if (!__get_cpuid(1, &a, &b, &c, &d) || !(c & (1u << 19)) || !(c & (1u << 9))) return false;  // SSE4.1, SSSE3
    return __get_cpuid_count(7, 0, &a, &b, &c, &d) && (b & (1u << 29));                          // SHA
}

const bool kShaNi = cpu_has_sha() && !getenv("BITZ_NO_SHANI");

void compress_blocks(uint32_t h[8], const uint8_t* p, size_t blocks) {
    if (kShaNi) {
        compress_shani(h, p, blocks);
        return;
    }
    for (; blocks; --blocks, p += 64) compress(h, p);
}
}  // namespace

Sha256::Sha256() {
    static const uint32_t iv[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                   0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    memcpy(h, iv, sizeof h);
}

void Sha256::update(const void* data, size_t len) {
    const uint8_t* p = static_cast<const uint8_t*>(data);
    total += len;
    if (fill) {
        const size_t take = len < 64 - fill ? len : 64 - fill;
        memcpy(buf + fill, p, take);
        fill += take;
        p += take;
        len -= take;
        if (fill < 64) return;
        compress_blocks(h, buf, 1);
        fill = 0;
    }
    if (len >= 64) {
        compress_blocks(h, p, len / 64);
        p += len & ~(size_t)63;
        len &= 63;
    }
    memcpy(buf, p, len);
    fill = len;
}

void Sha256::update_u64(uint64_t v) {
    uint8_t b[8];
    for (int i = 0; i < 8; ++i) b[i] = (uint8_t)(v >> (8 * i));
    update(b, 8);
}

void Sha256::finish(uint8_t out[32]) {
    const uint64_t bits = total * 8;
    uint8_t tail[128] = {0};
    memcpy(tail, buf, fill);
    tail[fill] = 0x80;
    const size_t tl = fill + 9 <= 64 ? 64 : 128;
    for (int i = 0; i < 8; ++i) tail[tl - 1 - i] = (uint8_t)(bits >> (8 * i));
    compress_blocks(h, tail, tl / 64);
    for (int i = 0; i < 8; ++i) {
        out[4 * i] = (uint8_t)(h[i] >> 24);
        out[4 * i + 1] = (uint8_t)(h[i] >> 16);
        out[4 * i + 2] = (uint8_t)(h[i] >> 8);
        out[4 * i + 3] = (uint8_t)h[i];
    }
}

void sha256(const uint8_t* msg, size_t len, uint8_t out[32]) {
    Sha256 s;
    s.update(msg, len);
    s.finish(out);
}

uint64_t splitmix_next(uint64_t& x) {
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void Transcript::absorb(int kind, const uint8_t* bytes, size_t len) {
    if (kind == 1) narg.insert(narg.end(), bytes, bytes + len);
    switch (type) {
        case Type::kCallbacks:
            cb_absorb(cb_ctx, kind, bytes, len);
            break;
        case Type::kSeeded:
            break;
        case Type::kKeccak:
            keccak_message(keccak, kind, bytes, len);
            break;
        case Type::kDuplex: {
            std::vector<uint8_t> buf(32 + 2 + 8 + len);
            memcpy(buf.data(), state, 32);
            buf[32] = 'A';
            buf[33] = (uint8_t)kind;
            const uint64_t l = len;
            memcpy(buf.data() + 34, &l, 8);
            if (len) memcpy(buf.data() + 42, bytes, len);
            sha256(buf.data(), buf.size(), state);
            break;
        }
    }
}

f128 Transcript::squeeze() {
    uint8_t out[16];
    switch (type) {
        case Type::kCallbacks:
            cb_squeeze(cb_ctx, out);
            break;
        case Type::kSeeded: {
            const uint64_t w0 = splitmix_next(seed), w1 = splitmix_next(seed);
            memcpy(out, &w0, 8);
            memcpy(out + 8, &w1, 8);
            break;
        }
        case Type::kKeccak: {
            uint64_t w[2];
            keccak_squeeze(keccak, w);
            memcpy(out, w, 16);
            break;
        }
        case Type::kDuplex: {
            uint8_t buf[33], h[32];
            memcpy(buf, state, 32);
            buf[32] = 'S';
            sha256(buf, 33, h);
            memcpy(out, h, 16);
            buf[32] = 'R';
            sha256(buf, 33, state);
            break;
        }
    }
    f128 r;
    memcpy(&r.lo, out, 8);
    memcpy(&r.hi, out + 8, 8);
    return r;
}

namespace {
// BLAKE3 (the published specification) for inputs of one block: compressed
// once from the IV, counter 0, block length = input length, flags
// CHUNK_START | CHUNK_END | ROOT.
const uint32_t kB3Iv[8] = {0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A,
                           0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19};
const uint8_t kB3Perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
inline void b3_mix(uint32_t* s, int a, int b, int c, int d, uint32_t x, uint32_t y) {
    s[a] += s[b] + x;
    s[d] = rotr(s[d] ^ s[a], 16);
    s[c] += s[d];
    s[b] = rotr(s[b] ^ s[c], 12);
    s[a] += s[b] + y;
    s[d] = rotr(s[d] ^ s[a], 8);
    s[c] += s[d];
    s[b] = rotr(s[b] ^ s[c], 7);
}
const uint8_t kPowHashTag[] = "bitz-pcs-pow-v1";              // pow.rs:5
const uint8_t kPowTranscriptTag[] = "bitz-transcript-pow-v1"; // pow.rs:6
}  // namespace

bool blake3_small(const uint8_t* in, size_t len, uint8_t out[32]) {
    if (len > 64) return false;
    uint8_t block[64] = {0};
    if (len) memcpy(block, in, len);
    uint32_t m[16], st[16];
    for (int i = 0; i < 16; ++i) memcpy(&m[i], block + 4 * i, 4);  // little-endian words
    memcpy(st, kB3Iv, 32);
    memcpy(st + 8, kB3Iv, 16);
    st[12] = 0;
    st[13] = 0;
    st[14] = (uint32_t)len;
    st[15] = 1u | 2u | 8u;
    for (int round = 0; round < 7; ++round) {
        b3_mix(st, 0, 4, 8, 12, m[0], m[1]);
        b3_mix(st, 1, 5, 9, 13, m[2], m[3]);
        b3_mix(st, 2, 6, 10, 14, m[4], m[5]);
        b3_mix(st, 3, 7, 11, 15, m[6], m[7]);
        b3_mix(st, 0, 5, 10, 15, m[8], m[9]);
        b3_mix(st, 1, 6, 11, 12, m[10], m[11]);
        b3_mix(st, 2, 7, 8, 13, m[12], m[13]);
        b3_mix(st, 3, 4, 9, 14, m[14], m[15]);
        uint32_t p[16];
        for (int i = 0; i < 16; ++i) p[i] = m[kB3Perm[i]];
        memcpy(m, p, sizeof m);
    }
    for (int i = 0; i < 8; ++i) {
        const uint32_t w = st[i] ^ st[i + 8];
        memcpy(out + 4 * i, &w, 4);
    }
    return true;
}

bool pow_valid(const uint8_t seed[16], uint64_t nonce, uint32_t bits) {
    if (bits > kMaxGrindingBits) return false;
    if (bits == 0) return nonce == 0;
    uint8_t msg[15 + 16 + 8], h[32];
    memcpy(msg, kPowHashTag, 15);
    memcpy(msg + 15, seed, 16);
    memcpy(msg + 31, &nonce, 8);  // nonce.to_le_bytes()
    blake3_small(msg, sizeof msg, h);
    const uint32_t first = (uint32_t)h[0] << 24 | (uint32_t)h[1] << 16 | (uint32_t)h[2] << 8 | h[3];
    return (first == 0 ? 32u : (uint32_t)__builtin_clz(first)) >= bits;
}

uint64_t pow_find(const uint8_t seed[16], uint32_t bits) {
    uint64_t nonce = 0;
    while (!pow_valid(seed, nonce, bits)) ++nonce;
    return nonce;
}

uint32_t grinding_bits(uint32_t security_bits, uint64_t coefficient) {
    if (coefficient == 0) return 0;
    uint32_t log_c = 0;
    while (((uint64_t)1 << log_c) < coefficient) ++log_c;
    return security_bits + log_c > 128 ? security_bits + log_c - 128 : 0;
}

void Transcript::grind(const uint8_t* label, size_t len, uint32_t bits) {
    if (bits == 0) return;
    if (type == Type::kCallbacks && cb_grind) {
        cb_grind(cb_ctx, label, len, bits);
        return;
    }
    uint8_t le[8];
    absorb(0, kPowTranscriptTag, sizeof kPowTranscriptTag - 1);
    const uint64_t l = len;
    memcpy(le, &l, 8);
    absorb(0, le, 8);
    absorb(0, label, len);
    memcpy(le, &bits, 4);
    absorb(0, le, 4);
    const f128 seed = squeeze();
    uint8_t sb[16];
    memcpy(sb, &seed.lo, 8);
    memcpy(sb + 8, &seed.hi, 8);
    const uint64_t nonce = pow_find(sb, bits);
    memcpy(le, &nonce, 8);
    absorb(1, le, 8);
}

}  // namespace bitz
