// Keccak-f[1600] (FIPS 202) and the transcript's Keccak duplex, one
// implementation for host code and device kernels.
//
// The duplex (a stand-in with the shape of spongefish's Keccak duplex sponge;
// spongefish's own framing is not reproduced): a 200-byte state, rate 136
// bytes, bytes XORed in at the current position with a permutation each time
// the rate fills.
//   start(domain):  absorb 'D' || le64(len) || domain
//   absorb(kind, m): absorb ('P' if kind == 0 else 'N') || le64(len) || m
//   squeeze():      absorb 'S', then pad (0x1F at the position, 0x80 at the
//                   rate's last byte), permute, output the state's first 16
//                   bytes and continue from position 0.
#pragma once

#include <stddef.h>
#include <stdint.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define BITZ_HD __host__ __device__
#else
#define BITZ_HD
#endif

namespace bitz {

constexpr int kKeccakRate = 136;

struct KeccakState {
    uint64_t a[25];  // lane (x, y) at a[x + 5 y]
    uint32_t pos;    // next byte of the rate to absorb into
    uint32_t pad_;
};

BITZ_HD inline uint64_t keccak_rotl(uint64_t x, int n) { return n ? (x << n) | (x >> (64 - n)) : x; }

// The 24 rounds of theta, rho, pi, chi, iota (FIPS 202 section 3.2).
BITZ_HD inline void keccak_f1600(uint64_t a[25]) {
    const uint64_t rc[24] = {
        0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
        0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
        0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
        0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
        0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
        0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};
    // rho offsets of lane x + 5 y
    const int rho[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
    for (int round = 0; round < 24; ++round) {
        uint64_t c[5], b[25];
        for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
        for (int x = 0; x < 5; ++x) {
            const uint64_t d = c[(x + 4) % 5] ^ keccak_rotl(c[(x + 1) % 5], 1);
            for (int y = 0; y < 5; ++y) a[x + 5 * y] ^= d;
        }
        // rho and pi: B[y, 2x + 3y] = rot(A[x, y], rho[x, y])
        for (int x = 0; x < 5; ++x)
            for (int y = 0; y < 5; ++y) b[y + 5 * ((2 * x + 3 * y) % 5)] = keccak_rotl(a[x + 5 * y], rho[x + 5 * y]);
        for (int y = 0; y < 5; ++y)
            for (int x = 0; x < 5; ++x)
                a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
        a[0] ^= rc[round];
    }
}

BITZ_HD inline void keccak_xor_byte(KeccakState& s, uint32_t i, uint8_t v) {
    s.a[i >> 3] ^= (uint64_t)v << (8 * (i & 7));
}

BITZ_HD inline void keccak_absorb(KeccakState& s, const uint8_t* p, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        keccak_xor_byte(s, s.pos, p[i]);
        if (++s.pos == kKeccakRate) {
            keccak_f1600(s.a);
            s.pos = 0;
        }
    }
}

BITZ_HD inline void keccak_absorb_framed(KeccakState& s, uint8_t tag, const uint8_t* p, size_t n) {
    uint8_t head[9];
    head[0] = tag;
    for (int i = 0; i < 8; ++i) head[1 + i] = (uint8_t)((uint64_t)n >> (8 * i));
    keccak_absorb(s, head, 9);
    keccak_absorb(s, p, n);
}

BITZ_HD inline void keccak_start(KeccakState& s, const uint8_t* domain, size_t n) {
    for (int i = 0; i < 25; ++i) s.a[i] = 0;
    s.pos = 0;
    s.pad_ = 0;
    keccak_absorb_framed(s, 'D', domain, n);
}

BITZ_HD inline void keccak_message(KeccakState& s, int kind, const uint8_t* p, size_t n) {
    keccak_absorb_framed(s, kind == 0 ? 'P' : 'N', p, n);
}

// 16 challenge bytes as two little-endian words (lo, hi).
BITZ_HD inline void keccak_squeeze(KeccakState& s, uint64_t out[2]) {
    const uint8_t tag = 'S';
    keccak_absorb(s, &tag, 1);
    keccak_xor_byte(s, s.pos, 0x1F);
    keccak_xor_byte(s, kKeccakRate - 1, 0x80);
    keccak_f1600(s.a);
    s.pos = 0;
    out[0] = s.a[0];
    out[1] = s.a[1];
}

// A sponge hash (pad byte 0x06 for SHA3, 0x1F for SHAKE) over one message:
// the permutation's check against FIPS 202 known answers.
BITZ_HD inline void keccak_hash(const uint8_t* msg, size_t n, int rate, uint8_t pad, uint8_t* out, size_t out_len) {
    uint64_t a[25] = {0};
    size_t pos = 0;
    for (size_t i = 0; i < n; ++i) {
        a[pos >> 3] ^= (uint64_t)msg[i] << (8 * (pos & 7));
        if (++pos == (size_t)rate) {
            keccak_f1600(a);
            pos = 0;
        }
    }
    a[pos >> 3] ^= (uint64_t)pad << (8 * (pos & 7));
    a[(rate - 1) >> 3] ^= (uint64_t)0x80 << (8 * ((rate - 1) & 7));
    keccak_f1600(a);
    size_t o = 0;
    while (o < out_len) {
        for (int i = 0; i < rate && o < out_len; ++i, ++o) out[o] = (uint8_t)(a[i >> 3] >> (8 * (i & 7)));
        if (o < out_len) keccak_f1600(a);
    }
}

}  // namespace bitz
