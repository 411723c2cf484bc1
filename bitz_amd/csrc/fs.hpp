// Fiat-Shamir on the device: the Keccak duplex of keccak.hpp run by one
// workgroup, with its state in LDS, so a prover can draw each round's
// challenge right after the round's sums without a trip to the host.
//
// Every function is called by all threads of the workgroup (the work inside
// is split or left to thread 0) and ends with a barrier; messages are framed
// and absorbed exactly as keccak_message / keccak_squeeze do on the host.
#pragma once

#include "gf128.hpp"
#include "keccak.hpp"

namespace bitz {

constexpr int kFsMsgBytes = 9 + 4 * 16;  // a framed message of up to 4 elements

struct FsLds {
    uint64_t a[25];
    uint32_t pos, pad_;
    uint8_t msg[kFsMsgBytes + 7];
    uint32_t out[4];  // a challenge drawn by wave 0, for every thread
};

static __constant__ uint64_t kKeccakRoundConstants[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};

// rho and pi together as one walk: pi moves lane p = x + 5 y to
// y + 5 ((2 x + 3 y) mod 5) and, apart from lane 0, is a single cycle of 24
// lanes, so the step can run in place carrying one lane at a time (each
// rotated by rho of the lane it comes from).  The walk from lane 1, computed
// at compile time from the two step mappings.
struct KeccakPiWalk {
    int lane[24];  // the lane written at step t
    int rot[24];   // rho of the lane its value comes from
};
constexpr KeccakPiWalk keccak_pi_walk() {
    constexpr int rho[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
    KeccakPiWalk w{};
    int p = 1;
    for (int t = 0; t < 24; ++t) {
        const int x = p % 5, y = p / 5;
        const int q = y + 5 * ((2 * x + 3 * y) % 5);
        w.lane[t] = q;
        w.rot[t] = rho[p];
        p = q;
    }
    return w;
}

// Keccak-f[1600] (FIPS 202 section 3.2) in place: theta, the rho-pi walk,
// chi row by row, iota; 32 live words.  Called by a whole wave on the same
// (uniform) state, it runs on the scalar unit: 64-bit scalar XOR, AND-NOT
// and shifts at one instruction per clock.
__device__ __forceinline__ void fs_permute_regs(uint64_t (&a)[25]) {
    constexpr KeccakPiWalk walk = keccak_pi_walk();
#pragma unroll 1
    for (int round = 0; round < 24; ++round) {
        uint64_t c[5];
#pragma unroll
        for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
#pragma unroll
        for (int x = 0; x < 5; ++x) {
            const uint64_t d = c[(x + 4) % 5] ^ keccak_rotl(c[(x + 1) % 5], 1);
#pragma unroll
            for (int y = 0; y < 5; ++y) a[x + 5 * y] ^= d;
        }
        uint64_t carry = a[1];
#pragma unroll
        for (int t = 0; t < 24; ++t) {
            const uint64_t next = a[walk.lane[t]];
            a[walk.lane[t]] = keccak_rotl(carry, walk.rot[t]);
            carry = next;
        }
#pragma unroll
        for (int y = 0; y < 5; ++y) {
            uint64_t row[5];
#pragma unroll
            for (int x = 0; x < 5; ++x) row[x] = a[x + 5 * y];
#pragma unroll
            for (int x = 0; x < 5; ++x) a[x + 5 * y] = row[x] ^ (~row[(x + 1) % 5] & row[(x + 2) % 5]);
        }
        a[0] ^= kKeccakRoundConstants[round];
    }
}

// A 64-bit value every lane of the wave holds, as a scalar.
__device__ __forceinline__ uint64_t fs_uniform(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ void fs_load(FsLds& f, const KeccakState* g) {
    if (threadIdx.x < 25) f.a[threadIdx.x] = g->a[threadIdx.x];
    if (threadIdx.x == 0) f.pos = g->pos;
    __syncthreads();
}

// Writes the state back (after the last call that changes it).
__device__ __forceinline__ void fs_save(KeccakState* g, const FsLds& f) {
    if (threadIdx.x < 25) g->a[threadIdx.x] = f.a[threadIdx.x];
    if (threadIdx.x == 0) {
        g->pos = f.pos;
        g->pad_ = 0;
    }
}

// The permutation by wave 0 on scalars; the wave writes the state back
// (every lane the same words).
__device__ __forceinline__ void fs_permute(FsLds& f) {
    if (threadIdx.x < 64) {
        uint64_t a[25];
#pragma unroll
        for (int i = 0; i < 25; ++i) a[i] = fs_uniform(f.a[i]);
        fs_permute_regs(a);
        asm volatile("" ::: "memory");  // the stores stay after the rounds
#pragma unroll
        for (int i = 0; i < 25; ++i) f.a[i] = a[i];
    }
    __syncthreads();
}

// Absorbs f.msg[0, len) at the current position: one thread per state word
// the bytes land in, a permutation whenever the rate fills.
__device__ __forceinline__ void fs_absorb_msg(FsLds& f, int len) {
    int done = 0;
    while (done < len) {
        const int pos = (int)f.pos;
        const int take = min(len - done, kKeccakRate - pos);
        const int w = (pos >> 3) + (int)threadIdx.x;
        if (w <= ((pos + take - 1) >> 3)) {
            uint64_t m = 0;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const int q = 8 * w + b;
                if (q >= pos && q < pos + take) m |= (uint64_t)f.msg[done + q - pos] << (8 * b);
            }
            f.a[w] ^= m;
        }
        __syncthreads();  // every thread has read pos
        if (threadIdx.x == 0) f.pos = (uint32_t)(pos + take);
        __syncthreads();
        if (pos + take == kKeccakRate) {
            fs_permute(f);
            if (threadIdx.x == 0) f.pos = 0;
            __syncthreads();
        }
        done += take;
    }
}

// Thread 0: frames `n` elements as a message of `kind` (0 public, 1 prover)
// in f.msg ('P' or 'N', le64 byte length, the elements' bytes); returns the
// framed length.  The caller syncs, then absorbs.
__device__ __forceinline__ int fs_frame(FsLds& f, int kind, const g4* e, int n) {
    f.msg[0] = kind == 0 ? 'P' : 'N';
    const uint64_t len = 16ull * (uint64_t)n;
#pragma unroll
    for (int i = 0; i < 8; ++i) f.msg[1 + i] = (uint8_t)(len >> (8 * i));
    for (int k = 0; k < n; ++k)
#pragma unroll
        for (int i = 0; i < 16; ++i) f.msg[9 + 16 * k + i] = (uint8_t)(e[k].w[i >> 2] >> (8 * (i & 3)));
    return 9 + 16 * n;
}

// Absorbs a message of n <= 4 elements held by thread 0.
__device__ __forceinline__ void fs_message(FsLds& f, int kind, const g4* e, int n) {
    if (threadIdx.x == 0) fs_frame(f, kind, e, n);
    __syncthreads();
    fs_absorb_msg(f, 9 + 16 * n);
}

// The duplex squeeze ('S', pad 0x1F at the position and 0x80 at the rate's
// last byte, permute, the first 16 bytes): every thread gets the challenge.
__device__ __forceinline__ g4 fs_squeeze(FsLds& f) {
    if (threadIdx.x == 0) f.msg[0] = 'S';
    __syncthreads();
    fs_absorb_msg(f, 1);
    if (threadIdx.x == 0) {
        const uint32_t p = f.pos;
        f.a[p >> 3] ^= (uint64_t)0x1F << (8 * (p & 7));
        f.a[(kKeccakRate - 1) >> 3] ^= (uint64_t)0x80 << (8 * ((kKeccakRate - 1) & 7));
    }
    __syncthreads();
    fs_permute(f);
    g4 r;
    r.w[0] = (uint32_t)f.a[0];
    r.w[1] = (uint32_t)(f.a[0] >> 32);
    r.w[2] = (uint32_t)f.a[1];
    r.w[3] = (uint32_t)(f.a[1] >> 32);
    __syncthreads();  // every thread has read the output
    if (threadIdx.x == 0) f.pos = 0;
    __syncthreads();
    return r;
}

// Absorbs the two-element message [m0, m1] of `kind` and squeezes.  At
// position 0 (after any squeeze) every byte's place is fixed: the framing
// 'P'/'N' and le64(32) at bytes 0..8, the 32 bytes at 9..40, 'S' at 41, the
// pad 0x1F at 42 and 0x80 at 135, so wave 0 XORs seven words into the
// state on scalars and permutes.  Elsewhere every thread takes the byte
// path.  Called by all threads; m0, m1 are read from wave 0 (from thread 0
// on the byte path).
__device__ __forceinline__ g4 fs_message2_squeeze(FsLds& f, int kind, const g4& m0, const g4& m1) {
    if (f.pos != 0) {
        const g4 m[2] = {m0, m1};
        fs_message(f, kind, m, 2);
        return fs_squeeze(f);
    }
    if (threadIdx.x < 64) {
        auto word = [](const g4& x, int k) { return (uint64_t)x.w[2 * k] | ((uint64_t)x.w[2 * k + 1] << 32); };
        const uint64_t lo0 = fs_uniform(word(m0, 0)), hi0 = fs_uniform(word(m0, 1));
        const uint64_t lo1 = fs_uniform(word(m1, 0)), hi1 = fs_uniform(word(m1, 1));
        uint64_t a[25];
#pragma unroll
        for (int i = 0; i < 25; ++i) a[i] = fs_uniform(f.a[i]);
        a[0] ^= (uint64_t)(kind == 0 ? 'P' : 'N') | (32ull << 8);
        a[1] ^= lo0 << 8;
        a[2] ^= (lo0 >> 56) | (hi0 << 8);
        a[3] ^= (hi0 >> 56) | (lo1 << 8);
        a[4] ^= (lo1 >> 56) | (hi1 << 8);
        a[5] ^= (hi1 >> 56) | ((uint64_t)'S' << 8) | (0x1Full << 16);
        a[(kKeccakRate - 1) >> 3] ^= 0x80ull << 56;
        fs_permute_regs(a);
        asm volatile("" ::: "memory");
#pragma unroll
        for (int i = 0; i < 25; ++i) f.a[i] = a[i];
        f.out[0] = (uint32_t)a[0];
        f.out[1] = (uint32_t)(a[0] >> 32);
        f.out[2] = (uint32_t)a[1];
        f.out[3] = (uint32_t)(a[1] >> 32);
    }
    __syncthreads();
    g4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.w[i] = f.out[i];
    return r;
}

__device__ __forceinline__ bool g_is_zero(const g4& a) { return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) == 0; }

// a * b for one product held by a whole wave (the same a and b in every
// lane; every lane gets the product).  Lane (h, j), h = lane / 32,
// j = lane % 32, adds a X^(32 w + j) for the words w = 2 h, 2 h + 1 of b whose
// bit j is set; an XOR across the wave sums the 64 partial products, then
// the reduction.  About 12 times fewer instructions on the product's path
// than the comb run by one lane.
__device__ __forceinline__ g4 g_mul_wave(const g4& a, const g4& b) {
    const int lane = (int)(threadIdx.x & 63), j = lane & 31, h = lane >> 5;
    g8 t = w_zero();
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        if ((w >> 1) != h || !((b.w[w] >> j) & 1u)) continue;
        // a << j over five words, placed at word w
        uint32_t s[5];
        s[0] = a.w[0] << j;
#pragma unroll
        for (int i = 1; i < 4; ++i) s[i] = j ? (a.w[i] << j) | (a.w[i - 1] >> (32 - j)) : a.w[i];
        s[4] = j ? a.w[3] >> (32 - j) : 0u;
#pragma unroll
        for (int i = 0; i < 5; ++i)
            if (w + i < 8) t.w[w + i] ^= s[i];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
        for (int i = 0; i < 8; ++i) t.w[i] ^= (uint32_t)__shfl_xor((int)t.w[i], off, 64);
    return g_reduce(t);
}

}  // namespace bitz
