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

// Keccak-f[1600] (FIPS 202 section 3.2) on one wave, with cross-lane moves
// that stay off LDS except for pi.  Lane t = 8 y + j holds word (x, y) of the
// state, x = j mod 5: lanes j = 5, 6, 7 of each group of eight repeat
// x = 0, 1, 2, so that a row's neighbours x + 1, x + 2 (chi) and x - 1,
// x + 1 (theta's D) are DPP moves within a 16-lane row; lanes 40..63 hold
// zero.  Per round: theta's column parity C[x] is the XOR over the lanes of
// equal j (a row rotation by 8, then the permlane32 and permlane16 swaps);
// D = C[x - 1] ^ rot(C[x + 1], 1) from the lanes four on / one back (x = 0)
// and one on; rho is each lane's own rotation; pi is one ds_bpermute from
// lane 8 y' + x' of the source word, which also refills the repeats; chi
// reads the next two lanes; iota goes to lanes 0 and 5 (word 0 and its
// repeat).  The repeats in lanes j = 6, 7 go stale in chi, but only their
// values after pi are ever read.  About 60 instructions and one LDS round
// trip a round, against 20 ds_bpermutes for one word per lane.
constexpr uint64_t keccak_rho_pack(int part) {
    constexpr int rho[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
    uint64_t v = 0;
    for (int i = 0; i < 10 && 10 * part + i < 25; ++i) v |= (uint64_t)rho[10 * part + i] << (6 * i);
    return v;
}

// The state word lane t holds (lanes 40..63 hold none).
__device__ __forceinline__ int keccak_lane_word(int t) { return (t & 7) % 5 + 5 * min(t >> 3, 4); }
__device__ __forceinline__ bool keccak_lane_owns(int t) { return t < 40 && (t & 7) < 5; }

template <int Ctrl>
__device__ __forceinline__ uint32_t dpp_move(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, Ctrl, 0xF, 0xF, false);
}
// XOR of v over the lanes t' with t' mod 8 == t mod 8.
__device__ __forceinline__ uint32_t keccak_column_xor(uint32_t v) {
    v ^= dpp_move<0x128>(v);  // row_ror:8, the other group of the row
    const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    v = p[0] ^ p[1];
    const auto q = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return q[0] ^ q[1];
}

// Permutes this lane's word (lo, hi) of the state in the layout above; the
// whole wave must be active.  DPP moves run for every lane and masks pick
// among them: a move under a branch would read the disabled lanes as zero.
__device__ __forceinline__ void keccak_wave(uint32_t& lo, uint32_t& hi) {
    const int t = (int)(threadIdx.x & 63), j = t & 7, x = j % 5, y = min(t >> 3, 4), word = x + 5 * y;
    const uint64_t rpack = word < 10 ? keccak_rho_pack(0) : word < 20 ? keccak_rho_pack(1) : keccak_rho_pack(2);
    const int n = (int)((rpack >> (6 * (word % 10))) & 63), m = n & 31;
    // rot(w, n) = alignbit pairs of w's halves, swapped first when n >= 32;
    // n = 0 takes the swap too and alignbit by 0 swaps them back
    const uint32_t swap = 0u - (uint32_t)((n >= 32) != (m == 0));
    const uint32_t s = (uint32_t)((32 - m) & 31);
    const int src = (x + 3 * y) % 5 + 5 * x;  // pi: B[x, y] = A[x + 3 y, x] (FIPS 202 3.2.3)
    const int addr = 4 * (8 * (src / 5) + src % 5);
    const uint32_t j0 = 0u - (uint32_t)(j == 0);
    const uint32_t iota = 0u - (uint32_t)(t == 0 || t == 5), live = 0u - (uint32_t)(t < 40);
#pragma unroll 1
    for (int round = 0; round < 24; ++round) {
        const uint32_t cl = keccak_column_xor(lo), ch = keccak_column_xor(hi);
        const uint32_t pl = dpp_move<0x101>(cl), ph = dpp_move<0x101>(ch);  // C[x + 1]: row_shl:1
        const uint32_t ml = (dpp_move<0x104>(cl) & j0) | (dpp_move<0x111>(cl) & ~j0);  // C[x - 1]
        const uint32_t mh = (dpp_move<0x104>(ch) & j0) | (dpp_move<0x111>(ch) & ~j0);
        lo ^= ml ^ __builtin_amdgcn_alignbit(pl, ph, 31);
        hi ^= mh ^ __builtin_amdgcn_alignbit(ph, pl, 31);
        const uint32_t u = (lo & swap) | (hi & ~swap), v = (hi & swap) | (lo & ~swap);
        const uint32_t rh = __builtin_amdgcn_alignbit(u, v, s), rl = __builtin_amdgcn_alignbit(v, u, s);
        const uint32_t bl = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)rl);
        const uint32_t bh = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)rh);
        const uint64_t rc = kKeccakRoundConstants[round];
        lo = ((bl ^ (~dpp_move<0x101>(bl) & dpp_move<0x102>(bl))) ^ ((uint32_t)rc & iota)) & live;
        hi = ((bh ^ (~dpp_move<0x101>(bh) & dpp_move<0x102>(bh))) ^ ((uint32_t)(rc >> 32) & iota)) & live;
    }
}

// Wave 0 XORs `m` (its lane's word of a block, zero for none) into the state
// and permutes it; the owners write it back.  Returns the lane's new word.
__device__ __forceinline__ uint64_t fs_permute_lds(FsLds& f, uint64_t m) {
    const int t = (int)(threadIdx.x & 63), word = keccak_lane_word(t);
    const uint64_t a = t < 40 ? f.a[word] ^ m : 0ull;
    uint32_t lo = (uint32_t)a, hi = (uint32_t)(a >> 32);
    keccak_wave(lo, hi);
    const uint64_t w = ((uint64_t)hi << 32) | lo;
    if (keccak_lane_owns(t)) f.a[word] = w;
    return w;
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

// The permutation by wave 0.
__device__ __forceinline__ void fs_permute(FsLds& f) {
    if (threadIdx.x < 64) (void)fs_permute_lds(f, 0ull);
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

// Absorbs the message m[0, n) (n <= 4 elements) of `kind` and squeezes.  At
// position 0 (after any squeeze) every byte's place is fixed: 'P'/'N' and
// le64(16 n) at bytes 0..8, the elements' bytes at 9 .. 8 + 16 n, 'S' and the
// pad 0x1F after them and 0x80 at byte 135, so lane l of wave 0 builds state
// word l of that block from the payload words P_(l-1), P_(l-2) (one byte
// apart) and permutes.  Elsewhere every thread takes the byte path.  Called
// by all threads; m is read from wave 0 (from thread 0 on the byte path).
// Lanes of wave 0 read their state word before any writes it back: a
// wave's LDS reads and writes complete in order.
__device__ __forceinline__ g4 fs_message_squeeze(FsLds& f, int kind, const g4* m, int n) {
    if (f.pos != 0) {
        fs_message(f, kind, m, n);
        return fs_squeeze(f);
    }
    if (threadIdx.x < 64) {
        const int l = keccak_lane_word((int)threadIdx.x);  // this lane's word of the block
        uint64_t cur = 0, prev = 0;  // P_(l-1), P_(l-2), zero outside [0, 2 n)
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (i < 2 * n) {
                const g4& e = m[i >> 1];
                const uint64_t p = (uint64_t)e.w[2 * (i & 1)] | ((uint64_t)e.w[2 * (i & 1) + 1] << 32);
                if (i == l - 1) cur = p;
                if (i == l - 2) prev = p;
            }
        }
        uint64_t w = (cur << 8) | (prev >> 56);
        if (l == 0) w = (uint64_t)(kind == 0 ? 'P' : 'N') | ((uint64_t)(16 * n) << 8);
        if (l == 2 * n + 1) w |= ((uint64_t)'S' << 8) | (0x1Full << 16);
        if (l == (kKeccakRate - 1) >> 3) w |= 0x80ull << 56;
        const uint64_t a = fs_permute_lds(f, w);
        if (threadIdx.x < 2) {  // lanes 0, 1 hold words 0, 1
            f.out[2 * threadIdx.x] = (uint32_t)a;
            f.out[2 * threadIdx.x + 1] = (uint32_t)(a >> 32);
        }
    }
    __syncthreads();
    g4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.w[i] = f.out[i];
    return r;
}

// The two-element message [m0, m1] of `kind`, then the squeeze.
__device__ __forceinline__ g4 fs_message2_squeeze(FsLds& f, int kind, const g4& m0, const g4& m1) {
    const g4 m[2] = {m0, m1};
    return fs_message_squeeze(f, kind, m, 2);
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

// a[i] * b[i] for i < 4 by g_mul_rows, row i taking product i; every lane
// gets all four products.  The operands are the same in every lane; a row
// picks its pair with XOR masks.  The whole wave must be active.
__device__ __forceinline__ void g_mul4(const g4 (&a)[4], const g4 (&b)[4], g4 (&out)[4]) {
    const uint32_t g = (threadIdx.x >> 4) & 3u;
    const uint32_t m1 = 0u - (uint32_t)(g == 1), m2 = 0u - (uint32_t)(g == 2), m3 = 0u - (uint32_t)(g == 3);
    g4 x, y;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        x.w[k] = a[0].w[k] ^ ((a[0].w[k] ^ a[1].w[k]) & m1) ^ ((a[0].w[k] ^ a[2].w[k]) & m2) ^
                 ((a[0].w[k] ^ a[3].w[k]) & m3);
        y.w[k] = b[0].w[k] ^ ((b[0].w[k] ^ b[1].w[k]) & m1) ^ ((b[0].w[k] ^ b[2].w[k]) & m2) ^
                 ((b[0].w[k] ^ b[3].w[k]) & m3);
    }
    const g4 p = g_mul_rows(x, y);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = g_row(p, i);
}

}  // namespace bitz
