// GF(2^128) = F2[X]/(X^128 + X^7 + X^2 + X + 1), non-reflected basis: bit i
// of lo||hi is the coefficient of X^i (crates/field/src/gf128.rs:1-8, :52).
//
// Device side: elements live in four 32-bit VGPRs (w0 = bits 0..31, ...).
// The carry-less product is a comb over the multiplier's bits with the
// 256-bit accumulator shifted one bit per step; each masked XOR is a single
// v_bitop3_b32.  Reduction folds the high 128 bits with X^128 = X^7+X^2+X+1.
//
// Host side: PCLMULQDQ products for the drivers' scalar work (round
// polynomial reconstruction, Horner evaluation, eq factors, inverses).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

struct f128 {
    uint64_t lo, hi;
};

#if defined(__HIPCC__)

// (a & m) ^ c in one v_bitop3_b32: LUT index = a*4 + m*2 + c.
__device__ __forceinline__ uint32_t xor_and(uint32_t c, uint32_t a, uint32_t m) {
    return __builtin_amdgcn_bitop3_b32(a, m, c, 0x6A);
}
// a ^ b ^ c in one v_bitop3_b32.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
// High 32 bits of (hi:lo) << s, s in [1, 31]: v_alignbit_b32(hi, lo, 32 - s).
__device__ __forceinline__ uint32_t shl_pair(uint32_t hi, uint32_t lo, int s) {
    return __builtin_amdgcn_alignbit(hi, lo, 32 - s);
}

struct g4 {
    uint32_t w[4];
};

__device__ __forceinline__ g4 g_load(const f128* p) {
    uint4 v = *reinterpret_cast<const uint4*>(p);
    g4 r;
    r.w[0] = v.x; r.w[1] = v.y; r.w[2] = v.z; r.w[3] = v.w;
    return r;
}
// Load from device memory through a global-address-space pointer: the
// compiler then emits global loads, which (unlike flat loads of pointers it
// cannot place) do not hold up waits on LDS traffic.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ g4 g_load_hbm(const f128* p) {
    const u32x4 v = *(const __attribute__((address_space(1))) u32x4*)(p);
    g4 r;
    r.w[0] = v.x; r.w[1] = v.y; r.w[2] = v.z; r.w[3] = v.w;
    return r;
}
__device__ __forceinline__ uint4 u4_load_hbm(const uint4* p) {
    const u32x4 v = *(const __attribute__((address_space(1))) u32x4*)(p);
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void g_store_hbm(f128* p, const g4& a) {
    u32x4 v;
    v.x = a.w[0]; v.y = a.w[1]; v.z = a.w[2]; v.w = a.w[3];
    *(__attribute__((address_space(1))) u32x4*)(p) = v;
}
__device__ __forceinline__ void g_store(f128* p, const g4& a) {
    *reinterpret_cast<uint4*>(p) = make_uint4(a.w[0], a.w[1], a.w[2], a.w[3]);
}
__device__ __forceinline__ g4 g_zero() {
    g4 r;
    r.w[0] = r.w[1] = r.w[2] = r.w[3] = 0;
    return r;
}
__device__ __forceinline__ g4 g_one() {
    g4 r = g_zero();
    r.w[0] = 1;
    return r;
}
__device__ __forceinline__ g4 g_from(f128 x) {
    g4 r;
    r.w[0] = (uint32_t)x.lo; r.w[1] = (uint32_t)(x.lo >> 32);
    r.w[2] = (uint32_t)x.hi; r.w[3] = (uint32_t)(x.hi >> 32);
    return r;
}
__device__ __forceinline__ g4 g_add(const g4& a, const g4& b) {
    g4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.w[i] = a.w[i] ^ b.w[i];
    return r;
}
__device__ __forceinline__ g4 g_add3(const g4& a, const g4& b, const g4& c) {
    g4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.w[i] = xor3(a.w[i], b.w[i], c.w[i]);
    return r;
}

// Unreduced 256-bit value: eight words, least significant first.
struct g8 {
    uint32_t w[8];
};
__device__ __forceinline__ g8 w_zero() {
    g8 r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.w[i] = 0;
    return r;
}
__device__ __forceinline__ void w_acc(g8& acc, const g8& x) {
#pragma unroll
    for (int i = 0; i < 8; ++i) acc.w[i] ^= x.w[i];
}

// Carry-less 128x128 -> 256 product, accumulated into `c` (c ^= a*b).
// Comb: for bit j = 31..0 of every multiplier word, shift the accumulator
// left one bit, then XOR in a shifted by 32*w wherever bit j of b.w[w] is set.
__device__ __forceinline__ void clmul_acc(g8& c, const g4& a, const g4& b) {
    g8 t = w_zero();
#pragma unroll 4
    for (int j = 31; j >= 0; --j) {
#pragma unroll
        for (int i = 7; i >= 1; --i) t.w[i] = shl_pair(t.w[i], t.w[i - 1], 1);
        t.w[0] <<= 1;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int)b.w[w], j, 1);
#pragma unroll
            for (int i = 0; i < 4; ++i) t.w[w + i] = xor_and(t.w[w + i], a.w[i], m);
        }
    }
    w_acc(c, t);
}

__device__ __forceinline__ g8 clmul(const g4& a, const g4& b) {
    g8 c = w_zero();
    clmul_acc(c, a, b);
    return c;
}

// Reduce modulo X^128 + X^7 + X^2 + X + 1 (the fold of portable.rs:50-57,
// in 32-bit words): H = high 128 bits, H*g = H ^ H<<1 ^ H<<2 ^ H<<7 spills at
// most 7 bits above X^128, folded once more into word 0.
__device__ __forceinline__ g4 g_reduce(const g8& c) {
    const uint32_t h0 = c.w[4], h1 = c.w[5], h2 = c.w[6], h3 = c.w[7];
    g4 r;
    r.w[0] = xor3(c.w[0] ^ h0, h0 << 1, xor3(h0 << 2, h0 << 7, 0u));
    r.w[1] = xor3(c.w[1] ^ h1, shl_pair(h1, h0, 1), xor3(shl_pair(h1, h0, 2), shl_pair(h1, h0, 7), 0u));
    r.w[2] = xor3(c.w[2] ^ h2, shl_pair(h2, h1, 1), xor3(shl_pair(h2, h1, 2), shl_pair(h2, h1, 7), 0u));
    r.w[3] = xor3(c.w[3] ^ h3, shl_pair(h3, h2, 1), xor3(shl_pair(h3, h2, 2), shl_pair(h3, h2, 7), 0u));
    const uint32_t s = xor3(h3 >> 31, h3 >> 30, h3 >> 25);  // < 2^7
    r.w[0] ^= xor3(s, s << 1, xor3(s << 2, s << 7, 0u));
    return r;
}

__device__ __forceinline__ g4 g_mul(const g4& a, const g4& b) { return g_reduce(clmul(a, b)); }

// x ^ (x of the lane `ctrl` names), a DPP move within a row of 16 lanes.
template <int Ctrl>
__device__ __forceinline__ uint32_t xor_dpp(uint32_t x) {
    return x ^ (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, Ctrl, 0xF, 0xF, false);
}

// Four products in one wave: the 16 lanes of row g (lanes 16 g .. 16 g + 15)
// multiply that row's a by its b, and every lane of the row gets the
// product.  Lane j of a row adds a X^(32 w + j) and a X^(32 w + j + 16) for
// the words w of b whose bits j, j + 16 are set (8 shifted copies of a);
// the row's XOR is four DPP steps (quad swaps, then half-row and row
// mirrors), with no trip through LDS; then the reduction.
__device__ __forceinline__ g4 g_mul_rows(const g4& a, const g4& b) {
    const int j = (int)(threadIdx.x & 15);
    g8 t = w_zero();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int sh = j + 16 * h;  // 0 .. 31, this lane's bit
        // a << sh over five words
        uint32_t s[5];
        s[0] = a.w[0] << sh;
#pragma unroll
        for (int i = 1; i < 4; ++i) s[i] = sh ? shl_pair(a.w[i], a.w[i - 1], sh) : a.w[i];
        s[4] = sh ? a.w[3] >> (32 - sh) : 0u;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int)b.w[w], (uint32_t)sh, 1u);  // bit sh of b.w[w], spread
#pragma unroll
            for (int i = 0; i < 5; ++i)
                if (w + i < 8) t.w[w + i] = xor_and(t.w[w + i], s[i], m);
        }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        uint32_t x = t.w[i];
        x = xor_dpp<0xB1>(x);   // quad_perm [1, 0, 3, 2]
        x = xor_dpp<0x4E>(x);   // quad_perm [2, 3, 0, 1]
        x = xor_dpp<0x141>(x);  // row_half_mirror: the other quad of the half row
        x = xor_dpp<0x140>(x);  // row_mirror: the other half of the row
        t.w[i] = x;
    }
    return g_reduce(t);
}

// Row g's value of v (lane 16 g), in every lane.
__device__ __forceinline__ g4 g_row(const g4& v, int g) {
    g4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.w[i] = (uint32_t)__builtin_amdgcn_readlane((int)v.w[i], 16 * g);
    return r;
}

#endif  // __HIPCC__

// ---------------------------------------------------------------- host -----
// Host-only scalar arithmetic.  The device compilation pass parses (but never
// emits) host code, so it sees a shift/XOR body instead of PCLMULQDQ.
#if !defined(__HIP_DEVICE_COMPILE__)
#include <wmmintrin.h>
#endif

namespace hostf {

inline f128 add(f128 a, f128 b) { return f128{a.lo ^ b.lo, a.hi ^ b.hi}; }
inline bool eq(f128 a, f128 b) { return a.lo == b.lo && a.hi == b.hi; }
inline bool is_zero(f128 a) { return (a.lo | a.hi) == 0; }
inline f128 one() { return f128{1, 0}; }

inline f128 reduce4(uint64_t r0, uint64_t r1, uint64_t r2, uint64_t r3) {
    auto times_g = [](uint64_t w, uint64_t& lo, uint64_t& hi) {
        lo = w ^ (w << 1) ^ (w << 2) ^ (w << 7);
        hi = (w >> 63) ^ (w >> 62) ^ (w >> 57);
    };
    uint64_t a0, a1, b0, b1;
    times_g(r2, a0, a1);
    times_g(r3, b0, b1);
    return f128{r0 ^ a0 ^ b1 ^ (b1 << 1) ^ (b1 << 2) ^ (b1 << 7), r1 ^ a1 ^ b0};
}

#if defined(__HIP_DEVICE_COMPILE__)
inline f128 mul(f128 a, f128 b) {
    uint64_t r[4] = {0, 0, 0, 0};
    for (int i = 0; i < 128; ++i)
        if (((i < 64 ? b.lo >> i : b.hi >> (i - 64)) & 1)) {
            const int w = i / 64, s = i % 64;
            r[w] ^= a.lo << s;
            r[w + 1] ^= (s ? a.lo >> (64 - s) : 0) ^ (a.hi << s);
            r[w + 2] ^= s ? a.hi >> (64 - s) : 0;
        }
    return reduce4(r[0], r[1], r[2], r[3]);
}
#else
__attribute__((target("pclmul,sse2"))) inline f128 mul(f128 a, f128 b) {
    __m128i x = _mm_set_epi64x((long long)a.hi, (long long)a.lo);
    __m128i y = _mm_set_epi64x((long long)b.hi, (long long)b.lo);
    __m128i l = _mm_clmulepi64_si128(x, y, 0x00);
    __m128i h = _mm_clmulepi64_si128(x, y, 0x11);
    __m128i m = _mm_xor_si128(_mm_clmulepi64_si128(x, y, 0x10), _mm_clmulepi64_si128(x, y, 0x01));
    uint64_t r0 = (uint64_t)_mm_cvtsi128_si64(l);
    uint64_t r1 = (uint64_t)_mm_cvtsi128_si64(_mm_unpackhi_epi64(l, l)) ^ (uint64_t)_mm_cvtsi128_si64(m);
    uint64_t r2 = (uint64_t)_mm_cvtsi128_si64(h) ^ (uint64_t)_mm_cvtsi128_si64(_mm_unpackhi_epi64(m, m));
    uint64_t r3 = (uint64_t)_mm_cvtsi128_si64(_mm_unpackhi_epi64(h, h));
    return reduce4(r0, r1, r2, r3);
}
#endif

inline f128 square(f128 a) { return mul(a, a); }

// a^-1 = a^(2^128 - 2) = (a^(2^127 - 1))^2 (a != 0), by Itoh and Tsujii's
// addition chain: with b_k = a^(2^k - 1), b_(i+j) = b_i^(2^j) b_j, walked
// along 1, 2, 3, 6, 7, 14, 15, 30, 31, 62, 63, 126, 127 (127 squarings, 12
// products).
inline f128 inv(f128 a) {
    auto sq_n = [](f128 x, int n) {
        for (int i = 0; i < n; ++i) x = square(x);
        return x;
    };
    f128 b = a;  // b_1
    int k = 1;
    static const int chain[] = {2, 3, 6, 7, 14, 15, 30, 31, 62, 63, 126, 127};
    for (int next : chain) {
        const int j = next - k;  // j == k (doubling) or j == 1
        b = mul(sq_n(b, j), j == k ? b : a);
        k = next;
    }
    return square(b);
}

// Horner from the top coefficient (crates/spartan/src/sumcheck.rs:784-794).
inline f128 horner(const f128* c, int n, f128 x) {
    f128 v{0, 0};
    for (int i = n - 1; i >= 0; --i) v = add(mul(v, x), c[i]);
    return v;
}

}  // namespace hostf

#if defined(__HIPCC__)
// ------------------------------------------------ multiply by a constant ---
// Multiplying by a fixed element c is F2-linear: c * x = XOR over the
// 4-bit windows x_w of T_w[x_w], with T_w[v] = c * v * X^(4w) reduced.

// c * X^j (j < 128): shift into 256 bits, then reduce.
__device__ __forceinline__ g4 g_mul_xpow(const g4& c, int j) {
    g8 t = w_zero();
    const int w = j >> 5, s = j & 31;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        // word i of c lands in words w+i (low part) and w+i+1 (carry)
        const uint32_t lo = s ? (c.w[i] << s) : c.w[i];
        const uint32_t hi = s ? (c.w[i] >> (32 - s)) : 0u;
        // runtime word index: select with a small unrolled switch
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k == w + i) t.w[k] ^= lo;
            if (k == w + i + 1) t.w[k] ^= hi;
        }
    }
    return g_reduce(t);
}

// 4-bit windows: 32 x 16 entries (8 KiB), entry (w, v) = c * v * X^(4w):
// cheap to copy into LDS once per round, and conflict-free (a window's 16
// entries cover the 64 banks once).
constexpr int kConstTable4Entries = 32 * 16;

// Builds c's 4-bit table in `tab` (uint4[512]), `basis` (uint4[128]) as
// scratch.  All threads of the block must call it; it ends with a barrier.
__device__ __forceinline__ void build_const_table4(uint4* tab, uint4* basis, const g4& c) {
    for (int j = threadIdx.x; j < 128; j += blockDim.x) {
        const g4 b = g_mul_xpow(c, j);
        basis[j] = make_uint4(b.w[0], b.w[1], b.w[2], b.w[3]);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kConstTable4Entries; e += blockDim.x) {
        const int w = e >> 4, v = e & 15;
        uint4 x = make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int bit = 0; bit < 4; ++bit)
            if ((v >> bit) & 1) {
                const uint4 y = basis[4 * w + bit];
                x.x ^= y.x; x.y ^= y.y; x.z ^= y.z; x.w ^= y.w;
            }
        tab[e] = x;
    }
    __syncthreads();
}

// c * X: one shift, X^128 folded back as X^7 + X^2 + X + 1.
__device__ __forceinline__ g4 g_mul_x(const g4& b) {
    g4 r;
    const uint32_t top = b.w[3] >> 31;
    r.w[3] = shl_pair(b.w[3], b.w[2], 1);
    r.w[2] = shl_pair(b.w[2], b.w[1], 1);
    r.w[1] = shl_pair(b.w[1], b.w[0], 1);
    r.w[0] = (b.w[0] << 1) ^ (0x87u & (0u - top));
    return r;
}

// c's 4-bit table (build_const_table4's layout) written by one wave straight
// to `out` (global or LDS), with no barrier: lane l builds window l / 2 from
// c X^(4 w) and its three doublings and writes entries 8 (l % 2) .. + 7 of
// it.  About 150 instructions against build_const_table4's two LDS passes.
__device__ __forceinline__ void const_table4_wave(uint4* out, const g4& c) {
    const int lane = (int)(threadIdx.x & 63), w = lane >> 1, h = lane & 1;
    g4 b[4];
    b[0] = g_mul_xpow(c, 4 * w);
    b[1] = g_mul_x(b[0]);
    b[2] = g_mul_x(b[1]);
    b[3] = g_mul_x(b[2]);
    const g4 base = h ? b[3] : g_zero();
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        g4 e = base;
        if (u & 1) e = g_add(e, b[0]);
        if (u & 2) e = g_add(e, b[1]);
        if (u & 4) e = g_add(e, b[2]);
        out[16 * w + 8 * h + u] = make_uint4(e.w[0], e.w[1], e.w[2], e.w[3]);
    }
}

// LDS accesses through 32-bit LDS byte addresses, so lookups can build
// their addresses with one VALU instruction each.
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
    return (uint32_t)(size_t)(const __attribute__((address_space(3))) char*)p;
}
__device__ __forceinline__ uint4 lds_ld16(uint32_t addr) {
    const u32x4 v = *(const lds_u32x4*)(size_t)addr;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void lds_st16(uint32_t addr, const g4& x) {
    u32x4 v;
    v.x = x.w[0]; v.y = x.w[1]; v.z = x.w[2]; v.w = x.w[3];
    *(lds_u32x4*)(size_t)addr = v;
}

// c * x through c's 4-bit table at `tab` (LDS, 256-byte aligned).  Window k
// (nibble k of x) reads tab + 256 k + 16 n: the masked words hold 16 n in
// byte j (even nibbles of a word from x << 4, odd ones from x), and one
// v_perm_b32 puts that byte under the table's upper address bytes; 256 k
// goes into the instruction's offset.
__device__ __forceinline__ g4 g_mul_table4(const uint4* tab, const g4& x) {
    const uint32_t tb = lds_addr(tab);
    g4 r = g_zero();
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint32_t hi = x.w[w] & 0xF0F0F0F0u, lo = (x.w[w] << 4) & 0xF0F0F0F0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k0 = 8 * w + 2 * j;
            const uint4 e0 = lds_ld16(__builtin_amdgcn_perm(tb, lo, 0x07060500u | j) + 256 * k0);
            const uint4 e1 = lds_ld16(__builtin_amdgcn_perm(tb, hi, 0x07060500u | j) + 256 * (k0 + 1));
            r.w[0] = xor3(r.w[0], e0.x, e1.x);
            r.w[1] = xor3(r.w[1], e0.y, e1.y);
            r.w[2] = xor3(r.w[2], e0.z, e1.z);
            r.w[3] = xor3(r.w[3], e0.w, e1.w);
        }
    }
    return r;
}

// ------------------------------------- general product through LDS ---
// a * b with a per-lane 3-bit window table of a in LDS: T[v] = a' v (v < 8)
// for a' = a without its top two bits (so every entry fits 128 bits); b is
// consumed three bits per lookup (eleven positions per word, the top one
// two bits wide), most significant position first, shifting the 256-bit
// accumulator by 3 between positions.  a's bits 126, 127 add
// (top * b) X^126.  `col` is this lane's column: entry v at col[64 v]
// (entry-major, one wave of 64 lanes per row), so a wave's lookups are one
// conflict-free ds_read_b128 each.  Entry 0 must hold zero (written once by
// the caller).  Three-bit windows keep a wave's table at 8 KiB, so twice as
// many waves fit in LDS as with four-bit windows, for 7 table writes (a
// ds_write_b128 costs three reads) and 44 lookups per product.
constexpr int kLaneTableEntries = 8;  // per lane; a wave's table is 8 KiB

__device__ __forceinline__ void lane_table_clear(uint4* col) { col[0] = make_uint4(0, 0, 0, 0); }

__device__ __forceinline__ g4 g_mul_lane(const g4& a, const g4& b, uint4* col) {
    const uint32_t ca = lds_addr(col);
    g4 al = a;
    al.w[3] &= 0x3FFFFFFFu;
    g4 s1, s2;
    s1.w[0] = al.w[0] << 1;
    s2.w[0] = al.w[0] << 2;
#pragma unroll
    for (int i = 1; i < 4; ++i) {
        s1.w[i] = shl_pair(al.w[i], al.w[i - 1], 1);
        s2.w[i] = shl_pair(al.w[i], al.w[i - 1], 2);
    }
    const g4 t3 = g_add(al, s1);
    lds_st16(ca + 1024 * 1, al);
    lds_st16(ca + 1024 * 2, s1);
    lds_st16(ca + 1024 * 3, t3);
    lds_st16(ca + 1024 * 4, s2);
    lds_st16(ca + 1024 * 5, g_add(al, s2));
    lds_st16(ca + 1024 * 6, g_add(s1, s2));
    lds_st16(ca + 1024 * 7, g_add(t3, s2));
    g8 acc = w_zero();
#pragma unroll
    for (int k = 10; k >= 0; --k) {
        if (k != 10) {
#pragma unroll
            for (int i = 7; i >= 1; --i) acc.w[i] = shl_pair(acc.w[i], acc.w[i - 1], 3);
            acc.w[0] <<= 3;
        }
        // Entry w lands on words w .. w+3: 16 words, XORed in 10 ops.
        uint4 e[4];
#pragma unroll
        for (int w = 0; w < 4; ++w)
            e[w] = lds_ld16((__builtin_amdgcn_ubfe(b.w[w], 3 * k, k == 10 ? 2 : 3) << 10) + ca);
        acc.w[0] ^= e[0].x;
        acc.w[1] = xor3(acc.w[1], e[0].y, e[1].x);
        acc.w[2] = xor3(xor3(acc.w[2], e[0].z, e[1].y), e[2].x, 0u);
        acc.w[3] = xor3(xor3(acc.w[3], e[0].w, e[1].z), e[2].y, e[3].x);
        acc.w[4] = xor3(xor3(acc.w[4], e[1].w, e[2].z), e[3].y, 0u);
        acc.w[5] = xor3(acc.w[5], e[2].w, e[3].z);
        acc.w[6] ^= e[3].w;
    }
    const uint32_t m0 = (uint32_t)__builtin_amdgcn_sbfe((int)a.w[3], 30, 1);
    const uint32_t m1 = (uint32_t)__builtin_amdgcn_sbfe((int)a.w[3], 31, 1);
    uint32_t u[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const uint32_t bi = i < 4 ? b.w[i] : 0u, bp = i > 0 ? b.w[i - 1] : 0u;
        u[i] = xor_and(bi & m0, shl_pair(bi, bp, 1), m1);
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) acc.w[3 + i] ^= __builtin_amdgcn_alignbit(u[i], i ? u[i - 1] : 0u, 2);
    return g_reduce(acc);
}

// a * b and a * c with one table of a: the two Horner chains run
// interleaved over the shared table.  Kept in this shape (the chains picked
// through a pointer array inside one position loop): written as two calls
// of shared helpers, the compiler scheduled further ahead and the round
// kernels spilled (up to 324 bytes of scratch, 60 more registers).
__device__ __forceinline__ void g_mul_lane2(const g4& a, const g4& b, const g4& c, uint4* col, g4& ab, g4& ac) {
    const uint32_t ca = lds_addr(col);
    g4 al = a;
    al.w[3] &= 0x3FFFFFFFu;
    g4 s1, s2;
    s1.w[0] = al.w[0] << 1;
    s2.w[0] = al.w[0] << 2;
#pragma unroll
    for (int i = 1; i < 4; ++i) {
        s1.w[i] = shl_pair(al.w[i], al.w[i - 1], 1);
        s2.w[i] = shl_pair(al.w[i], al.w[i - 1], 2);
    }
    const g4 t3 = g_add(al, s1);
    lds_st16(ca + 1024 * 1, al);
    lds_st16(ca + 1024 * 2, s1);
    lds_st16(ca + 1024 * 3, t3);
    lds_st16(ca + 1024 * 4, s2);
    lds_st16(ca + 1024 * 5, g_add(al, s2));
    lds_st16(ca + 1024 * 6, g_add(s1, s2));
    lds_st16(ca + 1024 * 7, g_add(t3, s2));
    g8 acc[2] = {w_zero(), w_zero()};
    const g4* ops[2] = {&b, &c};
#pragma unroll
    for (int k = 10; k >= 0; --k) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            g8& A = acc[q];
            if (k != 10) {
#pragma unroll
                for (int i = 7; i >= 1; --i) A.w[i] = shl_pair(A.w[i], A.w[i - 1], 3);
                A.w[0] <<= 3;
            }
            uint4 e[4];
#pragma unroll
            for (int w = 0; w < 4; ++w)
                e[w] = lds_ld16((__builtin_amdgcn_ubfe(ops[q]->w[w], 3 * k, k == 10 ? 2 : 3) << 10) + ca);
            A.w[0] ^= e[0].x;
            A.w[1] = xor3(A.w[1], e[0].y, e[1].x);
            A.w[2] = xor3(xor3(A.w[2], e[0].z, e[1].y), e[2].x, 0u);
            A.w[3] = xor3(xor3(A.w[3], e[0].w, e[1].z), e[2].y, e[3].x);
            A.w[4] = xor3(xor3(A.w[4], e[1].w, e[2].z), e[3].y, 0u);
            A.w[5] = xor3(A.w[5], e[2].w, e[3].z);
            A.w[6] ^= e[3].w;
        }
    }
    const uint32_t m0 = (uint32_t)__builtin_amdgcn_sbfe((int)a.w[3], 30, 1);
    const uint32_t m1 = (uint32_t)__builtin_amdgcn_sbfe((int)a.w[3], 31, 1);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const g4& bb = *ops[q];
        uint32_t u[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const uint32_t bi = i < 4 ? bb.w[i] : 0u, bp = i > 0 ? bb.w[i - 1] : 0u;
            u[i] = xor_and(bi & m0, shl_pair(bi, bp, 1), m1);
        }
#pragma unroll
        for (int i = 0; i < 5; ++i) acc[q].w[3 + i] ^= __builtin_amdgcn_alignbit(u[i], i ? u[i - 1] : 0u, 2);
    }
    ab = g_reduce(acc[0]);
    ac = g_reduce(acc[1]);
}
#endif  // __HIPCC__
