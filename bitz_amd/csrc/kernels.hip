// HIP kernels for the GF(2^128) sumcheck path on gfx950 (MI355X).
//
// Tables are arrays of 16-byte elements (the reference's F128 layout), read
// and written as one dwordx4 per lane so every wave moves 1 KiB contiguous.
// Round sums are XOR-reduced inside each wave with lane shuffles, across the
// waves of a workgroup through LDS, and across workgroups with one 32-bit
// atomic XOR per word into a per-round slot (XOR is associative and
// commutative, so the result does not depend on arrival order).
#include <hip/hip_runtime.h>

#include "gf128.hpp"
#include "keccak.hpp"
#include "fs.hpp"
#include "kernels.hpp"

namespace bitz {

// ---------------------------------------------------------------- helpers --
// Requires blockDim.x to be a multiple of 64 (whole waves), at most 1024.
template <int K>
__device__ __forceinline__ void block_xor_to_global(g4 (&v)[K], uint32_t* dst, int nthreads = blockDim.x) {
    __shared__ uint32_t red[1024 / 64][4 * K];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint32_t x = v[k].w[i];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) x ^= __shfl_xor(x, off, 64);
            v[k].w[i] = x;
        }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int i = 0; i < 4; ++i) red[wave][4 * k + i] = v[k].w[i];
    }
    __syncthreads();
    const int waves = nthreads >> 6;
    for (int j = threadIdx.x; j < 4 * K; j += nthreads) {
        uint32_t x = 0;
        for (int w = 0; w < waves; ++w) x ^= red[w][j];
        if (x)
            __scoped_atomic_fetch_xor((__attribute__((address_space(1))) uint32_t*)(dst + j), x, __ATOMIC_RELAXED,
                                      __MEMORY_SCOPE_DEVICE);
    }
}

// XOR of v over the workgroup, stored (not accumulated) at dst: for kernels
// that run as a single workgroup.  Whole waves only.
template <int K>
__device__ __forceinline__ void block_xor_store(g4 (&v)[K], uint32_t* dst, int nthreads = blockDim.x) {
    __shared__ uint32_t red[1024 / 64][4 * K];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint32_t x = v[k].w[i];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) x ^= __shfl_xor(x, off, 64);
            v[k].w[i] = x;
        }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int i = 0; i < 4; ++i) red[wave][4 * k + i] = v[k].w[i];
    }
    __syncthreads();
    const int waves = nthreads >> 6;
    for (int j = threadIdx.x; j < 4 * K; j += nthreads) {
        uint32_t x = 0;
        for (int w = 0; w < waves; ++w) x ^= red[w][j];
        dst[j] = x;
    }
}

// XOR of W words per thread over the workgroup, XORed into dst[0 .. W) with
// one atomic per word.  Inside a wave, each butterfly step sends half of the
// words a lane still carries to its partner and keeps the other half, so W
// words take W - 1 shuffles in all (instead of 6 W): after log2 W steps lane
// l holds the wave's XOR of one word, the index its lane bits chose, and the
// remaining offsets reduce that word over the lanes that share it.
// W a power of two <= 64; whole waves.
template <int C, int O, int W>
__device__ __forceinline__ void xor_scatter_step(uint32_t (&x)[W], int lane, int& idx) {
    if constexpr (C > 1) {
        const bool up = (lane & O) != 0;
#pragma unroll
        for (int i = 0; i < C / 2; ++i) {
            const uint32_t send = up ? x[i] : x[i + C / 2];
            const uint32_t keep = up ? x[i + C / 2] : x[i];
            x[i] = keep ^ (uint32_t)__shfl_xor((int)send, O, 64);
        }
        if (up) idx += C / 2;
        xor_scatter_step<C / 2, O / 2, W>(x, lane, idx);
    } else {
#pragma unroll
        for (int o = O; o >= 1; o >>= 1) x[0] ^= (uint32_t)__shfl_xor((int)x[0], o, 64);
    }
}
template <int W>
__device__ __forceinline__ void block_xor_words(uint32_t (&x)[W], uint32_t* dst) {
    static_assert(W >= 1 && W <= 64 && (W & (W - 1)) == 0, "power of two up to 64 words");
    __shared__ uint32_t red[1024 / 64][W];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int idx = 0;
    xor_scatter_step<W, 32, W>(x, lane, idx);
    if ((lane & (64 / W - 1)) == 0) red[wave][idx] = x[0];
    __syncthreads();
    const int waves = (int)blockDim.x >> 6;
    for (int j = threadIdx.x; j < W; j += blockDim.x) {
        uint32_t v = 0;
        for (int w = 0; w < waves; ++w) v ^= red[w][j];
        if (v)
            __scoped_atomic_fetch_xor((__attribute__((address_space(1))) uint32_t*)(dst + j), v, __ATOMIC_RELAXED,
                                      __MEMORY_SCOPE_DEVICE);
    }
}

__device__ __forceinline__ g4 interp(const g4& z, const g4& o, const g4& r) {
    return g_add(z, g_mul(r, g_add(o, z)));
}

// EqualityPairs::pair (crates/spartan/src/sumcheck.rs:835-871).
__device__ __forceinline__ void eq_pair(const f128* low, int low_bits, const f128* high,
                                        size_t p, g4& e0, g4& e1) {
    if (low_bits == 0) {
        const g4 s = g_load(low);
        e0 = g_mul(s, g_load(high + 2 * p));
        e1 = g_mul(s, g_load(high + 2 * p + 1));
        return;
    }
    const size_t lp = p & ((((size_t)1) << (low_bits - 1)) - 1);
    const g4 hw = g_load(high + (p >> (low_bits - 1)));
    e0 = g_mul(g_load(low + 2 * lp), hw);
    e1 = g_mul(g_load(low + 2 * lp + 1), hw);
}

// cubic_contribution (sumcheck.rs:724-744), accumulated unreduced.
__device__ __forceinline__ void cubic_acc(const g4& e0, const g4& e1, const g4& a0, const g4& a1,
                                          const g4& b0, const g4& b1, const g4& c0, const g4& c1,
                                          g8& s0, g8& s2, g8& s3) {
    const g4 ed = g_add(e1, e0), ad = g_add(a1, a0), bd = g_add(b1, b0), cd = g_add(c1, c0);
    const g4 p00 = g_mul(a0, b0);
    const g4 pdd = g_mul(ad, bd);
    const g4 p11 = g_mul(a1, b1);
    // a0*bd + ad*b0 = a1*b1 + a0*b0 + ad*bd (characteristic 2)
    const g4 r0 = g_add(p00, c0);
    const g4 rl = g_add3(p11, p00, g_add(pdd, cd));
    clmul_acc(s0, e0, r0);
    clmul_acc(s2, e0, pdd);
    clmul_acc(s2, ed, rl);
    clmul_acc(s3, ed, pdd);
}

// ------------------------------------------------------------ primitives --
__global__ __launch_bounds__(kBlock) void k_mul(const f128* a, const f128* b, f128* out, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        g_store(out + i, g_mul(g_load(a + i), g_load(b + i)));
}

// The LDS product primitives on their own (bitz_mul_batch's modes): a[i] b[i]
// through this lane's 3-bit table (two: g_mul_lane2, pairing element i with
// i ^ 1 so that both chains are checked), or a[i] b[0] through b[0]'s 4-bit
// table.  256 threads; k_mul_lane takes kLaneSmem256 of dynamic LDS.
__global__ __launch_bounds__(kBlock) void k_mul_lane(const f128* a, const f128* b, f128* out, size_t n, int two) {
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    uint4* col = smem_raw + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const g4 x = g_load(a + i), y = g_load(b + i);
        if (two) {
            const size_t j = (i ^ 1) < n ? (i ^ 1) : i;
            const g4 z = g_load(b + j);
            g4 p, q;
            if (i & 1)
                g_mul_lane2(x, z, y, col, q, p);
            else
                g_mul_lane2(x, y, z, col, p, q);
            g_store(out + i, p);
        } else {
            g_store(out + i, g_mul_lane(x, y, col));
        }
    }
}
// g_mul_rows on its own (bitz_mul_batch's BITZ_MUL_ROWS): the grid's 16-lane
// row q multiplies a[q] by b[q]; lane 0 of each wave writes its four rows'
// products, read with g_row.  The stride keeps a row's lanes together.
__global__ __launch_bounds__(kBlock) void k_mul_rows(const f128* a, const f128* b, f128* out, size_t n) {
    const size_t stride = ((size_t)gridDim.x * blockDim.x) >> 4;
    for (size_t q = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 4; q < n; q += stride) {
        const g4 p = g_mul_rows(g_load(a + q), g_load(b + q));
        const size_t q0 = q & ~(size_t)3;  // row 0 of this wave (always active when any row is)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const g4 v = g_row(p, g);
            if ((threadIdx.x & 63) == 0 && q0 + g < n) g_store(out + q0 + g, v);
        }
    }
}
__global__ __launch_bounds__(kBlock) void k_mul_const(const f128* a, const f128* b, f128* out, size_t n) {
    __shared__ __attribute__((aligned(256))) uint4 tab[kConstTable4Entries];
    __shared__ uint4 basis[128];
    build_const_table4(tab, basis, g_load(b));  // ends with a barrier
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        g_store(out + i, g_mul_table4(tab, g_load(a + i)));
}

// a[i] * b[0] on the int8 matrix cores (bitz_mul_batch's BITZ_MUL_MFMA, a
// measured alternative to k_mul_const): multiplying by c = b[0] is the
// 128x128 F2 matrix whose column j is c X^j, so 16 elements (the rows of A,
// one 0/1 byte per bit) times that matrix (B) on v_mfma_i32_16x16x64_i8
// counts, per output bit, the set terms; the count's parity is the bit.
// Tile t of the output holds bits 16t..16t+15, K-step s input bits
// 64s..64s+63: 8 tiles x 2 steps = 16 MFMAs per 16 elements.  Lane l feeds
// row or column l % 16 and 16 bytes of the K range to both operands; A and B
// take the same per-lane K order, so the products line up whatever order
// the hardware sums in.  Output row 4(l / 16) + i, column l % 16 is in
// register i of lane l; one ballot per register gathers a tile's bits of
// four rows, and lane m < 16 assembles element m from the ballots.
typedef int v4i32 __attribute__((ext_vector_type(4)));

// Bit b of v (b < 4) -> byte b, as 0 or 1: v (v0 + v1 2^8 + ...) through
// the carry-free product v * (1 + 2^7 + 2^14 + 2^21) masked to bytes.
__device__ __forceinline__ uint32_t nibble_to_bytes(uint32_t v) { return __umul24(v, 0x204081u) & 0x01010101u; }

__global__ __launch_bounds__(kBlock) void k_mul_mfma(const f128* a, const f128* b, f128* out, size_t n) {
    __shared__ uint4 basis[128];
    const g4 c = g_load(b);
    for (int j = threadIdx.x; j < 128; j += blockDim.x) {
        const g4 x = g_mul_xpow(c, j);
        basis[j] = make_uint4(x.w[0], x.w[1], x.w[2], x.w[3]);
    }
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63), q = lane >> 4, col = lane & 15;
    // B for tile t, step s: byte j of this lane = bit 16t + col of c X^(64s + 16q + j).
    v4i32 bm[8][2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        uint32_t words[16][4];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint4 e = basis[64 * s + 16 * q + j];
            words[j][0] = e.x; words[j][1] = e.y; words[j][2] = e.z; words[j][3] = e.w;
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int bit = 16 * t + col, w = bit >> 5, sh = bit & 31;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                uint32_t v = 0;
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) v |= ((words[4 * d + jj][w] >> sh) & 1u) << (8 * jj);
                bm[t][s][d] = (int)v;
            }
        }
    }
    const size_t wave = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6;
    const size_t waves = ((size_t)gridDim.x * blockDim.x) >> 6;
    const size_t groups = (n + 15) / 16;
    const int qq = col >> 2, ii = col & 3;  // the output row this lane writes (lanes < 16)
    for (size_t g = wave; g < groups; g += waves) {
        const size_t e = 16 * g + col;
        const g4 x = e < n ? g_load_hbm(a + e) : g_zero();
        v4i32 am[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const uint32_t h = (x.w[2 * s + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
#pragma unroll
            for (int d = 0; d < 4; ++d) am[s][d] = (int)nibble_to_bytes((h >> (4 * d)) & 15u);
        }
        uint64_t bal[8][4];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            v4i32 acc = {0, 0, 0, 0};
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(am[0], bm[t][0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(am[1], bm[t][1], acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) bal[t][i] = __builtin_amdgcn_ballot_w64((acc[i] & 1) != 0);
        }
        if (lane < 16 && e < n) {
            g4 y;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                uint64_t lo = bal[2 * w][0], hi = bal[2 * w + 1][0];
                lo = ii == 1 ? bal[2 * w][1] : lo;
                hi = ii == 1 ? bal[2 * w + 1][1] : hi;
                lo = ii == 2 ? bal[2 * w][2] : lo;
                hi = ii == 2 ? bal[2 * w + 1][2] : hi;
                lo = ii == 3 ? bal[2 * w][3] : lo;
                hi = ii == 3 ? bal[2 * w + 1][3] : hi;
                y.w[w] = (uint32_t)((lo >> (16 * qq)) & 0xFFFFu) | ((uint32_t)((hi >> (16 * qq)) & 0xFFFFu) << 16);
            }
            g_store_hbm(out + e, y);
        }
    }
}

// In-place level doubling of a small eq table by one workgroup
// (crates/poly/src/eq.rs:63-99).
// Both halves of a split eq table in one launch: workgroup 0 builds eq(r_<kl)
// into low, workgroup 1 eq(r_[kl, kl+kh)) into high.
__global__ __launch_bounds__(kBlock) void k_eq_small2(const f128* r, int kl, f128* low, int kh, f128* high) {
    const bool hi = blockIdx.x == 1;
    const f128* rr = hi ? r + kl : r;
    const int k = hi ? kh : kl;
    f128* table = hi ? high : low;
    if (threadIdx.x == 0) g_store(table, g_one());
    for (size_t j = 1 + threadIdx.x; j < ((size_t)1 << k); j += blockDim.x) g_store(table + j, g_zero());
    __syncthreads();
    for (int i = 0; i < k; ++i) {
        const size_t half = (size_t)1 << i;
        const g4 ri = g_load(rr + i);
        for (size_t j = threadIdx.x; j < half; j += blockDim.x) {
            const g4 parent = g_load(table + j);
            const g4 one = g_mul(parent, ri);
            g_store(table + j, g_add(parent, one));
            g_store(table + j + half, one);
        }
        __syncthreads();
    }
}

// eq(r, .) over k <= kEqRowsMaxVars variables in 1024-thread workgroups:
// each builds the low kl = k / 2 and high kh = k - kl variables' tables (at
// most 256 entries each) level by level side by side in LDS (threads 0..255
// the low one, 256..511 the high), then its share of the entries j =
// low[j mod 2^kl] * high[j >> kl], at most kEqRowsPerThread per thread.
// Lane-table products throughout (kLaneSmem1024 of dynamic LDS): a level
// costs one product's latency instead of a comb product's.  With rr (rho and
// rho^2 on the device) it writes the bind's row scales instead, k_row_scales'
// layout [w | rho w | rho^2 w] over 2^k entries.
__global__ __launch_bounds__(1024) void k_eq_rows(const f128* r, int k, f128* out, const f128* rr) {
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    __shared__ uint4 lo[256], hi[256];
    uint4* col = smem_raw + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    auto ld = [](const uint4* p) {
        const uint4 u = *p;
        g4 v;
        v.w[0] = u.x; v.w[1] = u.y; v.w[2] = u.z; v.w[3] = u.w;
        return v;
    };
    auto sto = [](uint4* p, const g4& v) { *p = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]); };
    const int kl = k / 2, kh = k - kl, t = (int)threadIdx.x;
    if (t == 0) {
        sto(lo, g_one());
        sto(hi, g_one());
    }
    __syncthreads();
    for (int i = 0; i < kh; ++i) {  // kh >= kl
        const int half = 1 << i, j = t & 255;
        const bool high = t >= 256;
        if (t < 512 && i < (high ? kh : kl) && j < half) {
            uint4* T = high ? hi : lo;
            const g4 parent = ld(T + j), one = g_mul_lane(parent, g_load(r + (high ? kl : 0) + i), col);
            sto(T + j, g_add(parent, one));  // bit i clear: 1 + r_i
            sto(T + j + half, one);          // bit i set: r_i
        }
        __syncthreads();
    }
    const size_t n = (size_t)1 << k, lmask = ((size_t)1 << kl) - 1;
    g4 rho = g_zero(), rho2 = g_zero();
    if (rr) {
        rho = g_load(rr);
        rho2 = g_load(rr + 1);
    }
    for (size_t j = blockIdx.x * (size_t)1024 + t; j < n; j += (size_t)gridDim.x * 1024) {
        const g4 w = g_mul_lane(ld(lo + (j & lmask)), ld(hi + (j >> kl)), col);
        g_store_hbm(out + j, w);
        if (rr) {
            g4 a, b;
            g_mul_lane2(w, rho, rho2, col, a, b);
            g_store_hbm(out + n + j, a);
            g_store_hbm(out + 2 * n + j, b);
        }
    }
}

// table[j] = low[j & mask] * high[j >> low_bits]: every entry is the product
// of its per-coordinate factors, split into a low and a high group.
__global__ __launch_bounds__(kBlock) void k_eq_combine(const f128* low, int low_bits, const f128* high, f128* out, size_t n) {
    const size_t mask = ((size_t)1 << low_bits) - 1;
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < n;
         j += (size_t)gridDim.x * blockDim.x)
        g_store(out + j, g_mul(g_load(low + (j & mask)), g_load(high + (j >> low_bits))));
}

// f'(b) = f(0,b) + r (f(1,b) - f(0,b)) on adjacent pairs (mle.rs:240-258).
__global__ __launch_bounds__(kBlock) void k_fold(const f128* in, f128* out, size_t out_len, f128 r) {
    const g4 rr = g_from(r);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < out_len;
         i += (size_t)gridDim.x * blockDim.x)
        g_store(out + i, interp(g_load(in + 2 * i), g_load(in + 2 * i + 1), rr));
}

// MLE folds by a fixed r (DenseMultilinearExtension::fold, mle.rs:136-156)
// through r's 4-bit table, built in LDS by one wave: one constant-table
// product per output.  k_fold2_t binds two variables a pass (r0, then r1 on
// the two half-folded values), reading four inputs per output: 20 B of HBM
// per input instead of 24 for two single passes, at 3 products per output.
__global__ __launch_bounds__(kBlock) void k_fold_t(const f128* in, f128* out, size_t out_len, f128 r) {
    __shared__ __attribute__((aligned(256))) uint4 tab[kConstTable4Entries];
    if (threadIdx.x < 64) const_table4_wave(tab, g_from(r));
    __syncthreads();
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < out_len;
         i += (size_t)gridDim.x * blockDim.x) {
        const g4 a = g_load_hbm(in + 2 * i), b = g_load_hbm(in + 2 * i + 1);
        g_store_hbm(out + i, g_add(a, g_mul_table4(tab, g_add(a, b))));
    }
}
__global__ __launch_bounds__(kBlock) void k_fold2_t(const f128* in, f128* out, size_t out_len, f128 r0, f128 r1) {
    __shared__ __attribute__((aligned(256))) uint4 t0[kConstTable4Entries];
    __shared__ __attribute__((aligned(256))) uint4 t1[kConstTable4Entries];
    if (threadIdx.x < 64)
        const_table4_wave(t0, g_from(r0));
    else if (threadIdx.x < 128)
        const_table4_wave(t1, g_from(r1));
    __syncthreads();
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < out_len;
         i += (size_t)gridDim.x * blockDim.x) {
        const g4 a0 = g_load_hbm(in + 4 * i), a1 = g_load_hbm(in + 4 * i + 1);
        const g4 a2 = g_load_hbm(in + 4 * i + 2), a3 = g_load_hbm(in + 4 * i + 3);
        const g4 x = g_add(a0, g_mul_table4(t0, g_add(a0, a1)));
        const g4 y = g_add(a2, g_mul_table4(t0, g_add(a2, a3)));
        g_store_hbm(out + i, g_add(x, g_mul_table4(t1, g_add(x, y))));
    }
}

__global__ __launch_bounds__(kBlock) void k_fold3(const f128* a, const f128* b, const f128* c, f128* ao, f128* bo, f128* co,
                        size_t out_len, f128 r) {
    const g4 rr = g_from(r);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < out_len;
         i += (size_t)gridDim.x * blockDim.x) {
        g_store(ao + i, interp(g_load(a + 2 * i), g_load(a + 2 * i + 1), rr));
        g_store(bo + i, interp(g_load(b + 2 * i), g_load(b + 2 * i + 1), rr));
        g_store(co + i, interp(g_load(c + 2 * i), g_load(c + 2 * i + 1), rr));
    }
}

// ------------------------------------------------------- outer sumcheck --
// Round sums [c0, c2, c3] over adjacent pairs of the current tables
// (compute_coefficients_without_linear, sumcheck.rs:874-889).
__global__ __launch_bounds__(kBlock) void k_outer_round0(const f128* a, const f128* b, const f128* c, size_t pairs,
                               const f128* low, int low_bits, const f128* high, uint32_t* acc) {
    g8 s0 = w_zero(), s2 = w_zero(), s3 = w_zero();
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < pairs;
         p += (size_t)gridDim.x * blockDim.x) {
        g4 e0, e1;
        eq_pair(low, low_bits, high, p, e0, e1);
        cubic_acc(e0, e1, g_load(a + 2 * p), g_load(a + 2 * p + 1), g_load(b + 2 * p),
                  g_load(b + 2 * p + 1), g_load(c + 2 * p), g_load(c + 2 * p + 1), s0, s2, s3);
    }
    g4 v[3] = {g_reduce(s0), g_reduce(s2), g_reduce(s3)};
    block_xor_to_global<3>(v, acc);
}

// Fold A, B, C at r and accumulate the next round's [c0, c2, c3] from the
// freshly folded pairs (fold_products_and_compute_next, sumcheck.rs:963-1025).
// One output pair = four input elements per table.
__global__ __launch_bounds__(kBlock) void k_outer_fold_next(const f128* a, const f128* b, const f128* c, f128* ao, f128* bo,
                                  f128* co, size_t out_pairs, f128 r, const f128* low,
                                  int low_bits, const f128* high, uint32_t* acc) {
    const g4 rr = g_from(r);
    g8 s0 = w_zero(), s2 = w_zero(), s3 = w_zero();
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < out_pairs;
         p += (size_t)gridDim.x * blockDim.x) {
        const g4 a0 = interp(g_load(a + 4 * p), g_load(a + 4 * p + 1), rr);
        const g4 a1 = interp(g_load(a + 4 * p + 2), g_load(a + 4 * p + 3), rr);
        const g4 b0 = interp(g_load(b + 4 * p), g_load(b + 4 * p + 1), rr);
        const g4 b1 = interp(g_load(b + 4 * p + 2), g_load(b + 4 * p + 3), rr);
        const g4 c0 = interp(g_load(c + 4 * p), g_load(c + 4 * p + 1), rr);
        const g4 c1 = interp(g_load(c + 4 * p + 2), g_load(c + 4 * p + 3), rr);
        g_store(ao + 2 * p, a0); g_store(ao + 2 * p + 1, a1);
        g_store(bo + 2 * p, b0); g_store(bo + 2 * p + 1, b1);
        g_store(co + 2 * p, c0); g_store(co + 2 * p + 1, c1);
        g4 e0, e1;
        eq_pair(low, low_bits, high, p, e0, e1);
        cubic_acc(e0, e1, a0, a1, b0, b1, c0, c1, s0, s2, s3);
    }
    g4 v[3] = {g_reduce(s0), g_reduce(s2), g_reduce(s3)};
    block_xor_to_global<3>(v, acc);
}

// The outer sumcheck's claim sum_x eq(x) (A(x) B(x) + C(x)) with
// eq(x) = low[x & mask] * high[x >> low_bits] (sumcheck.rs:367-370 checks the
// prover's claim against exactly this sum).
__global__ __launch_bounds__(kBlock) void k_outer_claim(const f128* a, const f128* b, const f128* c, size_t n,
                                                        const f128* low, int low_bits, const f128* high,
                                                        uint32_t* acc) {
    const size_t mask = ((size_t)1 << low_bits) - 1;
    g8 s = w_zero();
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < n;
         x += (size_t)gridDim.x * blockDim.x) {
        const g4 e = g_mul(g_load(low + (x & mask)), g_load(high + (x >> low_bits)));
        clmul_acc(s, e, g_add(g_mul(g_load(a + x), g_load(b + x)), g_load(c + x)));
    }
    g4 v[1] = {g_reduce(s)};
    block_xor_to_global<1>(v, acc);
}

// ------------------------------------------------------------------ gkr --
// Round sums of a dense layer (round_sums, lib.rs:164-219) for the round
// whose halves are [0, half) and [half, 2 half):
//   sum_b   = sum_x eq[x] L_b(x) R_b(x)     (b = 1 when at_one)
//   sum_inf = sum_x eq[x] dL(x) dR(x)
__global__ __launch_bounds__(kBlock) void k_gkr_sums(const f128* left, const f128* right, const f128* eq, size_t half,
                           int at_one, uint32_t* acc) {
    g8 sb = w_zero(), si = w_zero();
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < half;
         x += (size_t)gridDim.x * blockDim.x) {
        const g4 l0 = g_load(left + x), l1 = g_load(left + x + half);
        const g4 r0 = g_load(right + x), r1 = g_load(right + x + half);
        const g4 e = g_load(eq + x);
        clmul_acc(sb, g_mul(e, at_one ? l1 : l0), at_one ? r1 : r0);
        clmul_acc(si, g_mul(e, g_add(l1, l0)), g_add(r1, r0));
    }
    g4 v[2] = {g_reduce(sb), g_reduce(si)};
    block_xor_to_global<2>(v, acc);
}

// Bind the round with halves [0, 2h) x [2h, 4h) at r (lib.rs:223-243) and
// accumulate the next round's sums over halves [0, h), [h, 2h) of the bound
// arrays.  The next round's eq table is the current one halved:
// eq(z_<i-1, x) = eq(z_<i, x) + eq(z_<i, x + h), written in place.
__global__ __launch_bounds__(kBlock) void k_gkr_bind_next(f128* left, f128* right, f128* eq, size_t h, f128 r,
                                int at_one, uint32_t* acc) {
    const g4 rr = g_from(r);
    const size_t hh = 2 * h;
    g8 sb = w_zero(), si = w_zero();
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < h;
         x += (size_t)gridDim.x * blockDim.x) {
        const g4 l0 = interp(g_load(left + x), g_load(left + x + hh), rr);
        const g4 l1 = interp(g_load(left + x + h), g_load(left + x + h + hh), rr);
        const g4 r0 = interp(g_load(right + x), g_load(right + x + hh), rr);
        const g4 r1 = interp(g_load(right + x + h), g_load(right + x + h + hh), rr);
        g_store(left + x, l0); g_store(left + x + h, l1);
        g_store(right + x, r0); g_store(right + x + h, r1);
        const g4 e = g_add(g_load(eq + x), g_load(eq + x + h));
        g_store(eq + x, e);
        clmul_acc(sb, g_mul(e, at_one ? l1 : l0), at_one ? r1 : r0);
        clmul_acc(si, g_mul(e, g_add(l1, l0)), g_add(r1, r0));
    }
    g4 v[2] = {g_reduce(sb), g_reduce(si)};
    block_xor_to_global<2>(v, acc);
}

// Bind only (the layer's last round): l0 += r (l1 - l0), same for right.
__global__ __launch_bounds__(kBlock) void k_gkr_bind(f128* left, f128* right, size_t half, f128 r) {
    const g4 rr = g_from(r);
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < half;
         x += (size_t)gridDim.x * blockDim.x) {
        g_store(left + x, interp(g_load(left + x), g_load(left + x + half), rr));
        g_store(right + x, interp(g_load(right + x), g_load(right + x + half), rr));
    }
}

// --------------------------------------------------------- sparse gathers --
// out[i] = sum over row i's nonzeros of value * h[column] for i < rows, zero
// above (the reference test's multiply + padded_mle, crates/spartan/src/piop.rs:352-372).
__global__ __launch_bounds__(kBlock) void k_r1cs_rows(const uint32_t* row_starts, const uint32_t* cols,
                                                      const f128* vals, const f128* h, size_t rows, f128* out,
                                                      size_t padded) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < padded; i += (size_t)gridDim.x * blockDim.x) {
        if (i >= rows) {
            g_store(out + i, g_zero());
            continue;
        }
        g8 s = w_zero();
        for (uint32_t k = row_starts[i]; k < row_starts[i + 1]; ++k) clmul_acc(s, g_load(vals + k), g_load(h + cols[k]));
        g_store(out + i, g_reduce(s));
    }
}

// build_assignment_mle (matrix.rs:392-420): bit i of the packed words as 0 / 1,
// zeros up to the padded length.
__global__ __launch_bounds__(kBlock) void k_assignment_bits(const uint64_t* words, size_t bit_len, f128* out,
                                                            size_t padded) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < padded; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t b = i < bit_len ? (uint32_t)((words[i / 64] >> (i % 64)) & 1) : 0u;
        g4 v = g_zero();
        v.w[0] = b;
        g_store(out + i, v);
    }
}

__global__ __launch_bounds__(kStageBlock) void k_bind_products(const uint32_t* rows, const uint8_t* which,
                                                               const f128* values, const f128* row_scale,
                                                               size_t rows_padded, f128* prod, size_t nnz) {
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    uint4* col = smem_raw + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
#pragma unroll 1
    for (size_t k = blockIdx.x * (size_t)kStageBlock + threadIdx.x; k < nnz; k += (size_t)gridDim.x * kStageBlock) {
        const g4 w = g_load_hbm(row_scale + which[k] * rows_padded + rows[k]);
        g_store_hbm(prod + k, g_mul_lane(w, g_load_hbm(values + k), col));
    }
}

__global__ __launch_bounds__(kBlock) void k_bind_segments(const uint32_t* col_starts, const f128* prod, f128* out,
                                                         size_t cols) {
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < cols; j += (size_t)gridDim.x * blockDim.x) {
        g4 s = g_zero();
        const uint32_t lo = col_starts[j], hi = col_starts[j + 1];
        if (hi - lo <= kBindLongColumn)
            for (uint32_t k = lo; k < hi; ++k) s = g_add(s, g_load_hbm(prod + k));
        g_store_hbm(out + j, s);
    }
}

// One run {column, first, end} of a long column per workgroup: the run's XOR
// into out[column] (left zero by k_bind_segments).
__global__ __launch_bounds__(kBlock) void k_bind_long(const uint32_t* chunks, const f128* prod, f128* out) {
    const uint32_t col = chunks[3 * blockIdx.x], lo = chunks[3 * blockIdx.x + 1], hi = chunks[3 * blockIdx.x + 2];
    g4 v[1] = {g_zero()};
    for (uint32_t k = lo + threadIdx.x; k < hi; k += kBlock) v[0] = g_add(v[0], g_load_hbm(prod + k));
    block_xor_to_global<1>(v, reinterpret_cast<uint32_t*>(out + col));
}

// row_scale[m][i] = eq(i, r_x) * scale_m.
__global__ __launch_bounds__(kBlock) void k_row_scales(const f128* eq_rows, f128 rho, f128 rho2, f128* row_scale,
                             size_t rows_padded) {
    const g4 s1 = g_from(rho), s2 = g_from(rho2);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < rows_padded;
         i += (size_t)gridDim.x * blockDim.x) {
        const g4 w = g_load(eq_rows + i);
        g_store(row_scale + i, w);
        g_store(row_scale + rows_padded + i, g_mul(w, s1));
        g_store(row_scale + 2 * rows_padded + i, g_mul(w, s2));
    }
}

// r * M over GF(2^128): out[col] = XOR of challenges[row] over the column
// (MaterializedMTranspose::apply, matrix_transpose.rs:327-375).
// M^T (crates/circuit/src/matrix_transpose.rs:327-441), one thread per
// column of M; a column of more than kMtLongColumn entries (the constant
// column, read by hundreds of rows) is skipped here and summed by a whole
// workgroup in the _long kernels, so one column does not hold up the grid.
__global__ __launch_bounds__(kBlock) void k_mt_apply(const uint32_t* offsets, const uint32_t* rows, const f128* ch,
                           f128* out, size_t cols) {
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < cols;
         c += (size_t)gridDim.x * blockDim.x) {
        const uint32_t lo = offsets[c], hi = offsets[c + 1];
        if (hi - lo > kMtLongColumn) continue;
        g4 s = g_zero();
        for (uint32_t k = lo; k < hi; ++k) s = g_add(s, g_load_hbm(ch + rows[k]));
        g_store_hbm(out + c, s);
    }
}
__global__ __launch_bounds__(kBlock) void k_mt_apply_long(const uint32_t* offsets, const uint32_t* rows,
                                                          const f128* ch, const uint32_t* long_cols, f128* out) {
    const uint32_t c = long_cols[blockIdx.x];
    g4 v[1] = {g_zero()};
    for (uint32_t k = offsets[c] + threadIdx.x; k < offsets[c + 1]; k += kBlock)
        v[0] = g_add(v[0], g_load_hbm(ch + rows[k]));
    block_xor_store<1>(v, reinterpret_cast<uint32_t*>(out + c));
}
// transpose_inner: per entry row_w[i & mask] * col_w[i >> log_rows] through
// this lane's 3-bit table (kLaneSmem256 of dynamic LDS), XORed per column.
__global__ __launch_bounds__(kBlock) void k_mt_transpose_inner(const uint32_t* offsets, const uint32_t* rows,
                                     const f128* row_w, int log_rows, const f128* col_w,
                                     f128* out, size_t cols) {
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    uint4* col = smem_raw + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    const uint32_t mask = (1u << log_rows) - 1;
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < cols;
         c += (size_t)gridDim.x * blockDim.x) {
        const uint32_t lo = offsets[c], hi = offsets[c + 1];
        if (hi - lo > kMtLongColumn) continue;
        g4 s = g_zero();
        for (uint32_t k = lo; k < hi; ++k) {
            const uint32_t idx = rows[k];
            s = g_add(s, g_mul_lane(g_load_hbm(row_w + (idx & mask)), g_load_hbm(col_w + (idx >> log_rows)), col));
        }
        g_store_hbm(out + c, s);
    }
}
__global__ __launch_bounds__(kBlock) void k_mt_inner_long(const uint32_t* offsets, const uint32_t* rows,
                                                          const f128* row_w, int log_rows, const f128* col_w,
                                                          const uint32_t* long_cols, f128* out) {
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    uint4* col = smem_raw + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    const uint32_t mask = (1u << log_rows) - 1, c = long_cols[blockIdx.x];
    g4 v[1] = {g_zero()};
    for (uint32_t k = offsets[c] + threadIdx.x; k < offsets[c + 1]; k += kBlock) {
        const uint32_t idx = rows[k];
        v[0] = g_add(v[0], g_mul_lane(g_load_hbm(row_w + (idx & mask)), g_load_hbm(col_w + (idx >> log_rows)), col));
    }
    block_xor_store<1>(v, reinterpret_cast<uint32_t*>(out + c));
}

// 4-bit constant tables, one workgroup per constant, into global memory;
// workgroups that multiply by a constant copy its table into LDS.
// ------------------------------------------------- post-GKR sumcheck --
// The first round off the packed bits (crates/post_gkr/src/lib.rs:189-232):
// element e carries bits 128 (e mod 2^(t-7)) .. of column e >> (t-7); one
// wave per element, lane l the bit pair (2l, 2l + 1).  Per element the wave
// XORs the row weights of the set bits (the claim's weighted sum), of the set
// even bits (a0) and w0 + w1 where the pair's bits differ (a2); lane 0 weights
// the three sums by the column's weight.  acc: 3 elements.
__global__ __launch_bounds__(256) void k_pg_first(const f128* rows, int t, const f128* cols, const f128* packed,
                                                  size_t n_elems, uint32_t* acc) {
    const int lane = threadIdx.x & 63;
    const size_t wave = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const size_t waves = (size_t)gridDim.x * (blockDim.x >> 6);
    const size_t row_mask = ((size_t)1 << (t - 7)) - 1;
    g4 sums[3] = {g_zero(), g_zero(), g_zero()};
    for (size_t e = wave; e < n_elems; e += waves) {
        const size_t base = (e & row_mask) << 7;
        const f128 el = packed[e];
        const uint64_t word = lane < 32 ? el.lo : el.hi;
        const int sh = (2 * lane) & 63;
        const bool f0 = (word >> sh) & 1, f1 = (word >> (sh + 1)) & 1;
        const g4 r0 = g_load(rows + base + 2 * lane), r1 = g_load(rows + base + 2 * lane + 1);
        g4 v[3];
        v[0] = g_add(f0 ? r0 : g_zero(), f1 ? r1 : g_zero());
        v[1] = f0 ? r0 : g_zero();
        v[2] = (f0 != f1) ? g_add(r0, r1) : g_zero();
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                uint32_t x = v[k].w[i];
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) x ^= __shfl_xor(x, off, 64);
                v[k].w[i] = x;
            }
        if (lane == 0) {
            const g4 c = g_load(cols + (e >> (t - 7)));
#pragma unroll
            for (int k = 0; k < 3; ++k) sums[k] = g_add(sums[k], g_mul(c, v[k]));
        }
    }
    block_xor_to_global<3>(sums, acc);
}

// The tables after the first round (lib.rs:236-244, 265-284): W[j] =
// columns[j >> (t-1)] rf[j mod 2^(t-1)] (rf: the rows folded at r) and V[j]
// = the bit pair j bound at r: 0, 1 + r, r or 1.  Pair j is bits 2j, 2j + 1,
// in element j >> 6.
__global__ __launch_bounds__(kBlock) void k_pg_expand(const f128* rf, int t, const f128* cols, const f128* packed,
                                                      f128 r, f128* W, f128* V, size_t n) {
    const g4 rr = g_from(r);
    const g4 one = g_from(f128{1, 0});
    const g4 rr1 = g_add(rr, one);
    const size_t row_mask = ((size_t)1 << (t - 1)) - 1;
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x) {
        g_store(W + j, g_mul(g_load(cols + (j >> (t - 1))), g_load(rf + (j & row_mask))));
        const f128 el = packed[j >> 6];
        const int p = (int)(j & 63);
        const uint64_t word = p < 32 ? el.lo : el.hi;
        const unsigned q = (unsigned)(word >> ((2 * p) & 63)) & 3u;  // f0 | f1 << 1
        g_store(V + j, q == 0 ? g_zero() : q == 1 ? rr1 : q == 2 ? rr : one);
    }
}

__global__ void k_keccak_hash(const uint8_t* msg, size_t n, int rate, uint8_t pad, uint8_t* out, size_t out_len) {
    if (blockIdx.x == 0 && threadIdx.x == 0) keccak_hash(msg, n, rate, pad, out, out_len);
}

__global__ void k_zero_words(uint32_t* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

__global__ __launch_bounds__(512) void k_const_tables4(const f128* consts, uint4* out) {
    __shared__ uint4 tab[kConstTable4Entries];
    __shared__ uint4 basis[128];
    build_const_table4(tab, basis, g_load(consts + blockIdx.x));
    uint4* dst = out + (size_t)blockIdx.x * kConstTable4Entries;
    for (int e = threadIdx.x; e < kConstTable4Entries; e += blockDim.x) dst[e] = tab[e];
}

// ============================================ outer sumcheck, eq-factored ===
// Used when the eq factors are eq tables of a point tau.  Round k's
// polynomial is g_k(X) = S_k eq(tau_k, X) q_k(X) with
//   q_k(X) = sum_y E_k(y) (A_k B_k + C_k)(X, y),   E_k(y) = eq(tau_{>k}, y),
// a quadratic; the kernels return q(0), q(inf) and, when asked, q(1).
//
// The weights are read straight from the given factors, without tau: for a
// run R = [s, s + j) of variables inside one factor T, U(x) = T[x << s'] (s'
// = s minus the factor's first variable) equals eq(tau_R, x) times a
// constant (the product of (1 + tau_i) over T's other variables), which the
// host divides out of the round sums.  The pair index y = i * T + t
// (t = global thread, i = iteration) splits the weight as L(t) H(i): H(i) is
// the same for every thread of an iteration and is applied through an LDS
// constant table; L(t) = La(t_low) Lb(t_high) is applied once per thread.

__device__ __forceinline__ void eq_check_block(const f128* L, int ell, const f128* H, int m, uint32_t* out,
                                               unsigned blocks_l, unsigned bidx) {
    const bool in_h = bidx >= blocks_l;
    const f128* T = in_h ? H : L;
    const int bits = in_h ? m : ell;
    const size_t x = (size_t)(in_h ? bidx - blocks_l : bidx) * blockDim.x + threadIdx.x;
    const size_t size = (size_t)1 << bits;
    f128* probes = reinterpret_cast<f128*>(out) + (in_h ? ell + 1 : 0);
    uint32_t* rec = out + 4 * (ell + m + 2) + 8 * bidx;
    g4 v[1] = {g_zero()};
    int bad = 0;
    if (x < size) {
        const g4 tx = g_load(T + x);
        v[0] = tx;
        if (x == 0) {
            g_store(probes, tx);
        } else {
            const int top = 63 - __builtin_clzll((unsigned long long)x);
            const size_t bit = (size_t)1 << top;
            const g4 lhs = g_mul(tx, g_load(T));
            const g4 rhs = g_mul(g_load(T + (x ^ bit)), g_load(T + bit));
            bad = ((lhs.w[0] ^ rhs.w[0]) | (lhs.w[1] ^ rhs.w[1]) | (lhs.w[2] ^ rhs.w[2]) | (lhs.w[3] ^ rhs.w[3])) != 0;
            if (x == bit) g_store(probes + 1 + top, tx);
        }
    }
    const int any_bad = __syncthreads_or(bad);
    block_xor_store<1>(v, rec + 4);
    if (threadIdx.x == 0) rec[0] = (uint32_t)any_bad;
}

// Checks that a factor is an eq table: T[x] T[0] = T[x - 2^top] T[2^top]
// for every x > 0 (so T[x] = T[0] prod_{i in x} rho_i), and XOR-sums the
// entries (eq tables sum to 1; the host checks the sums and the probes).
// out: the probes T[0], T[1], T[2], T[4], ... of L, then of H (n + 2
// elements), then per workgroup [mismatch flag, 0, 0, 0, XOR of its
// entries] (8 words).  Plain stores only: out needs no clearing.
__global__ __launch_bounds__(1024) void k_eq_check(const f128* L, int ell, const f128* H, int m, uint32_t* out,
                                                   unsigned blocks_l) {
    eq_check_block(L, ell, H, m, out, blocks_l, blockIdx.x);
}

// H(i) tables (4-bit windows) of the multi-iteration rounds: table b belongs
// to the round j with first[j] <= b < first[j + 1]; its constant is the
// weight of iteration i = b - first[j] over the variables [start[j], n),
// read from the factors (a product when the run crosses from L into H).
__device__ __forceinline__ void h4_block(const f128* L, int ell, const f128* H, const HtabDesc& d, uint4* out,
                                         unsigned bidx, uint4* tab, uint4* basis) {
    int j = 0;
    while (j + 1 < d.rounds && bidx >= d.first[j + 1]) ++j;
    const uint32_t i = bidx - d.first[j];
    const int s = d.start[j];
    g4 c;
    if (s < ell) {
        const int lo_bits = ell - s;
        const uint32_t lo = i & ((1u << lo_bits) - 1), hi = i >> lo_bits;
        c = g_mul(g_load(L + ((size_t)lo << s)), g_load(H + hi));
    } else {
        c = g_load(H + ((size_t)i << (s - ell)));
    }
    build_const_table4(tab, basis, c);
    uint4* dst = out + (size_t)bidx * kConstTable4Entries;
    for (int e = threadIdx.x; e < kConstTable4Entries; e += blockDim.x) dst[e] = tab[e];
}

__global__ __launch_bounds__(1024) void k_htabs4(const f128* L, int ell, const f128* H, HtabDesc d, uint4* out) {
    __shared__ uint4 tab[kConstTable4Entries];
    __shared__ uint4 basis[128];
    h4_block(L, ell, H, d, out, blockIdx.x, tab, basis);
}

__global__ __launch_bounds__(1024) void k_const_table4_1(f128 c, uint4* out) {
    __shared__ uint4 tab[kConstTable4Entries];
    __shared__ uint4 basis[128];
    build_const_table4(tab, basis, g_from(c));
    for (int e = threadIdx.x; e < kConstTable4Entries; e += blockDim.x) out[e] = tab[e];
}

// Everything the pipelined prover needs before round 0, in one launch of
// 1024-thread workgroups: clearing the round-sum slots, the factor check,
// the H(i) tables and the 4-bit tables of every challenge.
__global__ __launch_bounds__(1024) void k_prologue(PrologueArgs A) {
    __shared__ uint4 tab[kConstTable4Entries];
    __shared__ uint4 basis[128];
    unsigned b = blockIdx.x;
    if (b == 0)
        for (unsigned e = threadIdx.x; e < A.zero_words; e += blockDim.x) A.zero[e] = 0;
    if (b < A.eq_blocks) {
        eq_check_block(A.L, A.ell, A.H, A.m, A.eq_out, A.eq_blocks_l, b);
        return;
    }
    b -= A.eq_blocks;
    if (b < A.h4_blocks) {
        h4_block(A.L, A.ell, A.H, A.hdesc, A.htabs4, b, tab, basis);
        return;
    }
    b -= A.h4_blocks;
    build_const_table4(tab, basis, g_from(A.chal[b]));
    uint4* dst = A.rtabs4 + (size_t)b * kConstTable4Entries;
    for (int e = threadIdx.x; e < kConstTable4Entries; e += blockDim.x) dst[e] = tab[e];
}

// Round 0 of the eq-factored outer sumcheck with the general products
// through per-lane LDS tables (g_mul_lane): workgroups of up to 1024
// threads, each wave's 8 KiB of lane tables first, then two H(i) 4-bit
// tables that iterations fill alternately (one barrier per iteration, as in
// outer_fold_body).  Thread t handles pairs y = i * threads + t for
// iterations i (the iteration index i is the high part of y, so H(i) is one
// constant per iteration and the thread's low part t fixes L(t)).
template <bool Q1>
__device__ __forceinline__ void outer_r0_body(const f128* a, const f128* b, const f128* c, size_t threads,
                                              size_t iters, const f128* la, int la_bits, int la_shift,
                                              const f128* lb, int lb_shift, const uint4* h_tabs4, uint32_t* acc,
                                              size_t block) {
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    const int waves = blockDim.x >> 6;
    uint4* col = smem_raw + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    uint4* hbase = smem_raw + waves * (64 * kLaneTableEntries);
    lane_table_clear(col);
    const size_t t = block * blockDim.x + threadIdx.x;
    g4 s0 = g_zero(), si = g_zero(), s1 = g_zero();
    for (size_t i = 0; i < iters; ++i) {
        uint4* htab = hbase + (i & 1) * kConstTable4Entries;
        if (iters > 1) {
            for (int e = threadIdx.x; e < kConstTable4Entries; e += blockDim.x)
                htab[e] = u4_load_hbm(h_tabs4 + i * kConstTable4Entries + e);
            __syncthreads();
        }
        if (t < threads) {
            const size_t y = i * threads + t;
            const g4 a0 = g_load_hbm(a + 2 * y), a1 = g_load_hbm(a + 2 * y + 1);
            const g4 b0 = g_load_hbm(b + 2 * y), b1 = g_load_hbm(b + 2 * y + 1);
            const g4 p0 = g_add(g_mul_lane(a0, b0, col), g_load_hbm(c + 2 * y));
            const g4 pi = g_mul_lane(g_add(a0, a1), g_add(b0, b1), col);
            g4 p1;
            if (Q1) p1 = g_add(g_mul_lane(a1, b1, col), g_load_hbm(c + 2 * y + 1));
            if (iters > 1) {
                s0 = g_add(s0, g_mul_table4(htab, p0));
                si = g_add(si, g_mul_table4(htab, pi));
                if (Q1) s1 = g_add(s1, g_mul_table4(htab, p1));
            } else {
                s0 = p0;
                si = pi;
                if (Q1) s1 = p1;
            }
        }
    }
    g4 v[3] = {g_zero(), g_zero(), g_zero()};
    if (t < threads) {
        const size_t mask = ((size_t)1 << la_bits) - 1;
        const g4 L = g_mul_lane(g_load_hbm(la + ((t & mask) << la_shift)), g_load_hbm(lb + ((t >> la_bits) << lb_shift)), col);
        g_mul_lane2(L, s0, si, col, v[0], v[1]);
        if (Q1) v[2] = g_mul_lane(L, s1, col);
    }
    block_xor_to_global<3>(v, acc);
}
template <bool Q1>
__global__ __launch_bounds__(kR0Block) void k_outer_r0(const f128* a, const f128* b, const f128* c, size_t threads,
                                                  size_t iters, const f128* la, int la_bits, int la_shift,
                                                  const f128* lb, int lb_shift, const uint4* h_tabs4,
                                                  uint32_t* acc) {
    outer_r0_body<Q1>(a, b, c, threads, iters, la, la_bits, la_shift, lb, lb_shift, h_tabs4, acc, blockIdx.x);
}

#define OUTER_R0_SIG                                                                                        \
    const f128*, const f128*, const f128*, size_t, size_t, const f128*, int, int, const f128*, int, const uint4*, \
        uint32_t*
template __global__ void k_outer_r0<false>(OUTER_R0_SIG);
template __global__ void k_outer_r0<true>(OUTER_R0_SIG);
#undef OUTER_R0_SIG

// Rounds k >= 1 with more pairs than threads: the fold by r_{k-1} through
// r's 4-bit table, the general products through per-lane LDS tables
// (g_mul_lane) and H(i) through 4-bit tables.  Workgroups of up to 1024
// threads: each wave's 8 KiB of lane tables first, then r's table and two
// H(i) tables, iteration i filling buffer i & 1 (its readers of iteration
// i - 2 have all passed the barrier of iteration i - 1), so one barrier per
// iteration suffices.
template <bool Q1>
__device__ __forceinline__ void outer_fold_body(const f128* a, const f128* b, const f128* c, f128* ao, f128* bo,
                                                f128* co, size_t threads, size_t iters, const uint4* r_tab4,
                                                const f128* la, int la_bits, int la_shift, const f128* lb,
                                                int lb_shift, const uint4* h_tabs4, uint32_t* acc, size_t block) {
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    const int waves = blockDim.x >> 6;
    uint4* col = smem_raw + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    uint4* rt = smem_raw + waves * (64 * kLaneTableEntries);
    lane_table_clear(col);
    for (int e = threadIdx.x; e < kConstTable4Entries; e += blockDim.x) rt[e] = u4_load_hbm(r_tab4 + e);
    const size_t t = block * blockDim.x + threadIdx.x;
    g4 s0 = g_zero(), si = g_zero(), s1 = g_zero();
    for (size_t i = 0; i < iters; ++i) {
        uint4* ht = rt + ((i & 1) + 1) * kConstTable4Entries;
        if (iters > 1)
            for (int e = threadIdx.x; e < kConstTable4Entries; e += blockDim.x)
                ht[e] = u4_load_hbm(h_tabs4 + i * kConstTable4Entries + e);
        __syncthreads();
        if (t < threads) {
            const size_t y = i * threads + t;
            g4 in[3][4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                in[0][j] = g_load_hbm(a + 4 * y + j);
                in[1][j] = g_load_hbm(b + 4 * y + j);
                in[2][j] = g_load_hbm(c + 4 * y + j);
            }
            g4 f[3][2];
#pragma unroll
            for (int u = 0; u < 3; ++u)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    f[u][j] = g_add(in[u][2 * j], g_mul_table4(rt, g_add(in[u][2 * j], in[u][2 * j + 1])));
            g_store_hbm(ao + 2 * y, f[0][0]);
            g_store_hbm(ao + 2 * y + 1, f[0][1]);
            g_store_hbm(bo + 2 * y, f[1][0]);
            g_store_hbm(bo + 2 * y + 1, f[1][1]);
            g_store_hbm(co + 2 * y, f[2][0]);
            g_store_hbm(co + 2 * y + 1, f[2][1]);
            const g4 p0 = g_add(g_mul_lane(f[0][0], f[1][0], col), f[2][0]);
            const g4 pi = g_mul_lane(g_add(f[0][0], f[0][1]), g_add(f[1][0], f[1][1]), col);
            g4 p1;
            if (Q1) p1 = g_add(g_mul_lane(f[0][1], f[1][1], col), f[2][1]);
            if (iters > 1) {
                s0 = g_add(s0, g_mul_table4(ht, p0));
                si = g_add(si, g_mul_table4(ht, pi));
                if (Q1) s1 = g_add(s1, g_mul_table4(ht, p1));
            } else {
                s0 = p0;
                si = pi;
                if (Q1) s1 = p1;
            }
        }
    }
    g4 v[3] = {g_zero(), g_zero(), g_zero()};
    if (t < threads) {
        const size_t mask = ((size_t)1 << la_bits) - 1;
        const g4 L =
            g_mul_lane(g_load_hbm(la + ((t & mask) << la_shift)), g_load_hbm(lb + ((t >> la_bits) << lb_shift)), col);
        g_mul_lane2(L, s0, si, col, v[0], v[1]);
        if (Q1) v[2] = g_mul_lane(L, s1, col);
    }
    block_xor_to_global<3>(v, acc);
}
template <bool Q1>
__global__ __launch_bounds__(kFoldBlock) void k_outer_fold(const f128* a, const f128* b, const f128* c, f128* ao,
                                                    f128* bo, f128* co, size_t threads, size_t iters,
                                                    const uint4* r_tab4, const f128* la, int la_bits,
                                                    int la_shift, const f128* lb, int lb_shift,
                                                    const uint4* h_tabs4, uint32_t* acc) {
    outer_fold_body<Q1>(a, b, c, ao, bo, co, threads, iters, r_tab4, la, la_bits, la_shift, lb, lb_shift, h_tabs4,
                        acc, blockIdx.x);
}

#define OUTER_FOLD_SIG                                                                                    \
    const f128*, const f128*, const f128*, f128*, f128*, f128*, size_t, size_t, const uint4*, const f128*, int, \
        int, const f128*, int, const uint4*, uint32_t*
template __global__ void k_outer_fold<false>(OUTER_FOLD_SIG);
template __global__ void k_outer_fold<true>(OUTER_FOLD_SIG);
#undef OUTER_FOLD_SIG

// ============================== rounds held in LDS: chunk, tail, batch ===
// A stage (LdsStage) runs rounds k_first .. k_first+s-1 of one eq-factored
// proof, with the challenges known in advance, on one chunk of the tables
// held in LDS, folding by r_k between rounds.  Round l of the stage has
// 2^(pl0-l) pairs per chunk (pl0 = log2 of the chunk's first-round pairs);
// pair y of chunk g is pair g 2^(pl0-l) + y of the round, so its weight
// eq(tau_[k+1,n), .) splits into W_l(y) over the variables [k+1, b) and the
// chunk's W_hi(g) over [b, n), with b = k_first + 1 + pl0 for every round of
// the stage.  W_l(y) = E(y 2^l) for the one table E(x) over
// [k_first+1, b): the entries of E whose low l variables are zero.  The
// factor entries carry constants (products of 1 + tau_i over the variables
// they do not cover, and E(y 2^l) the factors 1 + tau_i of its zero
// variables) that the host divides out of the round's sums.

// Product of the factor entries over the variables [a, b) of the weight,
// at index idx (bit j of idx is variable a + j): eq_low's part times
// eq_high's part, which is eq(tau_[a,b), idx) times a constant.  The run
// reads eq_low alone when b <= ell, eq_high alone when a >= ell (an empty run
// reads that factor's entry 0).
__device__ __forceinline__ g4 run_weight(const f128* L, int ell, const f128* H, int a, int b, size_t idx,
                                         uint4* col) {
    if (b <= ell) return g_load_hbm(L + (idx << a));
    if (a >= ell) return g_load_hbm(H + (idx << (a - ell)));
    const int lo_bits = ell - a;
    return g_mul_lane(g_load_hbm(L + ((idx & (((size_t)1 << lo_bits) - 1)) << a)), g_load_hbm(H + (idx >> lo_bits)), col);
}

// Offset of round l's weight table W_l (2^(pl0-l) entries) in the stage's
// weight area of 2^(pl0+1) entries.
__device__ __forceinline__ int stage_woff(int pl0, int l) { return (2 << pl0) - (2 << (pl0 - l)); }

// Builds W_0(x) = E(x) W_hi(g) for x < 2^pl0, barrier, then the strided
// copies W_l(y) = W_0(y 2^l) for l in [1, s).  The caller's barrier ends it.
template <int B = kStageBlock>
__device__ __forceinline__ void build_stage_weights(uint4* wt, int pl0, int s, const f128* L, int ell,
                                                    const f128* H, int a0, int n, size_t g, uint4* col) {
    const int b = a0 + pl0;
    const g4 wh = run_weight(L, ell, H, b, n, g, col);
    for (int x = threadIdx.x; x < (1 << pl0); x += B) {
        const g4 w = g_mul_lane(run_weight(L, ell, H, a0, b, (size_t)x, col), wh, col);
        wt[x] = make_uint4(w.w[0], w.w[1], w.w[2], w.w[3]);
    }
    __syncthreads();
    for (int l = 1; l < s; ++l)
        for (int y = threadIdx.x; y < (1 << (pl0 - l)); y += B) wt[stage_woff(pl0, l) + y] = wt[y << l];
}

// Round sums over the 2^pl pairs of the chunk in LDS (three tables at stride
// 2^cap_lg) with the round's weights W at wt: sum_y W(y) (a0 b0 + c0),
// sum_y W(y) (a0+a1)(b0+b1) and, with q1, sum_y W(y) (a1 b1 + c1), one
// (pair, term) item at a time through this lane's table col.  Stored at
// acc_k by a stage of one workgroup, XORed into it otherwise.
template <int NT = kStageBlock>
__device__ __forceinline__ void lds_round_sums(const uint4* lds, int cap_lg, int pl, const uint4* wt, bool q1,
                                               uint32_t* acc_k, bool store, uint4* col) {
    const size_t P = (size_t)1 << pl;
    const size_t items = (q1 ? 3 : 2) * P;
    auto ld = [&](const uint4* p) {
        const uint4 u = *p;
        g4 r;
        r.w[0] = u.x; r.w[1] = u.y; r.w[2] = u.z; r.w[3] = u.w;
        return r;
    };
    const uint4* A = lds;
    const uint4* B = lds + ((size_t)1 << cap_lg);
    const uint4* C = lds + ((size_t)2 << cap_lg);
    g4 s[3] = {g_zero(), g_zero(), g_zero()};
    for (size_t it = threadIdx.x; it < items; it += NT) {
        const size_t y = it & (P - 1);
        const int which = (int)(it >> pl);
        const g4 a0 = ld(A + 2 * y), a1 = ld(A + 2 * y + 1);
        const g4 b0 = ld(B + 2 * y), b1 = ld(B + 2 * y + 1);
        g4 p;
        if (which == 0)
            p = g_add(g_mul_lane(a0, b0, col), ld(C + 2 * y));
        else if (which == 1)
            p = g_mul_lane(g_add(a0, a1), g_add(b0, b1), col);
        else
            p = g_add(g_mul_lane(a1, b1, col), ld(C + 2 * y + 1));
        const g4 wp = g_mul_lane(ld(wt + y), p, col);
        if (which == 0) s[0] = g_add(s[0], wp);
        else if (which == 1) s[1] = g_add(s[1], wp);
        else s[2] = g_add(s[2], wp);
    }
    if (store)
        block_xor_store<3>(s, acc_k, NT);
    else
        block_xor_to_global<3>(s, acc_k, NT);
}

// In-place fold of the three LDS tables (2^(pl+1) entries each at stride
// 2^cap_lg) through the 4-bit table of r in rt: read everything, barrier,
// write, barrier.  At most kFoldPerThread outputs per thread.
template <int kFoldPerThread, int B = kStageBlock>
__device__ __forceinline__ void lds_fold(uint4* lds, int cap_lg, int pl, const uint4* rt) {
    const size_t half = (size_t)1 << pl;
    g4 nv[kFoldPerThread];
#pragma unroll
    for (int q = 0; q < kFoldPerThread; ++q) {
        const size_t e = threadIdx.x + (size_t)q * B;
        if (e < 3 * half) {
            const int t = (int)(e >> pl);
            const size_t j = e & (half - 1);
            const uint4 u0 = lds[((size_t)t << cap_lg) + 2 * j], u1 = lds[((size_t)t << cap_lg) + 2 * j + 1];
            g4 lo, d;
            lo.w[0] = u0.x; lo.w[1] = u0.y; lo.w[2] = u0.z; lo.w[3] = u0.w;
            d.w[0] = u0.x ^ u1.x; d.w[1] = u0.y ^ u1.y; d.w[2] = u0.z ^ u1.z; d.w[3] = u0.w ^ u1.w;
            nv[q] = g_add(lo, g_mul_table4(rt, d));
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kFoldPerThread; ++q) {
        const size_t e = threadIdx.x + (size_t)q * B;
        if (e < 3 * half) {
            const int t = (int)(e >> pl);
            const size_t j = e & (half - 1);
            lds[((size_t)t << cap_lg) + j] = make_uint4(nv[q].w[0], nv[q].w[1], nv[q].w[2], nv[q].w[3]);
        }
    }
    __syncthreads();
}

// A challenge's 4-bit table in two steps: the loads into registers, issued
// early, and the LDS stores once the last reader of the previous table is
// done.  256 threads.
constexpr int kRtabPer = kConstTable4Entries / kStageBlock;
template <int B = kStageBlock>
__device__ __forceinline__ void rtab_fetch(g4 (&v)[kConstTable4Entries / B], const uint4* __restrict__ src) {
#pragma unroll
    for (int q = 0; q < kConstTable4Entries / B; ++q)
        v[q] = g_load_hbm(reinterpret_cast<const f128*>(src + threadIdx.x + q * B));
}
template <int B = kStageBlock>
__device__ __forceinline__ void rtab_store(uint4* rt, const g4 (&v)[kConstTable4Entries / B]) {
#pragma unroll
    for (int q = 0; q < kConstTable4Entries / B; ++q)
        rt[threadIdx.x + q * B] = make_uint4(v[q].w[0], v[q].w[1], v[q].w[2], v[q].w[3]);
}

// Loads entries [off, off + 2^lg) of a, b and c into the LDS tables (stride
// 2^cap_lg).  Chunks of kPer 2^8 entries per table take one group of loads
// per thread, all issued before the first LDS store; smaller ones go entry
// by entry.
template <int kPer, int B = kStageBlock>
__device__ __forceinline__ void lds_load_rows(uint4* lds, int cap_lg, int lg, const f128* a, const f128* b,
                                              const f128* c, size_t off) {
    const size_t len = (size_t)1 << lg;
    if (len == (size_t)kPer * B) {
        g4 v[3][kPer];
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int q = 0; q < kPer; ++q)
                v[t][q] = g_load_hbm((t == 0 ? a : (t == 1 ? b : c)) + off + q * B + threadIdx.x);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int q = 0; q < kPer; ++q)
                lds[((size_t)t << cap_lg) + q * B + threadIdx.x] =
                    make_uint4(v[t][q].w[0], v[t][q].w[1], v[t][q].w[2], v[t][q].w[3]);
        return;
    }
    for (size_t e = threadIdx.x; e < 3 * len; e += B) {
        const int t = (int)(e >> lg);
        const g4 v = g_load_hbm((t == 0 ? a : (t == 1 ? b : c)) + off + (e & (len - 1)));
        lds[((size_t)t << cap_lg) + (e & (len - 1))] = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]);
    }
}

// The same for 2^lg folded entries per table, out[j] = in[2j] + r (in[2j] +
// in[2j+1]) over the 2^(lg+1) inputs from off, with r's table in rt.
template <int kPer, int B = kStageBlock>
__device__ __forceinline__ void lds_load_folded(uint4* lds, int cap_lg, int lg, const f128* a, const f128* b,
                                                const f128* c, size_t off, const uint4* rt) {
    const size_t len = (size_t)1 << lg;
    if (len == (size_t)kPer * B) {
        g4 lo[3][kPer], hi[3][kPer];
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                const f128* src = (t == 0 ? a : (t == 1 ? b : c)) + off + 2 * (q * B + threadIdx.x);
                lo[t][q] = g_load_hbm(src);
                hi[t][q] = g_load_hbm(src + 1);
            }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                const g4 v = g_add(lo[t][q], g_mul_table4(rt, g_add(lo[t][q], hi[t][q])));
                lds[((size_t)t << cap_lg) + q * B + threadIdx.x] = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]);
            }
        return;
    }
    for (size_t e = threadIdx.x; e < 3 * len; e += B) {
        const int t = (int)(e >> lg);
        const size_t j = e & (len - 1);
        const f128* src = (t == 0 ? a : (t == 1 ? b : c)) + off + 2 * j;
        const g4 lo = g_load_hbm(src), hi = g_load_hbm(src + 1);
        const g4 v = g_add(lo, g_mul_table4(rt, g_add(lo, hi)));
        lds[((size_t)t << cap_lg) + j] = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]);
    }
}

// One chunk of a stage, 256 threads.  Dynamic LDS (stage_smem): the three
// tables (2^cap_lg entries each), the weight tables (2^cap_lg entries) and
// 8 KiB of lane tables per wave.
template <int B>
__device__ __forceinline__ void lds_stage(const LdsStage& A, size_t g) {
    __shared__ __attribute__((aligned(256))) uint4 rt[kConstTable4Entries];
    extern __shared__ __attribute__((aligned(256))) uint4 lds_raw[];
    const int cap_lg = A.lg_in - A.fold_first;
    const int pl0 = cap_lg - 1;
    uint4* wt = lds_raw + ((size_t)3 << cap_lg);
    uint4* col = wt + ((size_t)1 << cap_lg) + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    g4 nx[kConstTable4Entries / B];
    if (A.fold_first) {
        rtab_fetch<B>(nx, A.rtabs4 + (size_t)(A.k_first - 1) * kConstTable4Entries);
        rtab_store<B>(rt, nx);
    }
    build_stage_weights<B>(wt, pl0, A.s, A.L, A.ell, A.H, A.k_first + 1, A.n, g, col);  // barrier inside
    constexpr int kPer = (1 << kStageMaxLog) / B;  // loads per thread and table of a full chunk
    if (A.fold_first)
        lds_load_folded<kPer, B>(lds_raw, cap_lg, cap_lg, A.a, A.b, A.c, g << A.lg_in, rt);
    else
        lds_load_rows<kPer, B>(lds_raw, cap_lg, cap_lg, A.a, A.b, A.c, g << A.lg_in);
    __syncthreads();
    int lg = cap_lg;
    for (int l = 0; l < A.s; ++l) {
        const int k = A.k_first + l;
        const int pl = lg - 1;
        const bool fold = l + 1 < A.s || A.final_stage;
        if (fold) rtab_fetch<B>(nx, A.rtabs4 + (size_t)k * kConstTable4Entries);
        lds_round_sums<B>(lds_raw, cap_lg, pl, wt + stage_woff(pl0, l), (A.q1mask >> k) & 1, A.acc + 12 * k,
                       A.store_sums, col);
        if (fold) {
            __syncthreads();  // every reader of rt is done
            rtab_store<B>(rt, nx);
            __syncthreads();
            lds_fold<(3 << (kStageMaxLog - 1)) / B, B>(lds_raw, cap_lg, pl, rt);
            lg = pl;
        }
    }
    if (A.final_stage) {
        if (threadIdx.x < 3) *reinterpret_cast<uint4*>(A.fin + threadIdx.x) = lds_raw[(size_t)threadIdx.x << cap_lg];
        return;
    }
    const size_t len = (size_t)1 << lg;
    for (size_t e = threadIdx.x; e < 3 * len; e += B) {
        const int t = (int)(e >> lg);
        const size_t j = e & (len - 1);
        f128* dst = (t == 0 ? A.ao : (t == 1 ? A.bo : A.co)) + (g << lg) + j;
        *reinterpret_cast<uint4*>(dst) = lds_raw[((size_t)t << cap_lg) + j];
    }
}

// Middle rounds of a single proof: workgroup g runs the stage on chunk g.
__global__ __launch_bounds__(kOuterStageBlock) void k_outer_chunk(LdsStage A) { lds_stage<kOuterStageBlock>(A, blockIdx.x); }

// The last rounds of a single proof in one workgroup (a final stage of one
// chunk): barriers instead of kernel launches between rounds.
__global__ __launch_bounds__(kOuterStageBlock) void k_outer_tail(LdsStage A) { lds_stage<kOuterStageBlock>(A, 0); }

// ======================================= batches of independent proofs ===
__global__ __launch_bounds__(1024) void k_batch_prologue(const BatchPrologue* items, const uint32_t* bmap) {
    __shared__ uint4 tab[kConstTable4Entries];
    __shared__ uint4 basis[128];
    const BatchPrologue& P = items[bmap[blockIdx.x]];
    const unsigned b = blockIdx.x - P.block_base;
    if (b == 0)
        for (unsigned e = threadIdx.x; e < P.zero_words; e += blockDim.x) P.zero[e] = 0;
    if (b < P.eq_blocks) {
        eq_check_block(P.L, P.ell, P.H, P.m, P.eq_out, P.eq_blocks_l, b);
        return;
    }
    const unsigned t = b - P.eq_blocks;
    if (t >= (unsigned)P.n) {
        h4_block(P.L, P.ell, P.H, P.hdesc, P.htabs4, t - P.n, tab, basis);
        return;
    }
    build_const_table4(tab, basis, g_load(P.chal + t));
    uint4* dst = P.rtabs4 + (size_t)t * kConstTable4Entries;
    for (int e = threadIdx.x; e < kConstTable4Entries; e += blockDim.x) dst[e] = tab[e];
}

// One workgroup per (proof, chunk) of a batch stage; bmap gives each
// workgroup's stage descriptor.
__global__ __launch_bounds__(kStageBlock) void k_outer_batch(const LdsStage* stages, const uint32_t* bmap) {
    const LdsStage A = stages[bmap[blockIdx.x]];
    lds_stage<kStageBlock>(A, blockIdx.x - A.chunk_base);
}

// Grid-wide rounds of a batch: one launch per round over every proof still
// in its grid-wide rounds, workgroup -> descriptor through bmap.
__global__ __launch_bounds__(kR0Block) void k_batch_r0(const RoundDesc* d, const uint32_t* bmap) {
    const RoundDesc& A = d[bmap[blockIdx.x]];
    const size_t block = blockIdx.x - A.block_base;
    if (A.q1)
        outer_r0_body<true>(A.a, A.b, A.c, A.threads, A.iters, A.la, A.la_bits, A.la_shift, A.lb, A.lb_shift,
                            A.h_tabs4, A.acc, block);
    else
        outer_r0_body<false>(A.a, A.b, A.c, A.threads, A.iters, A.la, A.la_bits, A.la_shift, A.lb, A.lb_shift,
                             A.h_tabs4, A.acc, block);
}
__global__ __launch_bounds__(kFoldBlock) void k_batch_fold(const RoundDesc* d, const uint32_t* bmap) {
    const RoundDesc& A = d[bmap[blockIdx.x]];
    const size_t block = blockIdx.x - A.block_base;
    if (A.q1)
        outer_fold_body<true>(A.a, A.b, A.c, A.ao, A.bo, A.co, A.threads, A.iters, A.r_tab4, A.la, A.la_bits,
                              A.la_shift, A.lb, A.lb_shift, A.h_tabs4, A.acc, block);
    else
        outer_fold_body<false>(A.a, A.b, A.c, A.ao, A.bo, A.co, A.threads, A.iters, A.r_tab4, A.la, A.la_bits,
                               A.la_shift, A.lb, A.lb_shift, A.h_tabs4, A.acc, block);
}

// ======================== inner sumcheck with challenges known in advance ===
// [c0, c2] of round k = [sum m0 w0, sum (m1+m0)(w1+w0)] over the round's pairs
// (sumcheck.rs:568-609), products through lane tables.
// With live_out, also the tables' live length: 1 + the last index where D or
// h is nonzero (an atomic max), so later rounds can skip a zero tail.
__device__ __forceinline__ void inner_r0_body(const f128* d, const f128* w, size_t pairs, uint32_t* acc,
                                              unsigned long long* live_out, size_t block, size_t blocks) {
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    uint4* col = smem_raw + kConstTable4Entries + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    g4 s0 = g_zero(), s2 = g_zero();
    unsigned long long live = 0;
    auto nonzero = [](const g4& a) { return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) != 0; };
#pragma unroll 1
    for (size_t p = block * kInnerBlock + threadIdx.x; p < pairs; p += blocks * kInnerBlock) {
        const g4 m0 = g_load_hbm(d + 2 * p), m1 = g_load_hbm(d + 2 * p + 1);
        const g4 w0 = g_load_hbm(w + 2 * p), w1 = g_load_hbm(w + 2 * p + 1);
        if (nonzero(m1) || nonzero(w1))
            live = 2 * p + 2;
        else if (nonzero(m0) || nonzero(w0))
            live = 2 * p + 1;
        s0 = g_add(s0, g_mul_lane(m0, w0, col));
        __builtin_amdgcn_sched_barrier(0);  // one lane product in flight at a time
        s2 = g_add(s2, g_mul_lane(g_add(m1, m0), g_add(w1, w0), col));
    }
    g4 v[2] = {s0, s2};
    block_xor_to_global<2>(v, acc);
    if (live_out) {
        for (int off = 32; off >= 1; off >>= 1) live = max(live, (unsigned long long)__shfl_xor(live, off, 64));
        if ((threadIdx.x & 63) == 0 && live) atomicMax(live_out, live);
    }
}

__global__ __launch_bounds__(kInnerBlock) void k_inner_r0_lane(const f128* d, const f128* w, size_t pairs,
                                                              uint32_t* acc, unsigned long long* live_out) {
    inner_r0_body(d, w, pairs, acc, live_out, blockIdx.x, gridDim.x);
}

// Fold by r_{k-1} on load (4 constant-table products per output pair) and
// round k's sums (fold_and_compute_next_inner_round_coefficients_without_linear,
// sumcheck.rs:641-688).
// Inputs at index live and above are zero (not read): a table whose tail is
// known to be zero is folded over its live prefix only.
__device__ __forceinline__ void inner_fold_body(const f128* d, const f128* w, f128* dout, f128* wout,
                                                size_t out_pairs, size_t live, const uint4* rtab4, uint32_t* acc,
                                                size_t block, size_t blocks,
                                                const unsigned long long* live0 = nullptr, int live_shift = 0) {
    if (live0) {  // the live length measured on the device by round 0
        const unsigned long long l0 = *live0;
        live = min(live, (size_t)((l0 + (1ull << live_shift) - 1) >> live_shift));
        out_pairs = min(out_pairs, ((live + 1) / 2 + 1) / 2);
    }
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    uint4* rt = smem_raw;
    uint4* col = smem_raw + kConstTable4Entries + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    for (int e = threadIdx.x; e < kConstTable4Entries; e += kInnerBlock) rt[e] = u4_load_hbm(rtab4 + e);
    __syncthreads();
    g4 s0 = g_zero(), s2 = g_zero();
    auto ip = [&](const g4& a, const g4& b) { return g_add(a, g_mul_table4(rt, g_add(a, b))); };
#pragma unroll 1
    for (size_t p = block * kInnerBlock + threadIdx.x; p < out_pairs; p += blocks * kInnerBlock) {
        g4 in[2][4];
        if (4 * p + 3 < live) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                in[0][j] = g_load_hbm(d + 4 * p + j);
                in[1][j] = g_load_hbm(w + 4 * p + j);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool in_live = 4 * p + j < live;
                in[0][j] = in_live ? g_load_hbm(d + 4 * p + j) : g_zero();
                in[1][j] = in_live ? g_load_hbm(w + 4 * p + j) : g_zero();
            }
        }
        const g4 m0 = ip(in[0][0], in[0][1]), m1 = ip(in[0][2], in[0][3]);
        const g4 w0 = ip(in[1][0], in[1][1]), w1 = ip(in[1][2], in[1][3]);
        g_store_hbm(dout + 2 * p, m0);
        g_store_hbm(dout + 2 * p + 1, m1);
        g_store_hbm(wout + 2 * p, w0);
        g_store_hbm(wout + 2 * p + 1, w1);
        s0 = g_add(s0, g_mul_lane(m0, w0, col));
        __builtin_amdgcn_sched_barrier(0);  // one lane product in flight at a time
        s2 = g_add(s2, g_mul_lane(g_add(m1, m0), g_add(w1, w0), col));
    }
    g4 v[2] = {s0, s2};
    block_xor_to_global<2>(v, acc);
}

__global__ __launch_bounds__(kInnerBlock) void k_inner_fold_lane(const f128* d, const f128* w, f128* dout, f128* wout,
                                                                size_t out_pairs, size_t live, const uint4* rtab4,
                                                                uint32_t* acc) {
    inner_fold_body(d, w, dout, wout, out_pairs, live, rtab4, acc, blockIdx.x, gridDim.x);
}

// The same rounds of many inner sumchecks in one launch: workgroup -> round
// descriptor through bmap, each descriptor's workgroups striding its pairs.
__global__ __launch_bounds__(kInnerBlock) void k_inner_r0_batch(const InnerRoundDesc* d, const uint32_t* bmap) {
    const InnerRoundDesc A = d[bmap[blockIdx.x]];
    inner_r0_body(A.d, A.w, A.pairs, A.acc, A.live_out, blockIdx.x - A.block_base, A.blocks);
}
__global__ __launch_bounds__(kInnerBlock) BITZ_FOUR_WAVES void k_inner_fold_batch(const InnerRoundDesc* d, const uint32_t* bmap) {
    const InnerRoundDesc A = d[bmap[blockIdx.x]];
    inner_fold_body(A.d, A.w, A.dout, A.wout, A.pairs, A.live, A.rtab4, A.acc, blockIdx.x - A.block_base, A.blocks,
                    A.live0, A.live_shift);
}

// Rounds k0 .. n-1 in one workgroup: the two tables (folded by r_{k0-1} on
// load when fold_first) in LDS, each round's sums stored at acc + 8 k, then a
// fold by r_k; D(r) and h(r) at fin.
__global__ __launch_bounds__(kStageBlock) void k_inner_tail(InnerTail A) {
    extern __shared__ __attribute__((aligned(256))) uint4 lds_raw[];
    constexpr size_t kCap = (size_t)1 << kInnerTailMaxLog;
    uint4* Ds = lds_raw;
    uint4* Ws = Ds + kCap;
    uint4* rt = Ws + kCap;
    uint4* col = rt + kConstTable4Entries + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    auto ld = [](const uint4* p) {
        const uint4 u = *p;
        g4 r;
        r.w[0] = u.x; r.w[1] = u.y; r.w[2] = u.z; r.w[3] = u.w;
        return r;
    };
    auto st = [](uint4* p, const g4& v) { *p = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]); };
    int lg = A.lg_in - A.fold_first;
    const size_t len = (size_t)1 << lg;
    const size_t base = (size_t)blockIdx.x << A.lg_in;  // this chunk's first input
    g4 nx[kRtabPer];
    if (A.fold_first) {
        rtab_fetch(nx, A.rtabs4 + (size_t)(A.k0 - 1) * kConstTable4Entries);
        rtab_store(rt, nx);
        __syncthreads();
        for (size_t e = threadIdx.x; e < 2 * len; e += kStageBlock) {
            const f128* src = e < len ? A.d + base + 2 * e : A.w + base + 2 * (e - len);
            const g4 a = g_load_hbm(src), b = g_load_hbm(src + 1);
            st((e < len ? Ds + e : Ws + (e - len)), g_add(a, g_mul_table4(rt, g_add(a, b))));
        }
    } else {
        for (size_t e = threadIdx.x; e < len; e += kStageBlock) {
            Ds[e] = u4_load_hbm(reinterpret_cast<const uint4*>(A.d + base + e));
            Ws[e] = u4_load_hbm(reinterpret_cast<const uint4*>(A.w + base + e));
        }
    }
    __syncthreads();
    for (int k = A.k0; k < A.k0 + A.s; ++k) {
        const size_t P = (size_t)1 << (lg - 1);
        rtab_fetch(nx, A.rtabs4 + (size_t)k * kConstTable4Entries);
        g4 s0 = g_zero(), s2 = g_zero();
        for (size_t p = threadIdx.x; p < P; p += kStageBlock) {
            const g4 m0 = ld(Ds + 2 * p), m1 = ld(Ds + 2 * p + 1);
            const g4 w0 = ld(Ws + 2 * p), w1 = ld(Ws + 2 * p + 1);
            s0 = g_add(s0, g_mul_lane(m0, w0, col));
            __builtin_amdgcn_sched_barrier(0);
            s2 = g_add(s2, g_mul_lane(g_add(m1, m0), g_add(w1, w0), col));
        }
        g4 v[2] = {s0, s2};
        if (A.dout)
            block_xor_to_global<2>(v, A.acc + 8 * (size_t)k, kStageBlock);
        else
            block_xor_store<2>(v, A.acc + 8 * (size_t)k, kStageBlock);
        __syncthreads();  // readers of rt done (and of the tables, before the fold writes)
        rtab_store(rt, nx);
        __syncthreads();
        // fold both tables in place: read, barrier, write (2P <= 2^kInnerTailMaxLog
        // outputs, at most kFold per thread)
        constexpr int kFold = (2 << (kInnerTailMaxLog - 1)) / kStageBlock;
        g4 nv[kFold];
#pragma unroll
        for (int q = 0; q < kFold; ++q) {
            const size_t e = threadIdx.x + (size_t)q * kStageBlock;
            if (e < 2 * P) {
                const uint4* T = e < P ? Ds : Ws;
                const size_t j = e < P ? e : e - P;
                const g4 a = ld(T + 2 * j), b = ld(T + 2 * j + 1);
                nv[q] = g_add(a, g_mul_table4(rt, g_add(a, b)));
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kFold; ++q) {
            const size_t e = threadIdx.x + (size_t)q * kStageBlock;
            if (e < 2 * P) {
                uint4* T = e < P ? Ds : Ws;
                st(T + (e < P ? e : e - P), nv[q]);
            }
        }
        __syncthreads();
        lg -= 1;
    }
    if (A.dout) {
        const size_t out = (size_t)1 << lg;
        for (size_t e = threadIdx.x; e < out; e += kStageBlock) {
            *reinterpret_cast<uint4*>(A.dout + ((size_t)blockIdx.x << lg) + e) = Ds[e];
            *reinterpret_cast<uint4*>(A.wout + ((size_t)blockIdx.x << lg) + e) = Ws[e];
        }
        return;
    }
    if (threadIdx.x < 2) reinterpret_cast<uint4*>(A.fin)[threadIdx.x] = threadIdx.x == 0 ? Ds[0] : Ws[0];
}

// ============================ inner sumcheck, Boolean assignment (bits) ===
// Round K over output pair p: D_K(2p), D_K(2p+1) (round 0: D's pair; round
// K >= 1: four entries of D_{K-1} folded by r_{K-1} on load and written), and
// the 2^(K+1) assignment bits under them, h(2^(K+1) p + i).  A subset-XOR
// transform turns each half (2^K bits) into c_S for every S; D_K(2p) is XORed
// into sum S of c0 where c_S(2p) is set, D_K(2p) + D_K(2p+1) into sum S of c2
// where c_S(2p) and c_S(2p+1) differ.  No GF(2^128) products besides the
// fold's two constant-table products (skipped by a wave whose differences
// are all zero).
template <int K>
__device__ __forceinline__ void inner_bits_body(const InnerBitsDesc& A, size_t block, size_t blocks) {
    constexpr int S = 1 << K;
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    uint4* rt = smem_raw;
    if constexpr (K > 0) {
        for (int e = threadIdx.x; e < kConstTable4Entries; e += kInnerBlock) rt[e] = u4_load_hbm(A.rtab4 + e);
        __syncthreads();
    }
    auto nonzero = [](const g4& a) { return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) != 0; };
    g4 s0[S], s2[S];
#pragma unroll
    for (int s = 0; s < S; ++s) s0[s] = s2[s] = g_zero();
    unsigned long long live = 0;
    size_t in_live = A.live, pairs = A.pairs;
    if (K > 0 && A.live0) {  // the live length measured on the device by round 0
        const unsigned long long l0 = *A.live0, sh = K - 1;
        in_live = min(in_live, (size_t)((l0 + (1ull << sh) - 1) >> sh));
        pairs = min(pairs, ((in_live + 1) / 2 + 1) / 2);
    }
    const f128* mono = A.mono_dev ? A.mono_dev : A.mono;
#pragma unroll 1
    for (size_t p = block * kInnerBlock + threadIdx.x; p < pairs; p += blocks * kInnerBlock) {
        g4 m0, m1;
        if constexpr (K == 0) {
            m0 = g_load_hbm(A.d + 2 * p);
            m1 = g_load_hbm(A.d + 2 * p + 1);
        } else {
            g4 in[4];
            if (4 * p + 3 < in_live) {
#pragma unroll
                for (int j = 0; j < 4; ++j) in[j] = g_load_hbm(A.d + 4 * p + j);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) in[j] = 4 * p + j < in_live ? g_load_hbm(A.d + 4 * p + j) : g_zero();
            }
            const g4 d01 = g_add(in[0], in[1]), d23 = g_add(in[2], in[3]);
            g4 t01 = g_zero(), t23 = g_zero();
            if (__builtin_amdgcn_ballot_w64(nonzero(d01) || nonzero(d23))) {
                t01 = g_mul_table4(rt, d01);
                t23 = g_mul_table4(rt, d23);
            }
            m0 = g_add(in[0], t01);
            m1 = g_add(in[2], t23);
            g_store_hbm(A.dout + 2 * p, m0);
            g_store_hbm(A.dout + 2 * p + 1, m1);
        }
        const size_t bit0 = p << (K + 1);
        uint32_t v = (uint32_t)(A.bits[bit0 >> 6] >> (bit0 & 63));
        v &= (2u << (2 * S - 1)) - 1;  // the 2S bits of this pair
        // Subset XOR within each half: bit x becomes XOR of h(y), y subset of x.
        if constexpr (K >= 1) v ^= (v << 1) & 0xAAAAAAAAu;
        if constexpr (K >= 2) v ^= (v << 2) & 0xCCCCCCCCu;
        if constexpr (K >= 3) v ^= (v << 4) & 0xF0F0F0F0u;
        const uint32_t vd = v ^ (v >> S);
        const g4 dm = g_add(m0, m1);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const uint32_t k0 = 0u - ((v >> s) & 1u), kd = 0u - ((vd >> s) & 1u);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                s0[s].w[i] = xor_and(s0[s].w[i], m0.w[i], k0);
                s2[s].w[i] = xor_and(s2[s].w[i], dm.w[i], kd);
            }
        }
        if (A.wout) {  // h_K(2p), h_K(2p+1) = sum_S r_S c_S
            g4 w0 = g_zero(), w1 = g_zero();
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const g4 r = g_from(mono[s]);
                const uint32_t k0 = 0u - ((v >> s) & 1u), k1 = 0u - ((v >> (S + s)) & 1u);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    w0.w[i] = xor_and(w0.w[i], r.w[i], k0);
                    w1.w[i] = xor_and(w1.w[i], r.w[i], k1);
                }
            }
            g_store_hbm(A.wout + 2 * p, w0);
            g_store_hbm(A.wout + 2 * p + 1, w1);
        }
        if constexpr (K == 0) {
            if (nonzero(m1) || (v & 2u))
                live = 2 * p + 2;
            else if (nonzero(m0) || (v & 1u))
                live = 2 * p + 1;
        }
    }
    uint32_t x[8 * S];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            x[4 * s + i] = s0[s].w[i];
            x[4 * S + 4 * s + i] = s2[s].w[i];
        }
    uint32_t* rep = A.acc + (block % kInnerBitsReplicas) * kInnerBitsReplicaWords;
    if constexpr (K == 0) {
        __shared__ unsigned long long lv[kInnerBlock / 64];
        for (int off = 32; off >= 1; off >>= 1) live = max(live, (unsigned long long)__shfl_xor(live, off, 64));
        if ((threadIdx.x & 63) == 0) lv[threadIdx.x >> 6] = live;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long m = 0;
            for (int w = 0; w < kInnerBlock / 64; ++w) m = max(m, lv[w]);
            if (m) atomicMax(reinterpret_cast<unsigned long long*>(rep + kInnerBitsLiveWord), m);
        }
    }
    block_xor_words<8 * S>(x, rep);
}

template <int K>
__global__ __launch_bounds__(kInnerBlock) void k_inner_bits(InnerBitsDesc A) {
    inner_bits_body<K>(A, blockIdx.x, A.blocks);
}
template <int K>
__global__ __launch_bounds__(kInnerBlock) void k_inner_bits_batch(const InnerBitsDesc* d, const uint32_t* bmap) {
    const InnerBitsDesc A = d[bmap[blockIdx.x]];
    inner_bits_body<K>(A, blockIdx.x - A.block_base, A.blocks);
}
// Bit rounds of different K in one launch (the PIOP engine's positions):
// each descriptor's K picks the body; kInnerBitsSmem of dynamic LDS.
__global__ __launch_bounds__(kInnerBlock) void k_inner_bits_mixed(const InnerBitsDesc* d, const uint32_t* bmap) {
    const InnerBitsDesc A = d[bmap[blockIdx.x]];
    const size_t block = blockIdx.x - A.block_base;
    switch (A.k) {
        case 0: inner_bits_body<0>(A, block, A.blocks); break;
        case 1: inner_bits_body<1>(A, block, A.blocks); break;
        case 2: inner_bits_body<2>(A, block, A.blocks); break;
        default: inner_bits_body<3>(A, block, A.blocks); break;
    }
}
template __global__ void k_inner_bits<0>(InnerBitsDesc);
template __global__ void k_inner_bits<1>(InnerBitsDesc);
template __global__ void k_inner_bits<2>(InnerBitsDesc);
template __global__ void k_inner_bits<3>(InnerBitsDesc);
template __global__ void k_inner_bits_batch<0>(const InnerBitsDesc*, const uint32_t*);
template __global__ void k_inner_bits_batch<1>(const InnerBitsDesc*, const uint32_t*);
template __global__ void k_inner_bits_batch<2>(const InnerBitsDesc*, const uint32_t*);
template __global__ void k_inner_bits_batch<3>(const InnerBitsDesc*, const uint32_t*);
static_assert(kInnerBitsMaxRound == 3, "the instantiations above cover rounds 0..3");

__global__ __launch_bounds__(kBlock) void k_r1cs_rows_bits(const uint32_t* row_starts, const uint32_t* cols,
                                                           const f128* vals, const uint64_t* bits, size_t rows,
                                                           f128* out, size_t padded) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < padded; i += (size_t)gridDim.x * blockDim.x) {
        g4 s = g_zero();
        if (i < rows)
            for (uint32_t k = row_starts[i]; k < row_starts[i + 1]; ++k) {
                const uint32_t c = cols[k];
                const uint32_t m = 0u - (uint32_t)((bits[c >> 6] >> (c & 63)) & 1u);
                const g4 a = g_load(vals + k);
#pragma unroll
                for (int w = 0; w < 4; ++w) s.w[w] = xor_and(s.w[w], a.w[w], m);
            }
        g_store(out + i, s);
    }
}

// ================================ GKR layers in parallel (known challenges) ===
__global__ __launch_bounds__(kStageBlock) void k_gkr_eq_items(const GkrEqItem* items) {
    const GkrEqItem& A = items[blockIdx.x];
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    uint4* col = smem_raw + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    f128* table = A.out;
    const int k = A.k;
    if (threadIdx.x == 0) g_store_hbm(table, g_one());
    __syncthreads();
    for (int i = 0; i < k; ++i) {
        const size_t half = (size_t)1 << i;
        const g4 ri = g_load_hbm(A.r + i);
        for (size_t j = threadIdx.x; j < half; j += kStageBlock) {
            const g4 parent = g_load_hbm(table + j);
            const g4 one = g_mul_lane(parent, ri, col);
            g_store_hbm(table + j, g_add(parent, one));
            g_store_hbm(table + j + half, one);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kStageBlock) void k_gkr_eq_combine_batch(const GkrCombine* d, const uint32_t* bmap) {
    const GkrCombine A = d[bmap[blockIdx.x]];
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    uint4* col = smem_raw + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    const size_t mask = ((size_t)1 << A.low_bits) - 1;
#pragma unroll 1
    for (size_t j = (blockIdx.x - A.block_base) * (size_t)kStageBlock + threadIdx.x; j < A.n;
         j += (size_t)A.blocks * kStageBlock)
        g_store_hbm(A.out + j,
                    g_mul_lane(g_load_hbm(A.low + (j & mask)), g_load_hbm(A.high + (j >> A.low_bits)), col));
}

// Dense leaves of a BitTable (grand_product, crates/gkr/src/lib.rs:785-802;
// the reference's own dense construction, lib.rs:1023-1037): leaf
// (b << s) | c is images[b] where bit (c << log_rows) | b of the packed table
// is set, one elsewhere.
__global__ __launch_bounds__(kBlock) void k_bittable_leaves(const uint64_t* packed, const f128* images, int log_rows,
                                                            int log_columns, f128* leaves) {
    const size_t n = (size_t)1 << (log_rows + log_columns);
    const size_t cmask = ((size_t)1 << log_columns) - 1;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t b = i >> log_columns, c = i & cmask;
        const size_t bit = (c << log_rows) | b;
        const bool set = (packed[bit >> 6] >> (bit & 63)) & 1;
        g_store_hbm(leaves + i, set ? g_load_hbm(images + b) : g_one());
    }
}

// next[k] = cur[k] * cur[k + half] (fold_layers, crates/gkr/src/lib.rs:874-894).
__global__ __launch_bounds__(kStageBlock) void k_gkr_layer_lane(const f128* cur, f128* cur_copy, f128* next,
                                                                size_t half) {
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    uint4* col = smem_raw + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
#pragma unroll 1
    for (size_t k = blockIdx.x * (size_t)kStageBlock + threadIdx.x; k < half; k += (size_t)gridDim.x * kStageBlock) {
        const g4 x = g_load_hbm(cur + k), y = g_load_hbm(cur + k + half);
        if (cur_copy) {
            g_store_hbm(cur_copy + k, x);
            g_store_hbm(cur_copy + k + half, y);
        }
        g_store_hbm(next + k, g_mul_lane(x, y, col));
    }
}

// The `count` layers after `layer` (2^log_len elements, at most
// 2^kGkrSmallTreeLog) in one workgroup: the current layer sits in LDS and is
// halved in place (thread k reads k and k + half, writes k); each new layer is
// also written right after the previous one in the store.
__global__ __launch_bounds__(kStageBlock) void k_gkr_layers_small(f128* layer, int log_len, int count) {
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    uint4* T = smem_raw;
    uint4* col = smem_raw + ((size_t)1 << kGkrSmallTreeLog) + (threadIdx.x >> 6) * (64 * kLaneTableEntries) +
                 (threadIdx.x & 63);
    lane_table_clear(col);
    auto ld = [](const uint4* p) {
        const uint4 u = *p;
        g4 r;
        r.w[0] = u.x;
        r.w[1] = u.y;
        r.w[2] = u.z;
        r.w[3] = u.w;
        return r;
    };
    size_t cur = (size_t)1 << log_len;
    for (size_t e = threadIdx.x; e < cur; e += kStageBlock)
        T[e] = u4_load_hbm(reinterpret_cast<const uint4*>(layer + e));
    __syncthreads();
    f128* dst = layer + cur;
    for (int j = 0; j < count; ++j) {
        const size_t half = cur / 2;
        for (size_t k = threadIdx.x; k < half; k += kStageBlock) {
            const g4 v = g_mul_lane(ld(T + k), ld(T + k + half), col);
            T[k] = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]);
            g_store_hbm(dst + k, v);
        }
        __syncthreads();
        dst += half;
        cur = half;
    }
}

// One layer round's share of workgroup `block` of A.blocks: the sums
// (binding by r_(i+1) first unless it is the layer's first round), XORed
// into A.acc.
__device__ __forceinline__ void gkr_round_body(const GkrRound& A, size_t block) {
    extern __shared__ __attribute__((aligned(256))) uint4 smem_raw[];
    uint4* rt = smem_raw;
    uint4* col = smem_raw + kConstTable4Entries + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    const size_t start = block * (size_t)kStageBlock + threadIdx.x;
    const size_t stride = (size_t)A.blocks * kStageBlock;
    const size_t h = A.h;
    const int at_one = A.z ? (int)g_is_zero(g_load_hbm(A.z)) : A.at_one;
    f128* left = A.left;
    f128* right = A.right;
    f128* eq = A.eq;
    g4 sb = g_zero(), si = g_zero();
    auto sums = [&](const g4& l0, const g4& l1, const g4& r0, const g4& r1, const g4& e) {
        sb = g_add(sb, g_mul_lane(g_mul_lane(e, at_one ? l1 : l0, col), at_one ? r1 : r0, col));
        __builtin_amdgcn_sched_barrier(0);  // one lane product chain in flight at a time
        si = g_add(si, g_mul_lane(g_mul_lane(e, g_add(l1, l0), col), g_add(r1, r0), col));
    };
    if (A.first) {
#pragma unroll 1
        for (size_t x = start; x < h; x += stride)
            sums(g_load_hbm(left + x), g_load_hbm(left + x + h), g_load_hbm(right + x), g_load_hbm(right + x + h),
                 g_load_hbm(eq + x));
    } else {
        for (int e = threadIdx.x; e < kConstTable4Entries; e += kStageBlock) rt[e] = u4_load_hbm(A.rt4 + e);
        __syncthreads();
        const size_t hh = 2 * h;
        auto ip = [&](const g4& u, const g4& v) { return g_add(u, g_mul_table4(rt, g_add(u, v))); };
#pragma unroll 1
        for (size_t x = start; x < h; x += stride) {
            const g4 l0 = ip(g_load_hbm(left + x), g_load_hbm(left + x + hh));
            const g4 l1 = ip(g_load_hbm(left + x + h), g_load_hbm(left + x + h + hh));
            const g4 r0 = ip(g_load_hbm(right + x), g_load_hbm(right + x + hh));
            const g4 r1 = ip(g_load_hbm(right + x + h), g_load_hbm(right + x + h + hh));
            g_store_hbm(left + x, l0);
            g_store_hbm(left + x + h, l1);
            g_store_hbm(right + x, r0);
            g_store_hbm(right + x + h, r1);
            const g4 e = g_add(g_load_hbm(eq + x), g_load_hbm(eq + x + h));
            g_store_hbm(eq + x, e);
            sums(l0, l1, r0, r1, e);
        }
    }
    g4 v[2] = {sb, si};
    block_xor_to_global<2>(v, A.acc, kStageBlock);
}

__global__ __launch_bounds__(kStageBlock) void k_gkr_round_batch(const GkrRound* d, const uint32_t* bmap) {
    const GkrRound A = d[bmap[blockIdx.x]];
    gkr_round_body(A, blockIdx.x - A.block_base);
}

// The layer's arrays in LDS: L and R (2^(i0+2) entries at most), E, the
// current round's r table, then the lane tables.  Round i (h = 2^i) binds
// the halves [0, 2h) x [2h, 4h) by r_(i+1) in place (thread x owns entries
// x, x+h, x+2h, x+3h) and halves E, then sums over the bound [0, h) x [h, 2h);
// the layer's first round sums without binding.
__global__ __launch_bounds__(kStageBlock) void k_gkr_tail(const GkrTail* d) {
    const GkrTail A = d[blockIdx.x];
    extern __shared__ __attribute__((aligned(256))) uint4 lds_raw[];
    constexpr size_t kArr = (size_t)4 << kGkrTailMax;
    uint4* Ls = lds_raw;
    uint4* Rs = Ls + kArr;
    uint4* Es = Rs + kArr;
    uint4* rt = Es + ((size_t)2 << kGkrTailMax);
    uint4* col = rt + kConstTable4Entries + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    const int n = A.n, i0 = A.i0;
    const bool whole = i0 == n - 1;
    const size_t len = whole ? (size_t)1 << n : (size_t)4 << i0;
    const size_t elen = whole ? (size_t)1 << (n - 1) : (size_t)2 << i0;
    auto ld = [](const uint4* p) {
        const uint4 u = *p;
        g4 r;
        r.w[0] = u.x; r.w[1] = u.y; r.w[2] = u.z; r.w[3] = u.w;
        return r;
    };
    auto st = [](uint4* p, const g4& v) { *p = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]); };
    for (size_t e = threadIdx.x; e < len; e += kStageBlock) {
        Ls[e] = u4_load_hbm(reinterpret_cast<const uint4*>(A.left + e));
        Rs[e] = u4_load_hbm(reinterpret_cast<const uint4*>(A.right + e));
    }
    for (size_t e = threadIdx.x; e < elen; e += kStageBlock) Es[e] = u4_load_hbm(reinterpret_cast<const uint4*>(A.eq + e));
    g4 nx[kRtabPer];
    if (!whole) rtab_fetch(nx, A.rtabs4 + (size_t)(n - 1 - (i0 + 1)) * kConstTable4Entries);
    __syncthreads();
    for (int i = i0; i >= 0; --i) {
        const size_t h = (size_t)1 << i;
        const bool bind = !(whole && i == n - 1);
        if (bind) {
            rtab_store(rt, nx);
            __syncthreads();
        }
        const int at_one = (A.at_one >> i) & 1;
        g4 sb = g_zero(), si = g_zero();
        for (size_t x = threadIdx.x; x < h; x += kStageBlock) {
            g4 l0, l1, r0, r1, e;
            if (bind) {
                const size_t hh = 2 * h;
                auto ip = [&](const uint4* p, size_t a, size_t b) {
                    const g4 u = ld(p + a), v = ld(p + b);
                    return g_add(u, g_mul_table4(rt, g_add(u, v)));
                };
                l0 = ip(Ls, x, x + hh);
                l1 = ip(Ls, x + h, x + h + hh);
                r0 = ip(Rs, x, x + hh);
                r1 = ip(Rs, x + h, x + h + hh);
                st(Ls + x, l0);
                st(Ls + x + h, l1);
                st(Rs + x, r0);
                st(Rs + x + h, r1);
                e = g_add(ld(Es + x), ld(Es + x + h));
                st(Es + x, e);
            } else {
                l0 = ld(Ls + x);
                l1 = ld(Ls + x + h);
                r0 = ld(Rs + x);
                r1 = ld(Rs + x + h);
                e = ld(Es + x);
            }
            g4 el, ed;
            g_mul_lane2(e, at_one ? l1 : l0, g_add(l1, l0), col, el, ed);
            sb = g_add(sb, g_mul_lane(el, at_one ? r1 : r0, col));
            si = g_add(si, g_mul_lane(ed, g_add(r1, r0), col));
        }
        // the next table: r_i binds round i-1 (or the final bind after round 0)
        rtab_fetch(nx, A.rtabs4 + (size_t)(n - 1 - i) * kConstTable4Entries);
        g4 v[2] = {sb, si};
        block_xor_store<2>(v, A.acc + 8 * (size_t)(n - 1 - i), kStageBlock);
        __syncthreads();  // every round-i reader of rt, L, R and E is done
    }
    // Final bind of the two remaining entries by r_0.
    rtab_store(rt, nx);
    __syncthreads();
    if (threadIdx.x < 2) {
        uint4* P = threadIdx.x == 0 ? Ls : Rs;
        const g4 u = ld(P), v = ld(P + 1);
        const g4 w = g_add(u, g_mul_table4(rt, g_add(u, v)));
        reinterpret_cast<uint4*>(A.evals)[threadIdx.x] = make_uint4(w.w[0], w.w[1], w.w[2], w.w[3]);
    }
}

// eq(z, r) for one coordinate: 1 + z + r in characteristic 2.
__device__ __forceinline__ g4 eq_factor(const g4& r, const g4& z) {
    g4 e = g_add(r, z);
    e.w[0] ^= 1u;
    return e;
}

// The round message [eq_bound sum_b, eq_bound sum_inf], one product per wave
// (waves 0 and 1), for every thread.
__device__ __forceinline__ void gkr_fs_messages(const g4& eb, const g4& s0, const g4& s1, uint4* mbuf, g4 (&m)[2]) {
    const int wave = (int)(threadIdx.x >> 6);
    if (wave < 2) {
        const g4 v = g_mul_wave(eb, wave == 0 ? s0 : s1);
        if ((threadIdx.x & 63) == 0) mbuf[wave] = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const uint4 u = mbuf[k];
        m[k].w[0] = u.x; m[k].w[1] = u.y; m[k].w[2] = u.z; m[k].w[3] = u.w;
    }
}

// lib.rs:142-150 after a grid-wide round: the message as a prover message,
// the challenge, eq_bound times eq(z, r), and r's table for the next round's
// bind.  Called by every thread of one workgroup with the round's sums.
__device__ __forceinline__ void gkr_fs_step(const GkrFsRound& A, const g4& s0, const g4& s1) {
    __shared__ FsLds f;
    __shared__ uint4 tab[kConstTable4Entries];
    __shared__ uint4 basis[128];
    __shared__ uint4 mbuf[2];
    fs_load(f, &A.st->ts);
    const g4 z = g_load_hbm(A.z);
    const g4 eb = g_load_hbm(&A.st->eq_bound);
    g4 m[2];
    gkr_fs_messages(eb, s0, s1, mbuf, m);
    if (threadIdx.x == 0) {
        g_store_hbm(A.msg_out, m[0]);
        g_store_hbm(A.msg_out + 1, m[1]);
    }
    const g4 r = fs_message2_squeeze(f, 1, m[0], m[1]);
    if (threadIdx.x < 64) {
        const g4 e = g_mul_wave(eb, eq_factor(r, z));
        if (threadIdx.x == 0) {
            g_store_hbm(A.r_out, r);
            g_store_hbm(&A.st->eq_bound, e);
        }
    }
    fs_save(&A.st->ts, f);
    build_const_table4(tab, basis, r);
    for (int e = threadIdx.x; e < kConstTable4Entries; e += kStageBlock) A.rtab_out[e] = tab[e];
}

// The challenge after a grid-wide round launched on its own (d[0]'s sums).
__global__ __launch_bounds__(kStageBlock) void k_gkr_fs_round(const GkrRound* d, GkrFsRound fs) {
    g4 s[2];
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k >> 2].w[k & 3] = d[0].acc[k];
    gkr_fs_step(fs, s[0], s[1]);
}

// A grid-wide round and its challenge in one launch: every workgroup adds
// its sums, then counts itself done; the last one to finish (its count
// comes after every other workgroup's sums, fenced device-wide) reads the
// sums and draws the challenge.  `done` starts at zero.
__global__ __launch_bounds__(kStageBlock) void k_gkr_round_fs(const GkrRound* d, GkrFsRound fs, uint32_t* done) {
    const GkrRound A = d[0];
    gkr_round_body(A, blockIdx.x);
    __shared__ int last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(done, 1u) == A.blocks - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    g4 s[2];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        s[k >> 2].w[k & 3] = __hip_atomic_load(A.acc + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    gkr_fs_step(fs, s[0], s[1]);
}

// k_gkr_tail with the transcript on the device: after each round's sums, the
// message, the challenge (written to the next point) and r's table, built in
// LDS for the next round's bind; after round 0 the final bind, [L(r), R(r)]
// as a prover message, the challenge t and the layer's claim
// L + t (R + L) (lib.rs:142-158).  The layer's coordinates come from its
// point on the device, so a zero coordinate takes the at-one sums.
__global__ __launch_bounds__(kStageBlock) void k_gkr_tail_fs(const GkrTail* d, GkrFsTail fs) {
    const GkrTail A = d[0];
    extern __shared__ __attribute__((aligned(256))) uint4 lds_raw[];
    __shared__ FsLds f;
    __shared__ uint4 basis[128];
    __shared__ uint32_t red[8];
    __shared__ uint4 lr[2];
    __shared__ uint4 mbuf[2];
    __shared__ uint4 zs[64];  // the layer's point
    constexpr size_t kArr = (size_t)4 << kGkrTailMax;
    uint4* Ls = lds_raw;
    uint4* Rs = Ls + kArr;
    uint4* Es = Rs + kArr;
    uint4* rt = Es + ((size_t)2 << kGkrTailMax);
    uint4* col = rt + kConstTable4Entries + (threadIdx.x >> 6) * (64 * kLaneTableEntries) + (threadIdx.x & 63);
    lane_table_clear(col);
    const int n = A.n, i0 = A.i0;
    const bool whole = i0 == n - 1;
    const size_t len = whole ? (size_t)1 << n : (size_t)4 << i0;
    const size_t elen = whole ? (size_t)1 << (n - 1) : (size_t)2 << i0;
    auto ld = [](const uint4* p) {
        const uint4 u = *p;
        g4 r;
        r.w[0] = u.x; r.w[1] = u.y; r.w[2] = u.z; r.w[3] = u.w;
        return r;
    };
    auto st = [](uint4* p, const g4& v) { *p = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]); };
    for (size_t e = threadIdx.x; e < len; e += kStageBlock) {
        Ls[e] = u4_load_hbm(reinterpret_cast<const uint4*>(A.left + e));
        Rs[e] = u4_load_hbm(reinterpret_cast<const uint4*>(A.right + e));
    }
    for (size_t e = threadIdx.x; e < elen; e += kStageBlock) Es[e] = u4_load_hbm(reinterpret_cast<const uint4*>(A.eq + e));
    if (!whole) {
        g4 nx[kRtabPer];
        rtab_fetch(nx, A.rtabs4);
        rtab_store(rt, nx);
    }
    if (threadIdx.x < (unsigned)n) zs[threadIdx.x] = u4_load_hbm(reinterpret_cast<const uint4*>(fs.point + threadIdx.x));
    fs_load(f, &fs.st->ts);  // ends with a barrier: the arrays, rt and the point are in place
    g4 eb = g_load_hbm(&fs.st->eq_bound);
    for (int i = i0; i >= 0; --i) {
        const size_t h = (size_t)1 << i;
        const bool bind = !(whole && i == n - 1);
        const g4 z = ld(zs + i);
        const bool at_one = g_is_zero(z);
        g4 sb = g_zero(), si = g_zero();
        for (size_t x = threadIdx.x; x < h; x += kStageBlock) {
            g4 l0, l1, r0, r1, e;
            if (bind) {
                const size_t hh = 2 * h;
                auto ip = [&](const uint4* p, size_t a, size_t b) {
                    const g4 u = ld(p + a), v = ld(p + b);
                    return g_add(u, g_mul_table4(rt, g_add(u, v)));
                };
                l0 = ip(Ls, x, x + hh);
                l1 = ip(Ls, x + h, x + h + hh);
                r0 = ip(Rs, x, x + hh);
                r1 = ip(Rs, x + h, x + h + hh);
                st(Ls + x, l0);
                st(Ls + x + h, l1);
                st(Rs + x, r0);
                st(Rs + x + h, r1);
                e = g_add(ld(Es + x), ld(Es + x + h));
                st(Es + x, e);
            } else {
                l0 = ld(Ls + x);
                l1 = ld(Ls + x + h);
                r0 = ld(Rs + x);
                r1 = ld(Rs + x + h);
                e = ld(Es + x);
            }
            g4 el, ed;
            g_mul_lane2(e, at_one ? l1 : l0, g_add(l1, l0), col, el, ed);
            sb = g_add(sb, g_mul_lane(el, at_one ? r1 : r0, col));
            si = g_add(si, g_mul_lane(ed, g_add(r1, r0), col));
        }
        g4 v[2] = {sb, si};
        block_xor_store<2>(v, red, kStageBlock);  // its barrier: every round-i reader of rt is done
        __syncthreads();
        g4 s0, s1, m[2];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            s0.w[k] = red[k];
            s1.w[k] = red[4 + k];
        }
        gkr_fs_messages(eb, s0, s1, mbuf, m);
        if (threadIdx.x == 0) {
            g_store_hbm(fs.msgs + 2 * (size_t)(n - 1 - i), m[0]);
            g_store_hbm(fs.msgs + 2 * (size_t)(n - 1 - i) + 1, m[1]);
        }
        const g4 r = fs_message2_squeeze(f, 1, m[0], m[1]);
        eb = g_mul_wave(eb, eq_factor(r, z));  // every wave keeps eq_bound
        if (threadIdx.x == 0) g_store_hbm(fs.next + i, r);
        build_const_table4(rt, basis, r);  // r_i binds round i - 1 (or the final bind)
    }
    // Final bind of the two remaining entries by r_0.
    if (threadIdx.x < 2) {
        uint4* P = threadIdx.x == 0 ? Ls : Rs;
        const g4 u = ld(P), v = ld(P + 1);
        st(lr + threadIdx.x, g_add(u, g_mul_table4(rt, g_add(u, v))));
    }
    __syncthreads();
    g4 e2[2] = {ld(lr), ld(lr + 1)};
    if (threadIdx.x == 0) {
        g_store_hbm(fs.msgs + 2 * (size_t)n, e2[0]);
        g_store_hbm(fs.msgs + 2 * (size_t)n + 1, e2[1]);
    }
    const g4 t = fs_message2_squeeze(f, 1, e2[0], e2[1]);
    if (threadIdx.x < 64) {
        const g4 claim = g_add(e2[0], g_mul_wave(t, g_add(e2[1], e2[0])));
        if (threadIdx.x == 0) {
            g_store_hbm(fs.next + n, t);
            g_store_hbm(&fs.st->claim, claim);
            g_store_hbm(&fs.st->eq_bound, g_one());
        }
    }
    fs_save(&fs.st->ts, f);
}

}  // namespace bitz

namespace bitz {

// ============================== prove_spartan_piop: transcript on the device ===
// One wave per proof.  Products come four at a time from g_mul4 (one 16-lane
// row each), grouped into levels of independent products; the message and
// squeeze go through the fixed-position duplex path; r's 4-bit table is
// written straight to the next round's fold by const_table4_wave.
__global__ __launch_bounds__(64) void k_piop_fs(const PiopFsDesc* descs) {
    const PiopFsDesc D = descs[blockIdx.x];
    PiopFsState* st = D.st;
    const int lane = (int)threadIdx.x;
    __shared__ FsLds f;
    __shared__ uint32_t words[8 << kInnerBitsMaxRound];
    auto sum_at = [&](int word) {
        g4 v;
#pragma unroll
        for (int i = 0; i < 4; ++i) v.w[i] = D.sums[word + i];
        return v;
    };
    const g4 Z = g_zero();
    if (D.fin_count) {  // the sumcheck's last fold, into the sums slot read below
        if (lane < D.fin_count) {
            const f128* t = D.fin_tab[lane];
            const g4 a = g_load(t), b = g_load(t + 1);
            g_store(reinterpret_cast<f128*>(const_cast<uint32_t*>(D.sums)) + lane,
                    g_add(a, g_mul(g_load(D.fin_r), g_add(a, b))));
        }
        __syncthreads();
    }
    g4 claim = g_load(&st->claim);
    if (D.kind == kFsInnerFinal) {  // D(r) and the final claim; the transcript is untouched
        if (lane == 0) {
            g_store(D.round_out, sum_at(0));
            g_store(D.final_out, claim);
        }
        return;
    }
    fs_load(f, &st->ts);
    if (D.kind == kFsOuterFinal) {
        // sumcheck.rs:372 then piop.rs:94-97: absorb [A, B, C](r), squeeze rho,
        // the inner claim A + rho B + rho^2 C.
        const g4 cl[3] = {sum_at(0), sum_at(4), sum_at(8)};
        const g4 rho = fs_message_squeeze(f, 0, cl, 3);
        g4 p[4], p2[4];
        g_mul4({rho, rho, rho, Z}, {cl[1], cl[2], rho, Z}, p);  // rho B, rho C, rho^2
        g_mul4({rho, Z, Z, Z}, {p[1], Z, Z, Z}, p2);          // rho^2 C
        claim = g_add3(cl[0], p[0], p2[0]);
        if (lane == 0) {
            for (int u = 0; u < 3; ++u) g_store(D.round_out + u, cl[u]);
            g_store(&st->rho, rho);
            g_store(&st->rho2, p[2]);
            g_store(&st->claim, claim);
        }
        fs_save(&st->ts, f);
        return;
    }
    g4 c[4];
    int deg;
    g4 q[3];
    if (D.kind == kFsOuterRound) {
        // The eq-factored round: q0 = raw0 scale, q2 = raw2 scale, q(1) = raw1 scale
        // or (sigma + (1 + tau) q0) / tau, q1 = q(1) + q0 + q2; c0 = S (1 + tau) q0,
        // c2 = S ((1 + tau) q2 + q1), c3 = S q2, c1 = claim + c2 + c3.
        const g4 scale = g_from(D.scale), st_ = g_from(D.scale_t);
        const g4 raw0 = sum_at(0), raw2 = sum_at(4);
        g4 p[4], p2[4];
        g_mul4({raw0, raw2, raw0, raw2}, {scale, scale, st_, st_}, p);  // q0, q2, (1+tau) q0, (1+tau) q2
        if (D.have_q1)
            g_mul4({sum_at(8), Z, Z, Z}, {scale, Z, Z, Z}, p2);
        else
            g_mul4({g_load(&st->sigma), raw0, Z, Z}, {g_from(D.tau_inv), g_from(D.scale_ti), Z, Z}, p2);
        q[0] = p[0];
        q[2] = p[1];
        q[1] = g_add3(g_add(p2[0], p2[1]), q[0], q[2]);
        const g4 S = g_load(&st->S);
        g4 s3[4];
        g_mul4({S, S, S, Z}, {p[2], g_add(p[3], q[1]), q[2], Z}, s3);
        c[0] = s3[0];
        c[2] = s3[1];
        c[3] = s3[2];
        c[1] = g_add3(claim, c[2], c[3]);
        deg = 3;
    } else {
        if (D.K >= 0) {
            // A bit round: the copies' XOR, then sum_S r_S (sum S of c0 / c2),
            // products S of c0 in rows of the first pass(es), then those of c2.
            const int S = 1 << D.K;
            if (lane < 8 * S) {
                uint32_t x = 0;
                for (int r = 0; r < kInnerBitsReplicas; ++r) x ^= D.sums[r * kInnerBitsReplicaWords + lane];
                words[lane] = x;
            }
            if (D.K == 0 && lane == 0) {
                unsigned long long lv = 0;
                for (int r = 0; r < kInnerBitsReplicas; ++r)
                    lv = max(lv, *reinterpret_cast<const unsigned long long*>(D.sums + r * kInnerBitsReplicaWords +
                                                                              kInnerBitsLiveWord));
                st->live = lv;
            }
            __syncthreads();
            g4 lo = Z, hi = Z;
            for (int base = 0; base < 2 * S; base += 4) {
                const int idx = base + (lane >> 4);  // this row's product
                g4 x = Z, y = Z;
                if (idx < 2 * S) {
                    x = g_load(&st->mono[idx & (S - 1)]);
#pragma unroll
                    for (int i = 0; i < 4; ++i) y.w[i] = words[4 * idx + i];
                }
                const g4 p = g_mul_rows(x, y);
                if (idx < S)
                    lo = g_add(lo, p);
                else
                    hi = g_add(hi, p);
            }
            c[0] = g_add(g_add(g_row(lo, 0), g_row(lo, 1)), g_add(g_row(lo, 2), g_row(lo, 3)));
            c[2] = g_add(g_add(g_row(hi, 0), g_row(hi, 1)), g_add(g_row(hi, 2), g_row(hi, 3)));
        } else {
            c[0] = sum_at(0);
            c[2] = sum_at(4);
        }
        c[1] = g_add(claim, c[2]);
        deg = 2;
    }
    // sumcheck.rs:803-829: absorb the full round polynomial, squeeze r, claim = g(r).
    const g4 r = fs_message_squeeze(f, 0, c, deg + 1);
    if (D.kind == kFsOuterRound) {
        // claim = c0 + c1 r + (c2 + c3 r) r^2; sigma_{k+1} = q_k(r) = q0 + q1 r + q2 r^2;
        // S_{k+1} = S_k eq(tau_k, r) = S_k (1 + tau_k + r)
        g4 p[4], p2[4];
        g_mul4({r, c[3], c[1], q[1]}, {r, r, r, r}, p);  // r^2, c3 r, c1 r, q1 r
        g4 e = g_add(g_from(D.tau), r);
        e.w[0] ^= 1u;
        g_mul4({g_add(c[2], p[1]), q[2], g_load(&st->S), Z}, {p[0], p[0], e, Z}, p2);
        claim = g_add3(c[0], p[2], p2[0]);
        const g4 sigma = g_add3(q[0], p[3], p2[1]);
        if (lane == 0) {
            g_store(&st->sigma, sigma);
            g_store(&st->S, p2[2]);
        }
    } else {
        // claim = c0 + c1 r + c2 r^2; with mono_next, r_S for the next bit round:
        // mono[half + i] = mono[i] r (S with bit k set), half = 2^k <= 4
        const int half = D.mono_next ? 1 << D.k : 0;
        const g4 m0 = half > 0 ? g_load(&st->mono[0]) : Z, m1 = half > 1 ? g_load(&st->mono[1]) : Z;
        const g4 m2 = half > 2 ? g_load(&st->mono[2]) : Z, m3 = half > 3 ? g_load(&st->mono[3]) : Z;
        g4 p[4], p2[4];
        g_mul4({r, c[1], m0, m1}, {r, r, r, r}, p);
        g_mul4({c[2], m2, m3, Z}, {p[0], r, r, Z}, p2);
        claim = g_add3(c[0], p[1], p2[0]);
        if (lane == 0) {
            if (half > 0) g_store(&st->mono[half], p[2]);
            if (half > 1) g_store(&st->mono[half + 1], p[3]);
            if (half > 2) g_store(&st->mono[half + 2], p2[1]);
            if (half > 3) g_store(&st->mono[half + 3], p2[2]);
        }
    }
    if (lane == 0) {
        for (int i = 0; i <= deg; ++i) g_store(D.round_out + i, c[i]);
        g_store(D.point_out, r);
        g_store(&st->claim, claim);
    }
    fs_save(&st->ts, f);
    const_table4_wave(D.rtab, r);
}

__global__ __launch_bounds__(kBlock) void k_row_scales_dev(const f128* eq_rows, const f128* rr, f128* row_scale,
                                                           size_t rows_padded) {
    const g4 rho = g_load(rr), rho2 = g_load(rr + 1);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < rows_padded; i += (size_t)gridDim.x * blockDim.x) {
        const g4 w = g_load(eq_rows + i);  // k_row_scales' layout: [w | rho w | rho^2 w]
        g_store(row_scale + i, w);
        g_store(row_scale + rows_padded + i, g_mul(w, rho));
        g_store(row_scale + 2 * rows_padded + i, g_mul(w, rho2));
    }
}

}  // namespace bitz
