// Kernel declarations shared by the kernel and driver translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "gf128.hpp"
#include "keccak.hpp"

namespace bitz {

constexpr int kBlock = 256;

__global__ void k_mul(const f128* a, const f128* b, f128* out, size_t n);
__global__ void k_mul_lane(const f128* a, const f128* b, f128* out, size_t n, int two);
__global__ void k_mul_const(const f128* a, const f128* b, f128* out, size_t n);
__global__ void k_mul_mfma(const f128* a, const f128* b, f128* out, size_t n);
__global__ void k_mul_rows(const f128* a, const f128* b, f128* out, size_t n);
// eq tables of up to 2^kEqRowsMaxVars entries in 1024-thread workgroups of
// kEqRowsPerThread entries per thread (kLaneSmem1024 of dynamic LDS); rr
// non-null: the bind's row scales.
constexpr int kEqRowsMaxVars = 16;
constexpr int kEqRowsPerThread = 4;
constexpr unsigned eq_rows_grid(int k) {
    return (unsigned)((((size_t)1 << k) + 1024 * kEqRowsPerThread - 1) / (1024 * kEqRowsPerThread));
}
constexpr size_t kLaneSmem1024 = sizeof(uint4) * 16 * 64 * kLaneTableEntries;
__global__ void k_eq_rows(const f128* r, int k, f128* out, const f128* rr);
__global__ void k_eq_small2(const f128* r, int kl, f128* low, int kh, f128* high);
__global__ void k_eq_combine(const f128* low, int low_bits, const f128* high, f128* out, size_t n);
__global__ void k_fold(const f128* in, f128* out, size_t out_len, f128 r);
// MLE folds through the challenges' constant tables: one variable, or two a pass.
__global__ void k_fold_t(const f128* in, f128* out, size_t out_len, f128 r);
__global__ void k_fold2_t(const f128* in, f128* out, size_t out_len, f128 r0, f128 r1);
__global__ void k_fold3(const f128* a, const f128* b, const f128* c, f128* ao, f128* bo, f128* co,
                        size_t out_len, f128 r);
__global__ void k_outer_round0(const f128* a, const f128* b, const f128* c, size_t pairs,
                               const f128* low, int low_bits, const f128* high, uint32_t* acc);
__global__ void k_outer_fold_next(const f128* a, const f128* b, const f128* c, f128* ao, f128* bo,
                                  f128* co, size_t out_pairs, f128 r, const f128* low,
                                  int low_bits, const f128* high, uint32_t* acc);
__global__ void k_outer_claim(const f128* a, const f128* b, const f128* c, size_t n, const f128* low,
                              int low_bits, const f128* high, uint32_t* acc);
__global__ void k_gkr_sums(const f128* left, const f128* right, const f128* eq, size_t half,
                           int at_one, uint32_t* acc);
__global__ void k_gkr_bind_next(f128* left, f128* right, f128* eq, size_t h, f128 r,
                                int at_one, uint32_t* acc);
__global__ void k_gkr_bind(f128* left, f128* right, size_t half, f128 r);
// bind_and_batch in two passes: every nonzero's product W[row] * value
// (lane tables, 256 threads), then each column's XOR over its contiguous run.
__global__ void k_bind_products(const uint32_t* rows, const uint8_t* which, const f128* values, const f128* row_scale,
                                size_t rows_padded, f128* prod, size_t nnz);
__global__ void k_bind_segments(const uint32_t* col_starts, const f128* prod, f128* out, size_t cols);
constexpr size_t kBindSmem = sizeof(uint4) * (4 * 64 * kLaneTableEntries);
// bind_and_batch: columns with more nonzeros than this are summed in runs of
// kBindLongRun entries, one workgroup per run (k_bind_long).
constexpr uint32_t kBindLongColumn = 64;
constexpr uint32_t kBindLongRun = 2048;
__global__ void k_bind_long(const uint32_t* chunks, const f128* prod, f128* out);
__global__ void k_row_scales(const f128* eq_rows, f128 rho, f128 rho2, f128* row_scale, size_t rows_padded);
__global__ void k_r1cs_rows(const uint32_t* row_starts, const uint32_t* cols, const f128* vals, const f128* h,
                            size_t rows, f128* out, size_t padded);
__global__ void k_assignment_bits(const uint64_t* words, size_t bit_len, f128* out, size_t padded);
// M^T: columns longer than this are summed by one workgroup each (_long).
constexpr uint32_t kMtLongColumn = 64;
__global__ void k_mt_apply(const uint32_t* offsets, const uint32_t* rows, const f128* ch,
                           f128* out, size_t cols);
__global__ void k_mt_apply_long(const uint32_t* offsets, const uint32_t* rows, const f128* ch,
                                const uint32_t* long_cols, f128* out);
// transpose_inner with lane-table products: kLaneSmem256 of dynamic LDS.
__global__ void k_mt_transpose_inner(const uint32_t* offsets, const uint32_t* rows,
                                     const f128* row_w, int log_rows, const f128* col_w,
                                     f128* out, size_t cols);
__global__ void k_mt_inner_long(const uint32_t* offsets, const uint32_t* rows, const f128* row_w, int log_rows,
                                const f128* col_w, const uint32_t* long_cols, f128* out);

__global__ void k_eq_check(const f128* L, int ell, const f128* H, int m, uint32_t* out, unsigned blocks_l);

// H(i) tables of the multi-iteration rounds (see k_htabs4).
constexpr int kMaxRounds = 36;
struct HtabDesc {
    int rounds;
    uint32_t first[kMaxRounds + 1];
    int start[kMaxRounds];
};

__global__ void k_const_tables4(const f128* consts, uint4* out);
// post-GKR sumcheck (crates/post_gkr): the first round off the packed bits
// (acc: [weighted sum, a0, a2]) and the tables W, V after it.
__global__ void k_pg_first(const f128* rows, int t, const f128* cols, const f128* packed, size_t n_elems,
                           uint32_t* acc);
__global__ void k_pg_expand(const f128* rf, int t, const f128* cols, const f128* packed, f128 r, f128* W, f128* V,
                            size_t n);
// keccak_hash (keccak.hpp) by one device thread: the device permutation's check.
__global__ void k_keccak_hash(const uint8_t* msg, size_t n, int rate, uint8_t pad, uint8_t* out, size_t out_len);
// p[0 .. n) = 0 (a kernel, where a memset node would not do: captured graphs).
__global__ void k_zero_words(uint32_t* p, size_t n);

struct PrologueArgs {
    const f128* L;
    const f128* H;
    int ell, m;
    unsigned eq_blocks, eq_blocks_l;  // factor check (k_eq_check layout): 1024 entries per workgroup
    uint32_t* eq_out;
    unsigned h4_blocks;  // 4-bit H(i) tables of the multi-iteration rounds
    HtabDesc hdesc;
    uint4* htabs4;
    uint32_t* zero;  // round-sum slots, cleared by workgroup 0
    unsigned zero_words;
    f128 chal[kMaxRounds];  // the challenges, by value
    unsigned t4_blocks;  // 4-bit tables of chal[0 ..)
    uint4* rtabs4;
};
__global__ void k_prologue(PrologueArgs args);
__global__ void k_htabs4(const f128* L, int ell, const f128* H, HtabDesc d, uint4* out);
__global__ void k_const_table4_1(f128 c, uint4* out);

template <bool Q1>
__global__ void k_outer_fold(const f128* a, const f128* b, const f128* c, f128* ao, f128* bo, f128* co,
                             size_t threads, size_t iters, const uint4* r_tab4, const f128* la, int la_bits,
                             int la_shift, const f128* lb, int lb_shift, const uint4* h_tabs4, uint32_t* acc);
constexpr int kFoldBlock = 1024;
// LDS of k_outer_fold / k_batch_fold with `block` threads: the lane tables,
// r's 4-bit table and two H(i) tables.
constexpr size_t fold_smem(size_t block) {
    return sizeof(uint4) * (3 * kConstTable4Entries + (block / 64) * 64 * kLaneTableEntries);
}
constexpr size_t kFoldSmem = fold_smem(kFoldBlock);

template <bool Q1>
__global__ void k_outer_r0(const f128* a, const f128* b, const f128* c, size_t threads, size_t iters,
                           const f128* la, int la_bits, int la_shift, const f128* lb, int lb_shift,
                           const uint4* h_tabs4, uint32_t* acc);
// Round 0 with lane tables: 1024 threads, 16 waves x 8 KiB + two 4-bit tables.
constexpr int kR0Block = 1024;
constexpr size_t kR0Smem = sizeof(uint4) * (2 * kConstTable4Entries + (kR0Block / 64) * 64 * kLaneTableEntries);
// Threads of a multi-iteration round: one 1024-thread workgroup per CU
// across the chip.
constexpr int kThreadBits = 18;
// Rounds whose pair count exceeds 2^kChunkFromBits run grid-wide; the
// following ones go to the chunked LDS stage.
constexpr int kChunkFromBits = 17;
// Occupancy target of the lane-table kernels: four waves per SIMD (at most
// 128 VGPRs), which their LDS use also allows.
#define BITZ_FOUR_WAVES __attribute__((amdgpu_waves_per_eu(4)))

// ---- rounds held in LDS: chunk, tail and batch stages -------------------
// One stage of one proof: rounds [k_first, k_first + s) on chunks of 2^lg_in
// inputs per table (the tables before the fold by r_{k_first - 1} when
// fold_first), chunk g at inputs [g 2^lg_in, (g + 1) 2^lg_in).  A final stage
// folds after its last round too and stores A(r), B(r), C(r) at fin; other
// stages store their 2^(lg_in - fold_first - s + 1) entries per chunk and
// table at ao/bo/co.  Round k's sums go to acc + 12 k: stored when
// store_sums (a stage of one chunk), XORed in otherwise.
struct LdsStage {
    const f128* a;
    const f128* b;
    const f128* c;
    f128* ao;
    f128* bo;
    f128* co;
    const f128* L;
    const f128* H;
    const uint4* rtabs4;  // 4-bit table of r_k at rtabs4 + 512 k
    uint32_t* acc;
    f128* fin;
    uint64_t q1mask;
    int ell, n, k_first, s, lg_in, fold_first, final_stage, store_sums;
    uint32_t chunk_base;  // batches: first workgroup of this stage
};
constexpr int kStageBlock = 256;
constexpr int kStageMaxLog = 10;  // at most 2^10 entries per table in LDS
constexpr int kChunkMaxLog = 11;  // chunk stage of a single proof: 2^11 inputs, folded on load
constexpr int kTailMaxLog = 11;
// Lane tables of a 256-thread workgroup (4 waves x 64 lanes x kLaneTableEntries entries).
constexpr size_t kLaneTables256 = 4 * 64 * kLaneTableEntries;
// Dynamic LDS of a stage with 2^cap_lg entries per table: the three tables,
// the weight tables (2^cap_lg entries) and the lane tables.
constexpr size_t stage_smem(int cap_lg) { return sizeof(uint4) * (((size_t)4 << cap_lg) + kLaneTables256); }
// The single proof's chunk and tail stages run 512-thread workgroups (eight
// waves of lane tables): their rounds with more pairs than 256 threads take
// fewer sequential items per thread.
constexpr int kOuterStageBlock = 512;
constexpr size_t outer_stage_smem(int cap_lg) {
    return sizeof(uint4) * (((size_t)4 << cap_lg) + (size_t)(kOuterStageBlock / 64) * 64 * kLaneTableEntries);
}
__global__ void k_outer_chunk(LdsStage args);
__global__ void k_outer_tail(LdsStage args);
__global__ void k_outer_batch(const LdsStage* stages, const uint32_t* bmap);

// One proof's grid-wide round in a batched launch: the arguments of
// k_outer_r0 / k_outer_fold, the first workgroup of the proof in the
// launch, and q1 (compute q(1) too).
struct RoundDesc {
    const f128* a;
    const f128* b;
    const f128* c;
    f128* ao;
    f128* bo;
    f128* co;
    const uint4* r_tab4;
    const f128* la;
    const f128* lb;
    const uint4* h_tabs4;
    uint32_t* acc;
    size_t threads, iters;
    int la_bits, la_shift, lb_shift, q1;
    uint32_t block_base;
};
__global__ void k_batch_r0(const RoundDesc* d, const uint32_t* bmap);
__global__ void k_batch_fold(const RoundDesc* d, const uint32_t* bmap);

// Prologue of a batch: per instance, clearing of its sums, the factor check
// (k_eq_check layout at eq_out) and the 4-bit tables of its challenges.
struct BatchPrologue {
    const f128* L;
    const f128* H;
    int ell, m, n;
    uint32_t* zero;
    unsigned zero_words;
    unsigned eq_blocks_l, eq_blocks;
    uint32_t* eq_out;
    const f128* chal;  // device array, n entries
    uint4* rtabs4;
    HtabDesc hdesc;  // H(i) tables of the multi-iteration grid-wide rounds
    uint4* htabs4;
    uint32_t block_base;  // first workgroup: eq check, then n challenge tables, then H(i) tables
};
__global__ void k_batch_prologue(const BatchPrologue* items, const uint32_t* bmap);

// ---- inner sumcheck with challenges known in advance ----------------------
// Round 0 (k_inner_r0_lane) and the grid-wide fold rounds (k_inner_fold_lane,
// fold by r_{k-1} on load through r's 4-bit table) with lane-table products,
// then the last rounds in one workgroup (k_inner_tail).
constexpr int kInnerBlock = 256;
constexpr size_t kInnerGridSmem = sizeof(uint4) * (kConstTable4Entries + 4 * 64 * kLaneTableEntries);
__global__ void k_inner_r0_lane(const f128* d, const f128* w, size_t pairs, uint32_t* acc,
                                unsigned long long* live_out);
// One inner sumcheck's round in a batched launch (k_inner_r0_batch: pairs of
// round 0; k_inner_fold_batch: output pairs, folding d, w by rtab4's challenge).
struct InnerRoundDesc {
    const f128* d;
    const f128* w;
    f128* dout;
    f128* wout;
    size_t pairs;
    size_t live;         // fold rounds: input entries below live (the rest is zero)
    unsigned long long* live_out;  // round 0: atomic max of 1 + the last nonzero index of D or h (or null)
    const uint4* rtab4;
    uint32_t* acc;
    uint32_t blocks, block_base;
    // Fold rounds, or null: `live` and `pairs` are bounds, and the live
    // length of the input tables is ceil(*live0 / 2^live_shift).
    const unsigned long long* live0;
    int live_shift;
};
__global__ void k_inner_r0_batch(const InnerRoundDesc* d, const uint32_t* bmap);
__global__ void k_inner_fold_batch(const InnerRoundDesc* d, const uint32_t* bmap);
__global__ void k_inner_fold_lane(const f128* d, const f128* w, f128* dout, f128* wout, size_t out_pairs,
                                  size_t live, const uint4* rtab4, uint32_t* acc);
constexpr int kInnerTailMaxLog = 10;  // at most 2^10 entries per table in LDS
// Rounds [k0, k0 + s) on chunks of 2^lg_in inputs (folded by r_{k0-1} on
// load when fold_first), chunk g = workgroup g at inputs [g 2^lg_in, ...).
// A final stage (dout null, one chunk) runs to the last round and stores
// D(r), h(r) at fin and its sums at acc; a chunk stage XORs its sums into
// acc and stores each chunk's folded entries at dout / wout.
struct InnerTail {
    const f128* d;
    const f128* w;
    const uint4* rtabs4;  // r_k's table at rtabs4 + 512 k
    uint32_t* acc;        // round k's [c0, c2] at acc + 8 k
    f128* fin;            // D(r), h(r)
    f128* dout;
    f128* wout;
    int n, k0, fold_first, lg_in, s;
};
// Rounds of the inner sumcheck's chunk stage: chunks of 2^11 inputs.
constexpr int kInnerChunkLog = 11;
constexpr int kInnerChunkRounds = 8;
constexpr size_t kInnerTailSmem =
    sizeof(uint4) * (((size_t)2 << kInnerTailMaxLog) + kConstTable4Entries + 4 * 64 * kLaneTableEntries);
__global__ void k_inner_tail(InnerTail args);

// ---- inner sumcheck against a Boolean assignment (packed bits) ------------
// With h in {0, 1} (bit j = bits[j / 64] >> (j % 64) & 1), round K's bound
// table is h_K(j) = sum_S r_S c_S(j) over the subsets S of {0..K-1}, r_S the
// product of r_i over S and c_S(j) the XOR of h(2^K j + x) over the subsets x
// of S: bits only.  Round K's sums split the same way,
//   c0 = sum_S r_S [sum_p c_S(2p) D_K(2p)],
//   c2 = sum_S r_S [sum_p (c_S(2p) + c_S(2p+1)) (D_K(2p) + D_K(2p+1))],
// so the kernel only XORs D_K under masks from the bits (2^(K+1) sums) and the
// host multiplies the sums by r_S.  Rounds 0 .. kInnerBitsMaxRound run so;
// the last of them also writes h_K as elements for the general kernels.
constexpr int kInnerBitsMaxRound = 3;
// The sums go to kInnerBitsReplicas copies (workgroup g adds into copy
// g mod kInnerBitsReplicas), which the reader XORs: thousands of workgroups
// that finish together would otherwise serialise on a few words' atomics.
// A copy: the 2^(K+1) sums (4 words each: 2^K for c0 with S ascending, then
// 2^K for c2) from word 0, and at words 64-65 (round 0) the maximum of 1 +
// the last index where D or h is nonzero.
constexpr int kInnerBitsReplicas = 32;
constexpr int kInnerBitsReplicaWords = 68;
constexpr int kInnerBitsLiveWord = 64;
constexpr size_t kInnerBitsAccWords = (size_t)kInnerBitsReplicas * kInnerBitsReplicaWords;
struct InnerBitsDesc {
    const f128* d;          // round 0: D; round K >= 1: D_{K-1}, folded by r_{K-1} on load
    const uint64_t* bits;   // the packed assignment, 2^n bits (zero past the assignment)
    f128* dout;             // K >= 1: D_K, 2 entries per output pair
    f128* wout;             // or null: h_K as elements, 2 entries per output pair
    size_t pairs;           // output pairs (pairs of D_K)
    size_t live;            // K >= 1: entries of D_{K-1} below live (the rest is zero)
    const uint4* rtab4;     // K >= 1: r_{K-1}'s 4-bit table
    uint32_t* acc;          // kInnerBitsAccWords words, zero on entry (the copies above)
    uint32_t blocks, block_base;
    f128 mono[1 << kInnerBitsMaxRound];  // wout: r_S for S < 2^K
    const f128* mono_dev;   // or null: wout's r_S read from the device instead
    // Or null: `live` and `pairs` are bounds, and the live length of D_{K-1}
    // is ceil(*live0 / 2^(K-1)) (live0: round 0's measure on the device).
    const unsigned long long* live0;
    int k;                  // the round K (k_inner_bits_mixed)
};
// Workgroups of a bit round: enough to stream at HBM rate, few enough that
// the sums' atomics stay cheap.
constexpr unsigned kInnerBitsMaxBlocks = 1024;
constexpr size_t kInnerBitsSmem = sizeof(uint4) * kConstTable4Entries;
template <int K> __global__ void k_inner_bits(InnerBitsDesc A);
template <int K> __global__ void k_inner_bits_batch(const InnerBitsDesc* d, const uint32_t* bmap);
__global__ void k_inner_bits_mixed(const InnerBitsDesc* d, const uint32_t* bmap);
// ---- prove_spartan_piop rounds with the transcript on the device ---------
// A proof's running values: the Keccak duplex, the claim, the eq-factored
// outer sumcheck's S_k = prod_{i<k} eq(tau_i, r_i) and sigma_k = q_{k-1}(r),
// the current q(0), q(1) + q(0) + q(inf), q(inf), rho and rho^2, the inner
// sumcheck's monomials r_S of its challenges so far (bit rounds) and its live
// length (round 0's measure).
struct PiopFsState {
    KeccakState ts;
    f128 claim, S, sigma, q[3], rho, rho2;
    f128 mono[1 << kInnerBitsMaxRound];
    unsigned long long live;
    uint64_t pad_;
};
constexpr int kFsOuterRound = 0, kFsOuterFinal = 1, kFsInnerRound = 2, kFsInnerFinal = 3;
// One proof's step after its round's kernels (one 64-thread workgroup each):
//   outer round k: c0, c2, c3 from the eq-factored sums (q(0), q(inf), q(1)
//     at sums + 0, 4, 8, times `scale`; without q(1) it comes from sigma),
//     c1 = claim + c2 + c3, absorb [c0..c3] (public), squeeze r, claim =
//     g(r), sigma, S; [c0..c3] at round_out, r at point_out, r's table;
//   outer final: A(r), B(r), C(r) at sums (12 words) to round_out, absorbed
//     (public), rho squeezed, claim = A + rho B + rho^2 C;
//   inner round k: [c0, c2] at sums (a bit round K >= 0: the copies of its
//     monomial sums, weighed by st->mono), c1 = claim + c2, absorb, squeeze,
//     claim, outputs and r's table; mono_next: st->mono grows by r;
//   inner final: D(r) (sums) to round_out, the claim to final_out.
struct PiopFsDesc {
    PiopFsState* st;
    const uint32_t* sums;
    f128* round_out;
    f128* point_out;
    uint4* rtab;
    f128* final_out;
    f128 tau, tau_inv, scale;  // outer round k: tau_k, 1 / tau_k, the sums' scale
    f128 scale_t, scale_ti;    // scale (1 + tau_k), scale (1 + tau_k) / tau_k (zero when tau_k is)
    int kind, k, K, have_q1, mono_next;
    // Finals: the last fold of the sumcheck's tables (fin_tab[t][0] + r (fin_tab[t][0] + fin_tab[t][1]),
    // t < fin_count, r at *fin_r on the device) into `sums`, before the step reads them.
    const f128* fin_tab[3];
    const f128* fin_r;
    int fin_count;
};
__global__ void k_piop_fs(const PiopFsDesc* d);
// k_row_scales with rho and rho^2 read from the device (rr[0], rr[1]).
__global__ void k_row_scales_dev(const f128* eq_rows, const f128* rr, f128* row_scale, size_t rows_padded);

// Az, Bz, Cz rows over a packed Boolean assignment: the XOR of the row's
// values at the columns whose bit is set (no products).
__global__ void k_r1cs_rows_bits(const uint32_t* row_starts, const uint32_t* cols, const f128* vals,
                                 const uint64_t* bits, size_t rows, f128* out, size_t padded);

// ---- GKR layers in parallel (challenges known in advance) ----------------
// With every challenge known, each layer's point is known, so the layers'
// sumchecks are independent: small eq tables and the low/high halves of
// large ones in one launch (k_gkr_eq_items, one workgroup each), the large
// ones combined in one launch, the grid-wide rounds of all large layers one
// launch per round index, and every layer's last rounds in one launch of
// LDS tails (one workgroup per layer).
struct GkrEqItem {
    const f128* r;  // k coordinates
    int k;          // at most kGkrEqSmallMax
    f128* out;      // 2^k entries
};
constexpr int kGkrEqSmallMax = 12;
__global__ void k_gkr_eq_items(const GkrEqItem* items);  // 256 threads, dynamic LDS: kLaneSmem256
constexpr size_t kLaneSmem256 = sizeof(uint4) * 4 * 64 * kLaneTableEntries;
// Product tree with lane-table products: one grid-wide launch per large
// layer, then every layer from 2^kGkrSmallTreeLog elements down in one
// workgroup (LDS), each layer written to its place in the store.
constexpr int kGkrSmallTreeLog = 12;
__global__ void k_bittable_leaves(const uint64_t* packed, const f128* images, int log_rows, int log_columns,
                                  f128* leaves);
// cur_copy, when not null, receives a copy of cur (the leaf layer read once).
__global__ void k_gkr_layer_lane(const f128* cur, f128* cur_copy, f128* next, size_t half);
// layer holds 2^log_len elements; the next `count` layers follow it in memory.
__global__ void k_gkr_layers_small(f128* layer, int log_len, int count);
struct GkrCombine {
    const f128* low;
    const f128* high;
    f128* out;
    size_t n;
    int low_bits;
    uint32_t block_base, blocks;
};
__global__ void k_gkr_eq_combine_batch(const GkrCombine* d, const uint32_t* bmap);
// One layer's round in a batched launch: the first round's sums (first) or
// bind by r then the next round's sums, as k_gkr_sums / k_gkr_bind_next.
struct GkrRound {
    f128* left;
    f128* right;
    f128* eq;
    size_t h;
    const uint4* rt4;  // bind rounds: r's 4-bit table
    uint32_t* acc;
    int first, at_one;
    const f128* z;     // when set, at_one is whether the coordinate *z is zero (read on the device)
    uint32_t block_base, blocks;
};
// 256 threads: r's table (bind rounds) and the lane tables in dynamic LDS.
constexpr size_t kGkrRoundSmem = sizeof(uint4) * (kConstTable4Entries + 4 * 64 * kLaneTableEntries);
__global__ void k_gkr_round_batch(const GkrRound* d, const uint32_t* bmap);
// A layer's rounds i0 .. 0 and the final bind in one workgroup.  Before
// round i0 the arrays hold 2^(i0+2) entries each and eq 2^(i0+1) (i0 < n-1),
// or the layer's 2^n children each and eq(z_<n-1) (i0 = n-1).
constexpr int kGkrTailMax = 8;  // tail rounds have h = 2^i <= 2^8: arrays of at most 2^10 entries
struct GkrTail {
    f128* left;
    f128* right;
    f128* eq;
    const uint4* rtabs4;  // 4-bit table of the layer's r_i at rtabs4 + 512 (n - 1 - i)
    uint32_t* acc;        // round i's sums at acc + 8 (n - 1 - i)
    uint32_t* evals;      // [L(r), R(r)]
    uint32_t at_one;      // bit i: the point's coordinate i is zero
    int n, i0;
};
constexpr size_t kGkrTailSmem =
    sizeof(uint4) * (2 * ((size_t)4 << kGkrTailMax) + ((size_t)2 << kGkrTailMax) + kConstTable4Entries +
                     4 * 64 * kLaneTableEntries);
__global__ void k_gkr_tail(const GkrTail* d);

// ---- GKR with the transcript on the device (the Keccak duplex, fs.hpp) ----
// Each round's message and challenge are drawn by the device right after the
// round's sums: by the last workgroup of a grid-wide round (k_gkr_round_fs),
// and k_gkr_tail_fs draws them between the LDS rounds of a layer's tail and
// ends the layer.
struct GkrFsState {
    KeccakState ts;  // the transcript
    f128 eq_bound;   // eq(z, r) over the layer's coordinates bound so far
    f128 claim;      // the claim of the last finished layer
};
struct GkrFsRound {
    GkrFsState* st;
    const f128* z;         // the point's coordinate of the round
    f128* r_out;           // the next point's coordinate
    uint4* rtab_out;       // r's 4-bit table (the next round binds by it)
    f128* msg_out;         // [eq_bound sum_b, eq_bound sum_inf]
};
// A grid-wide round (d[0], launched with d[0].blocks workgroups) whose last
// workgroup to finish draws the challenge; `done` starts at zero.  Each
// workgroup's device-scope fence costs more than a second launch once the
// round spans many workgroups: those rounds run k_gkr_round_batch, then
// k_gkr_fs_round (one workgroup) over d[0]'s sums.
__global__ void k_gkr_round_fs(const GkrRound* d, GkrFsRound fs, uint32_t* done);
__global__ void k_gkr_fs_round(const GkrRound* d, GkrFsRound fs);
constexpr uint32_t kGkrFusedMaxBlocks = 64;
// d's rtabs4 is r_(i0+1)'s table when the tail does not start the layer;
// acc, evals and at_one are unused.
struct GkrFsTail {
    GkrFsState* st;
    const f128* point;  // the layer's point (n coordinates)
    f128* next;         // the next layer's point (n + 1 coordinates)
    f128* msgs;         // round i's message at msgs + 2 (n - 1 - i), then [L(r), R(r)] at msgs + 2 n
};
__global__ void k_gkr_tail_fs(const GkrTail* d, GkrFsTail fs);

}  // namespace bitz
