/*
 * CPU oracle for the GF(2^128) sumcheck path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithms, used by tests/, the
 * smoke check and bench.py's cpu_baseline leg as the checker.  The product
 * (bitz_amd / libbitz_hip.so) never links or calls anything here.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the worldfnd/BitZ checkout).  Field arithmetic is exact, so the restatement
 * keeps the reference's definitions (pair order, variable binding order,
 * omitted coefficient, transcript order) and is free in everything else.
 */
#ifndef BITZ_ORACLE_H
#define BITZ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* crates/field/src/gf128.rs:60-68: bit i of lo||hi is the coefficient of X^i. */
typedef struct {
    uint64_t lo, hi;
} of128;

/* ---- field: crates/field/src/gf128/portable.rs, wide.rs, pow.rs ---------- */
/* 0 = bit-serial `portable` multiply (the kernel the reference selects on
 * x86, gf128.rs:40-49), 1 = PCLMULQDQ (same results, not the reference). */
void oracle_set_kernel(int kernel);
int oracle_get_kernel(void);
void oracle_clmul64(uint64_t a, uint64_t b, uint64_t out[2]);   /* portable.rs:11-24 */
void oracle_clmul128(of128 a, of128 b, uint64_t out[4]);        /* portable.rs:28-34 */
of128 oracle_reduce(const uint64_t r[4]);                       /* portable.rs:50-57 */
of128 oracle_mul(of128 a, of128 b);                             /* portable.rs:60-62 */
of128 oracle_square(of128 a);                                   /* portable.rs:67-71 */
of128 oracle_mul_x(of128 a);                                    /* portable.rs:114-120 */
of128 oracle_inv(of128 a);                                      /* pow.rs:87-103 (a^(2^128-2)) */
void oracle_mul_batch(const of128* a, const of128* b, of128* out, size_t n);

/* ---- transcript: challenge sources standing in for crates/transcript ------ */
/* A transcript is a pair of callbacks: absorb a message (kind 0 = public
 * message, 1 = prover message written to the narg string; prover.rs:44-62)
 * and squeeze 16 bytes (verifier_message::<u128>, challenge.rs:85-90). */
typedef struct {
    void* ctx;
    void (*absorb)(void* ctx, int kind, const uint8_t* bytes, size_t len);
    void (*squeeze)(void* ctx, uint8_t out[16]);
    /* ProverState::grind (prover.rs:29-41) handed over whole, or NULL: then
     * oracle_transcript_grind frames it over absorb and squeeze. */
    void (*grind)(void* ctx, const uint8_t* label, size_t len, uint32_t bits);
} oracle_transcript;
/* Grinding (crates/transcript/src/{prover.rs:29-41,pow.rs:11-52}); BLAKE3 of
 * at most 64 bytes; security.rs:25-32 grinding_bits. */
void oracle_transcript_grind(oracle_transcript* t, const uint8_t* label, size_t len, uint32_t bits);
int oracle_blake3_small(const uint8_t* in, size_t len, uint8_t out[32]);
int oracle_pow_valid(const uint8_t seed[16], uint64_t nonce, uint32_t bits);
uint64_t oracle_pow_find(const uint8_t seed[16], uint32_t bits);
uint32_t oracle_grinding_bits(uint32_t security_bits, uint64_t coefficient);

/* Seeded source: challenges are splitmix64 words (gkr/src/lib.rs:986-999),
 * F128::from([w0, w1]); absorbed messages are recorded but do not steer it. */
typedef struct {
    uint64_t state;
    uint8_t* narg;
    size_t narg_len, narg_cap;
    uint64_t absorbed_bytes;
} oracle_seeded;
void oracle_seeded_init(oracle_seeded* s, uint64_t seed);
void oracle_seeded_free(oracle_seeded* s);
oracle_transcript oracle_seeded_transcript(oracle_seeded* s);
uint64_t oracle_splitmix_next(uint64_t* state);

/* SHA-256 duplex: every absorbed message and every squeeze ratchets a
 * SHA-256 chaining state, so each challenge depends on every earlier message.
 * A message-binding stand-in for spongefish; it is not spongefish. */
typedef struct {
    uint8_t state[32];
    uint8_t* narg;
    size_t narg_len, narg_cap;
} oracle_duplex;
void oracle_duplex_init(oracle_duplex* d, const uint8_t* domain, size_t len);
void oracle_duplex_free(oracle_duplex* d);
oracle_transcript oracle_duplex_transcript(oracle_duplex* d);
void oracle_sha256(const uint8_t* msg, size_t len, uint8_t out[32]);

/* Keccak-f[1600] (FIPS 202), a one-message sponge for the known answers,
 * and the Keccak duplex transcript (rate 136 bytes): start 'D' || le64(len)
 * || domain; message ('P' public | 'N' narg) || le64(len) || bytes; squeeze
 * 'S', pad 0x1F at the position and 0x80 at byte 135, permute, output the
 * first 16 bytes.  A stand-in shaped like spongefish's duplex; it is not
 * spongefish. */
void oracle_keccak_f1600(uint64_t a[25]);
void oracle_keccak_hash(const uint8_t* msg, size_t n, int rate, uint8_t pad, uint8_t* out, size_t out_len);
typedef struct {
    uint64_t a[25];
    size_t pos;
    uint8_t* narg;
    size_t narg_len, narg_cap;
} oracle_keccak_duplex;
void oracle_keccak_duplex_init(oracle_keccak_duplex* d, const uint8_t* domain, size_t len);
void oracle_keccak_duplex_free(oracle_keccak_duplex* d);
oracle_transcript oracle_keccak_duplex_transcript(oracle_keccak_duplex* d);

/* ---- poly: crates/poly/src/{eq.rs,mle.rs,f128.rs} ------------------------ */
void oracle_eq_table(const of128* r, int k, of128* out);        /* eq.rs:63-99 */
of128 oracle_eq_eval(const of128* x, const of128* y, int n);    /* eq.rs:22-41 */
void oracle_mle_fold(const of128* in, size_t len, of128 r, of128* out); /* mle.rs:240-258 */
of128 oracle_mle_evaluate(const of128* evals, const of128* point, int n); /* mle.rs:174-238 */
of128 oracle_inner_product(const of128* a, const of128* b, size_t n);    /* f128.rs:22-28 */

/* ---- spartan: crates/spartan/src/sumcheck.rs ----------------------------- */
enum {
    ORACLE_OK = 0,
    ORACLE_ERR_PRODUCT_DIMS = -1,  /* SumcheckError::InvalidProductDimensions */
    ORACLE_ERR_EQUALITY_DIMS = -2, /* SumcheckError::InvalidEqualityDimensions */
    ORACLE_ERR_ROUND_CLAIM = -3,   /* SumcheckError::InvalidRoundClaim */
    ORACLE_ERR_TERMINAL = -4,      /* SumcheckError::InvalidTerminalClaim */
    ORACLE_ERR_ROW_POINT = -5,     /* SpartanMatrixError::InvalidRowPointLength */
    ORACLE_ERR_ROUND_COUNT = -6,   /* SumcheckError::InvalidRoundCount */
    ORACLE_ERR_ARG = -7,
    ORACLE_POST_GKR_WITNESS_LENGTH = -13, /* post_gkr ProveError::WitnessLengthMismatch */
    ORACLE_POST_GKR_CLAIM = -14,          /* post_gkr ProveError::ClaimDoesNotHold */
    ORACLE_POST_GKR_MISMATCH = -15        /* post_gkr VerifyError::EvaluationMismatch */
};

/* ---- post_gkr: crates/post_gkr/src/{lib.rs,sumcheck.rs} ------------------- */
/* prove_factors (lib.rs:125-148): rows 2^t (t >= 7), columns 2^s, packed
 * 2^(t-7+s) elements of 128 bits; point receives t + s challenges (the
 * first off the packed bits), *evaluation = MLE[f](point). */
int oracle_post_gkr_prove(oracle_transcript* tr, const of128* rows, int t, const of128* columns, int s,
                          of128 target, const of128* packed, size_t packed_len, uint32_t security_bits,
                          of128* point, of128* evaluation);
/* verify_factors (lib.rs:151-166) over the prover's (t + s) x [a0, a2]. */
int oracle_post_gkr_verify(oracle_transcript* tr, const of128* rows, int t, const of128* columns, int s,
                           of128 target, uint32_t security_bits, const of128* messages, of128 evaluation,
                           of128* point);

/* prove_outer_sumcheck, sumcheck.rs:252-384.  round_polys receives
 * num_vars x [c0,c1,c2,c3]; claims = [az, bz, cz](r_x). */
int oracle_outer_sumcheck(oracle_transcript* t, of128 initial_claim,
                          const of128* eq_low, int low_vars,
                          const of128* eq_high, int high_vars,
                          const of128* az, const of128* bz, const of128* cz,
                          int az_vars, int bz_vars, int cz_vars,
                          of128* round_polys, of128* eval_points,
                          of128* final_claim, of128 claims[3]);

/* prove_inner_sumcheck, sumcheck.rs:442-552.  round_polys: num_vars x [c0,c1,c2];
 * evals = [D(r_y), h(r_y)]. */
int oracle_inner_sumcheck(oracle_transcript* t, of128 initial_claim,
                          const of128* matrix, int matrix_vars,
                          const of128* witness, int witness_vars,
                          of128* round_polys, of128* eval_points,
                          of128* final_claim, of128 evals[2]);

/* SumcheckProof::verify (sumcheck.rs:63-111) for COEFFS = coeffs. */
int oracle_sumcheck_verify(oracle_transcript* t, of128 initial_claim,
                           const of128* round_polys, int rounds, int coeffs,
                           of128* eval_points, of128* final_claim);
/* OuterSumcheckProof::verify (sumcheck.rs:218-243). */
int oracle_outer_verify(oracle_transcript* t, of128 initial_claim,
                        const of128* tau, int n, const of128* round_polys,
                        const of128 claims[3], of128* eval_points);

/* ---- spartan matrices: crates/spartan/src/matrix.rs ---------------------- */
/* One sparse matrix in compressed-row form (matrix.rs:97-125). */
typedef struct {
    const uint64_t* row_starts; /* rows + 1 */
    const uint32_t* columns;
    const of128* values;
} oracle_csr;
/* bind_and_batch, matrix.rs:429-468: out has 2^col_vars entries. */
int oracle_bind_and_batch(const oracle_csr mats[3], size_t rows, int row_vars,
                          int col_vars, const of128* row_point, of128 rho,
                          of128* out);
/* The reference tests' `multiply` (crates/spartan/src/piop.rs:352-364):
 * out[i] = sum over row i's entries of coefficient * assignment[column], for
 * i < rows; out has `rows` entries (padding is the caller's). */
void oracle_csr_multiply(const oracle_csr* m, size_t rows, const of128* assignment, of128* out);
/* evaluate_batched, matrix.rs:477-529 (direct, not via the bound table). */
int oracle_evaluate_batched(const oracle_csr mats[3], size_t rows, int row_vars,
                            int col_vars, const of128* row_point, of128 rho,
                            const of128* col_point, of128* out);

/* ---- circuit: crates/circuit/src/matrix_transpose.rs ---------------------- */
/* MaterializedMTranspose::apply (matrix_transpose.rs:327-375). */
void oracle_mt_apply(const uint32_t* column_offsets, size_t columns,
                     const uint32_t* row_indices, const of128* challenges,
                     of128* out);
/* transpose_inner (matrix_transpose.rs:377-441): out[c] for every column c
 * (column 0 is M's constant column). */
int oracle_mt_transpose_inner(const uint32_t* column_offsets, size_t columns,
                              const uint32_t* row_indices,
                              const of128* row_weights, size_t n_row_weights,
                              const of128* col_weights, size_t n_col_weights,
                              of128* out);

/* ---- gkr: crates/gkr/src/lib.rs ------------------------------------------- */
/* fold_layers (lib.rs:874-894): layers[0] = leaves (len 2^L), layer j+1 =
 * left half x right half of layer j, until `groups` values remain.  `store`
 * receives every layer back to back (leaves first, output last); returns the
 * number of layers including leaves and output. */
int oracle_gkr_fold_layers(const of128* leaves, int log_len, int log_groups,
                           of128* store);
/* EqTables::new (lib.rs:663-693): tables for point[0..n-1]; storage 2^(n)-1. */
void oracle_gkr_eq_tables(const of128* point, int n, of128* storage);
/* prove_layer over a dense children layer (lib.rs:88-159): children has
 * 2^(n+1) entries, point has n coordinates.  next_point gets n+1
 * coordinates; returns the claim W(r, t). */
of128 oracle_gkr_prove_layer(oracle_transcript* t, const of128* point, int n,
                             const of128* children, of128* next_point,
                             of128* messages /* (n + 1) x 2 */, uint32_t grind_bits);
/* gpgkr_prove over dense layers: from the output point down to the leaves;
 * security_bits 100 or 128 (SecurityLevel, crates/common/src/security.rs):
 * each cubic round grinds grinding_bits(3) before its challenge
 * (gkr/src/lib.rs:144). */
of128 oracle_gkr_prove(oracle_transcript* t, const of128* point, int n_out,
                       const of128* store, int log_leaves, of128* leaf_point,
                       uint32_t security_bits);
/* gpgkr_verify (lib.rs:717-772); returns 0 on success. */
int oracle_gkr_verify(oracle_transcript* t, of128 claim, const of128* point,
                      int n_out, int layers, const of128* messages,
                      of128* leaf_point, of128* leaf_claim);

#ifdef __cplusplus
}
#endif
#endif
