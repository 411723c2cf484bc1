/*
 * bitz_hip.h — C ABI of the MI355X (gfx950) GF(2^128) sumcheck engine.
 *
 * The drop-in boundary for BitZ's Spartan and GKR provers.  Plain pointers and
 * sizes only; no exceptions cross it; every call returns a status (0 = ok).
 *
 * Element layout: bitz_f128 is byte-identical to the reference's
 * `#[repr(C, align(16))] struct F128 { lo: u64, hi: u64 }`
 * (crates/field/src/gf128.rs:60-68): bit i of lo||hi is the coefficient of
 * X^i in GF(2)[X]/(X^128 + X^7 + X^2 + X + 1).
 *
 * Pointers: unless a call's flags include BITZ_DEVICE, element arrays are
 * host memory owned by the caller (copied in / out).  With BITZ_DEVICE they
 * are device pointers from bitz_device_alloc (inputs stay resident in HBM).
 *
 * Fiat–Shamir stays with the caller: the full provers take a bitz_transcript
 * (two callbacks: absorb a message, squeeze 16 bytes), and the per-round
 * handle API (bitz_outer_*, bitz_inner_*) takes each challenge as an argument,
 * so a Rust wrapper can keep spongefish on its side of the boundary.
 *
 * Handles are not thread-safe; use one per host thread.
 */
#ifndef BITZ_HIP_H
#define BITZ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t lo, hi;
} bitz_f128;

/* Status codes.  The negative protocol errors mirror the reference's enums. */
enum {
    BITZ_OK = 0,
    BITZ_ERR_PRODUCT_DIMS = -1,  /* SumcheckError::InvalidProductDimensions (crates/spartan/src/sumcheck.rs:44) */
    BITZ_ERR_EQUALITY_DIMS = -2, /* SumcheckError::InvalidEqualityDimensions (sumcheck.rs:45) */
    BITZ_ERR_ROUND_CLAIM = -3,   /* SumcheckError::InvalidRoundClaim (sumcheck.rs:42) */
    BITZ_ERR_TERMINAL = -4,      /* SumcheckError::InvalidTerminalClaim (sumcheck.rs:43) */
    BITZ_ERR_ROW_POINT = -5,     /* SpartanMatrixError::InvalidRowPointLength (crates/spartan/src/matrix.rs:25) */
    BITZ_ERR_TOO_MANY = -6,      /* DenseMleError::TooManyChallenges (crates/poly/src/mle.rs:23) */
    BITZ_ERR_ARG = -7,           /* bad argument (null pointer, size out of range) */
    BITZ_ERR_STATE = -8,         /* handle used out of order */
    BITZ_ERR_POINT_LEN = -9,     /* GKR point length != tree bits (crates/gkr/src/lib.rs:39-43 panics) */
    BITZ_ERR_ASSIGNMENT_DIMS = -10,  /* SpartanError::InvalidAssignmentDimensions (crates/spartan/src/piop.rs:30, checked :80-82) */
    BITZ_ERR_ASSIGNMENT_LEN = -11,   /* SpartanMatrixError::InvalidAssignmentLength (matrix.rs:23) */
    BITZ_ERR_ASSIGNMENT_CONST = -12, /* SpartanMatrixError::InvalidAssignmentConstant (matrix.rs:24) */
    BITZ_ERR_WITNESS_LENGTH = -13,   /* post_gkr ProveError::WitnessLengthMismatch (crates/post_gkr/src/lib.rs:52-53) */
    BITZ_ERR_CLAIM = -14,            /* post_gkr ProveError::ClaimDoesNotHold (lib.rs:54-55) */
    /* ProofDecodingError (crates/host/src/wire_proof.rs:43-60) */
    BITZ_ERR_PROOF_SHORT = -20,      /* OutOfInput: the input ends inside the header */
    BITZ_ERR_PROOF_MAGIC = -21,      /* BadMagic */
    BITZ_ERR_PROOF_VERSION = -22,    /* UnsupportedVersion(version) */
    BITZ_ERR_PROOF_HEADER_LEN = -23, /* BadHeaderLen(len) */
    BITZ_ERR_PROOF_FLAGS = -24,      /* ReservedFlagsSet(flags) */
    BITZ_ERR_PROOF_OVERFLOW = -25,   /* LengthOverflow */
    BITZ_ERR_PROOF_LENGTH = -26,     /* LengthMismatch { declared, actual } */
    BITZ_ERR_DEVICE = -100,      /* HIP runtime error; see bitz_last_error() */
    BITZ_ERR_ALLOC = -101        /* device allocation failed */
};

#define BITZ_DEVICE 1u /* element arrays are device pointers */

/* ---- runtime --------------------------------------------------------------- */
int32_t bitz_init(int32_t device);
const char* bitz_last_error(void);
int32_t bitz_device_alloc(void** ptr, size_t bytes);
int32_t bitz_device_free(void* ptr);
int32_t bitz_memcpy_h2d(void* dst, const void* src, size_t bytes);
int32_t bitz_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* Device-to-device copy (a plain streaming copy kernel; also the known-size
 * reference for HBM traffic counters). */
int32_t bitz_memcpy_d2d(void* dst, const void* src, size_t bytes);
int32_t bitz_synchronize(void);
/* Device scratch is recycled across calls through a size-keyed cache; this
 * returns every cached block to the runtime. */
int32_t bitz_release_cache(void);

/* Kernel timing: with profiling on, the sumcheck kernels ("outer_round0",
 * "outer_fold_next", "inner_round0", "inner_fold_next", "gkr_sums",
 * "gkr_bind_next") are bracketed by HIP events on their own stream.  read
 * returns the summed duration, launch count and algorithmic bytes (the table
 * bytes each launch must read and write).  No reference counterpart: the
 * reference times stages with tracing spans (tooling/cli/src/span_stats.rs). */
int32_t bitz_profile_enable(int32_t on);
/* Restricts timing to launches of one name (NULL or "" = all): every timed
 * launch adds two event records to its stream. */
int32_t bitz_profile_only(const char* kernel);
int32_t bitz_profile_read(const char* kernel, double* total_ms, uint64_t* launches, double* algo_bytes);

/* ---- transcripts ----------------------------------------------------------- */
/* A transcript absorbs messages (kind 0 = public message, 1 = prover message
 * also appended to the narg string; crates/transcript/src/prover.rs:44-62) and
 * squeezes 16-byte challenges (verifier_message::<u128>, challenge.rs:85-90;
 * F128 challenges are F128::from(u128), little-endian lo first). */
typedef struct bitz_transcript bitz_transcript;
typedef void (*bitz_absorb_fn)(void* ctx, int32_t kind, const uint8_t* bytes, size_t len);
typedef void (*bitz_squeeze_fn)(void* ctx, uint8_t out[16]);
/* Caller-provided sponge (e.g. a Rust closure around transcript::ProverState). */
bitz_transcript* bitz_transcript_callbacks(void* ctx, bitz_absorb_fn absorb, bitz_squeeze_fn squeeze);
/* Challenges from the splitmix64 stream of the seed, F128::from([w0, w1])
 * (the generator of crates/gkr/src/lib.rs:986-999); messages are recorded,
 * not absorbed. */
bitz_transcript* bitz_transcript_seeded(uint64_t seed);
/* SHA-256 ratchet over every message: state = SHA256(state||'A'||kind||le64(len)||m),
 * challenge = SHA256(state||'S')[0..16], state = SHA256(state||'R'). */
bitz_transcript* bitz_transcript_duplex(const uint8_t* domain, size_t len);
/* Keccak duplex over Keccak-f[1600] (FIPS 202), rate 136 bytes: start
 * 'D' || le64(len) || domain; message ('P' public | 'N' prover) ||
 * le64(len) || bytes; squeeze: 'S', pad 0x1F at the position and 0x80 at
 * byte 135, permute, the state's first 16 bytes.  Shaped like spongefish's
 * duplex sponge over Keccak (crates/transcript/src/domain.rs:13-42), whose
 * own framing is not reproduced.  The same duplex runs on the device: a
 * prover given a Keccak transcript draws its challenges there (see
 * bitz_set_device_fiat_shamir) and hands the state back at the end. */
bitz_transcript* bitz_transcript_keccak(const uint8_t* domain, size_t len);
/* Whether provers draw a Keccak transcript's challenges on the device (on by
 * default; off: every round's message goes to the host transcript and the
 * challenge comes back).  Same transcript, messages and results either way;
 * no counterpart in the reference, whose transcript is host code.  Today:
 * bitz_gkr_prove without grinding (Bits100), and the PIOP engine's proofs
 * over a packed assignment (bitz_piop_engine_submit_packed / _assignment,
 * bitz_server_serve). */
int32_t bitz_set_device_fiat_shamir(int32_t on);
/* Rounds a PIOP engine step may advance each device-transcript proof by (1 to
 * 16), or 0 (the default): 6 while the engine holds fewer than 8 proofs, 4
 * from 8 on.  A tuning switch with no counterpart in the reference; the
 * proofs are the same for every value. */
int32_t bitz_set_engine_positions(int32_t positions);
/* One-message Keccak sponge (pad 0x06: SHA3-*, 0x1F: SHAKE*) on the host,
 * or with BITZ_DEVICE by one device thread: the permutation's check against
 * FIPS 202 answers. */
int32_t bitz_keccak_hash(const uint8_t* msg, size_t len, int32_t rate, uint8_t pad, uint8_t* out, size_t out_len,
                         uint32_t flags);
void bitz_transcript_free(bitz_transcript* t);
int32_t bitz_transcript_absorb(bitz_transcript* t, int32_t kind, const uint8_t* bytes, size_t len);
int32_t bitz_transcript_squeeze(bitz_transcript* t, bitz_f128* out);
/* Copies up to cap narg bytes; returns the total narg length. */
size_t bitz_transcript_narg(const bitz_transcript* t, uint8_t* out, size_t cap);
/* Grinding, ProverState::grind (crates/transcript/src/prover.rs:29-41): at
 * zero bits nothing; else, for a callback transcript with a grind hook, the
 * hook (the caller's own grind, e.g. ProverState::grind); otherwise over the
 * transcript's absorb and squeeze: public messages "bitz-transcript-pow-v1",
 * le64(label length), the label, le32(bits) (pow.rs:11-16), a squeezed seed,
 * and the first nonce n with BLAKE3("bitz-pcs-pow-v1" || seed || le64(n))
 * starting with `bits` zero bits (pow.rs:21-52) as an 8-byte prover message.
 * bits > 32 -> BITZ_ERR_ARG. */
typedef void (*bitz_grind_fn)(void* ctx, const uint8_t* label, size_t len, uint32_t bits);
int32_t bitz_transcript_set_grind(bitz_transcript* t, bitz_grind_fn grind);
int32_t bitz_transcript_grind(bitz_transcript* t, const uint8_t* label, size_t len, uint32_t bits);
/* pow.rs:21-52 find / valid, and SecurityLevel::grinding_bits
 * (crates/common/src/security.rs:25-32) for security_bits 100 or 128. */
int32_t bitz_pow_find(const uint8_t seed[16], uint32_t bits, uint64_t* nonce);
int32_t bitz_pow_valid(const uint8_t seed[16], uint64_t nonce, uint32_t bits);
uint32_t bitz_grinding_bits(uint32_t security_bits, uint64_t coefficient);

/* ---- proof bytes ------------------------------------------------------------ */
/* Proof { narg_string, hints } (crates/transcript/src/proof.rs:13-16): the
 * hint stream rides beside the narg string and is never absorbed
 * (ProverState::hint, prover.rs:64-68); built-in transcripts keep it, a
 * callback transcript keeps its own (BITZ_ERR_ARG). */
int32_t bitz_transcript_hint(bitz_transcript* t, const uint8_t* bytes, size_t len);
/* Copies up to cap hint bytes; returns the total hint length. */
size_t bitz_transcript_hints(const bitz_transcript* t, uint8_t* out, size_t cap);
/* The proof container (crates/host/src/wire_proof.rs): a 32-byte header
 * (magic "bitzPCS\0", wire_version 1, header_len 32, flags 0, le64 narg
 * length, le64 hint length, all little-endian), the narg bytes, the hint
 * bytes.  encode: *len = 32 + narg_len + hints_len (out NULL: length only;
 * cap too small: BITZ_ERR_ARG).  decode: the exact encoding or the first
 * failing check in the header's order (BITZ_ERR_PROOF_*; *detail = the
 * version, header length, flags or declared length of the error); on
 * success the two streams' offsets and lengths within `bytes`. */
int32_t bitz_proof_encode(const uint8_t* narg, size_t narg_len, const uint8_t* hints, size_t hints_len,
                          uint8_t* out, size_t cap, size_t* len);
int32_t bitz_proof_decode(const uint8_t* bytes, size_t len, size_t* narg_offset, size_t* narg_len,
                          size_t* hints_offset, size_t* hints_len, uint64_t* detail);
/* ProverState::finish then wire_proof::encode for a built-in transcript. */
int32_t bitz_transcript_proof(const bitz_transcript* t, uint8_t* out, size_t cap, size_t* len);

/* ---- field and multilinear primitives -------------------------------------- */
/* out[i] = a[i] * b[i] (F128 Mul, crates/field/src/gf128.rs:275-281).  The
 * product primitive is selectable so each can be checked on its own (no
 * counterpart in the reference): the default is the comb; BITZ_MUL_LANE the
 * per-lane LDS table product of the round kernels, BITZ_MUL_LANE2 its
 * two-operand form (one table of a[i], the second chain taking a neighbour's
 * b), BITZ_MUL_CONST out[i] = a[i] * b[0] through b[0]'s 4-bit table,
 * BITZ_MUL_ROWS the four-products-per-wave form (one 16-lane row per
 * product, the row's partial products XORed by DPP moves), BITZ_MUL_MFMA
 * out[i] = a[i] * b[0] as b[0]'s 128x128 F2 matrix on the int8 matrix cores
 * (counts, then parities).  With profiling on, the launch is timed as
 * "mul_comb", "mul_lane", "mul_lane2", "mul_const", "mul_rows" or
 * "mul_mfma". */
#define BITZ_MUL_LANE 0x100u
#define BITZ_MUL_LANE2 0x200u
#define BITZ_MUL_CONST 0x400u
#define BITZ_MUL_ROWS 0x800u
#define BITZ_MUL_MFMA 0x1000u
int32_t bitz_mul_batch(const bitz_f128* a, const bitz_f128* b, bitz_f128* out, size_t n, uint32_t flags);
/* poly::eq_table (crates/poly/src/eq.rs:63-99): out has 2^k entries. */
int32_t bitz_eq_table(const bitz_f128* r, uint32_t k, bitz_f128* out, uint32_t flags);
/* DenseMultilinearExtension::fold (crates/poly/src/mle.rs:136-156): binds the
 * lowest variables to r[0..m), in place; *num_vars is updated. */
int32_t bitz_mle_fold(bitz_f128* evals, uint32_t* num_vars, const bitz_f128* r, uint32_t m, uint32_t flags);
/* DenseMultilinearExtension::evaluate (mle.rs:158-172). */
int32_t bitz_mle_evaluate(const bitz_f128* evals, uint32_t num_vars, const bitz_f128* point,
                          uint32_t point_len, bitz_f128* out, uint32_t flags);

/* ---- Spartan outer sumcheck ------------------------------------------------- */
/* prove_outer_sumcheck (crates/spartan/src/sumcheck.rs:252-384).
 * eq factors are consumed, products borrowed (never modified).
 * Outputs: round_polys[num_vars][4] = [c0,c1,c2,c3], eval_points[num_vars],
 * *final_claim, claims[3] = [Az(r_x), Bz(r_x), Cz(r_x)] (absorbed last). */
int32_t bitz_prove_outer_sumcheck(bitz_transcript* t, bitz_f128 initial_claim,
                                  const bitz_f128* eq_low, uint32_t low_vars,
                                  const bitz_f128* eq_high, uint32_t high_vars,
                                  const bitz_f128* az, uint32_t az_vars,
                                  const bitz_f128* bz, uint32_t bz_vars,
                                  const bitz_f128* cz, uint32_t cz_vars, uint32_t flags,
                                  bitz_f128* round_polys, bitz_f128* eval_points,
                                  bitz_f128* final_claim, bitz_f128* claims);

/* The claim the outer sumcheck proves: sum_x eq(x) (Az(x) Bz(x) + Cz(x)) with
 * eq(x) = eq_low[x & (2^low_vars - 1)] * eq_high[x >> low_vars] — the identity
 * prove_outer_sumcheck asserts at its end (sumcheck.rs:367-370).  A satisfied
 * R1CS gives 0, which prove_spartan_piop passes (crates/spartan/src/piop.rs:91). */
int32_t bitz_outer_claim(const bitz_f128* eq_low, uint32_t low_vars, const bitz_f128* eq_high,
                         uint32_t high_vars, const bitz_f128* az, const bitz_f128* bz,
                         const bitz_f128* cz, uint32_t num_vars, uint32_t flags, bitz_f128* out);

/* The same prover against challenges fixed in advance, challenges[num_vars]
 * (a transcript whose challenges do not depend on the messages, such as the
 * seeded stream): all rounds are queued back to back with one read of the
 * round sums at the end.  eval_points receives the challenges. */
int32_t bitz_prove_outer_sumcheck_with_challenges(const bitz_f128* challenges, bitz_f128 initial_claim,
                                                  const bitz_f128* eq_low, uint32_t low_vars,
                                                  const bitz_f128* eq_high, uint32_t high_vars,
                                                  const bitz_f128* az, uint32_t az_vars,
                                                  const bitz_f128* bz, uint32_t bz_vars,
                                                  const bitz_f128* cz, uint32_t cz_vars, uint32_t flags,
                                                  bitz_f128* round_polys, bitz_f128* eval_points,
                                                  bitz_f128* final_claim, bitz_f128* claims);

/* A batch of independent outer sumchecks with challenges fixed in advance
 * (no counterpart in the reference, which proves one instance per call:
 * the batched mode behind many uneven instances on one GPU).  Every
 * instance's outputs equal bitz_prove_outer_sumcheck_with_challenges' on
 * the same inputs.  Tables and factors are device pointers (flags must
 * include BITZ_DEVICE); challenges and outputs are host memory. */
typedef struct {
    const bitz_f128* challenges; /* num_vars entries */
    bitz_f128 initial_claim;
    const bitz_f128* eq_low;
    uint32_t low_vars;
    const bitz_f128* eq_high;
    uint32_t high_vars;
    const bitz_f128* az;
    const bitz_f128* bz;
    const bitz_f128* cz;
    uint32_t num_vars;
    bitz_f128* round_polys; /* num_vars x [c0, c1, c2, c3] */
    bitz_f128* eval_points; /* num_vars */
    bitz_f128* final_claim; /* 1 */
    bitz_f128* claims;      /* [Az, Bz, Cz](r_x) */
} bitz_outer_instance;
int32_t bitz_prove_outer_batch(const bitz_outer_instance* instances, uint32_t count, uint32_t flags);
/* Batch flags (with BITZ_DEVICE).  By default a batch runs as a captured
 * launch graph (one launch per round over every instance), replayed while
 * the instance shapes repeat, with tile shapes picked by a timing autotuner
 * on the first batch of each shape. */
#define BITZ_BATCH_EAGER 2u   /* launch the kernels directly, no graph */
#define BITZ_BATCH_NO_TUNE 4u /* default tile shapes, no autotuning */
/* Tile shapes of the last batch: the grid-wide rounds' thread-grid cap
 * (log2) and the one-workgroup tail's size (log2 table length). */
int32_t bitz_outer_batch_config(uint32_t* thread_bits, uint32_t* tail_log);
/* Fixes the tile shapes of later batches (thread_bits in [6, 20], tail_log
 * in [1, 10]); thread_bits = 0 returns to the autotuner. */
int32_t bitz_outer_batch_set_config(uint32_t thread_bits, uint32_t tail_log);
/* Frees the cached batch plan (its device buffers and graph). */
void bitz_outer_batch_release(void);

/* Per-round form of the same prover for callers that keep the transcript:
 * begin -> round0 ([c0,c2,c3]) -> fold_next(r_k) per round (returns the next
 * round's [c0,c2,c3], zeros after the last) -> finish ([Az,Bz,Cz](r_x)).
 * The caller reconstructs c1 = claim - 2 c0 - c2 - c3, absorbs, squeezes
 * (recover_full_round_polynomial_and_sample_next_challenge, sumcheck.rs:803-829). */
typedef struct bitz_outer bitz_outer;
int32_t bitz_outer_begin(bitz_outer** h, const bitz_f128* eq_low, uint32_t low_vars,
                         const bitz_f128* eq_high, uint32_t high_vars, const bitz_f128* az,
                         uint32_t az_vars, const bitz_f128* bz, uint32_t bz_vars,
                         const bitz_f128* cz, uint32_t cz_vars, uint32_t flags);
/* 1 when the factors were recognised as eq tables of a point tau and the
 * eq-factored kernels run; 0 for the general per-pair eq kernels.  Waits for
 * the device's check of the factors if round0 has not read it yet. */
int32_t bitz_outer_eq_mode(bitz_outer* h);
int32_t bitz_outer_round0(bitz_outer* h, bitz_f128* coeffs3);
int32_t bitz_outer_fold_next(bitz_outer* h, bitz_f128 r, bitz_f128* coeffs3);
int32_t bitz_outer_finish(bitz_outer* h, bitz_f128* claims3);
/* Copies the current (folded) A, B, C tables, `len` entries each, to host. */
int32_t bitz_outer_tables(bitz_outer* h, bitz_f128* az, bitz_f128* bz, bitz_f128* cz, size_t* len);
void bitz_outer_free(bitz_outer* h);

/* ---- Spartan inner sumcheck ------------------------------------------------- */
/* prove_inner_sumcheck with the challenges known in advance (a transcript
 * whose challenges do not depend on the messages, such as the seeded
 * stream): all rounds queued back to back, one read-back; the outputs equal
 * bitz_prove_inner_sumcheck's under a transcript that squeezes the same
 * challenges.  No counterpart in the reference, which draws them round by
 * round (crates/spartan/src/sumcheck.rs:442-552). */
int32_t bitz_prove_inner_sumcheck_with_challenges(const bitz_f128* challenges, bitz_f128 initial_claim,
                                                  const bitz_f128* matrix, uint32_t matrix_vars,
                                                  const bitz_f128* witness, uint32_t witness_vars, uint32_t flags,
                                                  bitz_f128* round_polys, bitz_f128* eval_points,
                                                  bitz_f128* final_claim, bitz_f128* evals);
/* bitz_prove_inner_sumcheck_with_challenges without any allocation: no
 * hipMalloc, hipFree, pinned buffer or other allocating call.  No counterpart
 * in the reference; it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it).  matrix and
 * witness are device tables of at most 2^20 entries (num_vars <= 20, else
 * BITZ_ERR_ARG); the ping-pong tables, challenge tables and round sums live in
 * a static device array of 24.2 MiB that every GPU the library runs on holds
 * until the process ends.  The work runs on `stream` (a hipStream_t; null =
 * the default stream) and the call returns once it is done.  Calls that
 * overlap on one GPU share that array and race.  Outputs as
 * bitz_prove_inner_sumcheck_with_challenges. */
int32_t bitz_prove_inner_sumcheck_with_challenges_noalloc(const bitz_f128* challenges, bitz_f128 initial_claim,
                                                          const bitz_f128* matrix, uint32_t matrix_vars,
                                                          const bitz_f128* witness, uint32_t witness_vars,
                                                          void* stream, bitz_f128* round_polys,
                                                          bitz_f128* eval_points, bitz_f128* final_claim,
                                                          bitz_f128* evals);
/* prove_inner_sumcheck (sumcheck.rs:442-552): round_polys[num_vars][3],
 * evals[2] = [D(r_y), h(r_y)]. */
int32_t bitz_prove_inner_sumcheck(bitz_transcript* t, bitz_f128 initial_claim,
                                  const bitz_f128* matrix, uint32_t matrix_vars,
                                  const bitz_f128* witness, uint32_t witness_vars, uint32_t flags,
                                  bitz_f128* round_polys, bitz_f128* eval_points,
                                  bitz_f128* final_claim, bitz_f128* evals);
/* prove_inner_sumcheck (sumcheck.rs:442-552) with the assignment h given as
 * the packed bits of a Boolean table (bit j = bits[j / 64] >> (j % 64) & 1,
 * 2^assignment_vars bits, zero past the assignment; bitz_assignment_pack
 * writes them): the table build_assignment_mle (crates/spartan/src/matrix.rs:
 * 392-420) expands.  Outputs equal bitz_prove_inner_sumcheck's over that
 * table; the first rounds need no general products (h_k is a combination of
 * the monomials of r_0 .. r_{k-1} with bit coefficients).  BITZ_DEVICE:
 * matrix and bits are device arrays. */
int32_t bitz_prove_inner_sumcheck_packed(bitz_transcript* t, bitz_f128 initial_claim,
                                         const bitz_f128* matrix, uint32_t matrix_vars,
                                         const uint64_t* assignment_bits, uint32_t assignment_vars,
                                         uint32_t flags, bitz_f128* round_polys, bitz_f128* eval_points,
                                         bitz_f128* final_claim, bitz_f128* evals);
typedef struct bitz_inner bitz_inner;
int32_t bitz_inner_begin(bitz_inner** h, const bitz_f128* matrix, uint32_t matrix_vars,
                         const bitz_f128* witness, uint32_t witness_vars, uint32_t flags);
/* The per-round handle over a packed Boolean assignment (as
 * bitz_prove_inner_sumcheck_packed); round0 / fold_next / finish as above. */
int32_t bitz_inner_begin_packed(bitz_inner** h, const bitz_f128* matrix, uint32_t matrix_vars,
                                const uint64_t* assignment_bits, uint32_t assignment_vars, uint32_t flags);
int32_t bitz_inner_round0(bitz_inner* h, bitz_f128* coeffs2);
int32_t bitz_inner_fold_next(bitz_inner* h, bitz_f128 r, bitz_f128* coeffs2);
int32_t bitz_inner_finish(bitz_inner* h, bitz_f128* evals2);
void bitz_inner_free(bitz_inner* h);

/* ---- Spartan constraint matrices ------------------------------------------- */
/* PreparedConstraintMatrices (crates/spartan/src/matrix.rs:54-64, 71-125):
 * A, B, C in compressed-row form (row_starts[rows+1], columns, values). */
typedef struct bitz_matrices bitz_matrices;
int32_t bitz_matrices_prepare(bitz_matrices** h, size_t rows, uint32_t row_vars, uint32_t col_vars,
                              const uint64_t* const row_starts[3], const uint32_t* const columns[3],
                              const bitz_f128* const values[3]);
/* bind_and_batch (matrix.rs:429-468): out[2^col_vars] = D. */
int32_t bitz_bind_and_batch(bitz_matrices* h, const bitz_f128* row_point, uint32_t point_len,
                            bitz_f128 rho, bitz_f128* out, uint32_t flags);
/* Az, Bz, Cz = A h, B h, C h over the prepared matrices, each padded with
 * zeros to 2^row_vars: the F128 restatement of the reference test's
 * multiply + padded_mle (crates/spartan/src/piop.rs:352-372), the reference's
 * own producer build_product_mles (matrix.rs:336-370) being over a prime
 * field.  assignment: 2^col_vars entries; outputs: 2^row_vars each. */
int32_t bitz_r1cs_products(bitz_matrices* h, const bitz_f128* assignment, bitz_f128* az, bitz_f128* bz,
                           bitz_f128* cz, uint32_t flags);
/* build_assignment_mle (matrix.rs:392-420): the packed Boolean assignment
 * (bit i = words[i / 64] >> (i % 64) & 1, as PackedWitness::bit) as {0, 1}
 * elements padded with zeros to 2^ceil(log2 columns).  bit_len != columns ->
 * BITZ_ERR_ASSIGNMENT_LEN; columns == 0 or bit 0 clear -> BITZ_ERR_ASSIGNMENT_CONST. */
int32_t bitz_assignment_mle(const uint64_t* words, size_t bit_len, size_t columns, bitz_f128* out, uint32_t flags);
/* build_assignment_mle's checks (matrix.rs:392-420) over the packed
 * assignment, kept packed: out receives ceil(2^ceil(log2 columns) / 64)
 * words (at least one), the bits past bit_len cleared.  BITZ_DEVICE: out is
 * a device array.  Errors as bitz_assignment_mle's. */
int32_t bitz_assignment_pack(const uint64_t* words, size_t bit_len, size_t columns, uint64_t* out, uint32_t flags);
/* bitz_r1cs_products over a packed Boolean assignment (bitz_assignment_pack's
 * form, 2^col_vars bits): each row is the XOR of its values at the columns
 * whose bit is set.  BITZ_DEVICE: bits and outputs on the device. */
int32_t bitz_r1cs_products_packed(bitz_matrices* h, const uint64_t* assignment_bits, bitz_f128* az, bitz_f128* bz,
                                  bitz_f128* cz, uint32_t flags);
/* constraint_matrix_digest (matrix.rs:562-627): SHA-256 over the domain
 * "bitz/spartan/constraint-matrices/v1", M's positions (m_rows rows of
 * m_columns columns, CSR), then A, B, C row by row (rows x columns, CSR) with
 * every coefficient as its 16-byte encoding, lengths and indices as u64 LE. */
int32_t bitz_constraint_matrix_digest(size_t m_rows, size_t m_columns, const uint64_t* m_row_starts,
                                      const uint32_t* m_positions, size_t rows, size_t columns,
                                      const uint64_t* const row_starts[3], const uint32_t* const cols[3],
                                      const bitz_f128* const values[3], uint8_t digest[32]);
/* prove_spartan_piop (crates/spartan/src/piop.rs:67-112): absorb the digest,
 * squeeze tau (row_vars challenges), outer sumcheck from claim 0 over
 * eq(tau) = make_equality_factors(tau), squeeze rho, bind_and_batch at the
 * outer point, inner sumcheck from az + rho bz + rho^2 cz against the
 * assignment.  Products (2^row_vars each) and assignment (2^col_vars) on the
 * side flags names.  Outputs (host): outer_rounds (row_vars x [c0..c3]),
 * outer_claims [Az, Bz, Cz](r_x), inner_rounds (col_vars x [c0, c1, c2]),
 * inner_point r_y (col_vars), matrix_eval D(r_y), final_claim. */
int32_t bitz_prove_spartan_piop(bitz_transcript* t, bitz_matrices* h, const uint8_t digest[32],
                                const bitz_f128* az, const bitz_f128* bz, const bitz_f128* cz, uint32_t row_vars,
                                const bitz_f128* assignment, uint32_t col_vars, uint32_t flags,
                                bitz_f128* outer_rounds, bitz_f128* outer_claims, bitz_f128* inner_rounds,
                                bitz_f128* inner_point, bitz_f128* matrix_eval, bitz_f128* final_claim);
/* bitz_prove_spartan_piop with the assignment as packed bits (the form of
 * bitz_assignment_pack, 2^col_vars bits, on the side flags names): the same
 * outputs; the inner sumcheck runs as bitz_prove_inner_sumcheck_packed. */
int32_t bitz_prove_spartan_piop_packed(bitz_transcript* t, bitz_matrices* h, const uint8_t digest[32],
                                       const bitz_f128* az, const bitz_f128* bz, const bitz_f128* cz,
                                       uint32_t row_vars, const uint64_t* assignment_bits, uint32_t col_vars,
                                       uint32_t flags, bitz_f128* outer_rounds, bitz_f128* outer_claims,
                                       bitz_f128* inner_rounds, bitz_f128* inner_point, bitz_f128* matrix_eval,
                                       bitz_f128* final_claim);
void bitz_matrices_free(bitz_matrices* h);
/* Device bytes the handle holds (every buffer as allocated). */
int32_t bitz_matrices_device_bytes(const bitz_matrices* h, uint64_t* bytes);

/* ---- post-GKR inner-product sumcheck ----------------------------------------- */
/* post_gkr::prove (crates/post_gkr/src/lib.rs:85-148, sumcheck.rs:99-178):
 * the claim <rows (x) columns, f> = target over the committed bits f, 2^t x
 * 2^s (t = log_rows >= 7, bit (c << t) | b weighs rows[b] columns[c]),
 * packed 128 bits to an element column after column (packed_len =
 * 2^(t-7+s)); returns MLE[f](point) = *evaluation, point = t + s challenges
 * low coordinate first.  Per round: prover message [a0, a2], grind
 * grinding_bits(2) bits ("bitz/pcs/sumcheck/v1"; 1 at Bits128), challenge;
 * then the evaluation as a prover message.  Errors before the transcript:
 * BITZ_ERR_WITNESS_LENGTH, BITZ_ERR_CLAIM.  BITZ_DEVICE: rows, columns and
 * packed are device arrays. */
int32_t bitz_post_gkr_prove(bitz_transcript* t, const bitz_f128* rows, uint32_t log_rows, const bitz_f128* columns,
                            uint32_t log_columns, bitz_f128 target, const bitz_f128* packed, size_t packed_len,
                            uint32_t security_bits, uint32_t flags, bitz_f128* point, bitz_f128* evaluation);

/* ---- many PIOP proofs, rounds batched (no reference counterpart) -------- */
/* An engine advances every submitted prove_spartan_piop (piop.rs:67-112) by
 * one round per step: the rounds of one kind across all proofs run as one
 * launch, every proof's sums come back in one read, and each proof's own
 * transcript absorbs its round polynomial and squeezes its next challenge on
 * the host, in bitz_prove_spartan_piop's order.  A proof with a packed
 * assignment and a Keccak transcript (bitz_transcript_keccak) while device
 * Fiat-Shamir is on (bitz_set_device_fiat_shamir) draws its challenges on the
 * device instead and advances by up to 6 rounds per step; its transcript's
 * state is handed back when it finishes.  A proof's outputs equal
 * bitz_prove_spartan_piop's for the same transcript and inputs; proofs
 * submitted between steps join at the next step (continuous batching of a
 * proving server).  Products, assignment (device) and the transcript must
 * live until the proof is taken; a bitz_matrices handle in use by an engine
 * must not be used elsewhere meanwhile.  1 <= row_vars <= 32, col_vars >= 1.
 * Not thread-safe: one thread drives an engine. */
typedef struct bitz_piop_engine bitz_piop_engine;
int32_t bitz_piop_engine_create(bitz_piop_engine** e);
int32_t bitz_piop_engine_submit(bitz_piop_engine* e, uint64_t tag, bitz_transcript* t, bitz_matrices* m,
                                const uint8_t digest[32], const bitz_f128* az, const bitz_f128* bz,
                                const bitz_f128* cz, uint32_t row_vars, const bitz_f128* assignment,
                                uint32_t col_vars);
/* bitz_piop_engine_submit with the assignment as packed bits on the device
 * (bitz_assignment_pack's form, 2^col_vars bits); they must live until the
 * proof is taken.  Outputs as bitz_prove_spartan_piop_packed's. */
int32_t bitz_piop_engine_submit_packed(bitz_piop_engine* e, uint64_t tag, bitz_transcript* t, bitz_matrices* m,
                                       const uint8_t digest[32], const bitz_f128* az, const bitz_f128* bz,
                                       const bitz_f128* cz, uint32_t row_vars, const uint64_t* assignment_bits,
                                       uint32_t col_vars);
/* An engine proof from the packed assignment (PackedWitness words on the
 * host, bit_len bits; build_assignment_mle's checks, matrix.rs:392-420, and
 * 2^ceil(log2 columns) = the matrices' column domain): the engine uploads the
 * bits and computes Az, Bz, Cz itself (their XORs of values under the set
 * bits, on a stream of the proof's own), then proves as
 * bitz_piop_engine_submit_packed.  The caller keeps nothing alive but the
 * transcript and the matrices. */
int32_t bitz_piop_engine_submit_assignment(bitz_piop_engine* e, uint64_t tag, bitz_transcript* t, bitz_matrices* m,
                                           const uint8_t digest[32], const uint64_t* words, size_t bit_len,
                                           size_t columns);
/* One round of every proof in flight; *in_flight receives how many remain.
 * step = launch (enqueue the step's work on the engine's stream, no wait;
 * *launched = proofs in it) then finish (wait, then the host's part).  With
 * several engines, launching one while finishing another overlaps one's
 * host work with the other's device work. */
int32_t bitz_piop_engine_step(bitz_piop_engine* e, size_t* in_flight);
int32_t bitz_piop_engine_launch(bitz_piop_engine* e, size_t* launched);
int32_t bitz_piop_engine_finish(bitz_piop_engine* e, size_t* in_flight);
/* Background stepping: a thread of the engine steps while proofs are in
 * flight (submit, done, take and wait may then be called from other
 * threads; step/launch/finish may not).  wait blocks until a finished proof
 * is waiting to be taken or timeout_us passes; *finished = how many wait. */
int32_t bitz_piop_engine_run(bitz_piop_engine* e);
int32_t bitz_piop_engine_wait(bitz_piop_engine* e, uint32_t timeout_us, size_t* finished);
/* Tags of the finished proofs not yet taken (up to cap; *count = all). */
int32_t bitz_piop_engine_done(bitz_piop_engine* e, uint64_t* tags, size_t cap, size_t* count);
/* A finished proof's outputs (as bitz_prove_spartan_piop's; a null output is
 * not copied); releases it. */
/* Device bytes the proof `tag` holds in the engine, from its submission
 * (which allocates them all) until it is taken: what a serving budget must
 * set aside per request.  No counterpart in the reference. */
int32_t bitz_piop_engine_device_bytes(bitz_piop_engine* e, uint64_t tag, uint64_t* bytes);
int32_t bitz_piop_engine_take(bitz_piop_engine* e, uint64_t tag, bitz_f128* outer_rounds, bitz_f128* outer_claims,
                              bitz_f128* inner_rounds, bitz_f128* inner_point, bitz_f128* matrix_eval,
                              bitz_f128* final_claim);
void bitz_piop_engine_free(bitz_piop_engine* e);

/* ---- SHA-256 R1CS instances (crates/circuit/src/sha256.rs) ---------------- */
/* The constraint system (M, A, B, C) and witness (f, h) of BitZ's SHA-256
 * circuits, produced on the host by one gadget description run symbolically
 * (ConstraintGenerator, crates/circuit/src/constraints.rs:578-790) and
 * concretely (Witgen, crates/circuit/src/witgen.rs:425-500).
 *   COMPRESSION_CHAIN: inputs are the message words (bit j of a word is bit
 *     j of the big-endian u32) then the 256 chaining-value bits; block b
 *     starts from block b-1's output.  One block is compression_circuit
 *     (sha256.rs:476-490; 20457 x 7145 M, 184 rows, tooling/cli/tests/
 *     sha256_spartan.rs:44-49).  Matrices depend on the block count only.
 *   CHAIN: the CLI's sha256-chain statement (tooling/cli/src/circuits.rs:
 *     167-217): every message bit and every output bit is constrained to the
 *     statement (1 * bit = expected), compression from the FIPS IV.  The rows
 *     depend on the statement's message.
 *   BLOCK_ALIGNED: sha256_block_aligned_circuit (sha256.rs:503-550), the
 *     reference benchmark's circuit (tooling/cli/benches/sha256_spartan.rs:
 *     60-90): the message as a bit stream, then the padding block of
 *     constants; no statement rows. */
#define BITZ_SHA256_COMPRESSION_CHAIN 0u
#define BITZ_SHA256_CHAIN 1u
#define BITZ_SHA256_BLOCK_ALIGNED 2u
typedef struct bitz_sha256 bitz_sha256;
/* statement: 64 * blocks message bytes, for CHAIN only (NULL otherwise). */
int32_t bitz_sha256_circuit(bitz_sha256** out, uint32_t kind, size_t blocks, const uint8_t* statement);
/* rows of A/B/C; columns = |h| with the constant (= rows of M); witness_bits
 * = |f| (M has witness_bits + 1 columns); nonzeros of A, B, C; positions of M. */
int32_t bitz_sha256_shape(const bitz_sha256* c, size_t* rows, size_t* columns, size_t* witness_bits,
                          size_t nnz[3], size_t* m_positions);
/* CSR of A, B, C (rows + 1 starts, columns, integer coefficients, and their
 * images under the parity projection Z -> F2 in GF(2^128)), CSR of M
 * (columns + 1 starts, positions).  Any pointer may be NULL. */
int32_t bitz_sha256_export(const bitz_sha256* c, uint64_t* const row_starts[3], uint32_t* const cols[3],
                           int64_t* const coefficients[3], bitz_f128* const parity[3],
                           uint64_t* m_row_starts, uint32_t* m_positions);
/* The witness for `message` (64 * blocks bytes; block-aligned: the stream).
 * iv: the chaining value (COMPRESSION_CHAIN only; NULL = FIPS IV).  f_words
 * (ceil(witness_bits / 64)) and h_words (ceil(columns / 64)) are packed
 * least significant bit first (PackedWitness, witgen.rs:207-305); h bit 0 is
 * the constant one.  digest: the final state as 32 big-endian bytes
 * (block-aligned: the SHA-256 digest of the message).  satisfied: 1 when
 * every row holds over the integers for this witness.  Any output may be NULL. */
int32_t bitz_sha256_witness(const bitz_sha256* c, const uint8_t* message, const uint32_t iv[8], uint64_t* f_words,
                            uint64_t* h_words, uint8_t digest[32], int32_t* satisfied);
/* The parity projection of A, B, C prepared for the prover (as
 * bitz_matrices_prepare) and/or the statement digest over it (as
 * bitz_constraint_matrix_digest).  Either output may be NULL. */
int32_t bitz_sha256_prepare(const bitz_sha256* c, bitz_matrices** out, uint8_t digest[32]);
void bitz_sha256_free(bitz_sha256* c);

/* ---- M^T (crates/circuit/src/matrix_transpose.rs) ------------------------- */
typedef struct bitz_mt bitz_mt;
/* MaterializedMTranspose: CSC of M (column_offsets[columns+1], row_indices). */
int32_t bitz_mt_prepare(bitz_mt** h, size_t row_count, size_t columns, const uint32_t* column_offsets,
                        const uint32_t* row_indices);
/* apply (matrix_transpose.rs:327-375): out[columns] = r * M. */
int32_t bitz_mt_apply(bitz_mt* h, const bitz_f128* challenges, size_t n, bitz_f128* out, uint32_t flags);
/* transpose_inner (matrix_transpose.rs:377-441): out[columns], column 0 first. */
int32_t bitz_mt_transpose_inner(bitz_mt* h, const bitz_f128* row_weights, size_t n_row,
                                const bitz_f128* col_weights, size_t n_col, bitz_f128* out,
                                uint32_t flags);
void bitz_mt_free(bitz_mt* h);

/* ---- GKR grand product (crates/gkr/src/lib.rs) ---------------------------- */
typedef struct bitz_gkr bitz_gkr;
/* grand_product / fold_layers (lib.rs:785-802, 874-894) over dense leaves:
 * builds every layer on device; output[2^log_groups] = per-tree products
 * (tree c = leaves c, c + 2^log_groups, ...). */
int32_t bitz_gkr_build(bitz_gkr** h, const bitz_f128* leaves, uint32_t log_leaves,
                       uint32_t log_groups, bitz_f128* output, uint32_t flags);
/* grand_product over a BitTable (crates/gkr/src/lib.rs:785-802): 2^log_columns
 * trees of 2^log_rows leaves; leaf (b << s) | c (s = log_columns) is images[b]
 * where the table's bit (column c, row b) is set and one elsewhere, bit (c, b)
 * being bit (c << log_rows) | b of `packed` (BitTable::bit, crates/common/src/
 * table.rs; u64 words, least significant bit first).  The reference keeps
 * layers 0-2 as 16-entry lookup tables (LutLayers, lib.rs:274-302) and proves
 * their first rounds from them (prove_lut_layer, lib.rs:397-602), with the
 * same messages as the dense tree over these leaves (its tests build the
 * leaves so, lib.rs:1023-1037, 1173-1226); here the device expands the table
 * into the dense leaves and the dense tree follows, so bitz_gkr_prove and
 * bitz_gkr_prove_with_challenges take the handle as they take
 * bitz_gkr_build's.  log_columns < 4 -> BITZ_ERR_ARG (the reference panics).
 * images: 2^log_rows; packed: 2^(log_rows + log_columns) bits; output: the
 * 2^log_columns tree products.  BITZ_DEVICE: all three on the device. */
int32_t bitz_gkr_build_bittable(bitz_gkr** h, const bitz_f128* images, uint32_t log_rows, const uint64_t* packed,
                                uint32_t log_columns, bitz_f128* output, uint32_t flags);
/* gpgkr_prove (lib.rs:24-79) over the dense layers: point has log_groups
 * coordinates; leaf_point receives log_leaves coordinates and *leaf_claim
 * the claim on the leaves.  Consumes the layers (binds them in place).
 * security_bits is the SecurityLevel (100 = Bits100, 128 = Bits128; else
 * BITZ_ERR_ARG): every cubic round grinds grinding_bits(3) bits (0 at
 * Bits100, 2 at Bits128) under the label "gkr/cubic/v1" before its
 * challenge (lib.rs:144, 416), through bitz_transcript_grind.  With a Keccak
 * transcript and no grinding, every message and challenge is drawn on the
 * device (bitz_set_device_fiat_shamir). */
int32_t bitz_gkr_prove(bitz_gkr* h, bitz_transcript* t, const bitz_f128* point, uint32_t point_len,
                       uint32_t security_bits, bitz_f128* leaf_point, bitz_f128* leaf_claim);
/* Copies layer j (0 = leaves) to host; *len receives its length. */
/* gpgkr_prove with the challenges known in advance (a transcript whose
 * challenges do not depend on the messages, such as the seeded stream); no
 * counterpart in the reference, which draws them round by round.  All rounds
 * are queued back to back and read once.  challenges: per layer from the
 * top, its n round challenges in squeeze order, then t (sum over layers of
 * n + 1 values).  messages: the prover_message pairs [eq_bound sum_b,
 * eq_bound sum_inf] of every round in the same order; layer_evals: [L(r),
 * R(r)] per layer.  leaf_point / leaf_claim as bitz_gkr_prove.  Grinding is
 * the transcript's: the list holds the round challenges a transcript
 * squeezed after any grinding (at Bits100 there is none). */
int32_t bitz_gkr_prove_with_challenges(bitz_gkr* h, const bitz_f128* point, uint32_t point_len,
                                       const bitz_f128* challenges, size_t challenge_count, bitz_f128* messages,
                                       bitz_f128* layer_evals, bitz_f128* leaf_point, bitz_f128* leaf_claim);
int32_t bitz_gkr_layer(bitz_gkr* h, uint32_t j, bitz_f128* out, size_t cap, size_t* len);
void bitz_gkr_free(bitz_gkr* h);

/* ---- serving SHA-256 compression chains (no reference counterpart) -------- */
/* A proving server, native end to end: requests are admitted in arrival
 * order while a slot is free and their circuit's footprint fits the HBM
 * budget beside the requests in flight (a request larger than the budget
 * runs alone); preparer threads compute each admitted request's witness
 * (bitz_sha256_witness) and hand it to the least loaded of the server's PIOP
 * engines (bitz_piop_engine_submit_assignment, Keccak transcript with domain
 * "bitz/serve/<rid>"); harvester threads take finished proofs and free their
 * slot.  A request's result depends only on the request. */
typedef struct bitz_server bitz_server;
typedef struct {
    uint64_t rid;             /* names the transcript domain */
    uint32_t circuit;         /* bitz_server_add_circuit's id */
    uint32_t engine;          /* out: the engine that proved it */
    const uint8_t* message;   /* 64 bytes per block, read by a preparer */
    uint32_t iv[8];           /* the chaining value */
    double arrival_s;         /* seconds after the call starts */
    double admitted_s, done_s; /* out */
    uint8_t digest[32];       /* out: the chain's output state */
    int32_t status;           /* out */
    /* outputs as bitz_prove_spartan_piop's (row_vars x 4, 3, col_vars x 3,
     * col_vars, 1, 1 elements), each may be null */
    bitz_f128* outer_rounds;
    bitz_f128* outer_claims;
    bitz_f128* inner_rounds;
    bitz_f128* inner_point;
    bitz_f128* matrix_eval;
    bitz_f128* final_claim;
} bitz_serve_request;
int32_t bitz_server_create(bitz_server** s, uint32_t slots, uint32_t engines, uint32_t preparers,
                           uint64_t hbm_budget);
/* A compression-chain circuit with its prepared matrices (bitz_sha256_prepare)
 * and digest; footprint = device bytes a request holds at its peak.  The
 * circuit and matrices must outlive the server's use of them. */
int32_t bitz_server_add_circuit(bitz_server* s, const bitz_sha256* circuit, bitz_matrices* m,
                                const uint8_t digest[32], uint64_t footprint, uint32_t* id);
/* Serves every request and returns when all are done (or one failed).  The
 * requests share what the resident bytes leave of the budget. */
int32_t bitz_server_serve(bitz_server* s, bitz_serve_request* requests, size_t count);
/* Device bytes resident for the server's life: every added circuit's
 * matrices and, per engine, the bind buffers for it. */
int32_t bitz_server_resident_bytes(const bitz_server* s, uint64_t* bytes);
void bitz_server_free(bitz_server* s);

#ifdef __cplusplus
}
#endif
#endif
