//! Raw declarations of include/bitz_hip.h (the subset the prover wrappers
//! call).  Every function returns a status code (0 = ok, negative = the
//! error variants of the header's enum) unless noted.

use std::ffi::{c_char, c_void};

/// `bitz_f128`: the bytes of `field::F128` (lo, then hi; 16 bytes).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq, Eq)]
pub struct BitzF128 {
    pub lo: u64,
    pub hi: u64,
}

pub const BITZ_OK: i32 = 0;
pub const BITZ_ERR_PRODUCT_DIMS: i32 = -1;
pub const BITZ_ERR_EQUALITY_DIMS: i32 = -2;
pub const BITZ_ERR_WITNESS_LENGTH: i32 = -13;
pub const BITZ_ERR_CLAIM: i32 = -14;
/// Element arrays are device pointers (from `bitz_device_alloc`).
pub const BITZ_DEVICE: u32 = 1;

#[repr(C)]
pub struct Outer {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct Inner {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct Matrices {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct Gkr {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct Transcript {
    _opaque: [u8; 0],
}

pub type AbsorbFn = extern "C" fn(ctx: *mut c_void, kind: i32, bytes: *const u8, len: usize);
pub type SqueezeFn = extern "C" fn(ctx: *mut c_void, out: *mut u8);
pub type GrindFn = extern "C" fn(ctx: *mut c_void, label: *const u8, len: usize, bits: u32);

unsafe extern "C" {
    pub fn bitz_init(device: i32) -> i32;
    pub fn bitz_last_error() -> *const c_char;
    pub fn bitz_device_alloc(ptr: *mut *mut c_void, bytes: usize) -> i32;
    pub fn bitz_device_free(ptr: *mut c_void) -> i32;
    pub fn bitz_memcpy_h2d(dst: *mut c_void, src: *const c_void, bytes: usize) -> i32;
    pub fn bitz_memcpy_d2h(dst: *mut c_void, src: *const c_void, bytes: usize) -> i32;

    // Transcripts: the caller's own, through callbacks (void results).
    pub fn bitz_transcript_callbacks(ctx: *mut c_void, absorb: AbsorbFn, squeeze: SqueezeFn) -> *mut Transcript;
    pub fn bitz_transcript_set_grind(t: *mut Transcript, grind: GrindFn) -> i32;
    pub fn bitz_transcript_free(t: *mut Transcript);

    // crates/poly/src/eq.rs:63-99
    pub fn bitz_eq_table(r: *const BitzF128, k: u32, out: *mut BitzF128, flags: u32) -> i32;

    // crates/spartan/src/sumcheck.rs:252-384, per round: round0 and fold_next
    // give [c0, c2, c3] (c1 is the caller's, from the claim); finish gives
    // [Az(r), Bz(r), Cz(r)].
    pub fn bitz_outer_begin(
        h: *mut *mut Outer,
        eq_low: *const BitzF128,
        low_vars: u32,
        eq_high: *const BitzF128,
        high_vars: u32,
        az: *const BitzF128,
        az_vars: u32,
        bz: *const BitzF128,
        bz_vars: u32,
        cz: *const BitzF128,
        cz_vars: u32,
        flags: u32,
    ) -> i32;
    pub fn bitz_outer_round0(h: *mut Outer, coeffs3: *mut BitzF128) -> i32;
    pub fn bitz_outer_fold_next(h: *mut Outer, r: BitzF128, coeffs3: *mut BitzF128) -> i32;
    pub fn bitz_outer_finish(h: *mut Outer, claims3: *mut BitzF128) -> i32;
    pub fn bitz_outer_free(h: *mut Outer);

    // sumcheck.rs:442-552, per round: [c0, c2]; finish: [D(r), h(r)].
    pub fn bitz_inner_begin(
        h: *mut *mut Inner,
        matrix: *const BitzF128,
        matrix_vars: u32,
        witness: *const BitzF128,
        witness_vars: u32,
        flags: u32,
    ) -> i32;
    pub fn bitz_inner_round0(h: *mut Inner, coeffs2: *mut BitzF128) -> i32;
    pub fn bitz_inner_fold_next(h: *mut Inner, r: BitzF128, coeffs2: *mut BitzF128) -> i32;
    pub fn bitz_inner_finish(h: *mut Inner, evals2: *mut BitzF128) -> i32;
    pub fn bitz_inner_free(h: *mut Inner);

    // crates/spartan/src/matrix.rs:429-468
    pub fn bitz_matrices_prepare(
        h: *mut *mut Matrices,
        rows: usize,
        row_vars: u32,
        col_vars: u32,
        row_starts: *const *const u64,
        columns: *const *const u32,
        values: *const *const BitzF128,
    ) -> i32;
    pub fn bitz_bind_and_batch(
        h: *mut Matrices,
        row_point: *const BitzF128,
        point_len: u32,
        rho: BitzF128,
        out: *mut BitzF128,
        flags: u32,
    ) -> i32;
    pub fn bitz_matrices_free(h: *mut Matrices);

    // crates/gkr/src/lib.rs:24-79, 785-802, 874-894
    pub fn bitz_gkr_build(
        h: *mut *mut Gkr,
        leaves: *const BitzF128,
        log_leaves: u32,
        log_groups: u32,
        output: *mut BitzF128,
        flags: u32,
    ) -> i32;
    pub fn bitz_gkr_build_bittable(
        h: *mut *mut Gkr,
        images: *const BitzF128,
        log_rows: u32,
        packed: *const u64,
        log_columns: u32,
        output: *mut BitzF128,
        flags: u32,
    ) -> i32;
    pub fn bitz_gkr_prove(
        h: *mut Gkr,
        t: *mut Transcript,
        point: *const BitzF128,
        point_len: u32,
        security_bits: u32,
        leaf_point: *mut BitzF128,
        leaf_claim: *mut BitzF128,
    ) -> i32;
    pub fn bitz_gkr_free(h: *mut Gkr);

    // crates/post_gkr/src/lib.rs:85-148
    pub fn bitz_post_gkr_prove(
        t: *mut Transcript,
        rows: *const BitzF128,
        log_rows: u32,
        columns: *const BitzF128,
        log_columns: u32,
        target: BitzF128,
        packed: *const BitzF128,
        packed_len: usize,
        security_bits: u32,
        flags: u32,
        point: *mut BitzF128,
        evaluation: *mut BitzF128,
    ) -> i32;
}
