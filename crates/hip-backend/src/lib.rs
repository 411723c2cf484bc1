//! BitZ's GF(2^128) sumchecks and GKR layers on an MI355X (gfx950).
//!
//! Drop-ins for the prover functions on the hot path, with the reference's
//! signatures, argument meaning and errors:
//! - [`prove_outer_sumcheck`] for `spartan::prove_outer_sumcheck`
//!   (crates/spartan/src/sumcheck.rs:252-384), and
//!   [`prove_outer_sumcheck_device`] over tables already in HBM;
//! - [`prove_inner_sumcheck`] for `spartan::prove_inner_sumcheck`
//!   (sumcheck.rs:442-552), and [`prove_inner_sumcheck_device`];
//! - [`HipConstraintMatrices::bind_and_batch`] for
//!   `PreparedConstraintMatrices::bind_and_batch` (crates/spartan/src/matrix.rs:429-468);
//! - [`grand_product`] and [`gpgkr_prove`] for `gkr::grand_product` and
//!   `gkr::gpgkr_prove` (crates/gkr/src/lib.rs:24-79, 785-802);
//! - [`post_gkr_prove`] for `post_gkr::prove` (crates/post_gkr/src/lib.rs:85-148).
//!
//! The transcript stays here, in `transcript::ProverState`: the sumchecks run
//! through the library's per-round handles and this crate completes every
//! round as `recover_full_round_polynomial_and_sample_next_challenge` does
//! (sumcheck.rs:760-829), so the transcript sees the reference's calls in the
//! reference's order.  GKR runs behind a callback transcript that forwards
//! prover messages, challenges and grinding to the same `ProverState`.
//!
//! `F128` is `#[repr(C, align(16))] { lo: u64, hi: u64 }`
//! (crates/field/src/gf128.rs:60-68), the byte layout of `bitz_f128`, so
//! tables pass to the library as they are.  Failures the reference has no
//! variant for (a device error, an allocation failure) panic with the
//! library's message.

mod ffi;

use std::ffi::{CStr, c_void};
use std::ptr;
use std::sync::OnceLock;

use common::{BitTable, LinearClaim, OpeningQuery, SecurityLevel};
use field::F128;
use poly::DenseMultilinearExtension;
use spartan::{
    InnerSumcheckOutput, OuterSumcheckOutput, OuterSumcheckProof, R1csProductMles, SpartanMatrixError,
    SumcheckError, SumcheckProof, SumcheckProverOutput,
};
use transcript::{ProverState, SqueezableTranscript};

pub use ffi::BitzF128;

#[inline]
fn to_c(x: F128) -> BitzF128 {
    BitzF128 { lo: x.lo, hi: x.hi }
}

#[inline]
fn from_c(x: BitzF128) -> F128 {
    F128::new(x.lo, x.hi)
}

#[inline]
fn c_ptr(xs: &[F128]) -> *const BitzF128 {
    xs.as_ptr().cast()
}

/// A failure of the library that the reference's error types do not name.
#[derive(Clone, Debug, PartialEq, Eq)]
pub struct HipError {
    pub code: i32,
    pub message: String,
}

fn last_error(code: i32) -> HipError {
    // SAFETY: bitz_last_error returns this thread's NUL-terminated message.
    let message = unsafe { CStr::from_ptr(ffi::bitz_last_error()) }
        .to_string_lossy()
        .into_owned();
    HipError { code, message }
}

fn check(rc: i32) -> Result<(), HipError> {
    if rc == ffi::BITZ_OK { Ok(()) } else { Err(last_error(rc)) }
}

#[track_caller]
fn expect_ok(rc: i32) {
    if let Err(e) = check(rc) {
        panic!("bitz_hip: {} (status {})", e.message, e.code);
    }
}

/// The status of a sumcheck call as the reference's `SumcheckError`
/// (sumcheck.rs:39-47); the dimension checks come before any transcript call.
fn sumcheck_status(rc: i32) -> Result<(), SumcheckError> {
    match rc {
        ffi::BITZ_OK => Ok(()),
        ffi::BITZ_ERR_PRODUCT_DIMS => Err(SumcheckError::InvalidProductDimensions),
        ffi::BITZ_ERR_EQUALITY_DIMS => Err(SumcheckError::InvalidEqualityDimensions),
        other => {
            expect_ok(other);
            unreachable!()
        }
    }
}

static INIT: OnceLock<i32> = OnceLock::new();

/// Selects device 0 for the library on first use (bitz_init).
pub fn init() {
    // SAFETY: plain call; the status is checked.
    let rc = *INIT.get_or_init(|| unsafe { ffi::bitz_init(0) });
    expect_ok(rc);
}

/// A table of field elements resident in HBM (bitz_device_alloc).
pub struct DeviceTable {
    ptr: *mut c_void,
    len: usize,
}

// SAFETY: device memory is process-wide; the table owns its allocation.
unsafe impl Send for DeviceTable {}
unsafe impl Sync for DeviceTable {}

impl DeviceTable {
    /// Copies `values` into a new device table.
    pub fn from_host(values: &[F128]) -> Self {
        init();
        let mut ptr = ptr::null_mut();
        // SAFETY: out-pointer to a local; sizes in bytes.
        expect_ok(unsafe { ffi::bitz_device_alloc(&mut ptr, 16 * values.len().max(1)) });
        if !values.is_empty() {
            // SAFETY: ptr holds 16 * len bytes; values is len elements of 16 bytes.
            expect_ok(unsafe { ffi::bitz_memcpy_h2d(ptr, values.as_ptr().cast(), 16 * values.len()) });
        }
        Self { ptr, len: values.len() }
    }

    /// The MLE's evaluations in HBM (the layout `evaluations` has on the host).
    pub fn from_mle(mle: &DenseMultilinearExtension<F128>) -> Self {
        Self::from_host(&mle.evaluations)
    }

    pub fn len(&self) -> usize {
        self.len
    }

    pub fn is_empty(&self) -> bool {
        self.len == 0
    }

    /// log2 of the length; the tables of a sumcheck are powers of two.
    pub fn num_vars(&self) -> u32 {
        assert!(self.len.is_power_of_two(), "table length {} is not a power of two", self.len);
        self.len.trailing_zeros()
    }

    pub fn to_host(&self) -> Vec<F128> {
        let mut out = F128::zeroed_vec(self.len);
        if self.len > 0 {
            // SAFETY: both sides hold 16 * len bytes.
            expect_ok(unsafe { ffi::bitz_memcpy_d2h(out.as_mut_ptr().cast(), self.ptr, 16 * self.len) });
        }
        out
    }

    fn as_c(&self) -> *const BitzF128 {
        self.ptr.cast()
    }
}

impl Drop for DeviceTable {
    fn drop(&mut self) {
        // SAFETY: ptr came from bitz_device_alloc and is freed once.
        unsafe { ffi::bitz_device_free(self.ptr) };
    }
}

/// `R1csProductMles` with the three tables in HBM, so a proof copies nothing
/// (a 2^24 proof from host tables moves 768 MiB over PCIe first).
pub struct DeviceR1csProducts {
    pub az: DeviceTable,
    pub bz: DeviceTable,
    pub cz: DeviceTable,
}

impl DeviceR1csProducts {
    pub fn from_host(products: &R1csProductMles<F128>) -> Self {
        Self {
            az: DeviceTable::from_mle(&products.az),
            bz: DeviceTable::from_mle(&products.bz),
            cz: DeviceTable::from_mle(&products.cz),
        }
    }
}

/// One round's completion, as recover_full_round_polynomial_and_sample_next_challenge
/// (sumcheck.rs:760-829): c1 = claim - c0 - (c0 + c2 + ... + cD), absorb the
/// coefficients as a public message, squeeze r, claim <- g(r) by Horner.
fn recover_round<const IN: usize, const OUT: usize>(
    transcript: &mut ProverState,
    claim: &mut F128,
    without_linear: [F128; IN],
    rounds: &mut Vec<[F128; OUT]>,
    points: &mut Vec<F128>,
) -> F128 {
    let zero = F128::new(0, 0);
    let mut coefficients = [zero; OUT];
    coefficients[0] = without_linear[0];
    coefficients[2..].copy_from_slice(&without_linear[1..]);
    let at_one_without_c1 = coefficients.iter().fold(zero, |sum, &c| sum + c);
    coefficients[1] = *claim - coefficients[0] - at_one_without_c1;
    transcript.public_message(&coefficients);
    let r = transcript.squeeze::<F128>();
    *claim = coefficients.iter().rev().fold(zero, |value, &c| value * r + c);
    rounds.push(coefficients);
    points.push(r);
    r
}

/// sumcheck.rs:258-269: the products span one domain, the eq factors split it;
/// checked before the device or the transcript is touched.
fn outer_dims(low_vars: usize, high_vars: usize, a: usize, b: usize, c: usize) -> Result<(), SumcheckError> {
    if b != a || c != a {
        return Err(SumcheckError::InvalidProductDimensions);
    }
    if low_vars.checked_add(high_vars) != Some(a) {
        return Err(SumcheckError::InvalidEqualityDimensions);
    }
    Ok(())
}

struct OuterHandle(*mut ffi::Outer);

impl Drop for OuterHandle {
    fn drop(&mut self) {
        // SAFETY: a handle from bitz_outer_begin, freed once.
        unsafe { ffi::bitz_outer_free(self.0) };
    }
}

fn outer_rounds(
    transcript: &mut ProverState,
    initial_claim: F128,
    h: &OuterHandle,
    num_vars: usize,
) -> OuterSumcheckOutput<F128> {
    let mut co = [BitzF128::default(); 3];
    // SAFETY: h is live; co holds the three coefficients the call writes.
    expect_ok(unsafe { ffi::bitz_outer_round0(h.0, co.as_mut_ptr()) });
    let mut claim = initial_claim;
    let mut round_polynomials = Vec::with_capacity(num_vars);
    let mut eval_points = Vec::with_capacity(num_vars);
    for _ in 0..num_vars {
        let r = recover_round::<3, 4>(transcript, &mut claim, co.map(from_c), &mut round_polynomials, &mut eval_points);
        // SAFETY: as above.
        expect_ok(unsafe { ffi::bitz_outer_fold_next(h.0, to_c(r), co.as_mut_ptr()) });
    }
    let mut claims = [BitzF128::default(); 3];
    // SAFETY: as above; three claims.
    expect_ok(unsafe { ffi::bitz_outer_finish(h.0, claims.as_mut_ptr()) });
    let [az_mle_claim, bz_mle_claim, cz_mle_claim] = claims.map(from_c);
    transcript.public_message(&[az_mle_claim, bz_mle_claim, cz_mle_claim]); // sumcheck.rs:372
    OuterSumcheckOutput {
        proof: OuterSumcheckProof {
            sumcheck: SumcheckProof { round_polynomials },
            az_mle_claim,
            bz_mle_claim,
            cz_mle_claim,
        },
        eval_points,
        final_claim: claim,
    }
}

/// `spartan::prove_outer_sumcheck` on the GPU (host tables, copied in).
pub fn prove_outer_sumcheck(
    transcript: &mut ProverState,
    initial_claim: F128,
    (eq_low, eq_high): (DenseMultilinearExtension<F128>, DenseMultilinearExtension<F128>),
    products: &R1csProductMles<F128>,
) -> Result<OuterSumcheckOutput<F128>, SumcheckError> {
    outer_dims(
        eq_low.num_vars(),
        eq_high.num_vars(),
        products.az.num_vars(),
        products.bz.num_vars(),
        products.cz.num_vars(),
    )?;
    init();
    let mut h = ptr::null_mut();
    // SAFETY: every pointer covers 2^vars elements of its MLE.
    let rc = unsafe {
        ffi::bitz_outer_begin(
            &mut h,
            c_ptr(&eq_low.evaluations),
            eq_low.num_vars() as u32,
            c_ptr(&eq_high.evaluations),
            eq_high.num_vars() as u32,
            c_ptr(&products.az.evaluations),
            products.az.num_vars() as u32,
            c_ptr(&products.bz.evaluations),
            products.bz.num_vars() as u32,
            c_ptr(&products.cz.evaluations),
            products.cz.num_vars() as u32,
            0,
        )
    };
    sumcheck_status(rc)?;
    let handle = OuterHandle(h);
    Ok(outer_rounds(transcript, initial_claim, &handle, products.az.num_vars()))
}

/// The same proof over eq factors and products already in HBM.
pub fn prove_outer_sumcheck_device(
    transcript: &mut ProverState,
    initial_claim: F128,
    (eq_low, eq_high): (&DeviceTable, &DeviceTable),
    products: &DeviceR1csProducts,
) -> Result<OuterSumcheckOutput<F128>, SumcheckError> {
    outer_dims(
        eq_low.num_vars() as usize,
        eq_high.num_vars() as usize,
        products.az.num_vars() as usize,
        products.bz.num_vars() as usize,
        products.cz.num_vars() as usize,
    )?;
    init();
    let mut h = ptr::null_mut();
    // SAFETY: device tables of the stated lengths.
    let rc = unsafe {
        ffi::bitz_outer_begin(
            &mut h,
            eq_low.as_c(),
            eq_low.num_vars(),
            eq_high.as_c(),
            eq_high.num_vars(),
            products.az.as_c(),
            products.az.num_vars(),
            products.bz.as_c(),
            products.bz.num_vars(),
            products.cz.as_c(),
            products.cz.num_vars(),
            ffi::BITZ_DEVICE,
        )
    };
    sumcheck_status(rc)?;
    let handle = OuterHandle(h);
    Ok(outer_rounds(transcript, initial_claim, &handle, products.az.num_vars() as usize))
}

struct InnerHandle(*mut ffi::Inner);

impl Drop for InnerHandle {
    fn drop(&mut self) {
        // SAFETY: a handle from bitz_inner_begin, freed once.
        unsafe { ffi::bitz_inner_free(self.0) };
    }
}

fn inner_rounds(
    transcript: &mut ProverState,
    initial_claim: F128,
    h: &InnerHandle,
    num_vars: usize,
) -> InnerSumcheckOutput<F128> {
    let mut co = [BitzF128::default(); 2];
    // SAFETY: h is live; co holds [c0, c2].
    expect_ok(unsafe { ffi::bitz_inner_round0(h.0, co.as_mut_ptr()) });
    let mut claim = initial_claim;
    let mut round_polynomials = Vec::with_capacity(num_vars);
    let mut eval_points = Vec::with_capacity(num_vars);
    for _ in 0..num_vars {
        let r = recover_round::<2, 3>(transcript, &mut claim, co.map(from_c), &mut round_polynomials, &mut eval_points);
        // SAFETY: as above.
        expect_ok(unsafe { ffi::bitz_inner_fold_next(h.0, to_c(r), co.as_mut_ptr()) });
    }
    let mut evals = [BitzF128::default(); 2];
    // SAFETY: as above; [D(r), h(r)].
    expect_ok(unsafe { ffi::bitz_inner_finish(h.0, evals.as_mut_ptr()) });
    let [batched_matrix_evaluation, witness_evaluation] = evals.map(from_c);
    InnerSumcheckOutput {
        sumcheck: SumcheckProverOutput {
            proof: SumcheckProof { round_polynomials },
            eval_points,
            final_claim: claim,
        },
        batched_matrix_evaluation,
        witness_evaluation,
    }
}

/// `spartan::prove_inner_sumcheck` on the GPU (host tables, copied in).
pub fn prove_inner_sumcheck(
    transcript: &mut ProverState,
    initial_claim: F128,
    batched_matrix_mle: DenseMultilinearExtension<F128>,
    witness_mle: &DenseMultilinearExtension<F128>,
) -> Result<InnerSumcheckOutput<F128>, SumcheckError> {
    if witness_mle.num_vars() != batched_matrix_mle.num_vars() {
        return Err(SumcheckError::InvalidProductDimensions); // sumcheck.rs:449-451
    }
    init();
    let mut h = ptr::null_mut();
    // SAFETY: both pointers cover 2^vars elements of their MLEs.
    let rc = unsafe {
        ffi::bitz_inner_begin(
            &mut h,
            c_ptr(&batched_matrix_mle.evaluations),
            batched_matrix_mle.num_vars() as u32,
            c_ptr(&witness_mle.evaluations),
            witness_mle.num_vars() as u32,
            0,
        )
    };
    sumcheck_status(rc)?;
    let handle = InnerHandle(h);
    Ok(inner_rounds(transcript, initial_claim, &handle, batched_matrix_mle.num_vars()))
}

/// The same proof over D and the witness already in HBM.
pub fn prove_inner_sumcheck_device(
    transcript: &mut ProverState,
    initial_claim: F128,
    batched_matrix: &DeviceTable,
    witness: &DeviceTable,
) -> Result<InnerSumcheckOutput<F128>, SumcheckError> {
    if witness.num_vars() != batched_matrix.num_vars() {
        return Err(SumcheckError::InvalidProductDimensions);
    }
    init();
    let mut h = ptr::null_mut();
    // SAFETY: device tables of the stated lengths.
    let rc = unsafe {
        ffi::bitz_inner_begin(
            &mut h,
            batched_matrix.as_c(),
            batched_matrix.num_vars(),
            witness.as_c(),
            witness.num_vars(),
            ffi::BITZ_DEVICE,
        )
    };
    sumcheck_status(rc)?;
    let handle = InnerHandle(h);
    Ok(inner_rounds(transcript, initial_claim, &handle, batched_matrix.num_vars() as usize))
}

/// A, B, C prepared on the device for repeated `bind_and_batch` calls
/// (the counterpart of `PreparedConstraintMatrices`, matrix.rs:54-64).
pub struct HipConstraintMatrices {
    h: *mut ffi::Matrices,
    num_row_vars: usize,
    num_column_vars: usize,
}

// SAFETY: the library serialises calls on one matrices handle (its lock).
unsafe impl Send for HipConstraintMatrices {}
unsafe impl Sync for HipConstraintMatrices {}

/// One matrix in compressed-row form: `row_starts` (rows + 1 offsets), and
/// per entry its column and value (matrix.rs:97-125).
pub struct CsrMatrix<'a> {
    pub row_starts: &'a [u64],
    pub columns: &'a [u32],
    pub values: &'a [F128],
}

impl HipConstraintMatrices {
    pub fn new(
        rows: usize,
        num_row_vars: usize,
        num_column_vars: usize,
        [a, b, c]: [CsrMatrix<'_>; 3],
    ) -> Result<Self, HipError> {
        init();
        for m in [&a, &b, &c] {
            assert_eq!(m.row_starts.len(), rows + 1, "one offset per row and one past the last");
            assert_eq!(m.columns.len(), m.values.len(), "a column per value");
        }
        let starts = [a.row_starts.as_ptr(), b.row_starts.as_ptr(), c.row_starts.as_ptr()];
        let columns = [a.columns.as_ptr(), b.columns.as_ptr(), c.columns.as_ptr()];
        let values = [c_ptr(a.values), c_ptr(b.values), c_ptr(c.values)];
        let mut h = ptr::null_mut();
        // SAFETY: three CSR matrices of `rows` rows, checked above.
        check(unsafe {
            ffi::bitz_matrices_prepare(
                &mut h,
                rows,
                num_row_vars as u32,
                num_column_vars as u32,
                starts.as_ptr(),
                columns.as_ptr(),
                values.as_ptr(),
            )
        })?;
        Ok(Self { h, num_row_vars, num_column_vars })
    }

    pub fn num_row_vars(&self) -> usize {
        self.num_row_vars
    }

    pub fn num_column_vars(&self) -> usize {
        self.num_column_vars
    }

    /// `bind_and_batch` (matrix.rs:429-468): D(j) = sum_i eq(i, r_x)
    /// (A + rho B + rho^2 C)[i, j].
    pub fn bind_and_batch(
        &self,
        row_point: &[F128],
        rho: F128,
    ) -> Result<DenseMultilinearExtension<F128>, SpartanMatrixError> {
        if row_point.len() != self.num_row_vars {
            return Err(SpartanMatrixError::InvalidRowPointLength {
                expected: self.num_row_vars,
                actual: row_point.len(),
            });
        }
        let mut out = F128::zeroed_vec(1 << self.num_column_vars);
        // SAFETY: out holds 2^num_column_vars elements; the point is num_row_vars long.
        expect_ok(unsafe {
            ffi::bitz_bind_and_batch(
                self.h,
                c_ptr(row_point),
                row_point.len() as u32,
                to_c(rho),
                out.as_mut_ptr().cast(),
                0,
            )
        });
        Ok(DenseMultilinearExtension::from_evaluations(self.num_column_vars, out)
            .expect("2^num_column_vars evaluations"))
    }
}

impl Drop for HipConstraintMatrices {
    fn drop(&mut self) {
        // SAFETY: a handle from bitz_matrices_prepare, freed once.
        unsafe { ffi::bitz_matrices_free(self.h) };
    }
}

/// The product tree in HBM (the counterpart of `gkr::LayerWitnesses`):
/// every layer below the output, consumed by [`gpgkr_prove`].
pub struct HipLayerWitnesses {
    h: *mut ffi::Gkr,
    log_leaves: usize,
    log_trees: usize,
}

// SAFETY: the handle owns its device buffers and stream.
unsafe impl Send for HipLayerWitnesses {}

impl Drop for HipLayerWitnesses {
    fn drop(&mut self) {
        // SAFETY: a handle from bitz_gkr_build*, freed once.
        unsafe { ffi::bitz_gkr_free(self.h) };
    }
}

/// `gkr::grand_product` (lib.rs:785-802): tree c multiplies row_images[b]
/// over the rows b whose bit (c, b) is set.  The device expands the table
/// into the dense leaves (the layers' messages equal the LUT form's).
pub fn grand_product(row_images: &[F128], bittable: &BitTable) -> (Vec<F128>, HipLayerWitnesses) {
    init();
    let shape = bittable.shape();
    assert_eq!(row_images.len(), shape.rows(), "one image per row of the table");
    let (log_rows, log_columns) = (shape.log_rows(), shape.log_columns());
    // Bit (c << log_rows) | b of the table, packed 64 to a word.
    let mut packed = vec![0u64; (1usize << (log_rows + log_columns)).div_ceil(64)];
    for c in 0..shape.columns() {
        for b in 0..shape.rows() {
            if bittable.bit(c, b) {
                let index = (c << log_rows) | b;
                packed[index / 64] |= 1u64 << (index % 64);
            }
        }
    }
    let mut output = F128::zeroed_vec(shape.columns());
    let mut h = ptr::null_mut();
    // SAFETY: 2^log_rows images, 2^(log_rows + log_columns) bits, 2^log_columns outputs.
    expect_ok(unsafe {
        ffi::bitz_gkr_build_bittable(
            &mut h,
            c_ptr(row_images),
            log_rows as u32,
            packed.as_ptr(),
            log_columns as u32,
            output.as_mut_ptr().cast(),
            0,
        )
    });
    (output, HipLayerWitnesses { h, log_leaves: log_rows + log_columns, log_trees: log_columns })
}

/// The dense product tree (fold_layers, lib.rs:874-894) over 2^log_leaves
/// leaves and 2^log_trees trees (leaf (b << log_trees) | c is in tree c).
pub fn grand_product_dense(leaves: &[F128], log_trees: usize) -> (Vec<F128>, HipLayerWitnesses) {
    init();
    assert!(leaves.len().is_power_of_two(), "2^log_leaves leaves");
    let log_leaves = leaves.len().trailing_zeros() as usize;
    let mut output = F128::zeroed_vec(1 << log_trees);
    let mut h = ptr::null_mut();
    // SAFETY: 2^log_leaves leaves, 2^log_trees outputs.
    expect_ok(unsafe {
        ffi::bitz_gkr_build(
            &mut h,
            c_ptr(leaves),
            log_leaves as u32,
            log_trees as u32,
            output.as_mut_ptr().cast(),
            0,
        )
    });
    (output, HipLayerWitnesses { h, log_leaves, log_trees })
}

// The callback transcript of gpgkr_prove and post_gkr_prove: every absorb is
// a prover message, [F128; 2] (gkr lib.rs:142, 154; post_gkr lib.rs:141,
// sumcheck.rs:129) or one F128 (post_gkr sumcheck.rs:175), every squeeze a
// verifier message, every grind ProverState::grind (gkr lib.rs:144, post_gkr
// lib.rs:142).  A panic here aborts the process.
extern "C" fn gkr_absorb(ctx: *mut c_void, kind: i32, bytes: *const u8, len: usize) {
    // SAFETY: ctx is the &mut ProverState the prover lent; bytes holds len bytes.
    let (ps, data) = unsafe { (&mut *ctx.cast::<ProverState>(), std::slice::from_raw_parts(bytes, len)) };
    assert!(kind == 1 && (len == 32 || len == 16), "the provers absorb F128 prover messages");
    let element = |i: usize| F128::from_bytes(data[16 * i..16 * i + 16].try_into().expect("16 bytes"));
    if len == 32 {
        ps.prover_message(&[element(0), element(1)]);
    } else {
        ps.prover_message(&element(0));
    }
}

extern "C" fn gkr_squeeze(ctx: *mut c_void, out: *mut u8) {
    // SAFETY: as above; out holds 16 bytes.
    let ps = unsafe { &mut *ctx.cast::<ProverState>() };
    let r: F128 = ps.verifier_message();
    unsafe { ptr::copy_nonoverlapping(r.to_bytes().as_ptr(), out, 16) };
}

extern "C" fn gkr_grind(ctx: *mut c_void, label: *const u8, len: usize, bits: u32) {
    // SAFETY: as above; label holds len bytes.
    let (ps, label) = unsafe { (&mut *ctx.cast::<ProverState>(), std::slice::from_raw_parts(label, len)) };
    ps.grind(label, bits);
}

/// `gkr::gpgkr_prove` (lib.rs:24-79) on the GPU: the rounds of every layer
/// from the output down to the leaves, the transcript calls (and the
/// grinding of `security`) on `ps` in the reference's order.  Returns the
/// leaf point and the claim on the leaves.
pub fn gpgkr_prove(
    ps: &mut ProverState,
    point: &[F128],
    witnesses: HipLayerWitnesses,
    security: SecurityLevel,
) -> (Vec<F128>, F128) {
    assert_eq!(
        point.len(),
        witnesses.log_trees,
        "the output has one value per tree, so the point has a coordinate per tree bit"
    );
    let ctx: *mut ProverState = ps;
    // SAFETY: the callbacks run only inside bitz_gkr_prove below, while ps is lent.
    let t = unsafe { ffi::bitz_transcript_callbacks(ctx.cast(), gkr_absorb, gkr_squeeze) };
    assert!(!t.is_null(), "bitz_transcript_callbacks failed");
    // SAFETY: t is a callback transcript.
    expect_ok(unsafe { ffi::bitz_transcript_set_grind(t, gkr_grind) });
    let mut leaf_point = vec![BitzF128::default(); witnesses.log_leaves];
    let mut claim = BitzF128::default();
    // SAFETY: point has log_trees coordinates; leaf_point log_leaves.
    let rc = unsafe {
        ffi::bitz_gkr_prove(
            witnesses.h,
            t,
            c_ptr(point),
            point.len() as u32,
            security.bits(),
            leaf_point.as_mut_ptr(),
            &mut claim,
        )
    };
    // SAFETY: created above, freed once.
    unsafe { ffi::bitz_transcript_free(t) };
    expect_ok(rc);
    (leaf_point.into_iter().map(from_c).collect(), from_c(claim))
}

/// Why `post_gkr::prove` cannot run (crates/post_gkr/src/lib.rs:50-56).
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum PostGkrProveError {
    WitnessLengthMismatch,
    ClaimDoesNotHold,
}

/// `post_gkr::prove` (crates/post_gkr/src/lib.rs:85-148) on the GPU: the
/// claim's inner product against the committed bits (`packed`, 128 bits an
/// element, column after column) reduced to `MLE[f](point) = target`, the
/// transcript calls and grinding on `transcript` in the reference's order.
pub fn post_gkr_prove(
    claim: &LinearClaim<F128>,
    packed: &[F128],
    security: SecurityLevel,
    transcript: &mut ProverState,
) -> Result<OpeningQuery, PostGkrProveError> {
    init();
    let (rows, columns) = (claim.row_weights(), claim.column_weights());
    let (log_rows, log_columns) = (rows.len().trailing_zeros(), columns.len().trailing_zeros());
    let ctx: *mut ProverState = transcript;
    // SAFETY: the callbacks run only inside bitz_post_gkr_prove below.
    let t = unsafe { ffi::bitz_transcript_callbacks(ctx.cast(), gkr_absorb, gkr_squeeze) };
    assert!(!t.is_null(), "bitz_transcript_callbacks failed");
    // SAFETY: t is a callback transcript.
    expect_ok(unsafe { ffi::bitz_transcript_set_grind(t, gkr_grind) });
    let mut point = vec![BitzF128::default(); (log_rows + log_columns) as usize];
    let mut evaluation = BitzF128::default();
    // SAFETY: 2^t rows, 2^s columns, packed_len elements; point has t + s slots.
    let rc = unsafe {
        ffi::bitz_post_gkr_prove(
            t,
            c_ptr(rows),
            log_rows,
            c_ptr(columns),
            log_columns,
            to_c(claim.target()),
            c_ptr(packed),
            packed.len(),
            security.bits(),
            0,
            point.as_mut_ptr(),
            &mut evaluation,
        )
    };
    // SAFETY: created above, freed once.
    unsafe { ffi::bitz_transcript_free(t) };
    match rc {
        ffi::BITZ_ERR_WITNESS_LENGTH => Err(PostGkrProveError::WitnessLengthMismatch),
        ffi::BITZ_ERR_CLAIM => Err(PostGkrProveError::ClaimDoesNotHold),
        other => {
            expect_ok(other);
            Ok(OpeningQuery::Mle {
                point: point.into_iter().map(from_c).collect(),
                target: from_c(evaluation),
            })
        }
    }
}

#[cfg(test)]
mod tests {
    //! GPU parity with the reference provers on the same transcript: run on an
    //! MI355X with `cargo test -p hip-backend -- --ignored`.
    use super::*;

    fn elements(seed: u64, n: usize) -> Vec<F128> {
        // splitmix64, F128::from([w0, w1]) (the stream of crates/gkr/src/lib.rs:986-999)
        let mut x = seed;
        let mut next = move || {
            x = x.wrapping_add(0x9E37_79B9_7F4A_7C15);
            let mut z = x;
            z = (z ^ (z >> 30)).wrapping_mul(0xBF58_476D_1CE4_E5B9);
            z = (z ^ (z >> 27)).wrapping_mul(0x94D0_49BB_1331_11EB);
            z ^ (z >> 31)
        };
        (0..n).map(|_| F128::new(next(), next())).collect()
    }

    fn mle(values: Vec<F128>) -> DenseMultilinearExtension<F128> {
        let n = values.len().trailing_zeros() as usize;
        DenseMultilinearExtension::from_evaluations(n, values).unwrap()
    }

    #[test]
    #[ignore = "needs an MI355X"]
    fn outer_sumcheck_matches_the_reference() {
        for n in [1usize, 6, 13] {
            let products = R1csProductMles {
                az: mle(elements(1 + n as u64, 1 << n)),
                bz: mle(elements(2 + n as u64, 1 << n)),
                cz: mle(elements(3 + n as u64, 1 << n)),
            };
            let tau = elements(4, n);
            let claim = F128::new(7, n as u64);
            let mut ours = transcript::build_prover(b"hip".as_slice(), b"outer".as_slice());
            let mut theirs = transcript::build_prover(b"hip".as_slice(), b"outer".as_slice());
            let got = prove_outer_sumcheck(&mut ours, claim, poly::make_equality_factors(&tau).unwrap(), &products)
                .unwrap();
            let want =
                spartan::prove_outer_sumcheck(&mut theirs, claim, poly::make_equality_factors(&tau).unwrap(), &products)
                    .unwrap();
            assert_eq!(got, want);
            assert_eq!(ours.squeeze::<F128>(), theirs.squeeze::<F128>());
            let dev = DeviceR1csProducts::from_host(&products);
            let (lo, hi) = poly::make_equality_factors(&tau).unwrap();
            let (lo, hi) = (DeviceTable::from_mle(&lo), DeviceTable::from_mle(&hi));
            let mut again = transcript::build_prover(b"hip".as_slice(), b"outer".as_slice());
            assert_eq!(prove_outer_sumcheck_device(&mut again, claim, (&lo, &hi), &dev).unwrap(), want);
        }
    }

    #[test]
    #[ignore = "needs an MI355X"]
    fn inner_sumcheck_matches_the_reference() {
        for n in [0usize, 3, 12] {
            let d = mle(elements(10 + n as u64, 1 << n));
            let w = mle(elements(20 + n as u64, 1 << n));
            let claim = F128::new(3, 0);
            let mut ours = transcript::build_prover(b"hip".as_slice(), b"inner".as_slice());
            let mut theirs = transcript::build_prover(b"hip".as_slice(), b"inner".as_slice());
            let got = prove_inner_sumcheck(&mut ours, claim, d.clone(), &w).unwrap();
            let want = spartan::prove_inner_sumcheck(&mut theirs, claim, d, &w).unwrap();
            assert_eq!(got, want);
            assert_eq!(ours.finish().narg_string, theirs.finish().narg_string);
        }
    }

    #[test]
    fn dimension_errors_come_first() {
        let a = mle(elements(1, 8));
        let b = mle(elements(2, 4));
        let products = R1csProductMles { az: a.clone(), bz: b, cz: a };
        let mut ps = transcript::build_prover(b"hip".as_slice(), b"dims".as_slice());
        let tau = elements(3, 3);
        let err = prove_outer_sumcheck(&mut ps, F128::new(0, 0), poly::make_equality_factors(&tau).unwrap(), &products);
        assert_eq!(err, Err(SumcheckError::InvalidProductDimensions));
    }
}
