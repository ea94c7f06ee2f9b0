//! rblas-hip: rblas's public BLAS functions, with rblas's own signatures,
//! computed on an AMD Instinct MI355X (gfx950) by librblas_hip through its
//! C ABI (include/rblas_hip.h).
//!
//! NOT COMPILED HERE: the image this crate was written in has no Rust
//! toolchain (no cargo / rustc). tests/test_rust_shim.py of the repository
//! checks the `extern "C"` block below against the C header, entry point by
//! entry point (names, argument count, order and mapped types), and the
//! wrappers' parameter lists against rblas's.
//!
//! Switching: replace `rblas::level3::gemm` by `rblas_hip::level3::gemm`
//! (and likewise for the other routed functions); the argument lists,
//! panics and `try_gemm`'s `BlasError` values are rblas's. Routed:
//!
//! | rblas function | entry points |
//! |---|---|
//! | `level3::{gemm, try_gemm, syrk, symm, trsm}` (level3.rs:32, 68, 236, 301, 390) | `rblas_{s,d}{gemm, gemm_try, syrk, symm, trsm}` |
//! | `complex3::{gemm, symm, hemm, trsm}` (complex3.rs:27, 130, 151, 252) | `rblas_{c,z}{gemm, symm, hemm, trsm}` |
//! | `level1::{axpy, dot, nrm2}` (level1.rs:17, 45, 96) | `rblas_{s,d}{axpy, dot, nrm2}` |
//! | `level2::{gemv, ger}` (level2.rs:11, 65) | `rblas_{s,d}{gemv, ger}` |
//! | `complex2::{gemv, geru, gerc}` (complex2.rs:11, 80, 97) | `rblas_{c,z}{gemv, geru, gerc}` |
//!
//! The other public functions (`level1::{scal, asum, copy, swap}`,
//! `complex1`) are re-exported from rblas unchanged (CPU).
//!
//! Status handling: 0 returns; 10 (`RBLAS_PANIC`, raised exactly where
//! rblas panics) panics with the library's message; 1 / 2 / 3 are
//! `BlasError::{InvalidLeadingDim, BufferTooSmall, ZeroStride}` for
//! `try_gemm`; 11 / 12 / 20 / 21 (an enum out of range, a device-path
//! limit, a HIP error, a failed allocation) panic with the status and
//! message, since rblas's signatures have no error channel.
#![allow(clippy::too_many_arguments)]

use core::ffi::{c_char, CStr};

pub use rblas::{BlasError, Complex, Diag, Float, Layout, Result, Side, Transpose, Uplo, C32, C64};

/// The C ABI of librblas_hip, declaration for declaration.
pub mod ffi {
    pub use core::ffi::{c_char, c_void};
    pub use rblas::{C32, C64};

    pub const OK: i32 = 0;
    pub const INVALID_LEADING_DIM: i32 = 1;
    pub const BUFFER_TOO_SMALL: i32 = 2;
    pub const ZERO_STRIDE: i32 = 3;
    pub const PANIC: i32 = 10;
    pub const BAD_ENUM: i32 = 11;
    pub const BAD_ARG: i32 = 12;
    pub const HIP_ERROR: i32 = 20;
    pub const ALLOC_FAILED: i32 = 21;

    /// `rblas_error_t`: a `BlasError` as a C struct (`which` is the
    /// operand's letter, `bound` is `min` or `need`).
    #[repr(C)]
    #[derive(Clone, Copy, Debug, Default)]
    pub struct RblasError {
        pub kind: i32,
        pub which: c_char,
        pub got: u64,
        pub bound: u64,
    }

    /// `rblas_grouped_t` (opaque plan handle).
    #[repr(C)]
    pub struct Grouped {
        _opaque: [u8; 0],
    }

    /// `rblas_sgemm_problem_t`: one gemm call's arguments (f32).
    #[repr(C)]
    #[derive(Clone, Copy, Debug)]
    pub struct SgemmProblem {
        pub layout: i32,
        pub transa: i32,
        pub transb: i32,
        pub m: usize,
        pub n: usize,
        pub k: usize,
        pub alpha: f32,
        pub a: *const f32,
        pub a_len: usize,
        pub lda: usize,
        pub b: *const f32,
        pub b_len: usize,
        pub ldb: usize,
        pub beta: f32,
        pub c: *mut f32,
        pub c_len: usize,
        pub ldc: usize,
    }

    /// `rblas_dgemm_problem_t`: one gemm call's arguments (f64).
    #[repr(C)]
    #[derive(Clone, Copy, Debug)]
    pub struct DgemmProblem {
        pub layout: i32,
        pub transa: i32,
        pub transb: i32,
        pub m: usize,
        pub n: usize,
        pub k: usize,
        pub alpha: f64,
        pub a: *const f64,
        pub a_len: usize,
        pub lda: usize,
        pub b: *const f64,
        pub b_len: usize,
        pub ldb: usize,
        pub beta: f64,
        pub c: *mut f64,
        pub c_len: usize,
        pub ldc: usize,
    }

    /// `rblas_grouped_info_t`.
    #[repr(C)]
    #[derive(Clone, Copy, Debug, Default)]
    pub struct GroupedInfo {
        pub problems: usize,
        pub units: usize,
        pub grid: i32,
        pub tiles64: i32,
        pub tiles128: i32,
        pub split_problems: i32,
        pub scale_problems: i32,
        pub pad: i32,
        pub flops: f64,
        pub est_us: f64,
        pub workspace_bytes: usize,
    }

    /// `rblas_strsm_problem_t`: one trsm call's arguments (f32).
    #[repr(C)]
    #[derive(Clone, Copy, Debug)]
    pub struct StrsmProblem {
        pub side: i32,
        pub uplo: i32,
        pub trans: i32,
        pub diag: i32,
        pub m: usize,
        pub n: usize,
        pub alpha: f32,
        pub a: *const f32,
        pub a_len: usize,
        pub lda: usize,
        pub b: *mut f32,
        pub b_len: usize,
        pub ldb: usize,
    }

    /// `rblas_dtrsm_problem_t`: one trsm call's arguments (f64).
    #[repr(C)]
    #[derive(Clone, Copy, Debug)]
    pub struct DtrsmProblem {
        pub side: i32,
        pub uplo: i32,
        pub trans: i32,
        pub diag: i32,
        pub m: usize,
        pub n: usize,
        pub alpha: f64,
        pub a: *const f64,
        pub a_len: usize,
        pub lda: usize,
        pub b: *mut f64,
        pub b_len: usize,
        pub ldb: usize,
    }

    /// `rblas_serve_request_t`.
    #[repr(C)]
    #[derive(Clone, Copy, Debug, Default)]
    pub struct ServeRequest {
        pub kind: i32,
        pub dtype: i32,
        pub m: i32,
        pub n: i32,
        pub k: i32,
        pub pad: i32,
        pub seed: u64,
        pub arrival_s: f64,
    }

    /// `rblas_serve_config_t`.
    #[repr(C)]
    #[derive(Clone, Copy, Debug)]
    pub struct ServeConfig {
        pub mode: i32,
        pub max_inflight: i32,
        pub max_batch: i32,
        pub capture_count: i32,
        pub hash_col_step: i32,
        pub pad: i32,
        pub hbm_cap_bytes: f64,
        pub capture_ids: *const i32,
        pub capture_out: *const *mut c_void,
    }

    /// `rblas_serve_result_t`.
    #[repr(C)]
    #[derive(Clone, Copy, Debug, Default)]
    pub struct ServeResult {
        pub arrival_s: f64,
        pub admit_s: f64,
        pub enqueued_s: f64,
        pub done_s: f64,
        pub flops: f64,
        pub footprint_bytes: f64,
        pub hash: u64,
        pub batch: i32,
        pub pad: i32,
    }

    /// `rblas_serve_stats_t`.
    #[repr(C)]
    #[derive(Clone, Copy, Debug, Default)]
    pub struct ServeStats {
        pub wall_s: f64,
        pub flops: f64,
        pub p50_s: f64,
        pub p99_s: f64,
        pub peak_footprint_bytes: f64,
        pub batches: i32,
        pub max_inflight_seen: i32,
    }

    #[link(name = "rblas_hip")]
    extern "C" {
        pub fn rblas_last_error_message() -> *const c_char;
        pub fn rblas_sgemm(layout: i32, transa: i32, transb: i32, m: usize, n: usize, k: usize,
            alpha: f32, a: *const f32, a_len: usize, lda: usize, b: *const f32, b_len: usize,
            ldb: usize, beta: f32, c: *mut f32, c_len: usize, ldc: usize) -> i32;
        pub fn rblas_dgemm(layout: i32, transa: i32, transb: i32, m: usize, n: usize, k: usize,
            alpha: f64, a: *const f64, a_len: usize, lda: usize, b: *const f64, b_len: usize,
            ldb: usize, beta: f64, c: *mut f64, c_len: usize, ldc: usize) -> i32;
        pub fn rblas_sgemm_try(layout: i32, transa: i32, transb: i32, m: usize, n: usize, k: usize,
            alpha: f32, a: *const f32, a_len: usize, lda: usize, b: *const f32, b_len: usize,
            ldb: usize, beta: f32, c: *mut f32, c_len: usize, ldc: usize,
            err: *mut RblasError) -> i32;
        pub fn rblas_dgemm_try(layout: i32, transa: i32, transb: i32, m: usize, n: usize, k: usize,
            alpha: f64, a: *const f64, a_len: usize, lda: usize, b: *const f64, b_len: usize,
            ldb: usize, beta: f64, c: *mut f64, c_len: usize, ldc: usize,
            err: *mut RblasError) -> i32;
        pub fn rblas_sgemm_dev(layout: i32, transa: i32, transb: i32, m: usize, n: usize, k: usize,
            alpha: f32, a: *const f32, a_len: usize, lda: usize, b: *const f32, b_len: usize,
            ldb: usize, beta: f32, c: *mut f32, c_len: usize, ldc: usize,
            stream: *mut c_void) -> i32;
        pub fn rblas_dgemm_dev(layout: i32, transa: i32, transb: i32, m: usize, n: usize, k: usize,
            alpha: f64, a: *const f64, a_len: usize, lda: usize, b: *const f64, b_len: usize,
            ldb: usize, beta: f64, c: *mut f64, c_len: usize, ldc: usize,
            stream: *mut c_void) -> i32;
        pub fn rblas_dgemm_dev_noalloc(layout: i32, transa: i32, transb: i32, m: usize, n: usize,
            k: usize, alpha: f64, a: *const f64, a_len: usize, lda: usize, b: *const f64,
            b_len: usize, ldb: usize, beta: f64, c: *mut f64, c_len: usize, ldc: usize,
            stream: *mut c_void) -> i32;
        pub fn rblas_sgemm_ex(layout: i32, transa: i32, transb: i32, m: usize, n: usize, k: usize,
            alpha: f32, a: *const f32, a_len: usize, lda: usize, b: *const f32, b_len: usize,
            ldb: usize, beta: f32, c: *mut f32, c_len: usize, ldc: usize, compute: i32) -> i32;
        pub fn rblas_sgemm_ex_dev(layout: i32, transa: i32, transb: i32, m: usize, n: usize,
            k: usize, alpha: f32, a: *const f32, a_len: usize, lda: usize, b: *const f32,
            b_len: usize, ldb: usize, beta: f32, c: *mut f32, c_len: usize, ldc: usize,
            compute: i32, stream: *mut c_void) -> i32;
        pub fn rblas_cgemm(transa: i32, transb: i32, m: usize, n: usize, k: usize, alpha: C32,
            a: *const C32, a_len: usize, lda: usize, b: *const C32, b_len: usize, ldb: usize,
            beta: C32, c: *mut C32, c_len: usize, ldc: usize) -> i32;
        pub fn rblas_zgemm(transa: i32, transb: i32, m: usize, n: usize, k: usize, alpha: C64,
            a: *const C64, a_len: usize, lda: usize, b: *const C64, b_len: usize, ldb: usize,
            beta: C64, c: *mut C64, c_len: usize, ldc: usize) -> i32;
        pub fn rblas_cgemm_dev(transa: i32, transb: i32, m: usize, n: usize, k: usize, alpha: C32,
            a: *const C32, a_len: usize, lda: usize, b: *const C32, b_len: usize, ldb: usize,
            beta: C32, c: *mut C32, c_len: usize, ldc: usize, stream: *mut c_void) -> i32;
        pub fn rblas_zgemm_dev(transa: i32, transb: i32, m: usize, n: usize, k: usize, alpha: C64,
            a: *const C64, a_len: usize, lda: usize, b: *const C64, b_len: usize, ldb: usize,
            beta: C64, c: *mut C64, c_len: usize, ldc: usize, stream: *mut c_void) -> i32;
        pub fn rblas_ssyrk(uplo: i32, trans: i32, n: usize, k: usize, alpha: f32, a: *const f32,
            a_len: usize, lda: usize, beta: f32, c: *mut f32, c_len: usize, ldc: usize) -> i32;
        pub fn rblas_dsyrk(uplo: i32, trans: i32, n: usize, k: usize, alpha: f64, a: *const f64,
            a_len: usize, lda: usize, beta: f64, c: *mut f64, c_len: usize, ldc: usize) -> i32;
        pub fn rblas_ssyrk_dev(uplo: i32, trans: i32, n: usize, k: usize, alpha: f32, a: *const f32,
            a_len: usize, lda: usize, beta: f32, c: *mut f32, c_len: usize, ldc: usize,
            stream: *mut c_void) -> i32;
        pub fn rblas_dsyrk_dev(uplo: i32, trans: i32, n: usize, k: usize, alpha: f64, a: *const f64,
            a_len: usize, lda: usize, beta: f64, c: *mut f64, c_len: usize, ldc: usize,
            stream: *mut c_void) -> i32;
        pub fn rblas_strsm(side: i32, uplo: i32, trans: i32, diag: i32, m: usize, n: usize,
            alpha: f32, a: *const f32, a_len: usize, lda: usize, b: *mut f32, b_len: usize,
            ldb: usize) -> i32;
        pub fn rblas_dtrsm(side: i32, uplo: i32, trans: i32, diag: i32, m: usize, n: usize,
            alpha: f64, a: *const f64, a_len: usize, lda: usize, b: *mut f64, b_len: usize,
            ldb: usize) -> i32;
        pub fn rblas_strsm_dev(side: i32, uplo: i32, trans: i32, diag: i32, m: usize, n: usize,
            alpha: f32, a: *const f32, a_len: usize, lda: usize, b: *mut f32, b_len: usize,
            ldb: usize, stream: *mut c_void) -> i32;
        pub fn rblas_dtrsm_dev(side: i32, uplo: i32, trans: i32, diag: i32, m: usize, n: usize,
            alpha: f64, a: *const f64, a_len: usize, lda: usize, b: *mut f64, b_len: usize,
            ldb: usize, stream: *mut c_void) -> i32;
        pub fn rblas_ctrsm(side: i32, uplo: i32, trans: i32, diag: i32, m: usize, n: usize,
            alpha: C32, a: *const C32, a_len: usize, lda: usize, b: *mut C32, b_len: usize,
            ldb: usize) -> i32;
        pub fn rblas_ztrsm(side: i32, uplo: i32, trans: i32, diag: i32, m: usize, n: usize,
            alpha: C64, a: *const C64, a_len: usize, lda: usize, b: *mut C64, b_len: usize,
            ldb: usize) -> i32;
        pub fn rblas_ctrsm_dev(side: i32, uplo: i32, trans: i32, diag: i32, m: usize, n: usize,
            alpha: C32, a: *const C32, a_len: usize, lda: usize, b: *mut C32, b_len: usize,
            ldb: usize, stream: *mut c_void) -> i32;
        pub fn rblas_ztrsm_dev(side: i32, uplo: i32, trans: i32, diag: i32, m: usize, n: usize,
            alpha: C64, a: *const C64, a_len: usize, lda: usize, b: *mut C64, b_len: usize,
            ldb: usize, stream: *mut c_void) -> i32;
        pub fn rblas_ssymm(side: i32, uplo: i32, m: usize, n: usize, alpha: f32, a: *const f32,
            a_len: usize, lda: usize, b: *const f32, b_len: usize, ldb: usize, beta: f32,
            c: *mut f32, c_len: usize, ldc: usize) -> i32;
        pub fn rblas_dsymm(side: i32, uplo: i32, m: usize, n: usize, alpha: f64, a: *const f64,
            a_len: usize, lda: usize, b: *const f64, b_len: usize, ldb: usize, beta: f64,
            c: *mut f64, c_len: usize, ldc: usize) -> i32;
        pub fn rblas_ssymm_dev(side: i32, uplo: i32, m: usize, n: usize, alpha: f32, a: *const f32,
            a_len: usize, lda: usize, b: *const f32, b_len: usize, ldb: usize, beta: f32,
            c: *mut f32, c_len: usize, ldc: usize, stream: *mut c_void) -> i32;
        pub fn rblas_dsymm_dev(side: i32, uplo: i32, m: usize, n: usize, alpha: f64, a: *const f64,
            a_len: usize, lda: usize, b: *const f64, b_len: usize, ldb: usize, beta: f64,
            c: *mut f64, c_len: usize, ldc: usize, stream: *mut c_void) -> i32;
        pub fn rblas_csymm(side: i32, uplo: i32, m: usize, n: usize, alpha: C32, a: *const C32,
            a_len: usize, lda: usize, b: *const C32, b_len: usize, ldb: usize, beta: C32,
            c: *mut C32, c_len: usize, ldc: usize) -> i32;
        pub fn rblas_zsymm(side: i32, uplo: i32, m: usize, n: usize, alpha: C64, a: *const C64,
            a_len: usize, lda: usize, b: *const C64, b_len: usize, ldb: usize, beta: C64,
            c: *mut C64, c_len: usize, ldc: usize) -> i32;
        pub fn rblas_chemm(side: i32, uplo: i32, m: usize, n: usize, alpha: C32, a: *const C32,
            a_len: usize, lda: usize, b: *const C32, b_len: usize, ldb: usize, beta: C32,
            c: *mut C32, c_len: usize, ldc: usize) -> i32;
        pub fn rblas_zhemm(side: i32, uplo: i32, m: usize, n: usize, alpha: C64, a: *const C64,
            a_len: usize, lda: usize, b: *const C64, b_len: usize, ldb: usize, beta: C64,
            c: *mut C64, c_len: usize, ldc: usize) -> i32;
        pub fn rblas_csymm_dev(side: i32, uplo: i32, m: usize, n: usize, alpha: C32, a: *const C32,
            a_len: usize, lda: usize, b: *const C32, b_len: usize, ldb: usize, beta: C32,
            c: *mut C32, c_len: usize, ldc: usize, stream: *mut c_void) -> i32;
        pub fn rblas_zsymm_dev(side: i32, uplo: i32, m: usize, n: usize, alpha: C64, a: *const C64,
            a_len: usize, lda: usize, b: *const C64, b_len: usize, ldb: usize, beta: C64,
            c: *mut C64, c_len: usize, ldc: usize, stream: *mut c_void) -> i32;
        pub fn rblas_chemm_dev(side: i32, uplo: i32, m: usize, n: usize, alpha: C32, a: *const C32,
            a_len: usize, lda: usize, b: *const C32, b_len: usize, ldb: usize, beta: C32,
            c: *mut C32, c_len: usize, ldc: usize, stream: *mut c_void) -> i32;
        pub fn rblas_zhemm_dev(side: i32, uplo: i32, m: usize, n: usize, alpha: C64, a: *const C64,
            a_len: usize, lda: usize, b: *const C64, b_len: usize, ldb: usize, beta: C64,
            c: *mut C64, c_len: usize, ldc: usize, stream: *mut c_void) -> i32;
        pub fn rblas_sdot(n: usize, x: *const f32, x_len: usize, incx: usize, y: *const f32,
            y_len: usize, incy: usize, result: *mut f32) -> i32;
        pub fn rblas_ddot(n: usize, x: *const f64, x_len: usize, incx: usize, y: *const f64,
            y_len: usize, incy: usize, result: *mut f64) -> i32;
        pub fn rblas_snrm2(n: usize, x: *const f32, x_len: usize, incx: usize,
            result: *mut f32) -> i32;
        pub fn rblas_dnrm2(n: usize, x: *const f64, x_len: usize, incx: usize,
            result: *mut f64) -> i32;
        pub fn rblas_saxpy(n: usize, alpha: f32, x: *const f32, x_len: usize, incx: usize,
            y: *mut f32, y_len: usize, incy: usize) -> i32;
        pub fn rblas_daxpy(n: usize, alpha: f64, x: *const f64, x_len: usize, incx: usize,
            y: *mut f64, y_len: usize, incy: usize) -> i32;
        pub fn rblas_sdot_dev(n: usize, x: *const f32, x_len: usize, incx: usize, y: *const f32,
            y_len: usize, incy: usize, result: *mut f32, stream: *mut c_void) -> i32;
        pub fn rblas_ddot_dev(n: usize, x: *const f64, x_len: usize, incx: usize, y: *const f64,
            y_len: usize, incy: usize, result: *mut f64, stream: *mut c_void) -> i32;
        pub fn rblas_snrm2_dev(n: usize, x: *const f32, x_len: usize, incx: usize, result: *mut f32,
            stream: *mut c_void) -> i32;
        pub fn rblas_dnrm2_dev(n: usize, x: *const f64, x_len: usize, incx: usize, result: *mut f64,
            stream: *mut c_void) -> i32;
        pub fn rblas_saxpy_dev(n: usize, alpha: f32, x: *const f32, x_len: usize, incx: usize,
            y: *mut f32, y_len: usize, incy: usize, stream: *mut c_void) -> i32;
        pub fn rblas_daxpy_dev(n: usize, alpha: f64, x: *const f64, x_len: usize, incx: usize,
            y: *mut f64, y_len: usize, incy: usize, stream: *mut c_void) -> i32;
        pub fn rblas_sgemv(layout: i32, trans: i32, m: usize, n: usize, alpha: f32, a: *const f32,
            a_len: usize, lda: usize, x: *const f32, x_len: usize, incx: usize, beta: f32,
            y: *mut f32, y_len: usize, incy: usize) -> i32;
        pub fn rblas_dgemv(layout: i32, trans: i32, m: usize, n: usize, alpha: f64, a: *const f64,
            a_len: usize, lda: usize, x: *const f64, x_len: usize, incx: usize, beta: f64,
            y: *mut f64, y_len: usize, incy: usize) -> i32;
        pub fn rblas_cgemv(layout: i32, trans: i32, m: usize, n: usize, alpha: C32, a: *const C32,
            a_len: usize, lda: usize, x: *const C32, x_len: usize, incx: usize, beta: C32,
            y: *mut C32, y_len: usize, incy: usize) -> i32;
        pub fn rblas_zgemv(layout: i32, trans: i32, m: usize, n: usize, alpha: C64, a: *const C64,
            a_len: usize, lda: usize, x: *const C64, x_len: usize, incx: usize, beta: C64,
            y: *mut C64, y_len: usize, incy: usize) -> i32;
        pub fn rblas_sger(layout: i32, m: usize, n: usize, alpha: f32, x: *const f32, x_len: usize,
            incx: usize, y: *const f32, y_len: usize, incy: usize, a: *mut f32, a_len: usize,
            lda: usize) -> i32;
        pub fn rblas_dger(layout: i32, m: usize, n: usize, alpha: f64, x: *const f64, x_len: usize,
            incx: usize, y: *const f64, y_len: usize, incy: usize, a: *mut f64, a_len: usize,
            lda: usize) -> i32;
        pub fn rblas_cgeru(layout: i32, m: usize, n: usize, alpha: C32, x: *const C32, x_len: usize,
            incx: usize, y: *const C32, y_len: usize, incy: usize, a: *mut C32, a_len: usize,
            lda: usize) -> i32;
        pub fn rblas_zgeru(layout: i32, m: usize, n: usize, alpha: C64, x: *const C64, x_len: usize,
            incx: usize, y: *const C64, y_len: usize, incy: usize, a: *mut C64, a_len: usize,
            lda: usize) -> i32;
        pub fn rblas_cgerc(layout: i32, m: usize, n: usize, alpha: C32, x: *const C32, x_len: usize,
            incx: usize, y: *const C32, y_len: usize, incy: usize, a: *mut C32, a_len: usize,
            lda: usize) -> i32;
        pub fn rblas_zgerc(layout: i32, m: usize, n: usize, alpha: C64, x: *const C64, x_len: usize,
            incx: usize, y: *const C64, y_len: usize, incy: usize, a: *mut C64, a_len: usize,
            lda: usize) -> i32;
        pub fn rblas_sgemv_dev(layout: i32, trans: i32, m: usize, n: usize, alpha: f32,
            a: *const f32, a_len: usize, lda: usize, x: *const f32, x_len: usize, incx: usize,
            beta: f32, y: *mut f32, y_len: usize, incy: usize, stream: *mut c_void) -> i32;
        pub fn rblas_dgemv_dev(layout: i32, trans: i32, m: usize, n: usize, alpha: f64,
            a: *const f64, a_len: usize, lda: usize, x: *const f64, x_len: usize, incx: usize,
            beta: f64, y: *mut f64, y_len: usize, incy: usize, stream: *mut c_void) -> i32;
        pub fn rblas_cgemv_dev(layout: i32, trans: i32, m: usize, n: usize, alpha: C32,
            a: *const C32, a_len: usize, lda: usize, x: *const C32, x_len: usize, incx: usize,
            beta: C32, y: *mut C32, y_len: usize, incy: usize, stream: *mut c_void) -> i32;
        pub fn rblas_zgemv_dev(layout: i32, trans: i32, m: usize, n: usize, alpha: C64,
            a: *const C64, a_len: usize, lda: usize, x: *const C64, x_len: usize, incx: usize,
            beta: C64, y: *mut C64, y_len: usize, incy: usize, stream: *mut c_void) -> i32;
        pub fn rblas_sger_dev(layout: i32, m: usize, n: usize, alpha: f32, x: *const f32,
            x_len: usize, incx: usize, y: *const f32, y_len: usize, incy: usize, a: *mut f32,
            a_len: usize, lda: usize, stream: *mut c_void) -> i32;
        pub fn rblas_dger_dev(layout: i32, m: usize, n: usize, alpha: f64, x: *const f64,
            x_len: usize, incx: usize, y: *const f64, y_len: usize, incy: usize, a: *mut f64,
            a_len: usize, lda: usize, stream: *mut c_void) -> i32;
        pub fn rblas_cgeru_dev(layout: i32, m: usize, n: usize, alpha: C32, x: *const C32,
            x_len: usize, incx: usize, y: *const C32, y_len: usize, incy: usize, a: *mut C32,
            a_len: usize, lda: usize, stream: *mut c_void) -> i32;
        pub fn rblas_zgeru_dev(layout: i32, m: usize, n: usize, alpha: C64, x: *const C64,
            x_len: usize, incx: usize, y: *const C64, y_len: usize, incy: usize, a: *mut C64,
            a_len: usize, lda: usize, stream: *mut c_void) -> i32;
        pub fn rblas_cgerc_dev(layout: i32, m: usize, n: usize, alpha: C32, x: *const C32,
            x_len: usize, incx: usize, y: *const C32, y_len: usize, incy: usize, a: *mut C32,
            a_len: usize, lda: usize, stream: *mut c_void) -> i32;
        pub fn rblas_zgerc_dev(layout: i32, m: usize, n: usize, alpha: C64, x: *const C64,
            x_len: usize, incx: usize, y: *const C64, y_len: usize, incy: usize, a: *mut C64,
            a_len: usize, lda: usize, stream: *mut c_void) -> i32;
        pub fn rblas_sgemm_grouped_create(plan: *mut *mut Grouped, problems: *const SgemmProblem,
            count: usize, bad_index: *mut usize) -> i32;
        pub fn rblas_dgemm_grouped_create(plan: *mut *mut Grouped, problems: *const DgemmProblem,
            count: usize, bad_index: *mut usize) -> i32;
        pub fn rblas_strsm_grouped_create(plan: *mut *mut Grouped, problems: *const StrsmProblem,
            count: usize, bad_index: *mut usize) -> i32;
        pub fn rblas_dtrsm_grouped_create(plan: *mut *mut Grouped, problems: *const DtrsmProblem,
            count: usize, bad_index: *mut usize) -> i32;
        pub fn rblas_grouped_run(plan: *mut Grouped, stream: *mut c_void) -> i32;
        pub fn rblas_grouped_run_graph(plan: *mut Grouped, stream: *mut c_void) -> i32;
        pub fn rblas_grouped_info(plan: *const Grouped, info: *mut GroupedInfo) -> i32;
        pub fn rblas_grouped_destroy(plan: *mut Grouped) -> i32;
        pub fn rblas_grouped_set_tuning(dtype: i32, coeffs: *const f64) -> i32;
        pub fn rblas_grouped_get_tuning(dtype: i32, coeffs: *mut f64) -> i32;
        pub fn rblas_serve_run(requests: *const ServeRequest, count: usize,
            config: *const ServeConfig, results: *mut ServeResult, stats: *mut ServeStats) -> i32;
        pub fn rblas_malloc(ptr: *mut *mut c_void, bytes: usize) -> i32;
        pub fn rblas_free(ptr: *mut c_void) -> i32;
        pub fn rblas_memcpy_htod(dst: *mut c_void, src: *const c_void, bytes: usize,
            stream: *mut c_void) -> i32;
        pub fn rblas_memcpy_dtoh(dst: *mut c_void, src: *const c_void, bytes: usize,
            stream: *mut c_void) -> i32;
        pub fn rblas_memcpy_dtod(dst: *mut c_void, src: *const c_void, bytes: usize,
            stream: *mut c_void) -> i32;
        pub fn rblas_stream_create(stream: *mut *mut c_void) -> i32;
        pub fn rblas_stream_destroy(stream: *mut c_void) -> i32;
        pub fn rblas_stream_synchronize(stream: *mut c_void) -> i32;
        pub fn rblas_event_create(event: *mut *mut c_void) -> i32;
        pub fn rblas_event_destroy(event: *mut c_void) -> i32;
        pub fn rblas_event_record(event: *mut c_void, stream: *mut c_void) -> i32;
        pub fn rblas_event_elapsed_ms(ms: *mut f32, start: *mut c_void, stop: *mut c_void) -> i32;
        pub fn rblas_device_synchronize() -> i32;
        pub fn rblas_device_count() -> i32;
    }
}

/// rblas's enums are not `#[repr(C)]`; the C ABI encodes each in rblas's
/// declaration order (types.rs:9-72).
mod enc {
    use rblas::{Diag, Layout, Side, Transpose, Uplo};

    pub fn layout(v: Layout) -> i32 {
        match v {
            Layout::RowMajor => 0,
            Layout::ColMajor => 1,
        }
    }
    pub fn trans(v: Transpose) -> i32 {
        match v {
            Transpose::None => 0,
            Transpose::Trans => 1,
            Transpose::ConjTrans => 2,
        }
    }
    pub fn uplo(v: Uplo) -> i32 {
        match v {
            Uplo::Upper => 0,
            Uplo::Lower => 1,
        }
    }
    pub fn side(v: Side) -> i32 {
        match v {
            Side::Left => 0,
            Side::Right => 1,
        }
    }
    pub fn diag(v: Diag) -> i32 {
        match v {
            Diag::NonUnit => 0,
            Diag::Unit => 1,
        }
    }
}

/// The library's text for the calling thread's last non-OK status.
pub fn last_error() -> String {
    // SAFETY: the library returns a NUL-terminated string it owns (never null).
    unsafe { CStr::from_ptr(ffi::rblas_last_error_message()) }
        .to_string_lossy()
        .into_owned()
}

/// The panicking contract of rblas's unchecked functions.
fn check(status: i32, what: &str) {
    match status {
        ffi::OK => {}
        ffi::PANIC => panic!("{}", last_error()),
        s => panic!("rblas_hip: {what}: status {s}: {}", last_error()),
    }
}

fn operand(which: c_char) -> &'static str {
    match which as u8 {
        b'a' => "a",
        b'b' => "b",
        b'c' => "c",
        b'x' => "x",
        b'y' => "y",
        _ => "?",
    }
}

/// `try_*` statuses 1 / 2 / 3 as rblas's `BlasError` (error.rs:15-38).
fn to_result(status: i32, e: &ffi::RblasError, what: &str) -> rblas::Result<()> {
    match status {
        ffi::OK => Ok(()),
        ffi::INVALID_LEADING_DIM => Err(BlasError::InvalidLeadingDim {
            which: operand(e.which),
            got: e.got as usize,
            min: e.bound as usize,
        }),
        ffi::BUFFER_TOO_SMALL => Err(BlasError::BufferTooSmall {
            which: operand(e.which),
            got: e.got as usize,
            need: e.bound as usize,
        }),
        ffi::ZERO_STRIDE => Err(BlasError::ZeroStride {
            which: operand(e.which),
        }),
        s => {
            check(s, what);
            Ok(())
        }
    }
}

/// The element types librblas_hip computes in: `f32` routes to the `s` (and,
/// for `Complex<f32>`, `c`) entry points, `f64` to `d` (`z`). Each method is
/// one C entry point with every slice passed as `(ptr, len)`.
pub trait HipFloat: Float {
    #[doc(hidden)]
    unsafe fn gemm(l: i32, ta: i32, tb: i32, m: usize, n: usize, k: usize, alpha: Self,
                   a: *const Self, a_len: usize, lda: usize, b: *const Self, b_len: usize,
                   ldb: usize, beta: Self, c: *mut Self, c_len: usize, ldc: usize) -> i32;
    #[doc(hidden)]
    unsafe fn gemm_try(l: i32, ta: i32, tb: i32, m: usize, n: usize, k: usize, alpha: Self,
                       a: *const Self, a_len: usize, lda: usize, b: *const Self, b_len: usize,
                       ldb: usize, beta: Self, c: *mut Self, c_len: usize, ldc: usize,
                       err: *mut ffi::RblasError) -> i32;
    #[doc(hidden)]
    unsafe fn syrk(uplo: i32, trans: i32, n: usize, k: usize, alpha: Self, a: *const Self,
                   a_len: usize, lda: usize, beta: Self, c: *mut Self, c_len: usize,
                   ldc: usize) -> i32;
    #[doc(hidden)]
    unsafe fn symm(side: i32, uplo: i32, m: usize, n: usize, alpha: Self, a: *const Self,
                   a_len: usize, lda: usize, b: *const Self, b_len: usize, ldb: usize,
                   beta: Self, c: *mut Self, c_len: usize, ldc: usize) -> i32;
    #[doc(hidden)]
    unsafe fn trsm(side: i32, uplo: i32, trans: i32, diag: i32, m: usize, n: usize, alpha: Self,
                   a: *const Self, a_len: usize, lda: usize, b: *mut Self, b_len: usize,
                   ldb: usize) -> i32;
    #[doc(hidden)]
    unsafe fn dot(n: usize, x: *const Self, x_len: usize, incx: usize, y: *const Self,
                  y_len: usize, incy: usize, result: *mut Self) -> i32;
    #[doc(hidden)]
    unsafe fn nrm2(n: usize, x: *const Self, x_len: usize, incx: usize, result: *mut Self) -> i32;
    #[doc(hidden)]
    unsafe fn axpy(n: usize, alpha: Self, x: *const Self, x_len: usize, incx: usize,
                   y: *mut Self, y_len: usize, incy: usize) -> i32;
    #[doc(hidden)]
    unsafe fn gemv(l: i32, trans: i32, m: usize, n: usize, alpha: Self, a: *const Self,
                   a_len: usize, lda: usize, x: *const Self, x_len: usize, incx: usize,
                   beta: Self, y: *mut Self, y_len: usize, incy: usize) -> i32;
    #[doc(hidden)]
    unsafe fn ger(l: i32, m: usize, n: usize, alpha: Self, x: *const Self, x_len: usize,
                  incx: usize, y: *const Self, y_len: usize, incy: usize, a: *mut Self,
                  a_len: usize, lda: usize) -> i32;
    #[doc(hidden)]
    unsafe fn cgemm(ta: i32, tb: i32, m: usize, n: usize, k: usize, alpha: Complex<Self>,
                    a: *const Complex<Self>, a_len: usize, lda: usize, b: *const Complex<Self>,
                    b_len: usize, ldb: usize, beta: Complex<Self>, c: *mut Complex<Self>,
                    c_len: usize, ldc: usize) -> i32;
    #[doc(hidden)]
    unsafe fn csymm(side: i32, uplo: i32, m: usize, n: usize, alpha: Complex<Self>,
                    a: *const Complex<Self>, a_len: usize, lda: usize, b: *const Complex<Self>,
                    b_len: usize, ldb: usize, beta: Complex<Self>, c: *mut Complex<Self>,
                    c_len: usize, ldc: usize) -> i32;
    #[doc(hidden)]
    unsafe fn chemm(side: i32, uplo: i32, m: usize, n: usize, alpha: Complex<Self>,
                    a: *const Complex<Self>, a_len: usize, lda: usize, b: *const Complex<Self>,
                    b_len: usize, ldb: usize, beta: Complex<Self>, c: *mut Complex<Self>,
                    c_len: usize, ldc: usize) -> i32;
    #[doc(hidden)]
    unsafe fn ctrsm(side: i32, uplo: i32, trans: i32, diag: i32, m: usize, n: usize,
                    alpha: Complex<Self>, a: *const Complex<Self>, a_len: usize, lda: usize,
                    b: *mut Complex<Self>, b_len: usize, ldb: usize) -> i32;
    #[doc(hidden)]
    unsafe fn cgemv(l: i32, trans: i32, m: usize, n: usize, alpha: Complex<Self>,
                    a: *const Complex<Self>, a_len: usize, lda: usize, x: *const Complex<Self>,
                    x_len: usize, incx: usize, beta: Complex<Self>, y: *mut Complex<Self>,
                    y_len: usize, incy: usize) -> i32;
    #[doc(hidden)]
    unsafe fn cgeru(l: i32, m: usize, n: usize, alpha: Complex<Self>, x: *const Complex<Self>,
                    x_len: usize, incx: usize, y: *const Complex<Self>, y_len: usize,
                    incy: usize, a: *mut Complex<Self>, a_len: usize, lda: usize) -> i32;
    #[doc(hidden)]
    unsafe fn cgerc(l: i32, m: usize, n: usize, alpha: Complex<Self>, x: *const Complex<Self>,
                    x_len: usize, incx: usize, y: *const Complex<Self>, y_len: usize,
                    incy: usize, a: *mut Complex<Self>, a_len: usize, lda: usize) -> i32;
}

macro_rules! impl_hip_float {
    ($t:ty, $gemm:ident, $gemm_try:ident, $syrk:ident, $symm:ident, $trsm:ident, $dot:ident,
     $nrm2:ident, $axpy:ident, $gemv:ident, $ger:ident, $cgemm:ident, $csymm:ident,
     $chemm:ident, $ctrsm:ident, $cgemv:ident, $cgeru:ident, $cgerc:ident) => {
        impl HipFloat for $t {
            unsafe fn gemm(l: i32, ta: i32, tb: i32, m: usize, n: usize, k: usize, alpha: $t,
                           a: *const $t, a_len: usize, lda: usize, b: *const $t, b_len: usize,
                           ldb: usize, beta: $t, c: *mut $t, c_len: usize, ldc: usize) -> i32 {
                ffi::$gemm(l, ta, tb, m, n, k, alpha, a, a_len, lda, b, b_len, ldb, beta, c,
                           c_len, ldc)
            }
            unsafe fn gemm_try(l: i32, ta: i32, tb: i32, m: usize, n: usize, k: usize,
                               alpha: $t, a: *const $t, a_len: usize, lda: usize, b: *const $t,
                               b_len: usize, ldb: usize, beta: $t, c: *mut $t, c_len: usize,
                               ldc: usize, err: *mut ffi::RblasError) -> i32 {
                ffi::$gemm_try(l, ta, tb, m, n, k, alpha, a, a_len, lda, b, b_len, ldb, beta, c,
                               c_len, ldc, err)
            }
            unsafe fn syrk(uplo: i32, trans: i32, n: usize, k: usize, alpha: $t, a: *const $t,
                           a_len: usize, lda: usize, beta: $t, c: *mut $t, c_len: usize,
                           ldc: usize) -> i32 {
                ffi::$syrk(uplo, trans, n, k, alpha, a, a_len, lda, beta, c, c_len, ldc)
            }
            unsafe fn symm(side: i32, uplo: i32, m: usize, n: usize, alpha: $t, a: *const $t,
                           a_len: usize, lda: usize, b: *const $t, b_len: usize, ldb: usize,
                           beta: $t, c: *mut $t, c_len: usize, ldc: usize) -> i32 {
                ffi::$symm(side, uplo, m, n, alpha, a, a_len, lda, b, b_len, ldb, beta, c, c_len,
                           ldc)
            }
            unsafe fn trsm(side: i32, uplo: i32, trans: i32, diag: i32, m: usize, n: usize,
                           alpha: $t, a: *const $t, a_len: usize, lda: usize, b: *mut $t,
                           b_len: usize, ldb: usize) -> i32 {
                ffi::$trsm(side, uplo, trans, diag, m, n, alpha, a, a_len, lda, b, b_len, ldb)
            }
            unsafe fn dot(n: usize, x: *const $t, x_len: usize, incx: usize, y: *const $t,
                          y_len: usize, incy: usize, result: *mut $t) -> i32 {
                ffi::$dot(n, x, x_len, incx, y, y_len, incy, result)
            }
            unsafe fn nrm2(n: usize, x: *const $t, x_len: usize, incx: usize,
                           result: *mut $t) -> i32 {
                ffi::$nrm2(n, x, x_len, incx, result)
            }
            unsafe fn axpy(n: usize, alpha: $t, x: *const $t, x_len: usize, incx: usize,
                           y: *mut $t, y_len: usize, incy: usize) -> i32 {
                ffi::$axpy(n, alpha, x, x_len, incx, y, y_len, incy)
            }
            unsafe fn gemv(l: i32, trans: i32, m: usize, n: usize, alpha: $t, a: *const $t,
                           a_len: usize, lda: usize, x: *const $t, x_len: usize, incx: usize,
                           beta: $t, y: *mut $t, y_len: usize, incy: usize) -> i32 {
                ffi::$gemv(l, trans, m, n, alpha, a, a_len, lda, x, x_len, incx, beta, y, y_len,
                           incy)
            }
            unsafe fn ger(l: i32, m: usize, n: usize, alpha: $t, x: *const $t, x_len: usize,
                          incx: usize, y: *const $t, y_len: usize, incy: usize, a: *mut $t,
                          a_len: usize, lda: usize) -> i32 {
                ffi::$ger(l, m, n, alpha, x, x_len, incx, y, y_len, incy, a, a_len, lda)
            }
            unsafe fn cgemm(ta: i32, tb: i32, m: usize, n: usize, k: usize,
                            alpha: Complex<$t>, a: *const Complex<$t>, a_len: usize,
                            lda: usize, b: *const Complex<$t>, b_len: usize, ldb: usize,
                            beta: Complex<$t>, c: *mut Complex<$t>, c_len: usize,
                            ldc: usize) -> i32 {
                ffi::$cgemm(ta, tb, m, n, k, alpha, a, a_len, lda, b, b_len, ldb, beta, c, c_len,
                            ldc)
            }
            unsafe fn csymm(side: i32, uplo: i32, m: usize, n: usize, alpha: Complex<$t>,
                            a: *const Complex<$t>, a_len: usize, lda: usize,
                            b: *const Complex<$t>, b_len: usize, ldb: usize, beta: Complex<$t>,
                            c: *mut Complex<$t>, c_len: usize, ldc: usize) -> i32 {
                ffi::$csymm(side, uplo, m, n, alpha, a, a_len, lda, b, b_len, ldb, beta, c,
                            c_len, ldc)
            }
            unsafe fn chemm(side: i32, uplo: i32, m: usize, n: usize, alpha: Complex<$t>,
                            a: *const Complex<$t>, a_len: usize, lda: usize,
                            b: *const Complex<$t>, b_len: usize, ldb: usize, beta: Complex<$t>,
                            c: *mut Complex<$t>, c_len: usize, ldc: usize) -> i32 {
                ffi::$chemm(side, uplo, m, n, alpha, a, a_len, lda, b, b_len, ldb, beta, c,
                            c_len, ldc)
            }
            unsafe fn ctrsm(side: i32, uplo: i32, trans: i32, diag: i32, m: usize, n: usize,
                            alpha: Complex<$t>, a: *const Complex<$t>, a_len: usize, lda: usize,
                            b: *mut Complex<$t>, b_len: usize, ldb: usize) -> i32 {
                ffi::$ctrsm(side, uplo, trans, diag, m, n, alpha, a, a_len, lda, b, b_len, ldb)
            }
            unsafe fn cgemv(l: i32, trans: i32, m: usize, n: usize, alpha: Complex<$t>,
                            a: *const Complex<$t>, a_len: usize, lda: usize,
                            x: *const Complex<$t>, x_len: usize, incx: usize,
                            beta: Complex<$t>, y: *mut Complex<$t>, y_len: usize,
                            incy: usize) -> i32 {
                ffi::$cgemv(l, trans, m, n, alpha, a, a_len, lda, x, x_len, incx, beta, y, y_len,
                            incy)
            }
            unsafe fn cgeru(l: i32, m: usize, n: usize, alpha: Complex<$t>,
                            x: *const Complex<$t>, x_len: usize, incx: usize,
                            y: *const Complex<$t>, y_len: usize, incy: usize,
                            a: *mut Complex<$t>, a_len: usize, lda: usize) -> i32 {
                ffi::$cgeru(l, m, n, alpha, x, x_len, incx, y, y_len, incy, a, a_len, lda)
            }
            unsafe fn cgerc(l: i32, m: usize, n: usize, alpha: Complex<$t>,
                            x: *const Complex<$t>, x_len: usize, incx: usize,
                            y: *const Complex<$t>, y_len: usize, incy: usize,
                            a: *mut Complex<$t>, a_len: usize, lda: usize) -> i32 {
                ffi::$cgerc(l, m, n, alpha, x, x_len, incx, y, y_len, incy, a, a_len, lda)
            }
        }
    };
}

impl_hip_float!(f32, rblas_sgemm, rblas_sgemm_try, rblas_ssyrk, rblas_ssymm, rblas_strsm,
                rblas_sdot, rblas_snrm2, rblas_saxpy, rblas_sgemv, rblas_sger, rblas_cgemm,
                rblas_csymm, rblas_chemm, rblas_ctrsm, rblas_cgemv, rblas_cgeru, rblas_cgerc);
impl_hip_float!(f64, rblas_dgemm, rblas_dgemm_try, rblas_dsyrk, rblas_dsymm, rblas_dtrsm,
                rblas_ddot, rblas_dnrm2, rblas_daxpy, rblas_dgemv, rblas_dger, rblas_zgemm,
                rblas_zsymm, rblas_zhemm, rblas_ztrsm, rblas_zgemv, rblas_zgeru, rblas_zgerc);

/// `rblas::level3` (level3.rs) on the GPU.
pub mod level3 {
    use super::{check, enc, ffi, to_result, HipFloat};
    use rblas::{Diag, Layout, Side, Transpose, Uplo};

    /// level3.rs:32-62: `C := alpha * op(A) * op(B) + beta * C`.
    pub fn gemm<T: HipFloat>(
        layout: Layout,
        transa: Transpose,
        transb: Transpose,
        m: usize,
        n: usize,
        k: usize,
        alpha: T,
        a: &[T],
        lda: usize,
        b: &[T],
        ldb: usize,
        beta: T,
        c: &mut [T],
        ldc: usize,
    ) {
        // SAFETY: every slice is passed with its length; the library reads
        // and writes inside [ptr, ptr + len) only.
        let st = unsafe {
            T::gemm(enc::layout(layout), enc::trans(transa), enc::trans(transb), m, n, k, alpha,
                    a.as_ptr(), a.len(), lda, b.as_ptr(), b.len(), ldb, beta, c.as_mut_ptr(),
                    c.len(), ldc)
        };
        check(st, "gemm");
    }

    /// level3.rs:68-103: the checked gemm (a, then b, then c).
    pub fn try_gemm<T: HipFloat>(
        layout: Layout,
        transa: Transpose,
        transb: Transpose,
        m: usize,
        n: usize,
        k: usize,
        alpha: T,
        a: &[T],
        lda: usize,
        b: &[T],
        ldb: usize,
        beta: T,
        c: &mut [T],
        ldc: usize,
    ) -> rblas::Result<()> {
        let mut err = ffi::RblasError::default();
        // SAFETY: as in `gemm`; `err` outlives the call.
        let st = unsafe {
            T::gemm_try(enc::layout(layout), enc::trans(transa), enc::trans(transb), m, n, k,
                        alpha, a.as_ptr(), a.len(), lda, b.as_ptr(), b.len(), ldb, beta,
                        c.as_mut_ptr(), c.len(), ldc, &mut err)
        };
        to_result(st, &err, "try_gemm")
    }

    /// level3.rs:236-294: `C := alpha * op(A) * op(A)^T + beta * C` on the
    /// `uplo` triangle.
    pub fn syrk<T: HipFloat>(
        uplo: Uplo,
        trans: Transpose,
        n: usize,
        k: usize,
        alpha: T,
        a: &[T],
        lda: usize,
        beta: T,
        c: &mut [T],
        ldc: usize,
    ) {
        // SAFETY: slices passed with their lengths.
        let st = unsafe {
            T::syrk(enc::uplo(uplo), enc::trans(trans), n, k, alpha, a.as_ptr(), a.len(), lda,
                    beta, c.as_mut_ptr(), c.len(), ldc)
        };
        check(st, "syrk");
    }

    /// level3.rs:301-381: symmetric A (its `uplo` triangle) times B.
    pub fn symm<T: HipFloat>(
        side: Side,
        uplo: Uplo,
        m: usize,
        n: usize,
        alpha: T,
        a: &[T],
        lda: usize,
        b: &[T],
        ldb: usize,
        beta: T,
        c: &mut [T],
        ldc: usize,
    ) {
        // SAFETY: slices passed with their lengths.
        let st = unsafe {
            T::symm(enc::side(side), enc::uplo(uplo), m, n, alpha, a.as_ptr(), a.len(), lda,
                    b.as_ptr(), b.len(), ldb, beta, c.as_mut_ptr(), c.len(), ldc)
        };
        check(st, "symm");
    }

    /// level3.rs:390-520: solves op(A) X = alpha B or X op(A) = alpha B in
    /// place of B.
    pub fn trsm<T: HipFloat>(
        side: Side,
        uplo: Uplo,
        trans: Transpose,
        diag: Diag,
        m: usize,
        n: usize,
        alpha: T,
        a: &[T],
        lda: usize,
        b: &mut [T],
        ldb: usize,
    ) {
        // SAFETY: slices passed with their lengths.
        let st = unsafe {
            T::trsm(enc::side(side), enc::uplo(uplo), enc::trans(trans), enc::diag(diag), m, n,
                    alpha, a.as_ptr(), a.len(), lda, b.as_mut_ptr(), b.len(), ldb)
        };
        check(st, "trsm");
    }
}

/// `rblas::complex3` (complex3.rs) on the GPU.
pub mod complex3 {
    use super::{check, enc, HipFloat};
    use rblas::{Complex, Diag, Side, Transpose, Uplo};

    /// complex3.rs:27-74: column-major complex gemm (4M on the device).
    pub fn gemm<T: HipFloat>(
        transa: Transpose,
        transb: Transpose,
        m: usize,
        n: usize,
        k: usize,
        alpha: Complex<T>,
        a: &[Complex<T>],
        lda: usize,
        b: &[Complex<T>],
        ldb: usize,
        beta: Complex<T>,
        c: &mut [Complex<T>],
        ldc: usize,
    ) {
        // SAFETY: slices passed with their lengths (in complex elements);
        // Complex<T> is #[repr(C)] {re, im}, the C struct's layout.
        let st = unsafe {
            T::cgemm(enc::trans(transa), enc::trans(transb), m, n, k, alpha, a.as_ptr(), a.len(),
                     lda, b.as_ptr(), b.len(), ldb, beta, c.as_mut_ptr(), c.len(), ldc)
        };
        check(st, "complex gemm");
    }

    /// complex3.rs:130-149: complex symmetric A times B.
    pub fn symm<T: HipFloat>(
        side: Side,
        uplo: Uplo,
        m: usize,
        n: usize,
        alpha: Complex<T>,
        a: &[Complex<T>],
        lda: usize,
        b: &[Complex<T>],
        ldb: usize,
        beta: Complex<T>,
        c: &mut [Complex<T>],
        ldc: usize,
    ) {
        // SAFETY: slices passed with their lengths.
        let st = unsafe {
            T::csymm(enc::side(side), enc::uplo(uplo), m, n, alpha, a.as_ptr(), a.len(), lda,
                     b.as_ptr(), b.len(), ldb, beta, c.as_mut_ptr(), c.len(), ldc)
        };
        check(st, "complex symm");
    }

    /// complex3.rs:151-167: Hermitian A times B.
    pub fn hemm<T: HipFloat>(
        side: Side,
        uplo: Uplo,
        m: usize,
        n: usize,
        alpha: Complex<T>,
        a: &[Complex<T>],
        lda: usize,
        b: &[Complex<T>],
        ldb: usize,
        beta: Complex<T>,
        c: &mut [Complex<T>],
        ldc: usize,
    ) {
        // SAFETY: slices passed with their lengths.
        let st = unsafe {
            T::chemm(enc::side(side), enc::uplo(uplo), m, n, alpha, a.as_ptr(), a.len(), lda,
                     b.as_ptr(), b.len(), ldb, beta, c.as_mut_ptr(), c.len(), ldc)
        };
        check(st, "hemm");
    }

    /// complex3.rs:252-349: complex triangular solve in place of B.
    pub fn trsm<T: HipFloat>(
        side: Side,
        uplo: Uplo,
        trans: Transpose,
        diag: Diag,
        m: usize,
        n: usize,
        alpha: Complex<T>,
        a: &[Complex<T>],
        lda: usize,
        b: &mut [Complex<T>],
        ldb: usize,
    ) {
        // SAFETY: slices passed with their lengths.
        let st = unsafe {
            T::ctrsm(enc::side(side), enc::uplo(uplo), enc::trans(trans), enc::diag(diag), m, n,
                     alpha, a.as_ptr(), a.len(), lda, b.as_mut_ptr(), b.len(), ldb)
        };
        check(st, "complex trsm");
    }
}

/// `rblas::level1` (level1.rs): axpy, dot and nrm2 on the GPU; the rest is
/// rblas's CPU code.
pub mod level1 {
    use super::{check, HipFloat};
    pub use rblas::level1::{asum, copy, scal, swap};

    /// level1.rs:17-41: `y := alpha * x + y`.
    pub fn axpy<T: HipFloat>(n: usize, alpha: T, x: &[T], incx: usize, y: &mut [T], incy: usize) {
        // SAFETY: slices passed with their lengths.
        let st = unsafe {
            T::axpy(n, alpha, x.as_ptr(), x.len(), incx, y.as_mut_ptr(), y.len(), incy)
        };
        check(st, "axpy");
    }

    /// level1.rs:45-66: `sum x_i y_i`.
    pub fn dot<T: HipFloat>(n: usize, x: &[T], incx: usize, y: &[T], incy: usize) -> T {
        let mut r = T::ZERO;
        // SAFETY: slices passed with their lengths; `r` outlives the call.
        let st = unsafe {
            T::dot(n, x.as_ptr(), x.len(), incx, y.as_ptr(), y.len(), incy, &mut r)
        };
        check(st, "dot");
        r
    }

    /// level1.rs:96-123: the Euclidean norm (NaN for a NaN entry or two
    /// infinite entries, +Inf for one, as rblas's scaled sum of squares).
    pub fn nrm2<T: HipFloat>(n: usize, x: &[T], incx: usize) -> T {
        let mut r = T::ZERO;
        // SAFETY: slice passed with its length; `r` outlives the call.
        let st = unsafe { T::nrm2(n, x.as_ptr(), x.len(), incx, &mut r) };
        check(st, "nrm2");
        r
    }
}

/// `rblas::level2` (level2.rs) on the GPU.
pub mod level2 {
    use super::{check, enc, HipFloat};
    use rblas::{Layout, Transpose};

    /// level2.rs:11-61: `y := alpha * op(A) * x + beta * y`.
    pub fn gemv<T: HipFloat>(
        layout: Layout,
        trans: Transpose,
        m: usize,
        n: usize,
        alpha: T,
        a: &[T],
        lda: usize,
        x: &[T],
        incx: usize,
        beta: T,
        y: &mut [T],
        incy: usize,
    ) {
        // SAFETY: slices passed with their lengths.
        let st = unsafe {
            T::gemv(enc::layout(layout), enc::trans(trans), m, n, alpha, a.as_ptr(), a.len(), lda,
                    x.as_ptr(), x.len(), incx, beta, y.as_mut_ptr(), y.len(), incy)
        };
        check(st, "gemv");
    }

    /// level2.rs:65-88: `A := alpha * x * y^T + A`.
    pub fn ger<T: HipFloat>(
        layout: Layout,
        m: usize,
        n: usize,
        alpha: T,
        x: &[T],
        incx: usize,
        y: &[T],
        incy: usize,
        a: &mut [T],
        lda: usize,
    ) {
        // SAFETY: slices passed with their lengths.
        let st = unsafe {
            T::ger(enc::layout(layout), m, n, alpha, x.as_ptr(), x.len(), incx, y.as_ptr(),
                   y.len(), incy, a.as_mut_ptr(), a.len(), lda)
        };
        check(st, "ger");
    }
}

/// `rblas::complex2` (complex2.rs) on the GPU.
pub mod complex2 {
    use super::{check, enc, HipFloat};
    use rblas::{Complex, Layout, Transpose};

    /// complex2.rs:11-79: complex gemv (ConjTrans conjugates A).
    pub fn gemv<T: HipFloat>(
        layout: Layout,
        trans: Transpose,
        m: usize,
        n: usize,
        alpha: Complex<T>,
        a: &[Complex<T>],
        lda: usize,
        x: &[Complex<T>],
        incx: usize,
        beta: Complex<T>,
        y: &mut [Complex<T>],
        incy: usize,
    ) {
        // SAFETY: slices passed with their lengths.
        let st = unsafe {
            T::cgemv(enc::layout(layout), enc::trans(trans), m, n, alpha, a.as_ptr(), a.len(),
                     lda, x.as_ptr(), x.len(), incx, beta, y.as_mut_ptr(), y.len(), incy)
        };
        check(st, "complex gemv");
    }

    /// complex2.rs:80-93: `A := alpha * x * y^T + A`.
    pub fn geru<T: HipFloat>(
        layout: Layout,
        m: usize,
        n: usize,
        alpha: Complex<T>,
        x: &[Complex<T>],
        incx: usize,
        y: &[Complex<T>],
        incy: usize,
        a: &mut [Complex<T>],
        lda: usize,
    ) {
        // SAFETY: slices passed with their lengths.
        let st = unsafe {
            T::cgeru(enc::layout(layout), m, n, alpha, x.as_ptr(), x.len(), incx, y.as_ptr(),
                     y.len(), incy, a.as_mut_ptr(), a.len(), lda)
        };
        check(st, "geru");
    }

    /// complex2.rs:97-110: `A := alpha * x * y^H + A`.
    pub fn gerc<T: HipFloat>(
        layout: Layout,
        m: usize,
        n: usize,
        alpha: Complex<T>,
        x: &[Complex<T>],
        incx: usize,
        y: &[Complex<T>],
        incy: usize,
        a: &mut [Complex<T>],
        lda: usize,
    ) {
        // SAFETY: slices passed with their lengths.
        let st = unsafe {
            T::cgerc(enc::layout(layout), m, n, alpha, x.as_ptr(), x.len(), incx, y.as_ptr(),
                     y.len(), incy, a.as_mut_ptr(), a.len(), lda)
        };
        check(st, "gerc");
    }
}

/// `rblas::complex1` is not routed: re-exported unchanged (CPU).
pub use rblas::complex1;
