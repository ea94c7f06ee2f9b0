/*
 * rblas_hip.h — C ABI of the MI355X (gfx950) implementation of rblas's
 * Level-3 hot path.
 *
 * Every entry point below replaces one public rblas function (cited as
 * reference `file:line`), keeps its argument meaning and order, and adds the
 * two things a C boundary needs:
 *   - every Rust slice `&[T]` / `&mut [T]` becomes `(ptr, len)`; `len` is the
 *     slice length in elements, needed to reproduce rblas's bounds checks;
 *   - every call returns an `int32_t` status instead of panicking.
 *
 * Conventions
 *   - `usize` -> `size_t`; `f32`/`f64` -> `float`/`double`.
 *   - Enums are `int32_t` in rblas's declaration order (types.rs:9-72):
 *       Layout    { RowMajor = 0, ColMajor = 1 }
 *       Transpose { None = 0, Trans = 1, ConjTrans = 2 }
 *       Uplo      { Upper = 0, Lower = 1 }
 *       Side      { Left = 0, Right = 1 }
 *       Diag      { NonUnit = 0, Unit = 1 }
 *   - `Complex<T>` (complex.rs:17-24, `#[repr(C)] {re, im}`) is
 *     `rblas_complex32_t` / `rblas_complex64_t`, interleaved.
 *   - Host-slice entry points (no suffix) are synchronous like rblas: on
 *     return the output buffer (host memory) is final. They copy the touched
 *     extent of each operand to HBM, run the HIP kernels, and copy back.
 *   - `_dev` entry points take device pointers (HBM) and a `hipStream_t`
 *     passed as `void *` (NULL = the null stream). They only enqueue work on
 *     that stream; the caller synchronises.
 *   - Every entry point is thread-safe. Host-slice calls from different
 *     threads run on different streams and overlap on the device.
 */
#ifndef RBLAS_HIP_H
#define RBLAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- enums */
enum { RBLAS_ROW_MAJOR = 0, RBLAS_COL_MAJOR = 1 };
enum { RBLAS_NO_TRANS = 0, RBLAS_TRANS = 1, RBLAS_CONJ_TRANS = 2 };
enum { RBLAS_UPPER = 0, RBLAS_LOWER = 1 };
enum { RBLAS_LEFT = 0, RBLAS_RIGHT = 1 };
enum { RBLAS_NON_UNIT = 0, RBLAS_UNIT = 1 };

/* --------------------------------------------------------------- status
 * 1..3 are rblas's `BlasError` variants (error.rs:15-38); the rest are
 * conditions rblas signals by panicking or that only exist on a device. */
enum {
  RBLAS_OK = 0,
  RBLAS_INVALID_LEADING_DIM = 1, /* BlasError::InvalidLeadingDim {which, got, min} */
  RBLAS_BUFFER_TOO_SMALL = 2,    /* BlasError::BufferTooSmall {which, got, need}   */
  RBLAS_ZERO_STRIDE = 3,         /* BlasError::ZeroStride {which}                  */
  RBLAS_PANIC = 10,    /* a precondition whose violation panics in rblas          */
  RBLAS_BAD_ENUM = 11, /* an enum argument outside its declared range             */
  RBLAS_BAD_ARG = 12,  /* an argument of a device-only entry point is invalid     */
  RBLAS_HIP_ERROR = 20,   /* the HIP runtime reported an error                   */
  RBLAS_ALLOC_FAILED = 21 /* device memory allocation failed                      */
};

/* A `BlasError` (error.rs:15-38) as a C struct. `which` is the operand name
 * rblas uses ('a', 'b', 'c', 'x', 'y'); `bound` is `min` for
 * InvalidLeadingDim and `need` for BufferTooSmall. */
typedef struct rblas_error {
  int32_t kind;
  char which;
  uint64_t got;
  uint64_t bound;
} rblas_error_t;

typedef struct rblas_complex32 { float re, im; } rblas_complex32_t;
typedef struct rblas_complex64 { double re, im; } rblas_complex64_t;

/* Text of the last non-OK status on the calling thread (never NULL). */
const char *rblas_last_error_message(void);

/* ------------------------------------------------------------ level-3 gemm
 * C := alpha * op(A) * op(B) + beta * C.
 * Replaces `level3::gemm` (level3.rs:32-62). op(A) is m x k, op(B) is k x n,
 * C is m x n; lda/ldb/ldc are leading dimensions of the *stored* matrices.
 * Returns RBLAS_PANIC exactly where rblas's `gemm` would panic (the NN
 * asserts at level3.rs:164-166, the index checks of `materialize` at
 * level3.rs:179-222, the C slice in `apply_beta` at kernel/gemm.rs:327), and
 * also where rblas would write C out of bounds through a raw pointer.
 * RBLAS_BAD_ARG when C's columns overlap (after the RowMajor swap: ldc < m
 * with n > 1) and C is written, and for leading dimensions beyond the
 * device's 31-bit tile offsets (ld * 128 * sizeof(T) >= 2^31). */
int32_t rblas_sgemm(int32_t layout, int32_t transa, int32_t transb, size_t m,
                    size_t n, size_t k, float alpha, const float *a,
                    size_t a_len, size_t lda, const float *b, size_t b_len,
                    size_t ldb, float beta, float *c, size_t c_len, size_t ldc);
int32_t rblas_dgemm(int32_t layout, int32_t transa, int32_t transb, size_t m,
                    size_t n, size_t k, double alpha, const double *a,
                    size_t a_len, size_t lda, const double *b, size_t b_len,
                    size_t ldb, double beta, double *c, size_t c_len,
                    size_t ldc);

/* Checked gemm. Replaces `level3::try_gemm` (level3.rs:68-103): validates a,
 * then b, then c with `check_matrix` (level3.rs:107-140) and fills `*err`
 * (may be NULL) with the first failure; on success runs gemm. */
int32_t rblas_sgemm_try(int32_t layout, int32_t transa, int32_t transb,
                        size_t m, size_t n, size_t k, float alpha,
                        const float *a, size_t a_len, size_t lda,
                        const float *b, size_t b_len, size_t ldb, float beta,
                        float *c, size_t c_len, size_t ldc,
                        rblas_error_t *err);
int32_t rblas_dgemm_try(int32_t layout, int32_t transa, int32_t transb,
                        size_t m, size_t n, size_t k, double alpha,
                        const double *a, size_t a_len, size_t lda,
                        const double *b, size_t b_len, size_t ldb, double beta,
                        double *c, size_t c_len, size_t ldc,
                        rblas_error_t *err);

/* Device-pointer gemm (no rblas counterpart: the HBM-resident form of
 * `level3::gemm`). Same validation as the host form. */
int32_t rblas_sgemm_dev(int32_t layout, int32_t transa, int32_t transb,
                        size_t m, size_t n, size_t k, float alpha,
                        const float *a, size_t a_len, size_t lda,
                        const float *b, size_t b_len, size_t ldb, float beta,
                        float *c, size_t c_len, size_t ldc, void *stream);
int32_t rblas_dgemm_dev(int32_t layout, int32_t transa, int32_t transb,
                        size_t m, size_t n, size_t k, double alpha,
                        const double *a, size_t a_len, size_t lda,
                        const double *b, size_t b_len, size_t ldb,
                        double beta, double *c, size_t c_len, size_t ldc,
                        void *stream);

/* Device-pointer DGEMM that allocates nothing. No counterpart in rblas: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, validation, launch shape and result bits as rblas_dgemm_dev;
 * the k-split partial tiles and counters live in a fixed 64 MiB device array
 * that each GPU holds until the process ends, instead of the stream-ordered
 * pool. Its size covers every k-split shape the tile kernel picks for f64;
 * a call whose workspace would exceed it returns RBLAS_BAD_ARG. All calls on
 * one GPU share the array: calls that overlap on the device (on different
 * streams, or from different threads) race on it and must be serialized by
 * the caller, e.g. by issuing them on one stream. */
int32_t rblas_dgemm_dev_noalloc(int32_t layout, int32_t transa,
                                int32_t transb, size_t m, size_t n, size_t k,
                                double alpha, const double *a, size_t a_len,
                                size_t lda, const double *b, size_t b_len,
                                size_t ldb, double beta, double *c,
                                size_t c_len, size_t ldc, void *stream);

/* SGEMM with a selectable arithmetic (no rblas counterpart: rblas's SGEMM
 * is plain f32). `compute` is one of:
 *   RBLAS_COMPUTE_F32     plain f32 MFMA (what rblas_sgemm does)
 *   RBLAS_COMPUTE_BF16X3  each f32 split into 2 bf16 terms, 3 cross products
 *                         on bf16 MFMA with f32 accumulation
 *   RBLAS_COMPUTE_BF16X6  3 bf16 terms, 6 cross products (about f32 accuracy)
 * Arguments and validation otherwise as rblas_sgemm / rblas_sgemm_dev. The
 * split forms need a device workspace of 2 or 3 bf16 copies of op(A) and
 * op(B) (rounded up to 256 rows and 48 k, plus 32 k); the `_dev` form takes it from the
 * HIP stream-ordered pool of the current device. Non-finite inputs and
 * magnitudes above 3.38e38 (bf16's largest finite) are outside the split
 * forms' domain.
 * Accuracy: each BF16X3 product misses up to about 3 * 2^-16 of |a b|, a
 * representation error that does not shrink with k, while f32's
 * accumulation error grows with k. On N(0,1) operands the largest
 * |C - C_exact| / (|A||B|) over the whole C of a 512 x 512 product was
 * BF16X3 / F32 = 49x at k = 16, 15x at 64, 8.3x at 256, 3.5x at 1024, 2.3x
 * at 2048 and 2.7x at 4096; at k = 8192 (8192 x 8192, 48 sampled rows) it
 * was 1.25x and at k = 65536 (1024 x 1024) 1.45x. So BF16X3 stays within 2x
 * of F32's error only for k of about 8192 and more; below that use BF16X6,
 * whose error stayed within F32's at every k measured (0.36x - 0.90x). */
enum { RBLAS_COMPUTE_F32 = 0, RBLAS_COMPUTE_BF16X3 = 1, RBLAS_COMPUTE_BF16X6 = 2 };
int32_t rblas_sgemm_ex(int32_t layout, int32_t transa, int32_t transb, size_t m,
                       size_t n, size_t k, float alpha, const float *a,
                       size_t a_len, size_t lda, const float *b, size_t b_len,
                       size_t ldb, float beta, float *c, size_t c_len,
                       size_t ldc, int32_t compute);
int32_t rblas_sgemm_ex_dev(int32_t layout, int32_t transa, int32_t transb,
                           size_t m, size_t n, size_t k, float alpha,
                           const float *a, size_t a_len, size_t lda,
                           const float *b, size_t b_len, size_t ldb,
                           float beta, float *c, size_t c_len, size_t ldc,
                           int32_t compute, void *stream);

/* ------------------------------------------------------------ complex gemm
 * C := alpha * op(A) * op(B) + beta * C on interleaved complex matrices,
 * column-major (rblas's complex gemm has no Layout argument). Replaces
 * `complex3::gemm` (complex3.rs:27-74); op codes as Transpose, ConjTrans
 * conjugating. Lengths and leading dimensions count complex elements.
 * RBLAS_PANIC where rblas indexes past a slice (split_op's gathers of A and
 * B, the final update of C). Computed by the 4M decomposition on the device
 * (two real MFMA gemms with k concatenated, then the complex alpha/beta
 * update); beta == 0 does not read C, as in rblas. */
int32_t rblas_cgemm(int32_t transa, int32_t transb, size_t m, size_t n, size_t k,
                    rblas_complex32_t alpha, const rblas_complex32_t *a,
                    size_t a_len, size_t lda, const rblas_complex32_t *b,
                    size_t b_len, size_t ldb, rblas_complex32_t beta,
                    rblas_complex32_t *c, size_t c_len, size_t ldc);
int32_t rblas_zgemm(int32_t transa, int32_t transb, size_t m, size_t n, size_t k,
                    rblas_complex64_t alpha, const rblas_complex64_t *a,
                    size_t a_len, size_t lda, const rblas_complex64_t *b,
                    size_t b_len, size_t ldb, rblas_complex64_t beta,
                    rblas_complex64_t *c, size_t c_len, size_t ldc);
/* HBM-resident forms (no rblas counterpart); the 4M planes come from the
 * stream-ordered pool. */
int32_t rblas_cgemm_dev(int32_t transa, int32_t transb, size_t m, size_t n,
                        size_t k, rblas_complex32_t alpha,
                        const rblas_complex32_t *a, size_t a_len, size_t lda,
                        const rblas_complex32_t *b, size_t b_len, size_t ldb,
                        rblas_complex32_t beta, rblas_complex32_t *c,
                        size_t c_len, size_t ldc, void *stream);
int32_t rblas_zgemm_dev(int32_t transa, int32_t transb, size_t m, size_t n,
                        size_t k, rblas_complex64_t alpha,
                        const rblas_complex64_t *a, size_t a_len, size_t lda,
                        const rblas_complex64_t *b, size_t b_len, size_t ldb,
                        rblas_complex64_t beta, rblas_complex64_t *c,
                        size_t c_len, size_t ldc, void *stream);

/* -------------------------------------------------------------------- syrk
 * C := alpha * op(A) * op(A)^T + beta * C on the `uplo` triangle of the
 * symmetric n x n C, column-major; op(A) is n x k (stored n x k for
 * Transpose None, k x n otherwise). Replaces `level3::syrk`
 * (level3.rs:236-294). n == 0 returns at once. RBLAS_PANIC where rblas
 * indexes past a slice: the gathers of op(A) of its inner gemm, then the
 * triangle's largest C index (n-1)*(ldc+1). beta == 0 stores the triangle
 * without reading it. Only the tiles on or inside the triangle are computed.
 * RBLAS_BAD_ARG for ldc < n with n > 1 (overlapping C columns). */
int32_t rblas_ssyrk(int32_t uplo, int32_t trans, size_t n, size_t k, float alpha,
                    const float *a, size_t a_len, size_t lda, float beta,
                    float *c, size_t c_len, size_t ldc);
int32_t rblas_dsyrk(int32_t uplo, int32_t trans, size_t n, size_t k,
                    double alpha, const double *a, size_t a_len, size_t lda,
                    double beta, double *c, size_t c_len, size_t ldc);
/* HBM-resident forms (no rblas counterpart). */
int32_t rblas_ssyrk_dev(int32_t uplo, int32_t trans, size_t n, size_t k,
                        float alpha, const float *a, size_t a_len, size_t lda,
                        float beta, float *c, size_t c_len, size_t ldc,
                        void *stream);
int32_t rblas_dsyrk_dev(int32_t uplo, int32_t trans, size_t n, size_t k,
                        double alpha, const double *a, size_t a_len, size_t lda,
                        double beta, double *c, size_t c_len, size_t ldc,
                        void *stream);

/* -------------------------------------------------------------------- trsm
 * Solves op(A) X = alpha B (side Left, A m x m) or X op(A) = alpha B (Right,
 * A n x n) in place in B (m x n), column-major, A triangular (`uplo`, `diag`
 * Unit: the stored diagonal is not read for the solve). Replaces
 * `level3::trsm` (level3.rs:390-520). m == 0 or n == 0 returns at once.
 * RBLAS_PANIC where rblas indexes past a slice: B's largest index
 * (m-1) + (n-1)*ldb, A's diagonal (q-1)*(lda+1). No singularity check (a zero
 * diagonal gives Inf/NaN as in rblas). Blocked on the device: the 64 x 64
 * diagonal blocks of op(A) are inverted in FP64, the rest is MFMA gemm.
 * RBLAS_BAD_ARG for ldb < m with n > 1 (overlapping B columns). */
int32_t rblas_strsm(int32_t side, int32_t uplo, int32_t trans, int32_t diag,
                    size_t m, size_t n, float alpha, const float *a,
                    size_t a_len, size_t lda, float *b, size_t b_len,
                    size_t ldb);
int32_t rblas_dtrsm(int32_t side, int32_t uplo, int32_t trans, int32_t diag,
                    size_t m, size_t n, double alpha, const double *a,
                    size_t a_len, size_t lda, double *b, size_t b_len,
                    size_t ldb);
/* HBM-resident forms (no rblas counterpart); the workspace comes from the
 * stream-ordered pool. */
int32_t rblas_strsm_dev(int32_t side, int32_t uplo, int32_t trans, int32_t diag,
                        size_t m, size_t n, float alpha, const float *a,
                        size_t a_len, size_t lda, float *b, size_t b_len,
                        size_t ldb, void *stream);
int32_t rblas_dtrsm_dev(int32_t side, int32_t uplo, int32_t trans, int32_t diag,
                        size_t m, size_t n, double alpha, const double *a,
                        size_t a_len, size_t lda, double *b, size_t b_len,
                        size_t ldb, void *stream);

/* ------------------------------------------------------------ complex trsm
 * op(A) X = alpha B (side Left) or X op(A) = alpha B (Right), in place in B,
 * complex, column-major; op = None, Trans or ConjTrans (conjugating).
 * Replaces `complex3::trsm` (complex3.rs:252-349): the same checks as the
 * real trsm (nothing when m or n is 0; RBLAS_PANIC at B's largest index
 * (m-1) + (n-1)*ldb, then op(A)'s diagonal (q-1)*(lda+1)); RBLAS_BAD_ARG for
 * ldb < m with n > 1. alpha != 1 pre-scales B; NonUnit multiplies by
 * 1 / a_ii = conj(a_ii) / norm_sqr(a_ii), formed once per diagonal block,
 * where rblas's cdiv divides s conj(a_ii) by norm_sqr(a_ii) (the same
 * quotient up to rounding). Computed by recursive halving: real-view MFMA
 * gemm updates and 64-wide substitution leaves in f64.
 * Lengths and leading dimensions count complex elements. */
int32_t rblas_ctrsm(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m,
                    size_t n, rblas_complex32_t alpha, const rblas_complex32_t *a,
                    size_t a_len, size_t lda, rblas_complex32_t *b, size_t b_len, size_t ldb);
int32_t rblas_ztrsm(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m,
                    size_t n, rblas_complex64_t alpha, const rblas_complex64_t *a,
                    size_t a_len, size_t lda, rblas_complex64_t *b, size_t b_len, size_t ldb);
/* HBM-resident forms (no rblas counterpart). */
int32_t rblas_ctrsm_dev(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m,
                        size_t n, rblas_complex32_t alpha, const rblas_complex32_t *a,
                        size_t a_len, size_t lda, rblas_complex32_t *b, size_t b_len,
                        size_t ldb, void *stream);
int32_t rblas_ztrsm_dev(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m,
                        size_t n, rblas_complex64_t alpha, const rblas_complex64_t *a,
                        size_t a_len, size_t lda, rblas_complex64_t *b, size_t b_len,
                        size_t ldb, void *stream);

/* -------------------------------------------------------------- symm/hemm
 * C := alpha * A * B + beta * C (side Left, A m x m) or alpha * B * A +
 * beta * C (side Right, A n x n), B and C m x n, column-major, where A is
 * symmetric with only its `uplo` triangle stored (ld lda). Replaces
 * `level3::symm` (level3.rs:301-381): rblas builds the dense A from the
 * stored triangle, then calls gemm(ColMajor, N, N). RBLAS_PANIC where rblas
 * indexes past a slice: A's largest stored index (q-1)*(lda+1) first, then
 * that gemm's checks on B and C (including the `(k-1)` forms of the NN
 * path). Here the dense A is written in HBM by one launch and the product
 * runs on the MFMA gemm. */
int32_t rblas_ssymm(int32_t side, int32_t uplo, size_t m, size_t n, float alpha,
                    const float *a, size_t a_len, size_t lda, const float *b,
                    size_t b_len, size_t ldb, float beta, float *c, size_t c_len,
                    size_t ldc);
int32_t rblas_dsymm(int32_t side, int32_t uplo, size_t m, size_t n, double alpha,
                    const double *a, size_t a_len, size_t lda, const double *b,
                    size_t b_len, size_t ldb, double beta, double *c, size_t c_len,
                    size_t ldc);
/* HBM-resident forms (no rblas counterpart); the dense A and the gemm
 * workspace come from the stream-ordered pool. */
int32_t rblas_ssymm_dev(int32_t side, int32_t uplo, size_t m, size_t n, float alpha,
                        const float *a, size_t a_len, size_t lda, const float *b,
                        size_t b_len, size_t ldb, float beta, float *c, size_t c_len,
                        size_t ldc, void *stream);
int32_t rblas_dsymm_dev(int32_t side, int32_t uplo, size_t m, size_t n, double alpha,
                        const double *a, size_t a_len, size_t lda, const double *b,
                        size_t b_len, size_t ldb, double beta, double *c, size_t c_len,
                        size_t ldc, void *stream);
/* Complex symmetric (A = A^T) and Hermitian (A = A^H: the mirrored half is
 * conjugated and the diagonal's imaginary part is taken as zero) forms.
 * Replace `complex3::symm` / `complex3::hemm` (complex3.rs:127-247,
 * he_or_sy): nothing at all when m or n is 0; otherwise RBLAS_PANIC on A's
 * largest stored index, then complex3::gemm(N, N)'s checks on B and C. The
 * product runs on the 4M complex gemm. Lengths and leading dimensions count
 * complex elements. */
int32_t rblas_csymm(int32_t side, int32_t uplo, size_t m, size_t n, rblas_complex32_t alpha,
                    const rblas_complex32_t *a, size_t a_len, size_t lda,
                    const rblas_complex32_t *b, size_t b_len, size_t ldb,
                    rblas_complex32_t beta, rblas_complex32_t *c, size_t c_len, size_t ldc);
int32_t rblas_zsymm(int32_t side, int32_t uplo, size_t m, size_t n, rblas_complex64_t alpha,
                    const rblas_complex64_t *a, size_t a_len, size_t lda,
                    const rblas_complex64_t *b, size_t b_len, size_t ldb,
                    rblas_complex64_t beta, rblas_complex64_t *c, size_t c_len, size_t ldc);
int32_t rblas_chemm(int32_t side, int32_t uplo, size_t m, size_t n, rblas_complex32_t alpha,
                    const rblas_complex32_t *a, size_t a_len, size_t lda,
                    const rblas_complex32_t *b, size_t b_len, size_t ldb,
                    rblas_complex32_t beta, rblas_complex32_t *c, size_t c_len, size_t ldc);
int32_t rblas_zhemm(int32_t side, int32_t uplo, size_t m, size_t n, rblas_complex64_t alpha,
                    const rblas_complex64_t *a, size_t a_len, size_t lda,
                    const rblas_complex64_t *b, size_t b_len, size_t ldb,
                    rblas_complex64_t beta, rblas_complex64_t *c, size_t c_len, size_t ldc);
/* HBM-resident forms (no rblas counterpart). */
int32_t rblas_csymm_dev(int32_t side, int32_t uplo, size_t m, size_t n,
                        rblas_complex32_t alpha, const rblas_complex32_t *a, size_t a_len,
                        size_t lda, const rblas_complex32_t *b, size_t b_len, size_t ldb,
                        rblas_complex32_t beta, rblas_complex32_t *c, size_t c_len,
                        size_t ldc, void *stream);
int32_t rblas_zsymm_dev(int32_t side, int32_t uplo, size_t m, size_t n,
                        rblas_complex64_t alpha, const rblas_complex64_t *a, size_t a_len,
                        size_t lda, const rblas_complex64_t *b, size_t b_len, size_t ldb,
                        rblas_complex64_t beta, rblas_complex64_t *c, size_t c_len,
                        size_t ldc, void *stream);
int32_t rblas_chemm_dev(int32_t side, int32_t uplo, size_t m, size_t n,
                        rblas_complex32_t alpha, const rblas_complex32_t *a, size_t a_len,
                        size_t lda, const rblas_complex32_t *b, size_t b_len, size_t ldb,
                        rblas_complex32_t beta, rblas_complex32_t *c, size_t c_len,
                        size_t ldc, void *stream);
int32_t rblas_zhemm_dev(int32_t side, int32_t uplo, size_t m, size_t n,
                        rblas_complex64_t alpha, const rblas_complex64_t *a, size_t a_len,
                        size_t lda, const rblas_complex64_t *b, size_t b_len, size_t ldb,
                        rblas_complex64_t beta, rblas_complex64_t *c, size_t c_len,
                        size_t ldc, void *stream);

/* ----------------------------------------------------------------- level-1
 * Replace `level1::dot` (level1.rs:45), `level1::nrm2` (level1.rs:96) and
 * `level1::axpy` (level1.rs:17). Strides are rblas's `usize` increments.
 * RBLAS_PANIC where rblas panics: "slice shorter than n" on the contiguous
 * path (inc == 1 for every vector), "stride must be positive" and "x/y out of
 * bounds" on the strided path; axpy with n == 0 or alpha == 0 and dot/nrm2
 * with n == 0 return before any check, as in rblas. dot/nrm2 write the
 * result to `*result` (host memory; device memory for the `_dev` forms). */
int32_t rblas_sdot(size_t n, const float *x, size_t x_len, size_t incx,
                   const float *y, size_t y_len, size_t incy, float *result);
int32_t rblas_ddot(size_t n, const double *x, size_t x_len, size_t incx,
                   const double *y, size_t y_len, size_t incy, double *result);
int32_t rblas_snrm2(size_t n, const float *x, size_t x_len, size_t incx,
                    float *result);
int32_t rblas_dnrm2(size_t n, const double *x, size_t x_len, size_t incx,
                    double *result);
int32_t rblas_saxpy(size_t n, float alpha, const float *x, size_t x_len,
                    size_t incx, float *y, size_t y_len, size_t incy);
int32_t rblas_daxpy(size_t n, double alpha, const double *x, size_t x_len,
                    size_t incx, double *y, size_t y_len, size_t incy);
/* HBM-resident forms (no rblas counterpart), enqueued on `stream`. */
int32_t rblas_sdot_dev(size_t n, const float *x, size_t x_len, size_t incx,
                       const float *y, size_t y_len, size_t incy,
                       float *result, void *stream);
int32_t rblas_ddot_dev(size_t n, const double *x, size_t x_len, size_t incx,
                       const double *y, size_t y_len, size_t incy,
                       double *result, void *stream);
int32_t rblas_snrm2_dev(size_t n, const float *x, size_t x_len, size_t incx,
                        float *result, void *stream);
int32_t rblas_dnrm2_dev(size_t n, const double *x, size_t x_len, size_t incx,
                        double *result, void *stream);
int32_t rblas_saxpy_dev(size_t n, float alpha, const float *x, size_t x_len,
                        size_t incx, float *y, size_t y_len, size_t incy,
                        void *stream);
int32_t rblas_daxpy_dev(size_t n, double alpha, const double *x, size_t x_len,
                        size_t incx, double *y, size_t y_len, size_t incy,
                        void *stream);

/* ----------------------------------------------------------------- level-2
 * gemv: y := alpha op(A) x + beta y; ger: A := alpha x y^T + A. A is m x n
 * in `layout` with leading dimension lda ((rs, cs) = RowMajor (lda, 1),
 * ColMajor (1, lda), types.rs:20-25); x and y are strided by incx / incy.
 * rblas checks nothing up front: RBLAS_PANIC where its loops would index past
 * a slice (gemv: y over its length when beta != 1, real, or always, complex;
 * then, when alpha != 0, a and x for every output with a non-empty sum and
 * y; the real Trans loop reads x_i for every row. ger: nothing when
 * alpha == 0, else x_i for every row, then y_j and a_ij when n > 0).
 * RBLAS_BAD_ARG where the device path cannot follow rblas's sequential
 * order: gemv with incy == 0 and more than one output, ger with lda smaller
 * than the contiguous dimension (aliased elements), dimensions >= 2^31.
 * Real (level2.rs:11-104): beta == 1 leaves y, beta == 0 stores 0, then
 * y = fma(alpha, sum, y). Complex (complex2.rs:11-142): y = beta y always
 * (0 when beta == 0), then y + alpha sum with rblas's complex multiply;
 * ConjTrans conjugates A; gerc conjugates y. Sums run in a different order
 * from rblas's sequential loops (lanes, then splits). Lengths and strides
 * of complex operands count complex elements. */
int32_t rblas_sgemv(int32_t layout, int32_t trans, size_t m, size_t n, float alpha,
                    const float *a, size_t a_len, size_t lda, const float *x, size_t x_len,
                    size_t incx, float beta, float *y, size_t y_len, size_t incy);
int32_t rblas_dgemv(int32_t layout, int32_t trans, size_t m, size_t n, double alpha,
                    const double *a, size_t a_len, size_t lda, const double *x, size_t x_len,
                    size_t incx, double beta, double *y, size_t y_len, size_t incy);
int32_t rblas_cgemv(int32_t layout, int32_t trans, size_t m, size_t n, rblas_complex32_t alpha,
                    const rblas_complex32_t *a, size_t a_len, size_t lda,
                    const rblas_complex32_t *x, size_t x_len, size_t incx, rblas_complex32_t beta,
                    rblas_complex32_t *y, size_t y_len, size_t incy);
int32_t rblas_zgemv(int32_t layout, int32_t trans, size_t m, size_t n, rblas_complex64_t alpha,
                    const rblas_complex64_t *a, size_t a_len, size_t lda,
                    const rblas_complex64_t *x, size_t x_len, size_t incx, rblas_complex64_t beta,
                    rblas_complex64_t *y, size_t y_len, size_t incy);
/* level2::ger (level2.rs:65-88); complex2::geru / gerc (complex2.rs:80-142). */
int32_t rblas_sger(int32_t layout, size_t m, size_t n, float alpha, const float *x, size_t x_len,
                   size_t incx, const float *y, size_t y_len, size_t incy, float *a, size_t a_len,
                   size_t lda);
int32_t rblas_dger(int32_t layout, size_t m, size_t n, double alpha, const double *x,
                   size_t x_len, size_t incx, const double *y, size_t y_len, size_t incy,
                   double *a, size_t a_len, size_t lda);
int32_t rblas_cgeru(int32_t layout, size_t m, size_t n, rblas_complex32_t alpha,
                    const rblas_complex32_t *x, size_t x_len, size_t incx,
                    const rblas_complex32_t *y, size_t y_len, size_t incy, rblas_complex32_t *a,
                    size_t a_len, size_t lda);
int32_t rblas_zgeru(int32_t layout, size_t m, size_t n, rblas_complex64_t alpha,
                    const rblas_complex64_t *x, size_t x_len, size_t incx,
                    const rblas_complex64_t *y, size_t y_len, size_t incy, rblas_complex64_t *a,
                    size_t a_len, size_t lda);
int32_t rblas_cgerc(int32_t layout, size_t m, size_t n, rblas_complex32_t alpha,
                    const rblas_complex32_t *x, size_t x_len, size_t incx,
                    const rblas_complex32_t *y, size_t y_len, size_t incy, rblas_complex32_t *a,
                    size_t a_len, size_t lda);
int32_t rblas_zgerc(int32_t layout, size_t m, size_t n, rblas_complex64_t alpha,
                    const rblas_complex64_t *x, size_t x_len, size_t incx,
                    const rblas_complex64_t *y, size_t y_len, size_t incy, rblas_complex64_t *a,
                    size_t a_len, size_t lda);
/* HBM-resident forms (no rblas counterpart): the same checks, all pointers
 * device pointers, enqueued on `stream`. */
int32_t rblas_sgemv_dev(int32_t layout, int32_t trans, size_t m, size_t n, float alpha,
                        const float *a, size_t a_len, size_t lda, const float *x, size_t x_len,
                        size_t incx, float beta, float *y, size_t y_len, size_t incy,
                        void *stream);
int32_t rblas_dgemv_dev(int32_t layout, int32_t trans, size_t m, size_t n, double alpha,
                        const double *a, size_t a_len, size_t lda, const double *x, size_t x_len,
                        size_t incx, double beta, double *y, size_t y_len, size_t incy,
                        void *stream);
int32_t rblas_cgemv_dev(int32_t layout, int32_t trans, size_t m, size_t n,
                        rblas_complex32_t alpha, const rblas_complex32_t *a, size_t a_len,
                        size_t lda, const rblas_complex32_t *x, size_t x_len, size_t incx,
                        rblas_complex32_t beta, rblas_complex32_t *y, size_t y_len, size_t incy,
                        void *stream);
int32_t rblas_zgemv_dev(int32_t layout, int32_t trans, size_t m, size_t n,
                        rblas_complex64_t alpha, const rblas_complex64_t *a, size_t a_len,
                        size_t lda, const rblas_complex64_t *x, size_t x_len, size_t incx,
                        rblas_complex64_t beta, rblas_complex64_t *y, size_t y_len, size_t incy,
                        void *stream);
int32_t rblas_sger_dev(int32_t layout, size_t m, size_t n, float alpha, const float *x,
                       size_t x_len, size_t incx, const float *y, size_t y_len, size_t incy,
                       float *a, size_t a_len, size_t lda, void *stream);
int32_t rblas_dger_dev(int32_t layout, size_t m, size_t n, double alpha, const double *x,
                       size_t x_len, size_t incx, const double *y, size_t y_len, size_t incy,
                       double *a, size_t a_len, size_t lda, void *stream);
int32_t rblas_cgeru_dev(int32_t layout, size_t m, size_t n, rblas_complex32_t alpha,
                        const rblas_complex32_t *x, size_t x_len, size_t incx,
                        const rblas_complex32_t *y, size_t y_len, size_t incy,
                        rblas_complex32_t *a, size_t a_len, size_t lda, void *stream);
int32_t rblas_zgeru_dev(int32_t layout, size_t m, size_t n, rblas_complex64_t alpha,
                        const rblas_complex64_t *x, size_t x_len, size_t incx,
                        const rblas_complex64_t *y, size_t y_len, size_t incy,
                        rblas_complex64_t *a, size_t a_len, size_t lda, void *stream);
int32_t rblas_cgerc_dev(int32_t layout, size_t m, size_t n, rblas_complex32_t alpha,
                        const rblas_complex32_t *x, size_t x_len, size_t incx,
                        const rblas_complex32_t *y, size_t y_len, size_t incy,
                        rblas_complex32_t *a, size_t a_len, size_t lda, void *stream);
int32_t rblas_zgerc_dev(int32_t layout, size_t m, size_t n, rblas_complex64_t alpha,
                        const rblas_complex64_t *x, size_t x_len, size_t incx,
                        const rblas_complex64_t *y, size_t y_len, size_t incy,
                        rblas_complex64_t *a, size_t a_len, size_t lda, void *stream);

/* ------------------------------------------------------------ grouped gemm
 * No rblas counterpart: rblas runs one `level3::gemm` (level3.rs:32-62) per
 * call. A grouped plan holds many independent gemm problems, each with
 * exactly rblas's gemm arguments and meaning (any Layout / Transpose, alpha,
 * beta), on HBM-resident operands, and runs all of them in ONE persistent
 * launch. Creating a plan validates every problem as rblas_sgemm_dev would
 * (same RBLAS_PANIC / RBLAS_BAD_ENUM / RBLAS_BAD_ARG conditions; the first
 * failing problem's index goes to *bad_index and no plan is made), chooses
 * each problem's tile side and k splits from its shape alone (so a problem
 * gives the same bits in any group), orders the work units longest first and
 * uploads descriptors. Operand pointers are captured: a plan reruns on
 * whatever those buffers hold at run time. C blocks of different problems
 * must not overlap. One run of a plan at a time. */
typedef struct rblas_grouped rblas_grouped_t;

typedef struct {
  int32_t layout, transa, transb;
  size_t m, n, k;
  float alpha;
  const float *a;
  size_t a_len, lda;
  const float *b;
  size_t b_len, ldb;
  float beta;
  float *c;
  size_t c_len, ldc;
} rblas_sgemm_problem_t;

typedef struct {
  int32_t layout, transa, transb;
  size_t m, n, k;
  double alpha;
  const double *a;
  size_t a_len, lda;
  const double *b;
  size_t b_len, ldb;
  double beta;
  double *c;
  size_t c_len, ldc;
} rblas_dgemm_problem_t;

typedef struct {
  size_t problems;        /* problems in the plan                            */
  size_t units;           /* work units (C tile x k split) of the launch     */
  int32_t grid;           /* workgroups of the persistent launch            */
  int32_t tiles64;        /* 64 x 64 C tiles                                 */
  int32_t tiles128;       /* 128 x 128 C tiles                               */
  int32_t split_problems; /* problems whose k is split                      */
  int32_t scale_problems; /* problems that only scale C (k == 0, alpha == 0) */
  int32_t pad;
  double flops;           /* 2 m n k summed over product problems            */
  double est_us;          /* planner's cost-model estimate of one run        */
  size_t workspace_bytes; /* device memory held by the plan                 */
} rblas_grouped_info_t;

int32_t rblas_sgemm_grouped_create(rblas_grouped_t **plan,
                                   const rblas_sgemm_problem_t *problems,
                                   size_t count, size_t *bad_index);
int32_t rblas_dgemm_grouped_create(rblas_grouped_t **plan,
                                   const rblas_dgemm_problem_t *problems,
                                   size_t count, size_t *bad_index);
/* ------------------------------------------------------------ grouped trsm
 * No rblas counterpart: a plan of many independent `level3::trsm` problems
 * (level3.rs:390-520), each with rblas's arguments and meaning, on
 * HBM-resident operands. Creating it validates every problem as rblas_strsm
 * would (first failing index to *bad_index), and additionally requires the
 * order of A (m for Left, n for Right) to be at most 2048 (RBLAS_BAD_ARG
 * above; rblas_?trsm_dev has no such limit). A run is one batched launch
 * inverting every 64 x 64 diagonal block (f64), then one persistent launch
 * in which a workgroup solves one system for 16 (order <= 1024) or 8
 * right-hand sides held in LDS, in f64. B blocks of different problems must
 * not overlap. Run, replay, info and destroy as for grouped gemm plans; info
 * reports flops as the sum of m^2 n (Left) or m n^2 (Right). */
typedef struct {
  int32_t side, uplo, trans, diag;
  size_t m, n;
  float alpha;
  const float *a;
  size_t a_len, lda;
  float *b;
  size_t b_len, ldb;
} rblas_strsm_problem_t;

typedef struct {
  int32_t side, uplo, trans, diag;
  size_t m, n;
  double alpha;
  const double *a;
  size_t a_len, lda;
  double *b;
  size_t b_len, ldb;
} rblas_dtrsm_problem_t;

int32_t rblas_strsm_grouped_create(rblas_grouped_t **plan,
                                   const rblas_strsm_problem_t *problems,
                                   size_t count, size_t *bad_index);
int32_t rblas_dtrsm_grouped_create(rblas_grouped_t **plan,
                                   const rblas_dtrsm_problem_t *problems,
                                   size_t count, size_t *bad_index);

/* Enqueue one run on `stream` (not synchronised). Runs of one plan must not
 * overlap: they share the plan's work counters (and write the same outputs),
 * so enqueue them on one stream; different plans may run concurrently. */
int32_t rblas_grouped_run(rblas_grouped_t *plan, void *stream);
/* The same run, captured into a hipGraph on first use and replayed after;
 * needs a non-null stream. */
int32_t rblas_grouped_run_graph(rblas_grouped_t *plan, void *stream);
int32_t rblas_grouped_info(const rblas_grouped_t *plan, rblas_grouped_info_t *info);
int32_t rblas_grouped_destroy(rblas_grouped_t *plan);
/* The planner's cost model for dtype 0 (f32) or 1 (f64): coeffs =
 * {c0_64, c1_64, c0_128, c1_128, unit_cap_us}: a unit of a tile side costs
 * c0 + c1 * (its k / k-tile) microseconds; units above unit_cap_us get their k
 * split. Written by the autotuner; affects plans created afterwards. */
int32_t rblas_grouped_set_tuning(int32_t dtype, const double *coeffs);
int32_t rblas_grouped_get_tuning(int32_t dtype, double *coeffs);

/* ----------------------------------------------------------------- serving
 * No rblas counterpart. Replays a request stream against the wall clock on
 * the device: request i arrives arrival_s seconds after the start (arrivals
 * non-decreasing), joins a FIFO queue, is admitted when a stream slot is free
 * and the footprint of all admitted, unfinished requests stays within
 * hbm_cap_bytes, and completes when its batch's work has finished. Operands
 * are synthetic (workloads.py's counter-based splitmix64, uniform(-1, 1)).
 * The read-only A and B of gemm and complex requests are slices of an
 * operand pool per element type, generated in HBM once before the first
 * arrival: pool element i is draw i of seed RBLAS_SERVE_POOL_SEED, and the
 * slice of operand j (A: 1, B: 2) starts at element
 * mix(seed ^ j) mod (pool - len + 1), mix the splitmix64 finalizer, the pool
 * holding max(2^25, the largest such operand) elements. Written operands are
 * generated in HBM from `seed` at launch and freed on completion: gemm and
 * complex C from offset len(A) + len(B); a trsm A (lower triangular:
 * off-diagonal u / sqrt(m), diagonal 2 + |u|, zeros above) from 0 and B from
 * m * m. A request's footprint counts its written operands and workspace.
 * Kinds (column-major, NoTrans):
 *   RBLAS_SERVE_GEMM   C (m x n) := 1 * A (m x k) B (k x n) + 0.5 C, dtype 0 f32 / 1 f64
 *   RBLAS_SERVE_TRSM   A (m x m) X = B (m x n), Left, Lower, NonUnit, alpha 1
 *   RBLAS_SERVE_ZGEMM  complex f64: C := (1 + 0i) A B + (0.5 + 0i) C (4M)
 * Mode RBLAS_SERVE_BATCHED: continuous batching over max_inflight streams; a
 * batch is the queue's admitted head (at most max_batch requests): one
 * grouped gemm launch per dtype, one grouped trsm launch, 4M launches per
 * complex request. Mode RBLAS_SERVE_PLAIN: the per-call comparator, one
 * request at a time in FIFO order through the single-call kernels.
 * results[i] gets the request's times (seconds from the start), algorithmic
 * flops (2mnk, m^2 n, 8mnk), footprint and an order-independent 64-bit hash
 * of the result's bits (of every hash_col_step-th column); a request's hash
 * does not depend on what else ran (same batch mode). capture_ids / capture_out (optional, tests): the dense
 * result of those requests is copied to the given host buffers. Blocks until
 * every request has completed. */
enum { RBLAS_SERVE_GEMM = 0, RBLAS_SERVE_TRSM = 1, RBLAS_SERVE_ZGEMM = 2 };
enum { RBLAS_SERVE_BATCHED = 0, RBLAS_SERVE_PLAIN = 1 };
#define RBLAS_SERVE_POOL_SEED 0x5EEDB10C5EEDB10Cull

typedef struct {
  int32_t kind;
  int32_t dtype;    /* 0 f32, 1 f64 (complex requests are f64) */
  int32_t m, n, k;  /* trsm: m = order of A, n = right-hand sides, k unused */
  int32_t pad;
  uint64_t seed;
  double arrival_s;
} rblas_serve_request_t;

typedef struct {
  int32_t mode;
  int32_t max_inflight; /* batches on the device at once (batched mode)      */
  int32_t max_batch;    /* requests per batch (batched mode)                 */
  int32_t capture_count;
  int32_t hash_col_step; /* result columns hashed: every one (0 or 1) or
                            every hash_col_step-th                           */
  int32_t pad;
  double hbm_cap_bytes; /* admission cap on admitted, unfinished footprints  */
  const int32_t *capture_ids;
  void *const *capture_out;
} rblas_serve_config_t;

typedef struct {
  double arrival_s, admit_s, enqueued_s, done_s; /* enqueued: batch handed to the device */
  double flops;
  double footprint_bytes;
  uint64_t hash;
  int32_t batch;
  int32_t pad;
} rblas_serve_result_t;

typedef struct {
  double wall_s;              /* start to last completion                   */
  double flops;               /* sum of the requests' algorithmic flops     */
  double p50_s, p99_s;        /* latency (completion - arrival) percentiles */
  double peak_footprint_bytes;
  int32_t batches;
  int32_t max_inflight_seen;
} rblas_serve_stats_t;

int32_t rblas_serve_run(const rblas_serve_request_t *requests, size_t count,
                        const rblas_serve_config_t *config,
                        rblas_serve_result_t *results, rblas_serve_stats_t *stats);

/* ------------------------------------------------- device memory & streams
 * Plumbing for callers that keep operands resident in HBM (no rblas
 * counterpart). All return a status. */
int32_t rblas_malloc(void **ptr, size_t bytes);
int32_t rblas_free(void *ptr);
int32_t rblas_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
int32_t rblas_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
int32_t rblas_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream);
int32_t rblas_stream_create(void **stream);
int32_t rblas_stream_destroy(void *stream);
int32_t rblas_stream_synchronize(void *stream);
int32_t rblas_event_create(void **event);
int32_t rblas_event_destroy(void *event);
int32_t rblas_event_record(void *event, void *stream);
int32_t rblas_event_elapsed_ms(float *ms, void *start, void *stop);
int32_t rblas_device_synchronize(void);
/* Number of visible HIP devices (0 when there is none); never fails. */
int32_t rblas_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* RBLAS_HIP_H */
