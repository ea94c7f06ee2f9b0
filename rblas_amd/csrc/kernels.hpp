// Internal launch interface between the C-ABI host code (api.cpp) and the
// HIP kernels. Everything here is column-major: the host folds RowMajor into
// ColMajor with the A<->B, m<->n swap of rblas's level3.rs:56-60 before it
// calls any of these.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace rblas {

// op(A) is m x k, op(B) is k x n, C is m x n (column-major, leading dims of
// the stored matrices). ta/tb: operand stored transposed (Trans/ConjTrans).
template <typename T>
struct GemmProblem {
  bool ta, tb;
  int64_t m, n, k;
  T alpha, beta;
  const T *a;
  int64_t lda;
  const T *b;
  int64_t ldb;
  T *c;
  int64_t ldc;
  // 0: all of C. 1 / 2: only the upper (i <= j) / lower (i >= j) triangle of
  // a square C is computed and written (syrk); only tiles touching it run.
  int tri = 0;
};

// Compute units of one MI355X (8 XCDs x 32).
constexpr int kCUs = 256;

// Launch shape of one MFMA gemm: C tile side (64 or 128; 256 = the 128 x 256
// wide tile of gemm_wide.hip) and the number of k splits (1 = none). A pure
// function of the problem's shape; the k order of every C entry is the same
// for every tile, so a problem gives the same bits whenever and with
// whatever it runs.
struct GemmConfig {
  int tile;
  int splits;
};

// C := alpha * op(A) * op(B) + beta * C on MFMA tiles. Requires m, n > 0,
// k > 0 and alpha != 0 (the host routes the other cases to gemm_scale).
// beta == 0 never reads C (rblas stores exact zeros, kernel/gemm.rs:328-331).
// launch_gemm never splits k and needs no workspace; launch_gemm_cfg runs
// `cfg`, whose k splits need gemm_workspace_bytes(p, cfg) bytes at `ws`
// (partials and per-tile counters; free again once the stream has run it).
template <typename T>
hipError_t launch_gemm(const GemmProblem<T> &p, hipStream_t stream);
template <typename T>
GemmConfig gemm_config(const GemmProblem<T> &p);
template <typename T>
size_t gemm_workspace_bytes(const GemmProblem<T> &p, GemmConfig cfg);
template <typename T>
hipError_t launch_gemm_cfg(const GemmProblem<T> &p, GemmConfig cfg, void *ws,
                           hipStream_t stream);
// Whether every in-tile operand offset fits the 31-bit buffer addressing.
template <typename T>
bool gemm_fits(const GemmProblem<T> &p);
// The 128 x 256 tile (gemm_wide.hip): whether it applies (f64 NN, TN, NT;
// f32 NN, NT; no triangle, at least two tiles per CU, its in-tile offsets
// within 31 bits), and its launch.
template <typename T>
bool gemm_wide_ok(const GemmProblem<T> &p);
template <typename T>
hipError_t launch_gemm_wide(const GemmProblem<T> &p, hipStream_t stream);

// C := beta * C over an m x n block (rblas apply_beta, kernel/gemm.rs:322-338):
// beta == 0 stores exact zeros, beta == 1 is a no-op the host skips.
template <typename T>
hipError_t launch_scale(int64_t m, int64_t n, T beta, T *c, int64_t ldc,
                        hipStream_t stream);
// The same on one triangle of an n x n C (tri 1 = upper, 2 = lower).
template <typename T>
hipError_t launch_scale_tri(int64_t n, int tri, T beta, T *c, int64_t ldc, hipStream_t stream);

// SGEMM through bf16 splits (planes = 2: BF16x3, planes = 3: BF16x6). Same
// preconditions as launch_gemm. `workspace` holds the split planes and must
// have sgemm_split_workspace() bytes; it is free again once the work enqueued
// on `stream` completes.
size_t sgemm_split_workspace(int64_t m, int64_t n, int64_t k, int planes);
hipError_t launch_sgemm_split(const GemmProblem<float> &p, int planes, void *workspace,
                              hipStream_t stream);

// Complex gemm by 4M (interleaved {re, im} storage, column-major). op codes:
// 0 none, 1 transpose, 2 conjugate transpose. Requires m, n > 0. `ws` holds
// cgemm_workspace_elems(m, n, k) elements of T.
template <typename T>
struct CgemmProblem {
  int ta, tb;
  int64_t m, n, k;
  T alpha_re, alpha_im, beta_re, beta_im;
  const T *a;
  int64_t lda;
  const T *b;
  int64_t ldb;
  T *c;
  int64_t ldc;
};
size_t cgemm_workspace_elems(int64_t m, int64_t n, int64_t k);
template <typename T>
hipError_t launch_cgemm(const CgemmProblem<T> &p, T *ws, hipStream_t stream);

// Level-2 over a column-major storage view S of A: R x C, S(r, c) =
// a[r + c*ld] (complex: interleaved {re, im}; a RowMajor A is the view of its
// transpose). gemv: cols_shape ? y_c = f(sum_r S(r,c) v_r) : y_r =
// f(sum_c S(r,c) v_c), S conjugated when `conj`, f applying alpha and beta
// (real: fma(alpha, acc, beta y) with beta 1 / 0 special; complex: beta y +
// alpha acc). `ws` holds gemv_workspace_bytes(p) bytes (0: none needed).
template <typename T>
struct GemvProblem {
  bool cols_shape, cplx, conj;
  int64_t R, C;
  const T *a;
  int64_t ld;
  const T *v;
  int64_t incv;
  T *y;
  int64_t incy;
  T alpha_re, alpha_im, beta_re, beta_im;
};
template <typename T>
size_t gemv_workspace_bytes(const GemvProblem<T> &p);
template <typename T>
hipError_t launch_gemv(const GemvProblem<T> &p, void *ws, hipStream_t stream);

// Rank-1 update of the storage view: logical (i, j) = (r, c), or (c, r) when
// `rowmajor`; x is indexed by i, y by j (conjugated when `conj`).
template <typename T>
struct Rank1Problem {
  bool cplx, conj, rowmajor;
  int64_t R, C;
  T *a;
  int64_t ld;
  const T *x;
  int64_t incx;
  const T *y;
  int64_t incy;
  T alpha_re, alpha_im;
};
template <typename T>
hipError_t launch_rank1(const Rank1Problem<T> &p, hipStream_t stream);

// Complex trsm (interleaved complex B and A): op(A) X = alpha B (left) or
// X op(A) = alpha B, in place; `upper` is the stored triangle (uplo Upper),
// trans 0 / 1 / 2 = None / Trans / ConjTrans.
template <typename T>
struct CtrsmProblem {
  bool left, upper;
  int trans;
  bool unit;
  int64_t m, n;
  T alpha_re, alpha_im;
  const T *a;
  int64_t lda;
  T *b;
  int64_t ldb;
};
// Workspace of launch_ctrsm: ctrsm_workspace_elems(m, n, left) elements of T.
size_t ctrsm_workspace_elems(int64_t m, int64_t n, bool left);
template <typename T>
hipError_t launch_ctrsm(const CtrsmProblem<T> &p, T *ws, hipStream_t stream);

// out (q x q, ld q, dense) := the matrix whose `upper` (else lower) triangle
// is stored in a (ld lda), mirrored across the diagonal; complex_elems:
// interleaved complex elements (q x q of them), herm: the mirror is
// conjugated and the diagonal's imaginary part is zero (Hermitian). Reads
// only the stored triangle.
template <typename T>
hipError_t launch_sym_full(const T *a, int64_t lda, int64_t q, bool upper, bool complex_elems,
                           bool herm, T *out, hipStream_t stream);

// Triangular solve in place on B (m x n, column-major). `upper` is the
// effective triangle of op(A) (uplo flipped by trans). Requires the whole
// op(A) diagonal to be readable and m, n > 0 handled by the caller's checks.
// `ws` holds trsm_workspace_bytes() (the f64 inverses of the diagonal blocks).
template <typename T>
struct TrsmProblem {
  bool left, upper, trans, unit;
  int64_t m, n;
  T alpha;
  const T *a;
  int64_t lda;
  T *b;
  int64_t ldb;
};
size_t trsm_workspace_bytes(int64_t m, int64_t n, bool left);
template <typename T>
hipError_t launch_trsm(const TrsmProblem<T> &p, void *ws, hipStream_t stream);

// Level-1 (HBM-bound). `scratch` holds per-block partials, at least
// level1_scratch_bytes(); the result is written to device memory `out`.
size_t level1_scratch_bytes();
template <typename T>
hipError_t launch_dot(int64_t n, const T *x, int64_t incx, const T *y, int64_t incy, T *out,
                      void *scratch, hipStream_t stream);
template <typename T>
hipError_t launch_nrm2(int64_t n, const T *x, int64_t incx, T *out, void *scratch,
                       hipStream_t stream);
template <typename T>
hipError_t launch_axpy(int64_t n, T alpha, const T *x, int64_t incx, T *y, int64_t incy,
                       hipStream_t stream);

}  // namespace rblas
