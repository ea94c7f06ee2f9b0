// C-ABI syrk and trsm: level3::syrk (level3.rs:236-294) and level3::trsm
// (level3.rs:390-520), column-major, host slices and HBM-resident forms.
#include <climits>

#include "kernels.hpp"
#include "runtime.hpp"
#include "trsm_host.hpp"

namespace rblas {
namespace {

// ---------------------------------------------------------------- syrk

// Where rblas's syrk panics, in its order (n == 0 returns first):
//   the inner gemm (ColMajor, (N,T) or (T,N), n x n x k, beta 0, into an
//   n x n temporary) takes materialize's path and gathers A twice
//   (level3.rs:262-278, 193-220); then the fold indexes c[i + j*ldc] over
//   the uplo triangle (level3.rs:279-293), whose largest index is
//   (n-1)*(ldc+1) for either triangle.
template <typename T>
Check syrk_validate(int32_t uplo, int32_t trans, size_t n, size_t k, size_t a_len, size_t lda,
                    size_t c_len, size_t ldc) {
  if (!valid_uplo(uplo) || !valid_trans(trans)) return bad_enum("syrk: enum argument out of range");
  if (n == 0) return Check{};
  const bool t = is_transposed(trans);
  Check pc = gemm_panic_check(t, !t, n, n, k, a_len, lda, a_len, lda, n * n, n, false, false);
  if (!pc.ok()) return pc;
  size_t hi;
  const bool fits = max_index(n, n, 1, ldc, hi);  // (n-1) + (n-1)*ldc
  if (!fits || hi >= c_len) return oob(c_len, fits, hi);
  if (n > INT_MAX || k > INT_MAX) return bad_arg("syrk: n and k must be below 2^31 on the device");
  if (n > 1 && ldc < n)
    return bad_arg("syrk: ldc < n (overlapping columns of C) is not supported on the device");
  if (k > 0 && ld_too_large(lda, sizeof(T)))
    return bad_arg("syrk: lda above 2^31/(128*sizeof(T)) elements is not supported on the device");
  return Check{};
}

// The device work of a validated syrk (n > 0) on HBM operands.
template <typename T>
hipError_t syrk_enqueue(int32_t uplo, int32_t trans, int64_t n, int64_t k, T alpha, const T *a,
                        int64_t lda, T beta, T *c, int64_t ldc, hipStream_t st) {
  const int tri = uplo == RBLAS_UPPER ? 1 : 2;
  if (k == 0 || alpha == T(0)) {
    // The temporary is all zeros: c = 0 + (beta == 0 ? 0 : beta * c).
    if (beta == T(1)) return hipSuccess;
    return launch_scale_tri<T>(n, tri, beta, c, ldc, st);
  }
  const bool t = is_transposed(trans);
  GemmProblem<T> p{t, !t, n, n, k, alpha, beta, a, lda, a, lda, c, ldc, tri};
  return launch_gemm<T>(p, st);
}

template <typename T>
int32_t syrk_host(int32_t uplo, int32_t trans, size_t n, size_t k, T alpha, const T *a,
                  size_t a_len, size_t lda, T beta, T *c, size_t c_len, size_t ldc) {
  Check chk = syrk_validate<T>(uplo, trans, n, k, a_len, lda, c_len, ldc);
  if (!chk.ok()) return fail(chk, nullptr);
  if (n == 0) return RBLAS_OK;
  CtxLease lease;
  if (!lease.get()) return hip_fail(lease.status(), "stream create");
  Ctx &ctx = *lease.get();
  const bool t = is_transposed(trans);
  const bool need_a = k != 0 && alpha != T(0);
  HostXfer x(ctx, sizeof(T), ((need_a ? n * k : 0) + n * n) * sizeof(T));
  Staged sa, sc;
  if (need_a) sa.plan(t ? k : n, t ? n : k, lda, x.pinned());
  sc.plan(n, n, ldc, x.pinned());
  RBLAS_HIP_TRY(ctx.ensure(0, sa.elems * sizeof(T)), "device buffer A");
  RBLAS_HIP_TRY(ctx.ensure(2, sc.elems * sizeof(T)), "device buffer C");
  RBLAS_HIP_TRY(x.reserve(sa, sc), "pinned staging buffer");
  T *da = (T *)ctx.buf[0], *dc = (T *)ctx.buf[2];
  hipStream_t st = ctx.stream;
  if (need_a) RBLAS_HIP_TRY(x.in(sa, a, da), "copy A to device");
  RBLAS_HIP_TRY(x.in(sc, c, dc), "copy C to device");
  RBLAS_HIP_TRY(syrk_enqueue<T>(uplo, trans, (int64_t)n, (int64_t)k, alpha, da,
                                (int64_t)sa.dev_ld, beta, dc, (int64_t)sc.dev_ld, st),
                "syrk launch");
  RBLAS_HIP_TRY(x.out(sc, dc, c), "copy C to host");
  RBLAS_HIP_TRY(x.finish(), "syrk");
  return RBLAS_OK;
}

template <typename T>
int32_t syrk_dev(int32_t uplo, int32_t trans, size_t n, size_t k, T alpha, const T *a,
                 size_t a_len, size_t lda, T beta, T *c, size_t c_len, size_t ldc, void *stream) {
  Check chk = syrk_validate<T>(uplo, trans, n, k, a_len, lda, c_len, ldc);
  if (!chk.ok()) return fail(chk, nullptr);
  if (n == 0) return RBLAS_OK;
  RBLAS_HIP_TRY(syrk_enqueue<T>(uplo, trans, (int64_t)n, (int64_t)k, alpha, a, (int64_t)lda, beta,
                                c, (int64_t)ldc, (hipStream_t)stream),
                "syrk launch");
  return RBLAS_OK;
}

// ---------------------------------------------------------------- trsm

template <typename T>
TrsmProblem<T> trsm_problem(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m,
                            size_t n, T alpha, const T *a, size_t lda, T *b, size_t ldb) {
  const bool t = is_transposed(trans);
  // Effective triangle of op(A) (level3.rs:419-422).
  const bool eff_upper = (uplo == RBLAS_UPPER) != t;
  return TrsmProblem<T>{side == RBLAS_LEFT, eff_upper, t, diag == RBLAS_UNIT, (int64_t)m,
                        (int64_t)n, alpha, a, (int64_t)lda, b, (int64_t)ldb};
}

template <typename T>
int32_t trsm_host(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m, size_t n,
                  T alpha, const T *a, size_t a_len, size_t lda, T *b, size_t b_len, size_t ldb) {
  Check chk = trsm_validate<T>(side, uplo, trans, diag, m, n, a_len, lda, b_len, ldb);
  if (!chk.ok()) return fail(chk, nullptr);
  if (m == 0 || n == 0) return RBLAS_OK;
  const bool left = side == RBLAS_LEFT;
  const size_t q = left ? m : n;
  CtxLease lease;
  if (!lease.get()) return hip_fail(lease.status(), "stream create");
  Ctx &ctx = *lease.get();
  HostXfer x(ctx, sizeof(T), (q * q + m * n) * sizeof(T));
  Staged sa, sb;
  sa.plan(q, q, lda, x.pinned());
  sb.plan(m, n, ldb, x.pinned());
  RBLAS_HIP_TRY(ctx.ensure(0, sa.elems * sizeof(T)), "device buffer A");
  RBLAS_HIP_TRY(ctx.ensure(2, sb.elems * sizeof(T)), "device buffer B");
  RBLAS_HIP_TRY(ctx.ensure(3, trsm_workspace_bytes((int64_t)m, (int64_t)n, left)),
                "device workspace");
  RBLAS_HIP_TRY(x.reserve(sa, sb), "pinned staging buffer");
  T *da = (T *)ctx.buf[0], *db = (T *)ctx.buf[2];
  hipStream_t st = ctx.stream;
  RBLAS_HIP_TRY(x.in(sa, a, da), "copy A to device");
  RBLAS_HIP_TRY(x.in(sb, b, db), "copy B to device");
  TrsmProblem<T> p = trsm_problem<T>(side, uplo, trans, diag, m, n, alpha, da, sa.dev_ld, db,
                                     sb.dev_ld);
  RBLAS_HIP_TRY(launch_trsm<T>(p, ctx.buf[3], st), "trsm launch");
  RBLAS_HIP_TRY(x.out(sb, db, b), "copy B to host");
  RBLAS_HIP_TRY(x.finish(), "trsm");
  return RBLAS_OK;
}

template <typename T>
int32_t trsm_dev(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m, size_t n,
                 T alpha, const T *a, size_t a_len, size_t lda, T *b, size_t b_len, size_t ldb,
                 void *stream) {
  Check chk = trsm_validate<T>(side, uplo, trans, diag, m, n, a_len, lda, b_len, ldb);
  if (!chk.ok()) return fail(chk, nullptr);
  if (m == 0 || n == 0) return RBLAS_OK;
  hipStream_t st = (hipStream_t)stream;
  const bool left = side == RBLAS_LEFT;
  void *ws = nullptr;
  RBLAS_HIP_TRY(pool_alloc(&ws, trsm_workspace_bytes((int64_t)m, (int64_t)n, left), st),
                "trsm workspace");
  TrsmProblem<T> p = trsm_problem<T>(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb);
  hipError_t e = launch_trsm<T>(p, ws, st);
  e = pool_release(ws, st, e);
  RBLAS_HIP_TRY(e, "trsm launch");
  return RBLAS_OK;
}

}  // namespace
}  // namespace rblas

using namespace rblas;

extern "C" {

int32_t rblas_ssyrk(int32_t uplo, int32_t trans, size_t n, size_t k, float alpha, const float *a,
                    size_t a_len, size_t lda, float beta, float *c, size_t c_len, size_t ldc) {
  return syrk_host<float>(uplo, trans, n, k, alpha, a, a_len, lda, beta, c, c_len, ldc);
}

int32_t rblas_dsyrk(int32_t uplo, int32_t trans, size_t n, size_t k, double alpha,
                    const double *a, size_t a_len, size_t lda, double beta, double *c,
                    size_t c_len, size_t ldc) {
  return syrk_host<double>(uplo, trans, n, k, alpha, a, a_len, lda, beta, c, c_len, ldc);
}

int32_t rblas_ssyrk_dev(int32_t uplo, int32_t trans, size_t n, size_t k, float alpha,
                        const float *a, size_t a_len, size_t lda, float beta, float *c,
                        size_t c_len, size_t ldc, void *stream) {
  return syrk_dev<float>(uplo, trans, n, k, alpha, a, a_len, lda, beta, c, c_len, ldc, stream);
}

int32_t rblas_dsyrk_dev(int32_t uplo, int32_t trans, size_t n, size_t k, double alpha,
                        const double *a, size_t a_len, size_t lda, double beta, double *c,
                        size_t c_len, size_t ldc, void *stream) {
  return syrk_dev<double>(uplo, trans, n, k, alpha, a, a_len, lda, beta, c, c_len, ldc, stream);
}

int32_t rblas_strsm(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m, size_t n,
                    float alpha, const float *a, size_t a_len, size_t lda, float *b, size_t b_len,
                    size_t ldb) {
  return trsm_host<float>(side, uplo, trans, diag, m, n, alpha, a, a_len, lda, b, b_len, ldb);
}

int32_t rblas_dtrsm(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m, size_t n,
                    double alpha, const double *a, size_t a_len, size_t lda, double *b,
                    size_t b_len, size_t ldb) {
  return trsm_host<double>(side, uplo, trans, diag, m, n, alpha, a, a_len, lda, b, b_len, ldb);
}

int32_t rblas_strsm_dev(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m,
                        size_t n, float alpha, const float *a, size_t a_len, size_t lda, float *b,
                        size_t b_len, size_t ldb, void *stream) {
  return trsm_dev<float>(side, uplo, trans, diag, m, n, alpha, a, a_len, lda, b, b_len, ldb,
                         stream);
}

int32_t rblas_dtrsm_dev(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m,
                        size_t n, double alpha, const double *a, size_t a_len, size_t lda,
                        double *b, size_t b_len, size_t ldb, void *stream) {
  return trsm_dev<double>(side, uplo, trans, diag, m, n, alpha, a, a_len, lda, b, b_len, ldb,
                          stream);
}

}  // extern "C"
