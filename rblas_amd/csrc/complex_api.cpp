// C-ABI complex gemm: complex3::gemm (complex3.rs:27-74), column-major,
// interleaved Complex<T>, host slices and HBM-resident forms.
#include <climits>

#include "kernels.hpp"
#include "runtime.hpp"
#include "trsm_host.hpp"

namespace rblas {
namespace {

// Where complex3::gemm panics: split_op's gathers of op(A) (m x k) and op(B)
// (k x n) index past their slices (complex3.rs:111-118), or the final
// combine indexes C past its slice (complex3.rs:61-73). Lengths are in
// complex elements.
Check cgemm_panic_check(bool at, bool bt, size_t m, size_t n, size_t k, size_t a_len, size_t lda,
                        size_t b_len, size_t ldb, size_t c_len, size_t ldc) {
  if (m > 0 && k > 0) {
    size_t hi;
    bool fits = at ? max_index(m, k, lda, 1, hi) : max_index(m, k, 1, lda, hi);
    if (!fits || hi >= a_len)
      return panic("index out of bounds: the len is " + std::to_string(a_len) +
                   " but the index is " + (fits ? std::to_string(hi) : std::string("past usize")));
  }
  if (k > 0 && n > 0) {
    size_t hi;
    bool fits = bt ? max_index(k, n, ldb, 1, hi) : max_index(k, n, 1, ldb, hi);
    if (!fits || hi >= b_len)
      return panic("index out of bounds: the len is " + std::to_string(b_len) +
                   " but the index is " + (fits ? std::to_string(hi) : std::string("past usize")));
  }
  if (m > 0 && n > 0) {
    size_t hi;
    bool fits = max_index(m, n, 1, ldc, hi);
    if (!fits || hi >= c_len)
      return panic("index out of bounds: the len is " + std::to_string(c_len) +
                   " but the index is " + (fits ? std::to_string(hi) : std::string("past usize")));
  }
  return Check{};
}

template <typename T>
Check cgemm_validate(int32_t ta, int32_t tb, size_t m, size_t n, size_t k, size_t a_len,
                     size_t lda, size_t b_len, size_t ldb, size_t c_len, size_t ldc) {
  if (!valid_trans(ta) || !valid_trans(tb)) {
    Check bad;
    bad.status = RBLAS_BAD_ENUM;
    bad.err = {RBLAS_BAD_ENUM, 0, 0, 0};
    bad.msg = "complex gemm: enum argument out of range";
    return bad;
  }
  Check c = cgemm_panic_check(is_transposed(ta), is_transposed(tb), m, n, k, a_len, lda, b_len,
                              ldb, c_len, ldc);
  if (!c.ok()) return c;
  // complex3.rs:61-73 writes every C entry in column order, so overlapping
  // columns (ldc < m) see earlier columns' results; the device combine
  // writes them at the same time.
  if (m > 0 && n > 1 && ldc < m) {
    Check bad;
    bad.status = RBLAS_BAD_ARG;
    bad.err = {RBLAS_BAD_ARG, 0, 0, 0};
    bad.msg = "complex gemm: ldc < m (overlapping columns of C) is not supported on the device";
    return bad;
  }
  // The real gemms run with k' = 2k and leading dimensions m and 2k.
  if (m > INT_MAX || n > INT_MAX || 2 * k > INT_MAX ||
      ((m > 2 * k ? m : 2 * k) * 128 + 160) * sizeof(T) >= 0x7fffffe0ull) {
    Check bad;
    bad.status = RBLAS_BAD_ARG;
    bad.err = {RBLAS_BAD_ARG, 0, 0, 0};
    bad.msg = "complex gemm: dimensions too large for the device path";
    return bad;
  }
  return Check{};
}

template <typename T>
int32_t cgemm_host(int32_t ta, int32_t tb, size_t m, size_t n, size_t k, T ar, T ai, const T *a,
                   size_t a_len, size_t lda, const T *b, size_t b_len, size_t ldb, T br, T bi,
                   T *c, size_t c_len, size_t ldc) {
  Check chk = cgemm_validate<T>(ta, tb, m, n, k, a_len, lda, b_len, ldb, c_len, ldc);
  if (!chk.ok()) return fail(chk, nullptr);
  if (m == 0 || n == 0) return RBLAS_OK;
  const bool at = is_transposed(ta), bt = is_transposed(tb);
  CtxLease lease;
  if (!lease.get()) return hip_fail(lease.status(), "stream create");
  Ctx &ctx = *lease.get();
  const size_t cs = 2 * sizeof(T);
  HostXfer x(ctx, cs, (m * k + k * n + m * n) * cs);
  Staged sa, sb, sc;
  sa.plan(at ? k : m, at ? m : k, lda, x.pinned());
  sb.plan(bt ? n : k, bt ? k : n, ldb, x.pinned());
  sc.plan(m, n, ldc, x.pinned());
  RBLAS_HIP_TRY(ctx.ensure(0, sa.elems * cs), "device buffer A");
  RBLAS_HIP_TRY(ctx.ensure(1, sb.elems * cs), "device buffer B");
  RBLAS_HIP_TRY(ctx.ensure(2, sc.elems * cs), "device buffer C");
  RBLAS_HIP_TRY(ctx.ensure(3, cgemm_workspace_elems(m, n, k) * sizeof(T)), "device workspace");
  RBLAS_HIP_TRY(x.reserve(sa, sb, sc), "pinned staging buffer");
  hipStream_t st = ctx.stream;
  RBLAS_HIP_TRY(x.in(sa, a, ctx.buf[0]), "copy A");
  RBLAS_HIP_TRY(x.in(sb, b, ctx.buf[1]), "copy B");
  RBLAS_HIP_TRY(x.in(sc, c, ctx.buf[2]), "copy C");
  CgemmProblem<T> p{ta, tb, (int64_t)m, (int64_t)n, (int64_t)k, ar, ai, br, bi,
                    (const T *)ctx.buf[0], (int64_t)sa.dev_ld, (const T *)ctx.buf[1],
                    (int64_t)sb.dev_ld, (T *)ctx.buf[2], (int64_t)sc.dev_ld};
  RBLAS_HIP_TRY(launch_cgemm<T>(p, (T *)ctx.buf[3], st), "complex gemm launch");
  RBLAS_HIP_TRY(x.out(sc, ctx.buf[2], c), "copy C back");
  RBLAS_HIP_TRY(x.finish(), "complex gemm");
  return RBLAS_OK;
}

template <typename T>
int32_t cgemm_dev(int32_t ta, int32_t tb, size_t m, size_t n, size_t k, T ar, T ai, const T *a,
                  size_t a_len, size_t lda, const T *b, size_t b_len, size_t ldb, T br, T bi, T *c,
                  size_t c_len, size_t ldc, void *stream) {
  Check chk = cgemm_validate<T>(ta, tb, m, n, k, a_len, lda, b_len, ldb, c_len, ldc);
  if (!chk.ok()) return fail(chk, nullptr);
  if (m == 0 || n == 0) return RBLAS_OK;
  hipStream_t st = (hipStream_t)stream;
  void *ws = nullptr;
  RBLAS_HIP_TRY(pool_alloc(&ws, cgemm_workspace_elems(m, n, k) * sizeof(T), st),
                "complex gemm workspace");
  CgemmProblem<T> p{ta, tb, (int64_t)m, (int64_t)n, (int64_t)k, ar, ai, br, bi, a, (int64_t)lda,
                    b, (int64_t)ldb, c, (int64_t)ldc};
  hipError_t e = launch_cgemm<T>(p, (T *)ws, st);
  e = pool_release(ws, st, e);
  RBLAS_HIP_TRY(e, "complex gemm launch");
  return RBLAS_OK;
}

// --------------------------------------------------------- symm / hemm
// complex3::symm / hemm (complex3.rs:127-247, he_or_sy): nothing when m or n
// is 0; otherwise the dense A is built from its stored `uplo` triangle
// (Hermitian: mirrored half conjugated, diagonal imaginary part zero; every
// read a stored entry, the largest index (q-1)(lda+1)) and the product runs
// through complex3::gemm(N, N) with A on the left (Side Left) or right.
template <typename T>
Check hesy_validate(int32_t side, int32_t uplo, size_t m, size_t n, size_t a_len, size_t lda,
                    size_t b_len, size_t ldb, size_t c_len, size_t ldc) {
  if ((side != RBLAS_LEFT && side != RBLAS_RIGHT) || (uplo != RBLAS_UPPER && uplo != RBLAS_LOWER)) {
    Check bad;
    bad.status = RBLAS_BAD_ENUM;
    bad.err = {RBLAS_BAD_ENUM, 0, 0, 0};
    bad.msg = "symm/hemm: enum argument out of range";
    return bad;
  }
  if (m == 0 || n == 0) return Check{};
  const size_t q = side == RBLAS_LEFT ? m : n;
  size_t hi;
  const bool fits = max_index(q, q, 1, lda, hi);
  if (!fits || hi >= a_len)
    return panic("index out of bounds: the len is " + std::to_string(a_len) +
                 " but the index is " + (fits ? std::to_string(hi) : std::string("past usize")));
  if (side == RBLAS_LEFT)
    return cgemm_validate<T>(0, 0, m, n, q, q * q, q, b_len, ldb, c_len, ldc);
  return cgemm_validate<T>(0, 0, m, n, q, b_len, ldb, q * q, q, c_len, ldc);
}

// Dense-A elements of T (interleaved complex, 256-byte aligned).
inline size_t hesy_full_elems(size_t q) { return (2 * q * q + 31) / 32 * 32; }

template <typename T>
hipError_t hesy_enqueue(bool left, bool upper, bool herm, size_t m, size_t n, T ar, T ai,
                        const T *a, size_t lda, const T *b, size_t ldb, T br, T bi, T *c,
                        size_t ldc, T *ws, hipStream_t st) {
  const size_t q = left ? m : n;
  T *full = ws;
  hipError_t e = launch_sym_full<T>(a, (int64_t)lda, (int64_t)q, upper, true, herm, full, st);
  if (e != hipSuccess) return e;
  const T *pa = left ? full : b, *pb = left ? b : full;
  const size_t la = left ? q : ldb, lb = left ? ldb : q;
  CgemmProblem<T> p{0, 0, (int64_t)m, (int64_t)n, (int64_t)q, ar, ai, br, bi, pa, (int64_t)la,
                    pb, (int64_t)lb, c, (int64_t)ldc};
  return launch_cgemm<T>(p, ws + hesy_full_elems(q), st);
}

template <typename T>
int32_t hesy_host(bool herm, int32_t side, int32_t uplo, size_t m, size_t n, T ar, T ai,
                  const T *a, size_t a_len, size_t lda, const T *b, size_t b_len, size_t ldb, T br,
                  T bi, T *c, size_t c_len, size_t ldc) {
  Check chk = hesy_validate<T>(side, uplo, m, n, a_len, lda, b_len, ldb, c_len, ldc);
  if (!chk.ok()) return fail(chk, nullptr);
  if (m == 0 || n == 0) return RBLAS_OK;
  const bool left = side == RBLAS_LEFT;
  const size_t q = left ? m : n;
  CtxLease lease;
  if (!lease.get()) return hip_fail(lease.status(), "stream create");
  Ctx &ctx = *lease.get();
  const size_t cs = 2 * sizeof(T);
  HostXfer x(ctx, cs, (q * q + 2 * m * n) * cs);
  Staged sa, sb, sc;
  sa.plan(q, q, lda, x.pinned());
  sb.plan(m, n, ldb, x.pinned());
  sc.plan(m, n, ldc, x.pinned());
  RBLAS_HIP_TRY(ctx.ensure(0, sa.elems * cs), "device buffer A");
  RBLAS_HIP_TRY(ctx.ensure(1, sb.elems * cs), "device buffer B");
  RBLAS_HIP_TRY(ctx.ensure(2, sc.elems * cs), "device buffer C");
  RBLAS_HIP_TRY(ctx.ensure(3, (hesy_full_elems(q) + cgemm_workspace_elems(m, n, q)) * sizeof(T)),
                "device workspace");
  RBLAS_HIP_TRY(x.reserve(sa, sb, sc), "pinned staging buffer");
  RBLAS_HIP_TRY(x.in(sa, a, ctx.buf[0]), "copy A");
  RBLAS_HIP_TRY(x.in(sb, b, ctx.buf[1]), "copy B");
  RBLAS_HIP_TRY(x.in(sc, c, ctx.buf[2]), "copy C");
  RBLAS_HIP_TRY(hesy_enqueue<T>(left, uplo == RBLAS_UPPER, herm, m, n, ar, ai,
                                (const T *)ctx.buf[0], sa.dev_ld, (const T *)ctx.buf[1],
                                sb.dev_ld, br, bi, (T *)ctx.buf[2], sc.dev_ld, (T *)ctx.buf[3],
                                ctx.stream),
                "symm/hemm launch");
  RBLAS_HIP_TRY(x.out(sc, ctx.buf[2], c), "copy C back");
  RBLAS_HIP_TRY(x.finish(), "symm/hemm");
  return RBLAS_OK;
}

template <typename T>
int32_t hesy_dev(bool herm, int32_t side, int32_t uplo, size_t m, size_t n, T ar, T ai,
                 const T *a, size_t a_len, size_t lda, const T *b, size_t b_len, size_t ldb, T br,
                 T bi, T *c, size_t c_len, size_t ldc, void *stream) {
  Check chk = hesy_validate<T>(side, uplo, m, n, a_len, lda, b_len, ldb, c_len, ldc);
  if (!chk.ok()) return fail(chk, nullptr);
  if (m == 0 || n == 0) return RBLAS_OK;
  const bool left = side == RBLAS_LEFT;
  const size_t q = left ? m : n;
  hipStream_t st = (hipStream_t)stream;
  void *ws = nullptr;
  RBLAS_HIP_TRY(pool_alloc(&ws, (hesy_full_elems(q) + cgemm_workspace_elems(m, n, q)) * sizeof(T),
                           st),
                "symm/hemm workspace");
  hipError_t e = hesy_enqueue<T>(left, uplo == RBLAS_UPPER, herm, m, n, ar, ai, a, lda, b, ldb,
                                 br, bi, c, ldc, (T *)ws, st);
  e = pool_release(ws, st, e);
  RBLAS_HIP_TRY(e, "symm/hemm launch");
  return RBLAS_OK;
}

// ---------------------------------------------------------------- trsm
// complex3::trsm (complex3.rs:252-349) has the loops of level3::trsm, so its
// panics are at the same indices (lengths in complex elements): nothing
// when m or n is 0; B's largest index (m-1) + (n-1)*ldb, then op(A)'s
// diagonal (q-1)*(lda+1). The 4M updates add their own size limit.
template <typename T>
Check ctrsm_validate(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m, size_t n,
                     size_t a_len, size_t lda, size_t b_len, size_t ldb) {
  Check c = trsm_validate<T>(side, uplo, trans, diag, m, n, a_len, lda, b_len, ldb);
  if (!c.ok() || m == 0 || n == 0) return c;
  const size_t q = side == RBLAS_LEFT ? m : n;
  if ((2 * q * 128 + 160) * sizeof(T) >= 0x7fffffe0ull)
    return bad_arg("complex trsm: order too large for the device path");
  // The updates run as real gemms on B viewed with leading dimension 2 ldb.
  if ((2 * ldb * 128 + 160) * sizeof(T) >= 0x7fffffe0ull)
    return bad_arg("complex trsm: ldb too large for the device path");
  return Check{};
}

template <typename T>
CtrsmProblem<T> ctrsm_problem(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m,
                              size_t n, T ar, T ai, const T *a, size_t lda, T *b, size_t ldb) {
  return CtrsmProblem<T>{side == RBLAS_LEFT, uplo == RBLAS_UPPER, trans, diag == RBLAS_UNIT,
                         (int64_t)m, (int64_t)n, ar, ai, a, (int64_t)lda, b, (int64_t)ldb};
}

template <typename T>
int32_t ctrsm_host(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m, size_t n,
                   T ar, T ai, const T *a, size_t a_len, size_t lda, T *b, size_t b_len,
                   size_t ldb) {
  Check chk = ctrsm_validate<T>(side, uplo, trans, diag, m, n, a_len, lda, b_len, ldb);
  if (!chk.ok()) return fail(chk, nullptr);
  if (m == 0 || n == 0) return RBLAS_OK;
  const bool left = side == RBLAS_LEFT;
  const size_t q = left ? m : n;
  CtxLease lease;
  if (!lease.get()) return hip_fail(lease.status(), "stream create");
  Ctx &ctx = *lease.get();
  const size_t cs = 2 * sizeof(T);
  HostXfer x(ctx, cs, (q * q + m * n) * cs);
  Staged sa, sb;
  sa.plan(q, q, lda, x.pinned());
  sb.plan(m, n, ldb, x.pinned());
  RBLAS_HIP_TRY(ctx.ensure(0, sa.elems * cs), "device buffer A");
  RBLAS_HIP_TRY(ctx.ensure(2, sb.elems * cs), "device buffer B");
  RBLAS_HIP_TRY(ctx.ensure(3, ctrsm_workspace_elems((int64_t)m, (int64_t)n, left) * sizeof(T)),
                "device workspace");
  RBLAS_HIP_TRY(x.reserve(sa, sb), "pinned staging buffer");
  RBLAS_HIP_TRY(x.in(sa, a, ctx.buf[0]), "copy A");
  RBLAS_HIP_TRY(x.in(sb, b, ctx.buf[2]), "copy B");
  const CtrsmProblem<T> p = ctrsm_problem<T>(side, uplo, trans, diag, m, n, ar, ai,
                                             (const T *)ctx.buf[0], sa.dev_ld, (T *)ctx.buf[2],
                                             sb.dev_ld);
  RBLAS_HIP_TRY(launch_ctrsm<T>(p, (T *)ctx.buf[3], ctx.stream), "complex trsm launch");
  RBLAS_HIP_TRY(x.out(sb, ctx.buf[2], b), "copy B back");
  RBLAS_HIP_TRY(x.finish(), "complex trsm");
  return RBLAS_OK;
}

template <typename T>
int32_t ctrsm_dev(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m, size_t n,
                  T ar, T ai, const T *a, size_t a_len, size_t lda, T *b, size_t b_len,
                  size_t ldb, void *stream) {
  Check chk = ctrsm_validate<T>(side, uplo, trans, diag, m, n, a_len, lda, b_len, ldb);
  if (!chk.ok()) return fail(chk, nullptr);
  if (m == 0 || n == 0) return RBLAS_OK;
  hipStream_t st = (hipStream_t)stream;
  void *ws = nullptr;
  RBLAS_HIP_TRY(pool_alloc(&ws, ctrsm_workspace_elems((int64_t)m, (int64_t)n,
                                                      side == RBLAS_LEFT) * sizeof(T), st),
                "complex trsm workspace");
  const CtrsmProblem<T> p = ctrsm_problem<T>(side, uplo, trans, diag, m, n, ar, ai, a, lda, b,
                                             ldb);
  hipError_t e = launch_ctrsm<T>(p, (T *)ws, st);
  e = pool_release(ws, st, e);
  RBLAS_HIP_TRY(e, "complex trsm launch");
  return RBLAS_OK;
}

}  // namespace
}  // namespace rblas

using namespace rblas;

extern "C" {

int32_t rblas_cgemm(int32_t transa, int32_t transb, size_t m, size_t n, size_t k,
                    rblas_complex32_t alpha, const rblas_complex32_t *a, size_t a_len, size_t lda,
                    const rblas_complex32_t *b, size_t b_len, size_t ldb, rblas_complex32_t beta,
                    rblas_complex32_t *c, size_t c_len, size_t ldc) {
  return cgemm_host<float>(transa, transb, m, n, k, alpha.re, alpha.im, (const float *)a, a_len,
                           lda, (const float *)b, b_len, ldb, beta.re, beta.im, (float *)c, c_len,
                           ldc);
}

int32_t rblas_zgemm(int32_t transa, int32_t transb, size_t m, size_t n, size_t k,
                    rblas_complex64_t alpha, const rblas_complex64_t *a, size_t a_len, size_t lda,
                    const rblas_complex64_t *b, size_t b_len, size_t ldb, rblas_complex64_t beta,
                    rblas_complex64_t *c, size_t c_len, size_t ldc) {
  return cgemm_host<double>(transa, transb, m, n, k, alpha.re, alpha.im, (const double *)a, a_len,
                            lda, (const double *)b, b_len, ldb, beta.re, beta.im, (double *)c,
                            c_len, ldc);
}

int32_t rblas_cgemm_dev(int32_t transa, int32_t transb, size_t m, size_t n, size_t k,
                        rblas_complex32_t alpha, const rblas_complex32_t *a, size_t a_len,
                        size_t lda, const rblas_complex32_t *b, size_t b_len, size_t ldb,
                        rblas_complex32_t beta, rblas_complex32_t *c, size_t c_len, size_t ldc,
                        void *stream) {
  return cgemm_dev<float>(transa, transb, m, n, k, alpha.re, alpha.im, (const float *)a, a_len,
                          lda, (const float *)b, b_len, ldb, beta.re, beta.im, (float *)c, c_len,
                          ldc, stream);
}

int32_t rblas_zgemm_dev(int32_t transa, int32_t transb, size_t m, size_t n, size_t k,
                        rblas_complex64_t alpha, const rblas_complex64_t *a, size_t a_len,
                        size_t lda, const rblas_complex64_t *b, size_t b_len, size_t ldb,
                        rblas_complex64_t beta, rblas_complex64_t *c, size_t c_len, size_t ldc,
                        void *stream) {
  return cgemm_dev<double>(transa, transb, m, n, k, alpha.re, alpha.im, (const double *)a, a_len,
                           lda, (const double *)b, b_len, ldb, beta.re, beta.im, (double *)c,
                           c_len, ldc, stream);
}

int32_t rblas_csymm(int32_t side, int32_t uplo, size_t m, size_t n, rblas_complex32_t alpha,
                    const rblas_complex32_t *a, size_t a_len, size_t lda, const rblas_complex32_t *b, size_t b_len,
                    size_t ldb, rblas_complex32_t beta, rblas_complex32_t *c, size_t c_len, size_t ldc) {
  return hesy_host<float>(false, side, uplo, m, n, alpha.re, alpha.im, (const float *)a, a_len, lda,
                        (const float *)b, b_len, ldb, beta.re, beta.im, (float *)c, c_len, ldc);
}

int32_t rblas_csymm_dev(int32_t side, int32_t uplo, size_t m, size_t n, rblas_complex32_t alpha,
                        const rblas_complex32_t *a, size_t a_len, size_t lda, const rblas_complex32_t *b,
                        size_t b_len, size_t ldb, rblas_complex32_t beta, rblas_complex32_t *c, size_t c_len,
                        size_t ldc, void *stream) {
  return hesy_dev<float>(false, side, uplo, m, n, alpha.re, alpha.im, (const float *)a, a_len, lda,
                       (const float *)b, b_len, ldb, beta.re, beta.im, (float *)c, c_len, ldc, stream);
}

int32_t rblas_zsymm(int32_t side, int32_t uplo, size_t m, size_t n, rblas_complex64_t alpha,
                    const rblas_complex64_t *a, size_t a_len, size_t lda, const rblas_complex64_t *b, size_t b_len,
                    size_t ldb, rblas_complex64_t beta, rblas_complex64_t *c, size_t c_len, size_t ldc) {
  return hesy_host<double>(false, side, uplo, m, n, alpha.re, alpha.im, (const double *)a, a_len, lda,
                        (const double *)b, b_len, ldb, beta.re, beta.im, (double *)c, c_len, ldc);
}

int32_t rblas_zsymm_dev(int32_t side, int32_t uplo, size_t m, size_t n, rblas_complex64_t alpha,
                        const rblas_complex64_t *a, size_t a_len, size_t lda, const rblas_complex64_t *b,
                        size_t b_len, size_t ldb, rblas_complex64_t beta, rblas_complex64_t *c, size_t c_len,
                        size_t ldc, void *stream) {
  return hesy_dev<double>(false, side, uplo, m, n, alpha.re, alpha.im, (const double *)a, a_len, lda,
                       (const double *)b, b_len, ldb, beta.re, beta.im, (double *)c, c_len, ldc, stream);
}

int32_t rblas_chemm(int32_t side, int32_t uplo, size_t m, size_t n, rblas_complex32_t alpha,
                    const rblas_complex32_t *a, size_t a_len, size_t lda, const rblas_complex32_t *b, size_t b_len,
                    size_t ldb, rblas_complex32_t beta, rblas_complex32_t *c, size_t c_len, size_t ldc) {
  return hesy_host<float>(true, side, uplo, m, n, alpha.re, alpha.im, (const float *)a, a_len, lda,
                        (const float *)b, b_len, ldb, beta.re, beta.im, (float *)c, c_len, ldc);
}

int32_t rblas_chemm_dev(int32_t side, int32_t uplo, size_t m, size_t n, rblas_complex32_t alpha,
                        const rblas_complex32_t *a, size_t a_len, size_t lda, const rblas_complex32_t *b,
                        size_t b_len, size_t ldb, rblas_complex32_t beta, rblas_complex32_t *c, size_t c_len,
                        size_t ldc, void *stream) {
  return hesy_dev<float>(true, side, uplo, m, n, alpha.re, alpha.im, (const float *)a, a_len, lda,
                       (const float *)b, b_len, ldb, beta.re, beta.im, (float *)c, c_len, ldc, stream);
}

int32_t rblas_zhemm(int32_t side, int32_t uplo, size_t m, size_t n, rblas_complex64_t alpha,
                    const rblas_complex64_t *a, size_t a_len, size_t lda, const rblas_complex64_t *b, size_t b_len,
                    size_t ldb, rblas_complex64_t beta, rblas_complex64_t *c, size_t c_len, size_t ldc) {
  return hesy_host<double>(true, side, uplo, m, n, alpha.re, alpha.im, (const double *)a, a_len, lda,
                        (const double *)b, b_len, ldb, beta.re, beta.im, (double *)c, c_len, ldc);
}

int32_t rblas_zhemm_dev(int32_t side, int32_t uplo, size_t m, size_t n, rblas_complex64_t alpha,
                        const rblas_complex64_t *a, size_t a_len, size_t lda, const rblas_complex64_t *b,
                        size_t b_len, size_t ldb, rblas_complex64_t beta, rblas_complex64_t *c, size_t c_len,
                        size_t ldc, void *stream) {
  return hesy_dev<double>(true, side, uplo, m, n, alpha.re, alpha.im, (const double *)a, a_len, lda,
                       (const double *)b, b_len, ldb, beta.re, beta.im, (double *)c, c_len, ldc, stream);
}

int32_t rblas_ctrsm(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m,
                    size_t n, rblas_complex32_t alpha, const rblas_complex32_t *a, size_t a_len, size_t lda,
                    rblas_complex32_t *b, size_t b_len, size_t ldb) {
  return ctrsm_host<float>(side, uplo, trans, diag, m, n, alpha.re, alpha.im, (const float *)a, a_len,
                         lda, (float *)b, b_len, ldb);
}

int32_t rblas_ctrsm_dev(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m,
                        size_t n, rblas_complex32_t alpha, const rblas_complex32_t *a, size_t a_len,
                        size_t lda, rblas_complex32_t *b, size_t b_len, size_t ldb, void *stream) {
  return ctrsm_dev<float>(side, uplo, trans, diag, m, n, alpha.re, alpha.im, (const float *)a, a_len,
                        lda, (float *)b, b_len, ldb, stream);
}

int32_t rblas_ztrsm(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m,
                    size_t n, rblas_complex64_t alpha, const rblas_complex64_t *a, size_t a_len, size_t lda,
                    rblas_complex64_t *b, size_t b_len, size_t ldb) {
  return ctrsm_host<double>(side, uplo, trans, diag, m, n, alpha.re, alpha.im, (const double *)a, a_len,
                         lda, (double *)b, b_len, ldb);
}

int32_t rblas_ztrsm_dev(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m,
                        size_t n, rblas_complex64_t alpha, const rblas_complex64_t *a, size_t a_len,
                        size_t lda, rblas_complex64_t *b, size_t b_len, size_t ldb, void *stream) {
  return ctrsm_dev<double>(side, uplo, trans, diag, m, n, alpha.re, alpha.im, (const double *)a, a_len,
                        lda, (double *)b, b_len, ldb, stream);
}

}  // extern "C"
