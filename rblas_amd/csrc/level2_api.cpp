// C-ABI Level-2 entry points: level2::gemv / ger (level2.rs:11-104) and
// complex2::gemv / geru / gerc (complex2.rs:11-142), host slices and
// HBM-resident forms.
#include <string>

#include "kernels.hpp"
#include "runtime.hpp"
#include "trsm_host.hpp"

namespace rblas {
namespace {

constexpr size_t kDimLimit = size_t(1) << 31;

// x[t*inc] for t < cnt (cnt > 0) stays below len; rblas panics otherwise.
Check vec_check(const char *fn, char which, size_t cnt, size_t inc, size_t len) {
  size_t hi;
  if (__builtin_mul_overflow(cnt - 1, inc, &hi) || hi >= len)
    return panic(std::string(fn) + ": " + which + " index out of bounds");
  return Check{};
}

// a[i*rs + j*cs] for i < m, j < n (m, n > 0), (rs, cs) = layout.strides(lda)
// (types.rs:20-25).
Check mat_check(const char *fn, int32_t layout, size_t m, size_t n, size_t lda, size_t a_len) {
  const bool row = layout == RBLAS_ROW_MAJOR;
  size_t hi;
  if (!max_index(m, n, row ? lda : 1, row ? 1 : lda, hi) || hi >= a_len)
    return panic(std::string(fn) + ": a index out of bounds");
  return Check{};
}

// Where rblas's gemv panics. Real (level2.rs:25-61): y is scaled unless
// beta == 1; then, when alpha != 0, NoTrans touches a and x for every row
// with n > 0 and y for every row; Trans reads x_i for every row (also when
// n == 0), then a and y when n > 0. Complex (complex2.rs:25-76): y is always
// scaled; then, when alpha != 0, each output touches its a and x entries when
// the inner length is > 0. Device-path limits follow (RBLAS_BAD_ARG).
Check gemv_validate(bool cplx, int32_t layout, int32_t trans, size_t m, size_t n, bool alpha_zero,
                    bool beta_one, size_t a_len, size_t lda, size_t x_len, size_t incx,
                    size_t y_len, size_t incy) {
  if (!valid_layout(layout)) return bad_enum("gemv: layout out of range");
  if (!valid_trans(trans)) return bad_enum("gemv: trans out of range");
  const bool tr = is_transposed(trans);
  const size_t leny = tr ? n : m, lenx = tr ? m : n;
  Check c;
  if ((cplx || !beta_one) && leny > 0) {
    c = vec_check("gemv", 'y', leny, incy, y_len);
    if (!c.ok()) return c;
  }
  if (!alpha_zero) {
    if (!tr || cplx) {
      if (leny > 0) {
        c = vec_check("gemv", 'y', leny, incy, y_len);
        if (!c.ok()) return c;
        if (lenx > 0) {
          c = mat_check("gemv", layout, m, n, lda, a_len);
          if (!c.ok()) return c;
          c = vec_check("gemv", 'x', lenx, incx, x_len);
          if (!c.ok()) return c;
        }
      }
    } else if (m > 0) {
      c = vec_check("gemv", 'x', m, incx, x_len);
      if (!c.ok()) return c;
      if (n > 0) {
        c = vec_check("gemv", 'y', n, incy, y_len);
        if (!c.ok()) return c;
        c = mat_check("gemv", layout, m, n, lda, a_len);
        if (!c.ok()) return c;
      }
    }
  }
  if (m >= kDimLimit || n >= kDimLimit) return bad_arg("gemv: dimensions >= 2^31");
  const bool writes_y = leny > 0 && (cplx || !beta_one || !alpha_zero);
  if (writes_y && leny > 1 && incy == 0)
    return bad_arg("gemv: incy == 0 with more than one output (aliased outputs)");
  return Check{};
}

// Where rblas's ger / rank1 panics (level2.rs:77-87, complex2.rs:126-141):
// nothing when alpha == 0; x_i for every row (also when n == 0), then y_j
// and a_ij when n > 0. Writes that would alias (lda too small for distinct
// elements) are a device-path limit.
Check ger_validate(int32_t layout, size_t m, size_t n, bool alpha_zero, size_t x_len, size_t incx,
                   size_t y_len, size_t incy, size_t a_len, size_t lda) {
  if (!valid_layout(layout)) return bad_enum("ger: layout out of range");
  if (alpha_zero) return Check{};
  Check c;
  if (m > 0) {
    c = vec_check("ger", 'x', m, incx, x_len);
    if (!c.ok()) return c;
    if (n > 0) {
      c = vec_check("ger", 'y', n, incy, y_len);
      if (!c.ok()) return c;
      c = mat_check("ger", layout, m, n, lda, a_len);
      if (!c.ok()) return c;
    }
  }
  if (m >= kDimLimit || n >= kDimLimit) return bad_arg("ger: dimensions >= 2^31");
  const bool row = layout == RBLAS_ROW_MAJOR;
  const size_t inner = row ? n : m, outer = row ? m : n;
  if (inner > 0 && outer > 1 && lda < inner)
    return bad_arg("ger: lda smaller than the contiguous dimension (aliased writes)");
  return Check{};
}

// The storage view of A and the gemv shape (see level2.hip).
template <typename T>
GemvProblem<T> gemv_problem(bool cplx, int32_t layout, int32_t trans, size_t m, size_t n, T ar,
                            T ai, const T *a, size_t lda, const T *x, size_t incx, T br, T bi,
                            T *y, size_t incy) {
  const bool row = layout == RBLAS_ROW_MAJOR, tr = is_transposed(trans);
  GemvProblem<T> p;
  p.cols_shape = tr != row;
  p.cplx = cplx;
  p.conj = cplx && trans == RBLAS_CONJ_TRANS;
  p.R = (int64_t)(row ? n : m);
  p.C = (int64_t)(row ? m : n);
  p.a = a;
  p.ld = (int64_t)lda;
  p.v = x;
  p.incv = (int64_t)incx;
  p.y = y;
  p.incy = (int64_t)incy;
  p.alpha_re = ar;
  p.alpha_im = ai;
  p.beta_re = br;
  p.beta_im = bi;
  return p;
}

template <typename T>
int32_t gemv_host(bool cplx, int32_t layout, int32_t trans, size_t m, size_t n, T ar, T ai,
                  const T *a, size_t a_len, size_t lda, const T *x, size_t x_len, size_t incx,
                  T br, T bi, T *y, size_t y_len, size_t incy) {
  const bool alpha_zero = ar == T(0) && ai == T(0), beta_one = !cplx && br == T(1);
  Check chk = gemv_validate(cplx, layout, trans, m, n, alpha_zero, beta_one, a_len, lda, x_len,
                            incx, y_len, incy);
  if (!chk.ok()) return fail(chk, nullptr);
  const bool tr = is_transposed(trans), row = layout == RBLAS_ROW_MAJOR;
  const size_t leny = tr ? n : m, lenx = tr ? m : n;
  if (leny == 0 || (alpha_zero && beta_one)) return RBLAS_OK;
  const bool reads_a = !alpha_zero && lenx > 0;
  CtxLease lease;
  if (!lease.get()) return hip_fail(lease.status(), "stream create");
  Ctx &ctx = *lease.get();
  const size_t es = (cplx ? 2 : 1) * sizeof(T);
  HostXfer xf(ctx, es, ((reads_a ? m * n + lenx : 0) + leny) * es);
  Staged sa, sx, sy;
  if (reads_a) {
    sa.plan(row ? n : m, row ? m : n, lda, xf.pinned());
    sx.plan(1, lenx, incx, xf.pinned());
  }
  sy.plan(1, leny, incy, xf.pinned());
  RBLAS_HIP_TRY(ctx.ensure(0, sa.elems * es + 64), "device buffer a");
  RBLAS_HIP_TRY(ctx.ensure(1, sx.elems * es + 64), "device buffer x");
  RBLAS_HIP_TRY(ctx.ensure(2, sy.elems * es), "device buffer y");
  const GemvProblem<T> p = gemv_problem<T>(
      cplx, layout, trans, m, n, ar, ai, reads_a ? (const T *)ctx.buf[0] : nullptr, sa.dev_ld,
      reads_a ? (const T *)ctx.buf[1] : nullptr, sx.dev_ld, br, bi, (T *)ctx.buf[2], sy.dev_ld);
  RBLAS_HIP_TRY(ctx.ensure(3, gemv_workspace_bytes(p) + 64), "device workspace");
  RBLAS_HIP_TRY(xf.reserve(sa, sx, sy), "pinned staging buffer");
  RBLAS_HIP_TRY(xf.in(sa, a, ctx.buf[0]), "copy a");
  RBLAS_HIP_TRY(xf.in(sx, x, ctx.buf[1]), "copy x");
  RBLAS_HIP_TRY(xf.in(sy, y, ctx.buf[2]), "copy y");
  RBLAS_HIP_TRY(launch_gemv<T>(p, ctx.buf[3], ctx.stream), "gemv launch");
  RBLAS_HIP_TRY(xf.out(sy, ctx.buf[2], y), "copy y back");
  RBLAS_HIP_TRY(xf.finish(), "gemv");
  return RBLAS_OK;
}

template <typename T>
int32_t gemv_dev(bool cplx, int32_t layout, int32_t trans, size_t m, size_t n, T ar, T ai,
                 const T *a, size_t a_len, size_t lda, const T *x, size_t x_len, size_t incx, T br,
                 T bi, T *y, size_t y_len, size_t incy, void *stream) {
  const bool alpha_zero = ar == T(0) && ai == T(0), beta_one = !cplx && br == T(1);
  Check chk = gemv_validate(cplx, layout, trans, m, n, alpha_zero, beta_one, a_len, lda, x_len,
                            incx, y_len, incy);
  if (!chk.ok()) return fail(chk, nullptr);
  const size_t leny = is_transposed(trans) ? n : m;
  if (leny == 0 || (alpha_zero && beta_one)) return RBLAS_OK;
  hipStream_t st = (hipStream_t)stream;
  const GemvProblem<T> p =
      gemv_problem<T>(cplx, layout, trans, m, n, ar, ai, a, lda, x, incx, br, bi, y, incy);
  const size_t wsb = gemv_workspace_bytes(p);
  void *ws = nullptr;
  if (wsb > 0) RBLAS_HIP_TRY(pool_alloc(&ws, wsb, st), "gemv workspace");
  hipError_t e = launch_gemv<T>(p, ws, st);
  e = pool_release(ws, st, e);
  RBLAS_HIP_TRY(e, "gemv launch");
  return RBLAS_OK;
}

template <typename T>
Rank1Problem<T> rank1_problem(bool cplx, bool conj, int32_t layout, size_t m, size_t n, T ar, T ai,
                              const T *x, size_t incx, const T *y, size_t incy, T *a, size_t lda) {
  const bool row = layout == RBLAS_ROW_MAJOR;
  return Rank1Problem<T>{cplx,          conj,           row, (int64_t)(row ? n : m),
                         (int64_t)(row ? m : n), a, (int64_t)lda, x, (int64_t)incx, y,
                         (int64_t)incy, ar,             ai};
}

template <typename T>
int32_t rank1_host(bool cplx, bool conj, int32_t layout, size_t m, size_t n, T ar, T ai,
                   const T *x, size_t x_len, size_t incx, const T *y, size_t y_len, size_t incy,
                   T *a, size_t a_len, size_t lda) {
  const bool alpha_zero = ar == T(0) && ai == T(0);
  Check chk = ger_validate(layout, m, n, alpha_zero, x_len, incx, y_len, incy, a_len, lda);
  if (!chk.ok()) return fail(chk, nullptr);
  if (alpha_zero || m == 0 || n == 0) return RBLAS_OK;
  const bool row = layout == RBLAS_ROW_MAJOR;
  CtxLease lease;
  if (!lease.get()) return hip_fail(lease.status(), "stream create");
  Ctx &ctx = *lease.get();
  const size_t es = (cplx ? 2 : 1) * sizeof(T);
  HostXfer xf(ctx, es, (m * n + m + n) * es);
  Staged sa, sx, sy;
  sa.plan(row ? n : m, row ? m : n, lda, xf.pinned());
  sx.plan(1, m, incx, xf.pinned());
  sy.plan(1, n, incy, xf.pinned());
  RBLAS_HIP_TRY(ctx.ensure(0, sa.elems * es), "device buffer a");
  RBLAS_HIP_TRY(ctx.ensure(1, sx.elems * es), "device buffer x");
  RBLAS_HIP_TRY(ctx.ensure(2, sy.elems * es), "device buffer y");
  RBLAS_HIP_TRY(xf.reserve(sa, sx, sy), "pinned staging buffer");
  RBLAS_HIP_TRY(xf.in(sa, a, ctx.buf[0]), "copy a");
  RBLAS_HIP_TRY(xf.in(sx, x, ctx.buf[1]), "copy x");
  RBLAS_HIP_TRY(xf.in(sy, y, ctx.buf[2]), "copy y");
  const Rank1Problem<T> p =
      rank1_problem<T>(cplx, conj, layout, m, n, ar, ai, (const T *)ctx.buf[1], sx.dev_ld,
                       (const T *)ctx.buf[2], sy.dev_ld, (T *)ctx.buf[0], sa.dev_ld);
  RBLAS_HIP_TRY(launch_rank1<T>(p, ctx.stream), "rank-1 launch");
  RBLAS_HIP_TRY(xf.out(sa, ctx.buf[0], a), "copy a back");
  RBLAS_HIP_TRY(xf.finish(), "rank-1 update");
  return RBLAS_OK;
}

template <typename T>
int32_t rank1_dev(bool cplx, bool conj, int32_t layout, size_t m, size_t n, T ar, T ai,
                  const T *x, size_t x_len, size_t incx, const T *y, size_t y_len, size_t incy,
                  T *a, size_t a_len, size_t lda, void *stream) {
  const bool alpha_zero = ar == T(0) && ai == T(0);
  Check chk = ger_validate(layout, m, n, alpha_zero, x_len, incx, y_len, incy, a_len, lda);
  if (!chk.ok()) return fail(chk, nullptr);
  if (alpha_zero || m == 0 || n == 0) return RBLAS_OK;
  const Rank1Problem<T> p =
      rank1_problem<T>(cplx, conj, layout, m, n, ar, ai, x, incx, y, incy, a, lda);
  RBLAS_HIP_TRY(launch_rank1<T>(p, (hipStream_t)stream), "rank-1 launch");
  return RBLAS_OK;
}

}  // namespace
}  // namespace rblas

using namespace rblas;

extern "C" {

int32_t rblas_sgemv(int32_t layout, int32_t trans, size_t m, size_t n, float alpha,
                    const float *a, size_t a_len, size_t lda, const float *x, size_t x_len,
                    size_t incx, float beta, float *y, size_t y_len, size_t incy) {
  return gemv_host<float>(false, layout, trans, m, n, alpha, 0.f, a, a_len, lda, x, x_len, incx,
                          beta, 0.f, y, y_len, incy);
}

int32_t rblas_dgemv(int32_t layout, int32_t trans, size_t m, size_t n, double alpha,
                    const double *a, size_t a_len, size_t lda, const double *x, size_t x_len,
                    size_t incx, double beta, double *y, size_t y_len, size_t incy) {
  return gemv_host<double>(false, layout, trans, m, n, alpha, 0.0, a, a_len, lda, x, x_len, incx,
                           beta, 0.0, y, y_len, incy);
}

int32_t rblas_cgemv(int32_t layout, int32_t trans, size_t m, size_t n, rblas_complex32_t alpha,
                    const rblas_complex32_t *a, size_t a_len, size_t lda,
                    const rblas_complex32_t *x, size_t x_len, size_t incx, rblas_complex32_t beta,
                    rblas_complex32_t *y, size_t y_len, size_t incy) {
  return gemv_host<float>(true, layout, trans, m, n, alpha.re, alpha.im, (const float *)a, a_len,
                          lda, (const float *)x, x_len, incx, beta.re, beta.im, (float *)y, y_len,
                          incy);
}

int32_t rblas_zgemv(int32_t layout, int32_t trans, size_t m, size_t n, rblas_complex64_t alpha,
                    const rblas_complex64_t *a, size_t a_len, size_t lda,
                    const rblas_complex64_t *x, size_t x_len, size_t incx, rblas_complex64_t beta,
                    rblas_complex64_t *y, size_t y_len, size_t incy) {
  return gemv_host<double>(true, layout, trans, m, n, alpha.re, alpha.im, (const double *)a,
                           a_len, lda, (const double *)x, x_len, incx, beta.re, beta.im,
                           (double *)y, y_len, incy);
}

int32_t rblas_sgemv_dev(int32_t layout, int32_t trans, size_t m, size_t n, float alpha,
                        const float *a, size_t a_len, size_t lda, const float *x, size_t x_len,
                        size_t incx, float beta, float *y, size_t y_len, size_t incy,
                        void *stream) {
  return gemv_dev<float>(false, layout, trans, m, n, alpha, 0.f, a, a_len, lda, x, x_len, incx,
                         beta, 0.f, y, y_len, incy, stream);
}

int32_t rblas_dgemv_dev(int32_t layout, int32_t trans, size_t m, size_t n, double alpha,
                        const double *a, size_t a_len, size_t lda, const double *x, size_t x_len,
                        size_t incx, double beta, double *y, size_t y_len, size_t incy,
                        void *stream) {
  return gemv_dev<double>(false, layout, trans, m, n, alpha, 0.0, a, a_len, lda, x, x_len, incx,
                          beta, 0.0, y, y_len, incy, stream);
}

int32_t rblas_cgemv_dev(int32_t layout, int32_t trans, size_t m, size_t n,
                        rblas_complex32_t alpha, const rblas_complex32_t *a, size_t a_len,
                        size_t lda, const rblas_complex32_t *x, size_t x_len, size_t incx,
                        rblas_complex32_t beta, rblas_complex32_t *y, size_t y_len, size_t incy,
                        void *stream) {
  return gemv_dev<float>(true, layout, trans, m, n, alpha.re, alpha.im, (const float *)a, a_len,
                         lda, (const float *)x, x_len, incx, beta.re, beta.im, (float *)y, y_len,
                         incy, stream);
}

int32_t rblas_zgemv_dev(int32_t layout, int32_t trans, size_t m, size_t n,
                        rblas_complex64_t alpha, const rblas_complex64_t *a, size_t a_len,
                        size_t lda, const rblas_complex64_t *x, size_t x_len, size_t incx,
                        rblas_complex64_t beta, rblas_complex64_t *y, size_t y_len, size_t incy,
                        void *stream) {
  return gemv_dev<double>(true, layout, trans, m, n, alpha.re, alpha.im, (const double *)a, a_len,
                          lda, (const double *)x, x_len, incx, beta.re, beta.im, (double *)y,
                          y_len, incy, stream);
}

int32_t rblas_sger(int32_t layout, size_t m, size_t n, float alpha, const float *x, size_t x_len,
                   size_t incx, const float *y, size_t y_len, size_t incy, float *a, size_t a_len,
                   size_t lda) {
  return rank1_host<float>(false, false, layout, m, n, alpha, 0.f, x, x_len, incx, y, y_len, incy,
                           a, a_len, lda);
}

int32_t rblas_dger(int32_t layout, size_t m, size_t n, double alpha, const double *x,
                   size_t x_len, size_t incx, const double *y, size_t y_len, size_t incy,
                   double *a, size_t a_len, size_t lda) {
  return rank1_host<double>(false, false, layout, m, n, alpha, 0.0, x, x_len, incx, y, y_len,
                            incy, a, a_len, lda);
}

int32_t rblas_sger_dev(int32_t layout, size_t m, size_t n, float alpha, const float *x,
                       size_t x_len, size_t incx, const float *y, size_t y_len, size_t incy,
                       float *a, size_t a_len, size_t lda, void *stream) {
  return rank1_dev<float>(false, false, layout, m, n, alpha, 0.f, x, x_len, incx, y, y_len, incy,
                          a, a_len, lda, stream);
}

int32_t rblas_dger_dev(int32_t layout, size_t m, size_t n, double alpha, const double *x,
                       size_t x_len, size_t incx, const double *y, size_t y_len, size_t incy,
                       double *a, size_t a_len, size_t lda, void *stream) {
  return rank1_dev<double>(false, false, layout, m, n, alpha, 0.0, x, x_len, incx, y, y_len, incy,
                           a, a_len, lda, stream);
}

#define RBLAS_CRANK1(NAME, CT, T, CONJ)                                                          \
  int32_t rblas_##NAME(int32_t layout, size_t m, size_t n, CT alpha, const CT *x, size_t x_len, \
                       size_t incx, const CT *y, size_t y_len, size_t incy, CT *a, size_t a_len, \
                       size_t lda) {                                                            \
    return rank1_host<T>(true, CONJ, layout, m, n, alpha.re, alpha.im, (const T *)x, x_len,     \
                         incx, (const T *)y, y_len, incy, (T *)a, a_len, lda);                  \
  }                                                                                             \
  int32_t rblas_##NAME##_dev(int32_t layout, size_t m, size_t n, CT alpha, const CT *x,         \
                             size_t x_len, size_t incx, const CT *y, size_t y_len, size_t incy, \
                             CT *a, size_t a_len, size_t lda, void *stream) {                   \
    return rank1_dev<T>(true, CONJ, layout, m, n, alpha.re, alpha.im, (const T *)x, x_len,      \
                        incx, (const T *)y, y_len, incy, (T *)a, a_len, lda, stream);           \
  }

RBLAS_CRANK1(cgeru, rblas_complex32_t, float, false)
RBLAS_CRANK1(zgeru, rblas_complex64_t, double, false)
RBLAS_CRANK1(cgerc, rblas_complex32_t, float, true)
RBLAS_CRANK1(zgerc, rblas_complex64_t, double, true)

#undef RBLAS_CRANK1

}  // extern "C"
