// C-ABI Level-1 entry points: level1::dot / nrm2 / axpy (level1.rs:17-123),
// host slices and HBM-resident forms.
#include "kernels.hpp"
#include "runtime.hpp"

namespace rblas {
namespace {

// Largest index (n-1)*inc touched by a strided loop; false if it overflows
// (rblas would then panic on the first out-of-range index).
bool last_index(size_t n, size_t inc, size_t &out) {
  return !__builtin_mul_overflow(n - 1, inc, &out);
}

// rblas's checks for a vector argument in the strided branch
// (level1.rs:31-33, 53-55, 100-101).
Check strided_check(const char *fn, char which, size_t n, size_t inc, size_t len) {
  size_t hi;
  if (!last_index(n, inc, hi) || hi >= len)
    return panic(std::string(fn) + ": " + which + " out of bounds");
  return Check{};
}

// axpy (level1.rs:17-41): nothing to do when n == 0 or alpha == 0 (before
// any check); the contiguous branch asserts both slices hold n elements,
// the strided one checks strides then bounds.
template <typename T>
Check axpy_check(size_t n, T alpha, size_t x_len, size_t incx, size_t y_len, size_t incy,
                 bool &noop) {
  noop = n == 0 || alpha == T(0);
  if (noop) return Check{};
  if (incx == 1 && incy == 1) {
    if (!(x_len >= n && y_len >= n)) return panic("axpy: slice shorter than n");
    return Check{};
  }
  if (!(incx > 0 && incy > 0)) return panic("axpy: stride must be positive");
  Check c = strided_check("axpy", 'x', n, incx, x_len);
  if (!c.ok()) return c;
  return strided_check("axpy", 'y', n, incy, y_len);
}

// dot (level1.rs:45-66): 0 when n == 0.
Check dot_check(size_t n, size_t x_len, size_t incx, size_t y_len, size_t incy) {
  if (n == 0) return Check{};
  if (incx == 1 && incy == 1) {
    if (!(x_len >= n && y_len >= n)) return panic("dot: slice shorter than n");
    return Check{};
  }
  if (!(incx > 0 && incy > 0)) return panic("dot: stride must be positive");
  Check c = strided_check("dot", 'x', n, incx, x_len);
  if (!c.ok()) return c;
  return strided_check("dot", 'y', n, incy, y_len);
}

// nrm2 (level1.rs:96-101): 0 when n == 0.
Check nrm2_check(size_t n, size_t x_len, size_t incx) {
  if (n == 0) return Check{};
  if (!(incx > 0)) return panic("nrm2: stride must be positive");
  return strided_check("nrm2", 'x', n, incx, x_len);
}

template <typename T>
int32_t dot_host(size_t n, const T *x, size_t x_len, size_t incx, const T *y, size_t y_len,
                 size_t incy, T *result) {
  Check c = dot_check(n, x_len, incx, y_len, incy);
  if (!c.ok()) return fail(c, nullptr);
  if (!result) return fail(panic("dot: null result pointer"), nullptr);
  if (n == 0) {
    *result = T(0);
    return RBLAS_OK;
  }
  CtxLease lease;
  if (!lease.get()) return hip_fail(lease.status(), "stream create");
  Ctx &ctx = *lease.get();
  // A vector of stride inc is staged as a 1 x n operand with ld = inc.
  HostXfer xf(ctx, sizeof(T), 2 * n * sizeof(T));
  Staged sx, sy;
  sx.plan(1, n, incx, xf.pinned());
  sy.plan(1, n, incy, xf.pinned());
  RBLAS_HIP_TRY(ctx.ensure(0, sx.elems * sizeof(T)), "device buffer x");
  RBLAS_HIP_TRY(ctx.ensure(1, sy.elems * sizeof(T)), "device buffer y");
  RBLAS_HIP_TRY(ctx.ensure(3, level1_scratch_bytes() + 64), "device scratch");
  RBLAS_HIP_TRY(xf.reserve(sx, sy), "pinned staging buffer");
  hipStream_t st = ctx.stream;
  RBLAS_HIP_TRY(xf.in(sx, x, ctx.buf[0]), "copy x");
  RBLAS_HIP_TRY(xf.in(sy, y, ctx.buf[1]), "copy y");
  T *dres = (T *)((char *)ctx.buf[3] + level1_scratch_bytes());
  RBLAS_HIP_TRY(launch_dot<T>((int64_t)n, (const T *)ctx.buf[0], (int64_t)sx.dev_ld,
                              (const T *)ctx.buf[1], (int64_t)sy.dev_ld, dres, ctx.buf[3], st),
                "dot launch");
  RBLAS_HIP_TRY(hipMemcpyAsync(result, dres, sizeof(T), hipMemcpyDeviceToHost, st), "copy result");
  RBLAS_HIP_TRY(xf.finish(), "dot");
  return RBLAS_OK;
}

template <typename T>
int32_t nrm2_host(size_t n, const T *x, size_t x_len, size_t incx, T *result) {
  Check c = nrm2_check(n, x_len, incx);
  if (!c.ok()) return fail(c, nullptr);
  if (!result) return fail(panic("nrm2: null result pointer"), nullptr);
  if (n == 0) {
    *result = T(0);
    return RBLAS_OK;
  }
  CtxLease lease;
  if (!lease.get()) return hip_fail(lease.status(), "stream create");
  Ctx &ctx = *lease.get();
  HostXfer xf(ctx, sizeof(T), n * sizeof(T));
  Staged sx;
  sx.plan(1, n, incx, xf.pinned());
  RBLAS_HIP_TRY(ctx.ensure(0, sx.elems * sizeof(T)), "device buffer x");
  RBLAS_HIP_TRY(ctx.ensure(3, level1_scratch_bytes() + 64), "device scratch");
  RBLAS_HIP_TRY(xf.reserve(sx), "pinned staging buffer");
  hipStream_t st = ctx.stream;
  RBLAS_HIP_TRY(xf.in(sx, x, ctx.buf[0]), "copy x");
  T *dres = (T *)((char *)ctx.buf[3] + level1_scratch_bytes());
  RBLAS_HIP_TRY(launch_nrm2<T>((int64_t)n, (const T *)ctx.buf[0], (int64_t)sx.dev_ld, dres,
                               ctx.buf[3], st),
                "nrm2 launch");
  RBLAS_HIP_TRY(hipMemcpyAsync(result, dres, sizeof(T), hipMemcpyDeviceToHost, st), "copy result");
  RBLAS_HIP_TRY(xf.finish(), "nrm2");
  return RBLAS_OK;
}

template <typename T>
int32_t axpy_host(size_t n, T alpha, const T *x, size_t x_len, size_t incx, T *y, size_t y_len,
                  size_t incy) {
  bool noop;
  Check c = axpy_check(n, alpha, x_len, incx, y_len, incy, noop);
  if (!c.ok()) return fail(c, nullptr);
  if (noop) return RBLAS_OK;
  CtxLease lease;
  if (!lease.get()) return hip_fail(lease.status(), "stream create");
  Ctx &ctx = *lease.get();
  HostXfer xf(ctx, sizeof(T), 2 * n * sizeof(T));
  Staged sx, sy;
  sx.plan(1, n, incx, xf.pinned());
  sy.plan(1, n, incy, xf.pinned());
  RBLAS_HIP_TRY(ctx.ensure(0, sx.elems * sizeof(T)), "device buffer x");
  RBLAS_HIP_TRY(ctx.ensure(1, sy.elems * sizeof(T)), "device buffer y");
  RBLAS_HIP_TRY(xf.reserve(sx, sy), "pinned staging buffer");
  hipStream_t st = ctx.stream;
  RBLAS_HIP_TRY(xf.in(sx, x, ctx.buf[0]), "copy x");
  RBLAS_HIP_TRY(xf.in(sy, y, ctx.buf[1]), "copy y");
  RBLAS_HIP_TRY(launch_axpy<T>((int64_t)n, alpha, (const T *)ctx.buf[0], (int64_t)sx.dev_ld,
                               (T *)ctx.buf[1], (int64_t)sy.dev_ld, st),
                "axpy launch");
  RBLAS_HIP_TRY(xf.out(sy, ctx.buf[1], y), "copy y");
  RBLAS_HIP_TRY(xf.finish(), "axpy");
  return RBLAS_OK;
}

template <typename T>
int32_t dot_dev(size_t n, const T *x, size_t x_len, size_t incx, const T *y, size_t y_len,
                size_t incy, T *result, void *stream) {
  Check c = dot_check(n, x_len, incx, y_len, incy);
  if (!c.ok()) return fail(c, nullptr);
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) {
    // A store kernel, not hipMemsetAsync: callers may capture this call in
    // a HIP graph, where memset nodes are avoided.
    RBLAS_HIP_TRY(launch_scale<T>(1, 1, T(0), result, 1, st), "dot zero result");
    return RBLAS_OK;
  }
  void *scratch = nullptr;
  RBLAS_HIP_TRY(pool_alloc(&scratch, level1_scratch_bytes(), st), "dot scratch");
  hipError_t e = launch_dot<T>((int64_t)n, x, (int64_t)incx, y, (int64_t)incy, result, scratch, st);
  e = pool_release(scratch, st, e);
  RBLAS_HIP_TRY(e, "dot launch");
  return RBLAS_OK;
}

template <typename T>
int32_t nrm2_dev(size_t n, const T *x, size_t x_len, size_t incx, T *result, void *stream) {
  Check c = nrm2_check(n, x_len, incx);
  if (!c.ok()) return fail(c, nullptr);
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) {
    RBLAS_HIP_TRY(launch_scale<T>(1, 1, T(0), result, 1, st), "nrm2 zero result");
    return RBLAS_OK;
  }
  void *scratch = nullptr;
  RBLAS_HIP_TRY(pool_alloc(&scratch, level1_scratch_bytes(), st), "nrm2 scratch");
  hipError_t e = launch_nrm2<T>((int64_t)n, x, (int64_t)incx, result, scratch, st);
  e = pool_release(scratch, st, e);
  RBLAS_HIP_TRY(e, "nrm2 launch");
  return RBLAS_OK;
}

template <typename T>
int32_t axpy_dev(size_t n, T alpha, const T *x, size_t x_len, size_t incx, T *y, size_t y_len,
                 size_t incy, void *stream) {
  bool noop;
  Check c = axpy_check(n, alpha, x_len, incx, y_len, incy, noop);
  if (!c.ok()) return fail(c, nullptr);
  if (noop) return RBLAS_OK;
  RBLAS_HIP_TRY(launch_axpy<T>((int64_t)n, alpha, x, (int64_t)incx, y, (int64_t)incy,
                               (hipStream_t)stream),
                "axpy launch");
  return RBLAS_OK;
}

}  // namespace
}  // namespace rblas

using namespace rblas;

extern "C" {

int32_t rblas_sdot(size_t n, const float *x, size_t x_len, size_t incx, const float *y,
                   size_t y_len, size_t incy, float *result) {
  return dot_host<float>(n, x, x_len, incx, y, y_len, incy, result);
}
int32_t rblas_ddot(size_t n, const double *x, size_t x_len, size_t incx, const double *y,
                   size_t y_len, size_t incy, double *result) {
  return dot_host<double>(n, x, x_len, incx, y, y_len, incy, result);
}
int32_t rblas_snrm2(size_t n, const float *x, size_t x_len, size_t incx, float *result) {
  return nrm2_host<float>(n, x, x_len, incx, result);
}
int32_t rblas_dnrm2(size_t n, const double *x, size_t x_len, size_t incx, double *result) {
  return nrm2_host<double>(n, x, x_len, incx, result);
}
int32_t rblas_saxpy(size_t n, float alpha, const float *x, size_t x_len, size_t incx, float *y,
                    size_t y_len, size_t incy) {
  return axpy_host<float>(n, alpha, x, x_len, incx, y, y_len, incy);
}
int32_t rblas_daxpy(size_t n, double alpha, const double *x, size_t x_len, size_t incx,
                    double *y, size_t y_len, size_t incy) {
  return axpy_host<double>(n, alpha, x, x_len, incx, y, y_len, incy);
}
int32_t rblas_sdot_dev(size_t n, const float *x, size_t x_len, size_t incx, const float *y,
                       size_t y_len, size_t incy, float *result, void *stream) {
  return dot_dev<float>(n, x, x_len, incx, y, y_len, incy, result, stream);
}
int32_t rblas_ddot_dev(size_t n, const double *x, size_t x_len, size_t incx, const double *y,
                       size_t y_len, size_t incy, double *result, void *stream) {
  return dot_dev<double>(n, x, x_len, incx, y, y_len, incy, result, stream);
}
int32_t rblas_snrm2_dev(size_t n, const float *x, size_t x_len, size_t incx, float *result,
                        void *stream) {
  return nrm2_dev<float>(n, x, x_len, incx, result, stream);
}
int32_t rblas_dnrm2_dev(size_t n, const double *x, size_t x_len, size_t incx, double *result,
                        void *stream) {
  return nrm2_dev<double>(n, x, x_len, incx, result, stream);
}
int32_t rblas_saxpy_dev(size_t n, float alpha, const float *x, size_t x_len, size_t incx,
                        float *y, size_t y_len, size_t incy, void *stream) {
  return axpy_dev<float>(n, alpha, x, x_len, incx, y, y_len, incy, stream);
}
int32_t rblas_daxpy_dev(size_t n, double alpha, const double *x, size_t x_len, size_t incx,
                        double *y, size_t y_len, size_t incy, void *stream) {
  return axpy_dev<double>(n, alpha, x, x_len, incx, y, y_len, incy, stream);
}

}  // extern "C"
