// Level-1 dot / nrm2 / axpy on gfx950. All three are HBM-bound: one pass
// over the vectors (16-byte vector loads on the contiguous, aligned path:
// for dot and axpy each workgroup streams one contiguous chunk, four
// block-strides in flight; nrm2 and strided vectors use grid-stride loops),
// wavefront reductions through lane shuffles, and
// a second single-block kernel that folds the per-block partials in a fixed
// order (results do not depend on scheduling).
//
// dot   (level1.rs:45-66): sum of x_i*y_i, accumulated in T per lane.
// axpy  (level1.rs:17-41): y_i = fma(alpha, x_i, y_i), as rblas computes it.
// nrm2  (level1.rs:96-123): rblas uses the scaled sum of squares (one
//       division per element). Here: f32 sums x_i^2 in f64, where no
//       overflow or underflow is possible, and rounds sqrt to f32; f64 uses
//       Blue's three-accumulator scaling (the constants and final combination
//       of Anderson, "Algorithm 978: Safe scaling in the Level 1 BLAS", ACM
//       TOMS 44(1), 2017): squares of small, medium and large magnitudes are
//       summed in separately scaled accumulators, which add across threads.
//       Non-finite input gives rblas's result: its running scale / ssq pair
//       turns any NaN into NaN, and a second infinite entry makes inf / inf
//       (NaN) at level1.rs:116, while a single infinity gives +Inf. So the
//       partial kernels also count infinite entries, and the finish kernels
//       return NaN when the count reaches two (a NaN entry already
//       propagates through the sums).
#include "kernels.hpp"
#include "fma.hpp"

namespace rblas {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 2048;
constexpr int kUnroll = 4;  // block-strides per operand in flight per thread

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum of v (all threads return the total).
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T t = T(0);
#pragma unroll
  for (int i = 0; i < kThreads / 64; ++i) t += sh[i];
  __syncthreads();
  return t;
}

template <typename T>
struct Vec16;
template <>
struct Vec16<float> { typedef float type __attribute__((ext_vector_type(4))); static constexpr int n = 4; };
template <>
struct Vec16<double> { typedef double type __attribute__((ext_vector_type(2))); static constexpr int n = 2; };

__device__ __forceinline__ bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// This workgroup's contiguous share [c0, c1) of nv items: gridDim.x chunks
// of whole block-strides (kThreads items), the last one shorter.
__device__ __forceinline__ void chunk_of(int64_t nv, int64_t &c0, int64_t &c1) {
  const int64_t strides = (nv + kThreads - 1) / kThreads;
  const int64_t per = (strides + gridDim.x - 1) / gridDim.x * kThreads;
  c0 = blockIdx.x * per;
  c1 = c0 + per < nv ? c0 + per : nv;
}

// ------------------------------------------------------------------- dot
template <typename T>
__global__ __launch_bounds__(kThreads) void dot_partial(int64_t n, const T *x, int64_t incx,
                                                        const T *y, int64_t incy, T *partial) {
  __shared__ T sh[kThreads / 64];
  T acc0 = T(0), acc1 = T(0);
  const int64_t tid = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  if (incx == 1 && incy == 1 && aligned16(x) && aligned16(y)) {
    typedef typename Vec16<T>::type V;
    constexpr int W = Vec16<T>::n;
    const int64_t nv = n / W;
    const V *xv = (const V *)x, *yv = (const V *)y;
    // Each workgroup streams one contiguous chunk of the vectors, kUnroll
    // block-strides (kUnroll x 4 KiB per operand) in flight per trip.
    int64_t c0, c1;
    chunk_of(nv, c0, c1);
    int64_t i = c0 + threadIdx.x;
    for (; i + (kUnroll - 1) * kThreads < c1; i += kUnroll * kThreads) {
      V a[kUnroll], b[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        a[u] = xv[i + u * kThreads];
        b[u] = yv[i + u * kThreads];
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u)
#pragma unroll
        for (int e = 0; e < W; ++e)
          (e & 1 ? acc1 : acc0) = fma_t(a[u][e], b[u][e], e & 1 ? acc1 : acc0);
    }
    for (; i < c1; i += kThreads) {
      const V a = xv[i], b = yv[i];
#pragma unroll
      for (int e = 0; e < W; ++e) (e & 1 ? acc1 : acc0) = fma_t(a[e], b[e], e & 1 ? acc1 : acc0);
    }
    for (int64_t r = nv * W + tid; r < n; r += stride) acc0 = fma_t(x[r], y[r], acc0);
  } else {
    for (int64_t i = tid; i < n; i += stride) acc0 = fma_t(x[i * incx], y[i * incy], acc0);
  }
  const T s = block_sum(acc0 + acc1, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void sum_partials(int nparts, const T *partial, T *out) {
  __shared__ T sh[kThreads / 64];
  T acc = T(0);
  for (int i = threadIdx.x; i < nparts; i += kThreads) acc += partial[i];
  const T s = block_sum(acc, sh);
  if (threadIdx.x == 0) *out = s;
}

// ------------------------------------------------------------------ axpy
template <typename T>
__global__ __launch_bounds__(kThreads) void axpy_kernel(int64_t n, T alpha, const T *x,
                                                        int64_t incx, T *y, int64_t incy) {
  const int64_t tid = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  if (incx == 1 && incy == 1 && aligned16(x) && aligned16(y)) {
    typedef typename Vec16<T>::type V;
    constexpr int W = Vec16<T>::n;
    const int64_t nv = n / W;
    const V *xv = (const V *)x;
    V *yv = (V *)y;
    int64_t c0, c1;
    chunk_of(nv, c0, c1);
    int64_t i = c0 + threadIdx.x;
    for (; i + (kUnroll - 1) * kThreads < c1; i += kUnroll * kThreads) {
      V a[kUnroll], b[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        a[u] = xv[i + u * kThreads];
        b[u] = yv[i + u * kThreads];
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
        for (int e = 0; e < W; ++e) b[u][e] = fma_t(alpha, a[u][e], b[u][e]);
        yv[i + u * kThreads] = b[u];
      }
    }
    for (; i < c1; i += kThreads) {
      const V a = xv[i];
      V b = yv[i];
#pragma unroll
      for (int e = 0; e < W; ++e) b[e] = fma_t(alpha, a[e], b[e]);
      yv[i] = b;
    }
    for (int64_t r = nv * W + tid; r < n; r += stride) y[r] = fma_t(alpha, x[r], y[r]);
  } else {
    for (int64_t i = tid; i < n; i += stride)
      y[i * incy] = fma_t(alpha, x[i * incx], y[i * incy]);
  }
}

// ------------------------------------------------------------------ nrm2
// 1 for an infinite entry, else 0 (NaN compares unequal to infinity).
__device__ __forceinline__ double inf_count(double v) {
  return __builtin_fabs(v) == __builtin_inf() ? 1.0 : 0.0;
}

// f32: sum of squares in f64. Partials: [2 b] = sum of squares, [2 b + 1] =
// number of infinite entries.
__device__ __forceinline__ void sq_add(double &acc, double &ninf, double v) {
  acc = __builtin_fma(v, v, acc);
  ninf += inf_count(v);
}

__global__ __launch_bounds__(kThreads) void snrm2_partial(int64_t n, const float *x, int64_t incx,
                                                          double *partial) {
  __shared__ double sh[kThreads / 64];
  double acc = 0.0, ninf = 0.0;
  const int64_t tid = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  if (incx == 1 && aligned16(x)) {
    typedef float V __attribute__((ext_vector_type(4)));
    const int64_t nv = n / 4;
    const V *xv = (const V *)x;
    for (int64_t i = tid; i < nv; i += stride) {
      const V a = xv[i];
#pragma unroll
      for (int e = 0; e < 4; ++e) sq_add(acc, ninf, (double)a[e]);
    }
    for (int64_t i = nv * 4 + tid; i < n; i += stride) sq_add(acc, ninf, (double)x[i]);
  } else {
    for (int64_t i = tid; i < n; i += stride) sq_add(acc, ninf, (double)x[i * incx]);
  }
  const double s = block_sum(acc, sh);
  const double c = block_sum(ninf, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s;
    partial[2 * blockIdx.x + 1] = c;
  }
}

__global__ __launch_bounds__(kThreads) void snrm2_finish(int nparts, const double *partial,
                                                         float *out) {
  __shared__ double sh[kThreads / 64];
  double acc = 0.0, ninf = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kThreads) {
    acc += partial[2 * i];
    ninf += partial[2 * i + 1];
  }
  const double s = block_sum(acc, sh);
  const double c = block_sum(ninf, sh);
  if (threadIdx.x == 0) *out = c >= 2.0 ? __builtin_nanf("") : (float)__builtin_sqrt(s);
}

// f64: Blue's scaling. Thresholds and scale factors for IEEE double:
//   tsml = 2^-511, tbig = 2^486, ssml = 2^537, sbig = 2^-538.
struct Blue {
  double sml, med, big;
};
constexpr double kTsml = 0x1p-511, kTbig = 0x1p486, kSsml = 0x1p537, kSbig = 0x1p-538;

__device__ __forceinline__ void blue_add(Blue &b, double &ninf, double v) {
  const double a = __builtin_fabs(v);
  ninf += inf_count(v);
  if (a > kTbig) {
    const double t = a * kSbig;
    b.big = __builtin_fma(t, t, b.big);
  } else if (a < kTsml) {
    const double t = a * kSsml;
    b.sml = __builtin_fma(t, t, b.sml);
  } else {
    b.med = __builtin_fma(a, a, b.med);  // NaN lands here and propagates
  }
}

__global__ __launch_bounds__(kThreads) void dnrm2_partial(int64_t n, const double *x,
                                                          int64_t incx, double *partial) {
  __shared__ double sh[kThreads / 64];
  Blue b{0.0, 0.0, 0.0};
  double ninf = 0.0;
  const int64_t tid = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  if (incx == 1 && aligned16(x)) {
    typedef double V __attribute__((ext_vector_type(2)));
    const int64_t nv = n / 2;
    const V *xv = (const V *)x;
    for (int64_t i = tid; i < nv; i += stride) {
      const V a = xv[i];
      blue_add(b, ninf, a[0]);
      blue_add(b, ninf, a[1]);
    }
    for (int64_t i = nv * 2 + tid; i < n; i += stride) blue_add(b, ninf, x[i]);
  } else {
    for (int64_t i = tid; i < n; i += stride) blue_add(b, ninf, x[i * incx]);
  }
  const double s0 = block_sum(b.sml, sh);
  const double s1 = block_sum(b.med, sh);
  const double s2 = block_sum(b.big, sh);
  const double s3 = block_sum(ninf, sh);
  if (threadIdx.x == 0) {
    partial[4 * blockIdx.x] = s0;
    partial[4 * blockIdx.x + 1] = s1;
    partial[4 * blockIdx.x + 2] = s2;
    partial[4 * blockIdx.x + 3] = s3;
  }
}

__global__ __launch_bounds__(kThreads) void dnrm2_finish(int nparts, const double *partial,
                                                         double *out) {
  __shared__ double sh[kThreads / 64];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kThreads) {
    a0 += partial[4 * i];
    a1 += partial[4 * i + 1];
    a2 += partial[4 * i + 2];
    a3 += partial[4 * i + 3];
  }
  double asml = block_sum(a0, sh), amed = block_sum(a1, sh), abig = block_sum(a2, sh);
  const double ninf = block_sum(a3, sh);
  if (threadIdx.x != 0) return;
  if (ninf >= 2.0) {  // rblas: inf / inf in the scaled sum of squares
    *out = __builtin_nan("");
    return;
  }
  // Combine (Anderson 2017): the large accumulator dominates when present;
  // otherwise the small one is merged with the medium one without overflow.
  double scl, sumsq;
  if (abig > 0.0) {
    if (amed > 0.0 || amed != amed) abig += (amed * kSbig) * kSbig;
    scl = 1.0 / kSbig;
    sumsq = abig;
  } else if (asml > 0.0) {
    if (amed > 0.0 || amed != amed) {
      amed = __builtin_sqrt(amed);
      asml = __builtin_sqrt(asml) / kSsml;
      const double ymin = asml > amed ? amed : asml, ymax = asml > amed ? asml : amed;
      scl = 1.0;
      sumsq = ymax * ymax * (1.0 + (ymin / ymax) * (ymin / ymax));
    } else {
      scl = 1.0 / kSsml;
      sumsq = asml;
    }
  } else {
    scl = 1.0;
    sumsq = amed;
  }
  *out = scl * __builtin_sqrt(sumsq);
}

int blocks_for(int64_t n) {
  int64_t b = (n + kThreads * 8 - 1) / (kThreads * 8);
  if (b < 1) b = 1;
  if (b > kMaxBlocks) b = kMaxBlocks;
  return (int)b;
}

}  // namespace

size_t level1_scratch_bytes() { return (size_t)kMaxBlocks * 4 * sizeof(double); }

template <typename T>
hipError_t launch_dot(int64_t n, const T *x, int64_t incx, const T *y, int64_t incy, T *out,
                      void *scratch, hipStream_t st) {
  const int b = blocks_for(n);
  hipLaunchKernelGGL((dot_partial<T>), dim3(b), dim3(kThreads), 0, st, n, x, incx, y, incy,
                     (T *)scratch);
  hipLaunchKernelGGL((sum_partials<T>), dim3(1), dim3(kThreads), 0, st, b, (const T *)scratch, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_axpy(int64_t n, T alpha, const T *x, int64_t incx, T *y, int64_t incy,
                       hipStream_t st) {
  hipLaunchKernelGGL((axpy_kernel<T>), dim3(blocks_for(n)), dim3(kThreads), 0, st, n, alpha, x,
                     incx, y, incy);
  return hipGetLastError();
}

template <>
hipError_t launch_nrm2<float>(int64_t n, const float *x, int64_t incx, float *out, void *scratch,
                              hipStream_t st) {
  const int b = blocks_for(n);
  hipLaunchKernelGGL(snrm2_partial, dim3(b), dim3(kThreads), 0, st, n, x, incx, (double *)scratch);
  hipLaunchKernelGGL(snrm2_finish, dim3(1), dim3(kThreads), 0, st, b, (const double *)scratch, out);
  return hipGetLastError();
}

template <>
hipError_t launch_nrm2<double>(int64_t n, const double *x, int64_t incx, double *out,
                               void *scratch, hipStream_t st) {
  const int b = blocks_for(n);
  hipLaunchKernelGGL(dnrm2_partial, dim3(b), dim3(kThreads), 0, st, n, x, incx, (double *)scratch);
  hipLaunchKernelGGL(dnrm2_finish, dim3(1), dim3(kThreads), 0, st, b, (const double *)scratch, out);
  return hipGetLastError();
}

template hipError_t launch_dot<float>(int64_t, const float *, int64_t, const float *, int64_t,
                                      float *, void *, hipStream_t);
template hipError_t launch_dot<double>(int64_t, const double *, int64_t, const double *, int64_t,
                                       double *, void *, hipStream_t);
template hipError_t launch_axpy<float>(int64_t, float, const float *, int64_t, float *, int64_t,
                                       hipStream_t);
template hipError_t launch_axpy<double>(int64_t, double, const double *, int64_t, double *,
                                        int64_t, hipStream_t);

}  // namespace rblas
