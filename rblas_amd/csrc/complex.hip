// Complex gemm (C := alpha op(A) op(B) + beta C, interleaved Complex<T>,
// column-major) by the 4M decomposition of complex3.rs:27-125, on the device:
//
//   P_re = Ar Br - Ai Bi = [Ar | Ai] [Br ; -Bi]
//   P_im = Ar Bi + Ai Br = [Ar | Ai] [Bi ;  Br]
//
// cplx_split_a writes op(A) (transposition and conjugation folded in, as
// complex3::split_op does) as the real m x 2k matrix [Ar | Ai]; cplx_split_b
// writes op(B) as the two real 2k x n matrices above. Two calls of the real
// MFMA gemm (k concatenated, so 8mnk real flops as in 4M) produce P_re and
// P_im, and cplx_combine applies C = alpha*P + (beta == 0 ? 0 : beta*C) with
// rblas's complex multiply (complex.rs:101-111: re = fma(a.re, b.re,
// -(a.im*b.im)), im = fma(a.re, b.im, a.im*b.re)).
#include "kernels.hpp"
#include "fma.hpp"

namespace rblas {
namespace {

// Element (r, c) of op(X) for a stored column-major complex matrix.
template <typename T>
__device__ __forceinline__ void op_elem(const T *x, int64_t ld, bool tr, bool cj, int64_t r,
                                        int64_t c, T &re, T &im) {
  const int64_t idx = tr ? c + r * ld : r + c * ld;
  re = x[2 * idx];
  im = x[2 * idx + 1];
  if (cj) im = -im;
}

// [Ar | Ai]: out[i + p*m] = Re, out[i + (k+p)*m] = Im, for op(A) m x k.
template <typename T>
__global__ void cplx_split_a(const T *a, int64_t lda, bool tr, bool cj, int64_t m, int64_t k,
                             T *out) {
  const int64_t total = m * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = e / m, i = e - p * m;
    T re, im;
    op_elem(a, lda, tr, cj, i, p, re, im);
    out[i + p * m] = re;
    out[i + (k + p) * m] = im;
  }
}

// [Br ; -Bi] and [Bi ; Br], each 2k x n with leading dimension 2k.
template <typename T>
__global__ void cplx_split_b(const T *b, int64_t ldb, bool tr, bool cj, int64_t k, int64_t n,
                             T *b1, T *b2) {
  const int64_t total = k * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / k, p = e - j * k;
    T re, im;
    op_elem(b, ldb, tr, cj, p, j, re, im);
    const int64_t col = j * 2 * k;
    b1[col + p] = re;
    b1[col + k + p] = -im;
    b2[col + p] = im;
    b2[col + k + p] = re;
  }
}

// C = alpha*P + (beta == 0 ? 0 : beta*C), P = pr + i*pi (m x n, ld m).
template <typename T>
__global__ void cplx_combine(int64_t m, int64_t n, T ar, T ai, const T *pr, const T *pi, T br,
                             T bi, T *c, int64_t ldc) {
  const int64_t total = m * n;
  const bool beta_zero = br == T(0) && bi == T(0);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / m, i = e - j * m;
    const T p_re = pr[e], p_im = pi[e];
    T s_re = T(0), s_im = T(0);
    T *cp = c + 2 * (i + j * ldc);
    if (!beta_zero) {
      const T c_re = cp[0], c_im = cp[1];
      s_re = fma_t(br, c_re, -(bi * c_im));
      s_im = fma_t(br, c_im, bi * c_re);
    }
    const T q_re = fma_t(ar, p_re, -(ai * p_im));
    const T q_im = fma_t(ar, p_im, ai * p_re);
    cp[0] = q_re + s_re;
    cp[1] = q_im + s_im;
  }
}

unsigned grid_for(int64_t total) {
  int64_t b = (total + 255) / 256;
  if (b > 16384) b = 16384;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

size_t cgemm_workspace_elems(int64_t m, int64_t n, int64_t k) {
  return (size_t)(2 * m * k + 4 * k * n + 2 * m * n);
}

template <typename T>
hipError_t launch_cgemm(const CgemmProblem<T> &p, T *ws, hipStream_t st) {
  const int64_t m = p.m, n = p.n, k = p.k;
  T *ap = ws;                 // m x 2k
  T *b1 = ap + 2 * m * k;     // 2k x n
  T *b2 = b1 + 2 * k * n;     // 2k x n
  T *pr = b2 + 2 * k * n;     // m x n
  T *pi = pr + m * n;         // m x n
  if (k > 0) {
    hipLaunchKernelGGL((cplx_split_a<T>), dim3(grid_for(m * k)), dim3(256), 0, st, p.a, p.lda,
                       p.ta != 0, p.ta == 2, m, k, ap);
    hipLaunchKernelGGL((cplx_split_b<T>), dim3(grid_for(k * n)), dim3(256), 0, st, p.b, p.ldb,
                       p.tb != 0, p.tb == 2, k, n, b1, b2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    GemmProblem<T> g1{false, false, m, n, 2 * k, T(1), T(0), ap, m, b1, 2 * k, pr, m};
    e = launch_gemm<T>(g1, st);
    if (e != hipSuccess) return e;
    GemmProblem<T> g2{false, false, m, n, 2 * k, T(1), T(0), ap, m, b2, 2 * k, pi, m};
    e = launch_gemm<T>(g2, st);
    if (e != hipSuccess) return e;
  } else {
    // k == 0: rblas's real gemms leave P = 0 (beta 0 stores zeros). pr and
    // pi are adjacent m x n blocks: clear them as one m x 2n matrix.
    hipError_t e = launch_scale<T>(m, 2 * n, T(0), pr, m, st);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((cplx_combine<T>), dim3(grid_for(m * n)), dim3(256), 0, st, m, n, p.alpha_re,
                     p.alpha_im, (const T *)pr, (const T *)pi, p.beta_re, p.beta_im, p.c, p.ldc);
  return hipGetLastError();
}

template hipError_t launch_cgemm<float>(const CgemmProblem<float> &, float *, hipStream_t);
template hipError_t launch_cgemm<double>(const CgemmProblem<double> &, double *, hipStream_t);

}  // namespace rblas
