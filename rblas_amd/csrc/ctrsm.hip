// Complex triangular solve (complex3::trsm, complex3.rs:252-349): op(A) X =
// alpha B (Left) or X op(A) = alpha B (Right), in place in B, op = None,
// Trans or ConjTrans.
//
//   1. B := alpha B when alpha != 1 (rblas's complex multiply, every entry).
//   2. The diagonal range is solved by recursive halving: solve the half that
//      comes first in substitution order, subtract its contribution from the
//      other half with one 4M complex gemm (alpha -1, beta 1), solve the other
//      half.
//   3. A range of up to 64 is solved by substitution, as rblas does it
//      (s = s - a*x, then cdiv(s, a_ii) = s conj(a_ii) / norm_sqr(a_ii)), in
//      f64: one thread per right-hand side (a column of B for Left, a row for
//      Right), with the block of op(A) and the right-hand sides in LDS. A
//      non-finite B entry reaches exactly the entries rblas's substitution
//      reaches.
#include <cmath>

#include "kernels.hpp"

namespace rblas {
namespace {

constexpr int kNB = 64;

struct Cd {
  double re, im;
};

// complex.rs:101-111: (a+bi)(c+di) with rblas's fma pattern.
__device__ __forceinline__ Cd cmul(Cd a, Cd b) {
  return Cd{__builtin_fma(a.re, b.re, 0.0 - a.im * b.im), __builtin_fma(a.re, b.im, a.im * b.re)};
}

// complex3.rs:344-349: s * conj(d) / (fma(d.re, d.re, d.im * d.im)).
__device__ __forceinline__ Cd cdiv(Cd s, Cd d) {
  const double n = __builtin_fma(d.re, d.re, d.im * d.im);
  const Cd num = cmul(s, Cd{d.re, 0.0 - d.im});
  return Cd{num.re / n, num.im / n};
}

template <typename T>
__global__ void cscale_kernel(int64_t m, int64_t n, T ar, T ai, T *b, int64_t ldb) {
  const int64_t total = m * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / m, i = e - j * m;
    T *x = b + 2 * (i + j * ldb);
    const T xr = x[0], xi = x[1];
    x[0] = __builtin_fma(ar, xr, T(0) - ai * xi);
    x[1] = __builtin_fma(ar, xi, ai * xr);
  }
}

// op(A)(r, c), conjugated for ConjTrans.
template <typename T>
__device__ __forceinline__ Cd aop(const T *a, int64_t lda, int trans, int64_t r, int64_t c) {
  const int64_t idx = trans ? c + r * lda : r + c * lda;
  return Cd{(double)a[2 * idx], trans == 2 ? 0.0 - (double)a[2 * idx + 1] : (double)a[2 * idx + 1]};
}

// Diagonal block of op(A) at (lo, lo), nb x nb, into As[r][c] (only the
// effective triangle is read).
template <typename T>
__device__ __forceinline__ void load_block(const T *a, int64_t lda, int trans, bool upper,
                                           int64_t lo, int nb, Cd (*As)[kNB + 1]) {
  for (int e = threadIdx.x; e < kNB * kNB; e += blockDim.x) {
    const int r = e % kNB, c = e / kNB;
    const bool in = r < nb && c < nb && (upper ? r <= c : r >= c);
    As[r][c] = in ? aop(a, lda, trans, lo + r, lo + c) : Cd{0.0, 0.0};
  }
}

// Left: rows [lo, lo + nb) of B, column j = blockIdx.x * 64 + thread.
template <typename T>
__global__ __launch_bounds__(kNB) void ctrsm_leaf_left(const T *a, int64_t lda, int trans,
                                                       int upper, int unit, int64_t lo, int nb,
                                                       int64_t n, T *b, int64_t ldb) {
  __shared__ Cd As[kNB][kNB + 1];
  __shared__ Cd X[kNB][kNB];  // X[row][thread]
  load_block(a, lda, trans, upper, lo, nb, As);
  const int t = threadIdx.x;
  const int64_t j = blockIdx.x * (int64_t)kNB + t;
  const bool on = j < n;
  T *col = b + 2 * (lo + j * ldb);
  if (on)
    for (int r = 0; r < nb; ++r) X[r][t] = Cd{(double)col[2 * r], (double)col[2 * r + 1]};
  __syncthreads();
  if (!on) return;
  for (int s = 0; s < nb; ++s) {
    const int i = upper ? nb - 1 - s : s;
    Cd acc = X[i][t];
    const int k0 = upper ? i + 1 : 0, k1 = upper ? nb : i;
    for (int kk = k0; kk < k1; ++kk) {
      const Cd p = cmul(As[i][kk], X[kk][t]);
      acc = Cd{acc.re - p.re, acc.im - p.im};
    }
    X[i][t] = unit ? acc : cdiv(acc, As[i][i]);
  }
  for (int r = 0; r < nb; ++r) {
    col[2 * r] = (T)X[r][t].re;
    col[2 * r + 1] = (T)X[r][t].im;
  }
}

// Right: columns [lo, lo + nb) of B, row i = blockIdx.x * 64 + thread.
template <typename T>
__global__ __launch_bounds__(kNB) void ctrsm_leaf_right(const T *a, int64_t lda, int trans,
                                                        int upper, int unit, int64_t lo, int nb,
                                                        int64_t m, T *b, int64_t ldb) {
  __shared__ Cd As[kNB][kNB + 1];
  __shared__ Cd X[kNB][kNB];  // X[column][thread]
  load_block(a, lda, trans, upper, lo, nb, As);
  const int t = threadIdx.x;
  const int64_t i = blockIdx.x * (int64_t)kNB + t;
  const bool on = i < m;
  if (on)
    for (int c = 0; c < nb; ++c) {
      const T *x = b + 2 * (i + (lo + c) * ldb);
      X[c][t] = Cd{(double)x[0], (double)x[1]};
    }
  __syncthreads();
  if (!on) return;
  for (int s = 0; s < nb; ++s) {
    const int j = upper ? s : nb - 1 - s;
    Cd acc = X[j][t];
    const int k0 = upper ? 0 : j + 1, k1 = upper ? j : nb;
    for (int kk = k0; kk < k1; ++kk) {
      const Cd p = cmul(X[kk][t], As[kk][j]);
      acc = Cd{acc.re - p.re, acc.im - p.im};
    }
    X[j][t] = unit ? acc : cdiv(acc, As[j][j]);
  }
  for (int c = 0; c < nb; ++c) {
    T *x = b + 2 * (i + (lo + c) * ldb);
    x[0] = (T)X[c][t].re;
    x[1] = (T)X[c][t].im;
  }
}

unsigned grid_for(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 16384) g = 16384;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// Offset (in complex elements) of op(A)(r0, c0) as a gemm operand.
inline int64_t sub_off(int64_t lda, bool trans, int64_t r0, int64_t c0) {
  return trans ? c0 + r0 * lda : r0 + c0 * lda;
}

template <typename T>
struct CSolver {
  const CtrsmProblem<T> &p;
  T *ws;  // complex gemm workspace
  hipStream_t st;
  bool upper, forward;

  hipError_t range(int64_t lo, int64_t len) {
    if (len <= kNB) {
      if (p.left)
        hipLaunchKernelGGL((ctrsm_leaf_left<T>), dim3((unsigned)((p.n + kNB - 1) / kNB)),
                           dim3(kNB), 0, st, p.a, p.lda, p.trans, (int)upper, (int)p.unit, lo,
                           (int)len, p.n, p.b, p.ldb);
      else
        hipLaunchKernelGGL((ctrsm_leaf_right<T>), dim3((unsigned)((p.m + kNB - 1) / kNB)),
                           dim3(kNB), 0, st, p.a, p.lda, p.trans, (int)upper, (int)p.unit, lo,
                           (int)len, p.m, p.b, p.ldb);
      return hipGetLastError();
    }
    const int64_t h = ((len / 2 + kNB - 1) / kNB) * kNB;
    int64_t f_lo = lo, f_len = h, s_lo = lo + h, s_len = len - h;
    if (!forward) {
      f_lo = lo + h;
      f_len = len - h;
      s_lo = lo;
      s_len = h;
    }
    hipError_t e = range(f_lo, f_len);
    if (e != hipSuccess) return e;
    const bool tr = p.trans != 0;
    if (p.left) {
      // B[second rows] -= op(A)[second, first] X[first rows]
      CgemmProblem<T> u{p.trans, 0, s_len, p.n, f_len, T(-1), T(0), T(1), T(0),
                        p.a + 2 * sub_off(p.lda, tr, s_lo, f_lo), p.lda, p.b + 2 * f_lo, p.ldb,
                        p.b + 2 * s_lo, p.ldb};
      e = launch_cgemm<T>(u, ws, st);
    } else {
      // B[:, second] -= X[:, first] op(A)[first, second]
      CgemmProblem<T> u{0, p.trans, p.m, s_len, f_len, T(-1), T(0), T(1), T(0),
                        p.b + 2 * f_lo * p.ldb, p.ldb, p.a + 2 * sub_off(p.lda, tr, f_lo, s_lo),
                        p.lda, p.b + 2 * s_lo * p.ldb, p.ldb};
      e = launch_cgemm<T>(u, ws, st);
    }
    if (e != hipSuccess) return e;
    return range(s_lo, s_len);
  }
};

}  // namespace

size_t ctrsm_workspace_elems(int64_t m, int64_t n, bool left) {
  const int64_t q = left ? m : n;
  return left ? cgemm_workspace_elems(q, n, q) : cgemm_workspace_elems(m, q, q);
}

template <typename T>
hipError_t launch_ctrsm(const CtrsmProblem<T> &p, T *ws, hipStream_t st) {
  if (p.m == 0 || p.n == 0) return hipSuccess;
  if (!(p.alpha_re == T(1) && p.alpha_im == T(0)))
    hipLaunchKernelGGL((cscale_kernel<T>), dim3(grid_for(p.m * p.n)), dim3(256), 0, st, p.m, p.n,
                       p.alpha_re, p.alpha_im, p.b, p.ldb);
  const bool tr = p.trans != 0;
  const bool upper = p.upper != tr;  // effective triangle of op(A)
  // Left+lower and Right+upper substitute from index 0 up, the others down.
  CSolver<T> s{p, ws, st, upper, p.left ? !upper : upper};
  return s.range(0, p.left ? p.m : p.n);
}

template hipError_t launch_ctrsm<float>(const CtrsmProblem<float> &, float *, hipStream_t);
template hipError_t launch_ctrsm<double>(const CtrsmProblem<double> &, double *, hipStream_t);

}  // namespace rblas
