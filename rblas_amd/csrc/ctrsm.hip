// Complex triangular solve (complex3::trsm, complex3.rs:252-349): op(A) X =
// alpha B (Left) or X op(A) = alpha B (Right), in place in B, op = None,
// Trans or ConjTrans.
//
//   1. B := alpha B when alpha != 1 (rblas's complex multiply, every entry).
//   2. The diagonal range is solved by recursive halving: solve the half that
//      comes first in substitution order, subtract its contribution from the
//      other half, solve the other half, by real MFMA gemms on the real view
//      of B (its rows alternating re / im, ld 2 ldb). Left: one gemm with the
//      real representation of op(A) (realrep_kernel, formed once per call);
//      Right: X re(op A) + (i X) im(op A), two gemms with op(A)'s planes
//      (formed once) and a rotated copy of X (rot_kernel, per update).
//   3. A range of up to 64 is solved by substitution in f64, 16 lanes per
//      right-hand side (a column of B for Left, a row for Right), the block
//      of op(A) in LDS: s = s - a*x with rblas's complex multiply, then
//      s * (1 / a_ii) with 1 / a_ii = conj(a_ii) / norm_sqr(a_ii) formed once
//      per block where rblas divides each s by norm_sqr(a_ii)
//      (complex3.rs:344-349): the same quotient up to rounding. A non-finite
//      B entry reaches exactly the entries rblas's substitution reaches.
#include <cmath>

#include "kernels.hpp"
#include "fma.hpp"

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

// Componentwise select: a conditional on two Cd lvalues would select
// between their addresses and force both into memory.
__device__ __forceinline__ Cd csel(bool c, Cd a, Cd b) {
  return Cd{c ? a.re : b.re, c ? a.im : b.im};
}

template <typename T>
__global__ void cscale_kernel(int64_t m, int64_t n, T ar, T ai, T *b, int64_t ldb) {
  const int64_t total = m * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / m, i = e - j * m;
    T *x = b + 2 * (i + j * ldb);
    const T xr = x[0], xi = x[1];
    x[0] = fma_t(ar, xr, T(0) - ai * xi);
    x[1] = fma_t(ar, xi, ai * xr);
  }
}

// Leaves: one 16-lane DPP row per right-hand side (a column of B for Left,
// a row for Right), each lane holding entries r = g + 16 jj (g its lane in
// the row, jj < 4) in registers. The solve is right-looking: at step i the
// lane that holds entry i scales it by 1 / a_ii, the result x_i is broadcast
// across the row, and every later entry r subtracts coef(i, r) x_i. An entry
// receives its subtractions in rblas's order (k ascending) in a forward solve
// and in descending k in a backward one; the entries a product with a
// non-finite x_i reaches are the ones rblas's sums reach (every k in the
// triangle, zero coefficients included). Steps are unrolled so that the
// owning lane and register are compile-time constants.
constexpr int kRowLanes = 16;                 // lanes per right-hand side (a DPP row)
constexpr int kPer = kNB / kRowLanes;         // entries held per lane
constexpr int kLeafThreads = 256;
constexpr int kRhsPerBlock = kLeafThreads / kRowLanes;

// Value of lane S of each row of 16 lanes, for every lane of the row.
template <int S>
__device__ __forceinline__ int row_bcast_imm(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x150 + S, 0xf, 0xf, false);
}

template <int S>
__device__ __forceinline__ double row_bcast(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)row_bcast_imm<S>((int)(unsigned)u);
  const unsigned hi = (unsigned)row_bcast_imm<S>((int)(unsigned)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The four entries a lane holds, addressed by compile-time index only so
// that they stay in registers.
struct Held {
  Cd v0, v1, v2, v3;
  template <int J>
  __device__ __forceinline__ Cd &at() {
    if constexpr (J == 0) return v0;
    else if constexpr (J == 1) return v1;
    else if constexpr (J == 2) return v2;
    else return v3;
  }
};
static_assert(kPer == 4, "Held keeps four entries per lane");

// The coefficients step I needs, loaded one step ahead: U[i][g + 16 J]
// for the entries the step updates, and 1 / op(A)(i, i).
struct StepCoef {
  Cd c0, c1, c2, c3, dinv;
  template <int J>
  __device__ __forceinline__ const Cd &at() const {
    if constexpr (J == 0) return c0;
    else if constexpr (J == 1) return c1;
    else if constexpr (J == 2) return c2;
    else return c3;
  }
  template <int J>
  __device__ __forceinline__ Cd &at() {
    if constexpr (J == 0) return c0;
    else if constexpr (J == 1) return c1;
    else if constexpr (J == 2) return c2;
    else return c3;
  }
};

template <bool FORWARD, int I, int J>
__device__ __forceinline__ void load_coef_j(const Cd (*U)[kNB], int g, StepCoef &k) {
  constexpr int i = FORWARD ? I : kNB - 1 - I;
  constexpr int jo = i / kRowLanes;
  if constexpr (FORWARD ? J >= jo : J <= jo) k.at<J>() = U[i][g + kRowLanes * J];
}

template <bool FORWARD, bool UNIT, int I>
__device__ __forceinline__ void load_coef(const Cd (*U)[kNB], const Cd *dinv, int g,
                                          StepCoef &k) {
  if constexpr (I < kNB) {
    load_coef_j<FORWARD, I, 0>(U, g, k);
    load_coef_j<FORWARD, I, 1>(U, g, k);
    load_coef_j<FORWARD, I, 2>(U, g, k);
    load_coef_j<FORWARD, I, 3>(U, g, k);
    if constexpr (!UNIT) k.dinv = dinv[FORWARD ? I : kNB - 1 - I];
  }
}

// Entry r = g + 16 J after x_i is known.
template <bool LEFT, bool FORWARD, int I, int J>
__device__ __forceinline__ void leaf_update(const StepCoef &k, int g, Cd xi, Held &h) {
  constexpr int i = FORWARD ? I : kNB - 1 - I;
  constexpr int jo = i / kRowLanes;
  if constexpr (FORWARD ? J >= jo : J <= jo) {
    const int r = g + kRowLanes * J;
    const Cd c = k.at<J>();
    const Cd p = LEFT ? cmul(c, xi) : cmul(xi, c);
    Cd &e = h.at<J>();
    const Cd upd{e.re - p.re, e.im - p.im};
    if constexpr (J == jo) {
      const bool later = FORWARD ? r > i : r < i;
      e = csel(later, upd, csel(r == i, xi, e));
    } else {
      e = upd;
    }
  }
}

// Step I of the solve in substitution order: entry i = FORWARD ? I : 63 - I.
template <bool LEFT, bool FORWARD, bool FULL, bool UNIT, int I>
__device__ __forceinline__ void leaf_step(const StepCoef &k, int nb, int g, Held &h) {
  constexpr int i = FORWARD ? I : kNB - 1 - I;
  constexpr int go = i % kRowLanes, jo = i / kRowLanes;
  if constexpr (!FULL) {
    if (i >= nb) return;  // uniform: the block is shorter than 64
  }
  const Cd v = h.at<jo>();
  Cd xi = v;
  if constexpr (!UNIT) xi = cmul(v, k.dinv);
  xi.re = row_bcast<go>(xi.re);
  xi.im = row_bcast<go>(xi.im);
  leaf_update<LEFT, FORWARD, I, 0>(k, g, xi, h);
  leaf_update<LEFT, FORWARD, I, 1>(k, g, xi, h);
  leaf_update<LEFT, FORWARD, I, 2>(k, g, xi, h);
  leaf_update<LEFT, FORWARD, I, 3>(k, g, xi, h);
}

// Steps I .. 63, step I's coefficients in `cur`. FULL (a 64-wide block) and
// UNIT (unit diagonal) are compile-time so that the unrolled steps are
// straight-line code.
template <bool LEFT, bool FORWARD, bool FULL, bool UNIT, int I>
struct LeafRun {
  __device__ __forceinline__ static void go(const Cd (*U)[kNB], const Cd *dinv, int nb, int g,
                                            Held &h, const StepCoef &cur) {
    StepCoef nxt;
    load_coef<FORWARD, UNIT, I + 1>(U, dinv, g, nxt);
    leaf_step<LEFT, FORWARD, FULL, UNIT, I>(cur, nb, g, h);
    LeafRun<LEFT, FORWARD, FULL, UNIT, I + 1>::go(U, dinv, nb, g, h, nxt);
  }
};
template <bool LEFT, bool FORWARD, bool FULL, bool UNIT>
struct LeafRun<LEFT, FORWARD, FULL, UNIT, kNB> {
  __device__ __forceinline__ static void go(const Cd (*)[kNB], const Cd *, int, int, Held &,
                                            const StepCoef &) {}
};

// B entry r of right-hand side `rhs`: row lo + r of column rhs (Left) or
// column lo + r of row rhs (Right).
template <bool LEFT>
__device__ __forceinline__ int64_t b_index(int64_t lo, int r, int64_t rhs, int64_t ldb) {
  return LEFT ? lo + r + rhs * ldb : rhs + (lo + r) * ldb;
}

template <typename T, bool LEFT, int J>
__device__ __forceinline__ void leaf_load(const T *b, int64_t lo, int nb, int64_t rhs, bool on,
                                          int64_t ldb, int64_t safe, int g, Held &h) {
  const int r = g + kRowLanes * J;
  const bool ok = on && r < nb;
  const int64_t idx = ok ? b_index<LEFT>(lo, r, rhs, ldb) : safe;
  const T re = b[2 * idx], im = b[2 * idx + 1];
  h.at<J>() = Cd{ok ? (double)re : 0.0, ok ? (double)im : 0.0};
}

template <typename T, bool LEFT, int J>
__device__ __forceinline__ void leaf_store(T *b, int64_t lo, int nb, int64_t rhs, bool on,
                                           int64_t ldb, int g, Held &h) {
  const int r = g + kRowLanes * J;
  if (on && r < nb) {
    T *o = b + 2 * b_index<LEFT>(lo, r, rhs, ldb);
    o[0] = (T)h.at<J>().re;
    o[1] = (T)h.at<J>().im;
  }
}

// Solves the diagonal block [lo, lo + nb) of op(A) for `nrhs` right-hand
// sides. FORWARD: Left + effective lower, or Right + effective upper.
template <typename T, bool LEFT, bool FORWARD, bool FULL, bool UNIT>
__global__ __launch_bounds__(kLeafThreads) void ctrsm_leaf(const T *a, int64_t lda, int trans,
                                                           int64_t lo, int nb, int64_t nrhs, T *b,
                                                           int64_t ldb) {
  // U[i][r]: coefficient of x_i in the update of entry r (op(A)(r, i) for
  // Left, op(A)(i, r) for Right), for r after i in substitution order;
  // U[i][i] the diagonal; zero elsewhere.
  __shared__ Cd U[kNB][kNB];
  const int64_t diag = lo + lo * lda;  // an element every block reads
#pragma unroll
  for (int t = 0; t < kNB * kNB / kLeafThreads; ++t) {
    const int e = threadIdx.x + kLeafThreads * t;
    const int r = e % kNB, i = e / kNB;
    const bool in = r < nb && i < nb && (r == i || (FORWARD ? r > i : r < i));
    const int64_t row = LEFT ? lo + r : lo + i, col = LEFT ? lo + i : lo + r;
    const int64_t idx = in ? (trans ? col + row * lda : row + col * lda) : diag;
    const T re = a[2 * idx], im = a[2 * idx + 1];
    const double imv = trans == 2 ? 0.0 - (double)im : (double)im;
    U[i][r] = Cd{in ? (double)re : 0.0, in ? imv : 0.0};
  }
  // 1 / op(A)(i, i) = conj(a_ii) / norm_sqr(a_ii): a step multiplies by it
  // where rblas divides (complex3.rs:344-349).
  __shared__ Cd dinv[kNB];
  if constexpr (!UNIT) {
    __syncthreads();
    if (threadIdx.x < kNB) {
      const Cd d = U[threadIdx.x][threadIdx.x];
      const double n = __builtin_fma(d.re, d.re, d.im * d.im);
      dinv[threadIdx.x] = Cd{d.re / n, (0.0 - d.im) / n};
    }
  }
  const int g = threadIdx.x % kRowLanes;
  const int64_t rhs = blockIdx.x * (int64_t)kRhsPerBlock + threadIdx.x / kRowLanes;
  const bool on = rhs < nrhs;
  const int64_t safe = LEFT ? lo : lo * ldb;  // B(lo, 0) for Left, B(0, lo) for Right
  Held h;
  leaf_load<T, LEFT, 0>(b, lo, nb, rhs, on, ldb, safe, g, h);
  leaf_load<T, LEFT, 1>(b, lo, nb, rhs, on, ldb, safe, g, h);
  leaf_load<T, LEFT, 2>(b, lo, nb, rhs, on, ldb, safe, g, h);
  leaf_load<T, LEFT, 3>(b, lo, nb, rhs, on, ldb, safe, g, h);
  __syncthreads();
  StepCoef first;
  load_coef<FORWARD, UNIT, 0>(U, dinv, g, first);
  LeafRun<LEFT, FORWARD, FULL, UNIT, 0>::go(U, dinv, nb, g, h, first);
  leaf_store<T, LEFT, 0>(b, lo, nb, rhs, on, ldb, g, h);
  leaf_store<T, LEFT, 1>(b, lo, nb, rhs, on, ldb, g, h);
  leaf_store<T, LEFT, 2>(b, lo, nb, rhs, on, ldb, g, h);
  leaf_store<T, LEFT, 3>(b, lo, nb, rhs, on, ldb, g, h);
}

unsigned grid_for(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 16384) g = 16384;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// PLANES: op(A)'s real and imaginary parts as two q x q real matrices (ld q,
// the imaginary one after the real one). Otherwise the real representation
// of op(A) (q x q complex) as a 2q x 2q real column-major matrix R (ld 2q): R(2r + a, 2c + b) = [[re, -im], [im, re]][a][b]
// of op(A)(r, c), conjugated for ConjTrans. An interleaved complex B is the
// real 2m x n matrix of its rows (re, im alternating, ld 2 ldb), and
// R_sub X_view is then exactly the real view of op(A)_sub X: one real gemm
// per update, in place on B. 32 x 32 tiles of op(A); for Trans / ConjTrans
// the source tile is read along A's columns and transposed through LDS.
constexpr int kRepTile = 32, kRepRows = 8;

template <typename T, bool PLANES>
__global__ __launch_bounds__(kRepTile *kRepRows) void realrep_kernel(const T *a, int64_t lda,
                                                                     int trans, int64_t q, T *out) {
  __shared__ Cd s[kRepTile][kRepTile + 1];
  const int64_t r0 = (int64_t)blockIdx.x * kRepTile, c0 = (int64_t)blockIdx.y * kRepTile;
  const int tx = threadIdx.x, ty = threadIdx.y;
  if (trans) {
    // A(c, r) for c = c0 + tx (contiguous in A), r = r0 + ty + 8k.
#pragma unroll
    for (int k = 0; k < kRepTile / kRepRows; ++k) {
      const int64_t c = c0 + tx, r = r0 + ty + kRepRows * k;
      if (c < q && r < q) {
        const T *v = a + 2 * (c + r * lda);
        s[ty + kRepRows * k][tx] = Cd{(double)v[0], (double)v[1]};
      }
    }
    __syncthreads();
  }
  const int64_t ld = 2 * q;
#pragma unroll
  for (int k = 0; k < kRepTile / kRepRows; ++k) {
    const int64_t r = r0 + tx, c = c0 + ty + kRepRows * k;
    if (r >= q || c >= q) continue;
    T re, im;
    if (trans) {
      const Cd v = s[tx][ty + kRepRows * k];
      re = (T)v.re;
      im = trans == 2 ? T(0) - (T)v.im : (T)v.im;
    } else {
      const T *v = a + 2 * (r + c * lda);
      re = v[0];
      im = v[1];
    }
    if constexpr (PLANES) {
      out[r + c * q] = re;
      out[q * q + r + c * q] = im;
    } else {
      T *o = out + 2 * r + 2 * c * ld;
      o[0] = re;
      o[1] = im;
      o[ld] = T(0) - im;
      o[ld + 1] = re;
    }
  }
}

// W (real 2m x cols, ld 2m) := the real view of i X for X = B's columns
// [c0, c0 + cols): W(2i, j) = -im X(i, j), W(2i + 1, j) = re X(i, j).
template <typename T>
__global__ void rot_kernel(int64_t m, int64_t cols, const T *x, int64_t ldx, T *w) {
  const int64_t total = m * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / m, i = e - j * m;
    const T *v = x + 2 * (i + j * ldx);
    T *o = w + 2 * (i + j * m);
    const T re = v[0], im = v[1];
    o[0] = T(0) - im;
    o[1] = re;
  }
}

template <typename T>
struct CSolver {
  const CtrsmProblem<T> &p;
  T *ws;  // Left: the real representation of op(A); Right: op(A)'s planes, then i X
  hipStream_t st;
  bool upper, forward;

  template <bool LEFT, bool FORWARD, bool FULL, bool UNIT>
  hipError_t leaf_launch(int64_t lo, int nb, int64_t nrhs) {
    hipLaunchKernelGGL((ctrsm_leaf<T, LEFT, FORWARD, FULL, UNIT>),
                       dim3((unsigned)((nrhs + kRhsPerBlock - 1) / kRhsPerBlock)),
                       dim3(kLeafThreads), 0, st, p.a, p.lda, p.trans, lo, nb, nrhs, p.b, p.ldb);
    return hipGetLastError();
  }

  template <bool LEFT, bool FORWARD>
  hipError_t leaf_as(int64_t lo, int nb, int64_t nrhs) {
    if (nb == kNB)
      return p.unit ? leaf_launch<LEFT, FORWARD, true, true>(lo, nb, nrhs)
                    : leaf_launch<LEFT, FORWARD, true, false>(lo, nb, nrhs);
    return p.unit ? leaf_launch<LEFT, FORWARD, false, true>(lo, nb, nrhs)
                  : leaf_launch<LEFT, FORWARD, false, false>(lo, nb, nrhs);
  }

  hipError_t leaf(int64_t lo, int nb) {
    if (p.left)
      return forward ? leaf_as<true, true>(lo, nb, p.n) : leaf_as<true, false>(lo, nb, p.n);
    return forward ? leaf_as<false, true>(lo, nb, p.m) : leaf_as<false, false>(lo, nb, p.m);
  }

  hipError_t range(int64_t lo, int64_t len) {
    if (len <= kNB) return leaf(lo, (int)len);
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
    if (p.left) {
      // B[second rows] -= op(A)[second, first] X[first rows], on real views:
      // rows 2 s_lo.. of B -= R[2 s_lo.., 2 f_lo..] rows 2 f_lo.. of B.
      const int64_t q2 = 2 * p.m;
      GemmProblem<T> u{false, false, 2 * s_len, p.n, 2 * f_len, T(-1), T(1),
                       ws + 2 * s_lo + 2 * f_lo * q2, q2, p.b + 2 * f_lo, 2 * p.ldb,
                       p.b + 2 * s_lo, 2 * p.ldb};
      e = launch_gemm<T>(u, st);
    } else {
      // B[:, second] -= X[:, first] op(A)[first, second] on real views (rows
      // of B alternating re / im): X re(op A) + (i X) im(op A).
      const int64_t q = p.n, m2 = 2 * p.m;
      const T *pre = ws, *pim = ws + q * q;
      T *wrot = ws + 2 * q * q;
      const T *x = p.b + 2 * f_lo * p.ldb;
      hipLaunchKernelGGL((rot_kernel<T>), dim3(grid_for(p.m * f_len)), dim3(256), 0, st, p.m,
                         f_len, x, p.ldb, wrot);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
      T *c = p.b + 2 * s_lo * p.ldb;
      GemmProblem<T> g1{false, false, m2, s_len, f_len, T(-1), T(1),
                        x, 2 * p.ldb, pre + f_lo + s_lo * q, q, c, 2 * p.ldb};
      e = launch_gemm<T>(g1, st);
      if (e != hipSuccess) return e;
      GemmProblem<T> g2{false, false, m2, s_len, f_len, T(-1), T(1),
                        wrot, m2, pim + f_lo + s_lo * q, q, c, 2 * p.ldb};
      e = launch_gemm<T>(g2, st);
    }
    if (e != hipSuccess) return e;
    return range(s_lo, s_len);
  }
};

}  // namespace

size_t ctrsm_workspace_elems(int64_t m, int64_t n, bool left) {
  return left ? (size_t)(4 * m * m) : (size_t)(2 * n * n + 2 * m * n);
}

template <typename T>
hipError_t launch_ctrsm(const CtrsmProblem<T> &p, T *ws, hipStream_t st) {
  if (p.m == 0 || p.n == 0) return hipSuccess;
  if (!(p.alpha_re == T(1) && p.alpha_im == T(0)))
    hipLaunchKernelGGL((cscale_kernel<T>), dim3(grid_for(p.m * p.n)), dim3(256), 0, st, p.m, p.n,
                       p.alpha_re, p.alpha_im, p.b, p.ldb);
  const int64_t q = p.left ? p.m : p.n;
  if (q > kNB) {
    const unsigned t = (unsigned)((q + kRepTile - 1) / kRepTile);
    if (p.left)
      hipLaunchKernelGGL((realrep_kernel<T, false>), dim3(t, t), dim3(kRepTile, kRepRows), 0, st,
                         p.a, p.lda, p.trans, q, ws);
    else
      hipLaunchKernelGGL((realrep_kernel<T, true>), dim3(t, t), dim3(kRepTile, kRepRows), 0, st,
                         p.a, p.lda, p.trans, q, ws);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const bool tr = p.trans != 0;
  const bool upper = p.upper != tr;  // effective triangle of op(A)
  // Left+lower and Right+upper substitute from index 0 up, the others down.
  CSolver<T> s{p, ws, st, upper, p.left ? !upper : upper};
  return s.range(0, p.left ? p.m : p.n);
}

template hipError_t launch_ctrsm<float>(const CtrsmProblem<float> &, float *, hipStream_t);
template hipError_t launch_ctrsm<double>(const CtrsmProblem<double> &, double *, hipStream_t);

}  // namespace rblas
