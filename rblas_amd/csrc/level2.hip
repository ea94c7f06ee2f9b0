// Level-2: gemv (level2.rs:11-61, complex2.rs:11-76) and rank-1 updates
// (level2.rs:65-88 ger, complex2.rs:112-142 geru / gerc). Both stream A once
// from HBM; the vectors are small beside it and stay in the caches.
//
// A is addressed through a column-major storage view S (R x C, S(r, c) =
// a[r + c*ld], complex interleaved): a ColMajor m x n A is S itself
// (R = m, C = n), a RowMajor one is the view of its transpose (R = n, C = m).
// gemv then takes one of two shapes:
//   rows: out_r = sum_c S(r, c) v_c (ColMajor NoTrans, RowMajor Trans). One
//         thread per r, consecutive lanes on consecutive r (coalesced),
//         8 columns in flight; the columns split over workgroups when R alone
//         does not fill the chip.
//   cols: out_c = sum_r S(r, c) v_r (ColMajor Trans, RowMajor NoTrans). One
//         wave per 4 columns, lanes along r (coalesced, v read once for the
//         4), then a wave reduction; the rows split over workgroups when C
//         alone does not fill the chip.
// Split partial sums go to a workspace and are added in split order by
// gemv_finish, which applies alpha and beta as the unsplit kernels do:
//   real:    y = fma(alpha, acc, ys), ys = beta == 1 ? y : beta == 0 ? 0 : beta y
//   complex: y = ys + alpha acc,      ys = beta == 0 ? 0 : beta y (always)
// with rblas's complex multiply (complex.rs:101-111).
#include "fma.hpp"
#include "kernels.hpp"

namespace rblas {
namespace {

constexpr int kThreads = 256;
constexpr int kRowsUnroll = 8;   // columns in flight per thread (rows shape)
constexpr int kColsPerWave = 4;  // columns per wave (cols shape)
constexpr int kColsUnroll = 2;   // 64-row steps in flight (cols shape)
constexpr int kRank1Cols = 8;    // columns per thread (rank-1)
constexpr int64_t kTargetBlocks = 2048;

template <typename T>
struct Cx {
  T re, im;
};

template <typename T, bool CPLX>
__device__ __forceinline__ Cx<T> ld_elem(const T *p, int64_t idx) {
  if constexpr (CPLX) return Cx<T>{p[2 * idx], p[2 * idx + 1]};
  else return Cx<T>{p[idx], T(0)};
}

template <typename T, bool CPLX>
__device__ __forceinline__ void st_elem(T *p, int64_t idx, Cx<T> v) {
  if constexpr (CPLX) {
    p[2 * idx] = v.re;
    p[2 * idx + 1] = v.im;
  } else {
    p[idx] = v.re;
  }
}

// complex.rs:101-111
template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b) {
  return Cx<T>{fma_t(a.re, b.re, T(0) - a.im * b.im), fma_t(a.re, b.im, a.im * b.re)};
}

// acc + a x: rblas's real fma chain, or its complex multiply then add.
template <typename T, bool CPLX>
__device__ __forceinline__ Cx<T> madd(Cx<T> acc, Cx<T> a, Cx<T> x) {
  if constexpr (CPLX) {
    const Cx<T> p = cmul(a, x);
    return Cx<T>{acc.re + p.re, acc.im + p.im};
  } else {
    return Cx<T>{fma_t(a.re, x.re, acc.re), T(0)};
  }
}

template <typename T>
struct Scalars {
  T ar, ai, br, bi;
};

// y := beta y (see the file comment), then + alpha acc when `with_acc`.
template <typename T, bool CPLX>
__device__ __forceinline__ Cx<T> combine(Cx<T> y, Cx<T> acc, Scalars<T> s, bool with_acc) {
  Cx<T> ys;
  if constexpr (CPLX) {
    ys = (s.br == T(0) && s.bi == T(0)) ? Cx<T>{T(0), T(0)} : cmul(Cx<T>{s.br, s.bi}, y);
    if (!with_acc) return ys;
    const Cx<T> p = cmul(Cx<T>{s.ar, s.ai}, acc);
    return Cx<T>{ys.re + p.re, ys.im + p.im};
  } else {
    ys.re = s.br == T(1) ? y.re : (s.br == T(0) ? T(0) : s.br * y.re);
    ys.im = T(0);
    if (!with_acc) return ys;
    return Cx<T>{fma_t(s.ar, acc.re, ys.re), T(0)};
  }
}

template <typename T, bool CPLX>
__device__ __forceinline__ Cx<T> conj_if(Cx<T> v, bool conj) {
  if constexpr (CPLX) return Cx<T>{v.re, conj ? T(0) - v.im : v.im};
  else return v;
}

// rows shape. Block b: row block b % nrb, column split b / nrb.
template <typename T, bool CPLX>
__global__ __launch_bounds__(kThreads) void gemv_rows(int64_t R, int64_t C, const T *a, int64_t ld,
                                                      int conj, const T *v, int64_t incv,
                                                      int64_t nrb, int64_t cper, T *part, T *y,
                                                      int64_t incy, Scalars<T> s) {
  const int64_t rb = blockIdx.x % nrb, split = blockIdx.x / nrb;
  const int64_t r = rb * kThreads + threadIdx.x;
  const int64_t rr = r < R ? r : R - 1;  // rows past R read row R-1, store nothing
  const int64_t c0 = split * cper, c1 = c0 + cper < C ? c0 + cper : C;
  Cx<T> acc{T(0), T(0)};
  int64_t c = c0;
  for (; c + kRowsUnroll <= c1; c += kRowsUnroll) {
    Cx<T> av[kRowsUnroll], xv[kRowsUnroll];
#pragma unroll
    for (int u = 0; u < kRowsUnroll; ++u) {
      av[u] = ld_elem<T, CPLX>(a, rr + (c + u) * ld);
      xv[u] = ld_elem<T, CPLX>(v, (c + u) * incv);
    }
#pragma unroll
    for (int u = 0; u < kRowsUnroll; ++u) acc = madd<T, CPLX>(acc, conj_if<T, CPLX>(av[u], conj), xv[u]);
  }
  for (; c < c1; ++c)
    acc = madd<T, CPLX>(acc, conj_if<T, CPLX>(ld_elem<T, CPLX>(a, rr + c * ld), conj),
                        ld_elem<T, CPLX>(v, c * incv));
  if (r >= R) return;
  if (part) {
    st_elem<T, CPLX>(part, split * R + r, acc);
  } else {
    const Cx<T> yv = ld_elem<T, CPLX>(y, r * incy);
    st_elem<T, CPLX>(y, r * incy, combine<T, CPLX>(yv, acc, s, true));
  }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// cols shape. Block b: column block b % ncb (4 waves x 4 columns), row
// split b / ncb.
template <typename T, bool CPLX>
__global__ __launch_bounds__(kThreads) void gemv_cols(int64_t R, int64_t C, const T *a, int64_t ld,
                                                      int conj, const T *v, int64_t incv,
                                                      int64_t ncb, int64_t rper, T *part, T *y,
                                                      int64_t incy, Scalars<T> s) {
  constexpr int kColsPerBlock = kColsPerWave * (kThreads / 64);
  const int64_t cb = blockIdx.x % ncb, split = blockIdx.x / ncb;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t cbase = cb * kColsPerBlock + wave * kColsPerWave;
  int64_t col[kColsPerWave];
#pragma unroll
  for (int q = 0; q < kColsPerWave; ++q) col[q] = cbase + q < C ? cbase + q : C - 1;
  const int64_t r0 = split * rper, r1 = r0 + rper < R ? r0 + rper : R;
  Cx<T> acc[kColsPerWave];
#pragma unroll
  for (int q = 0; q < kColsPerWave; ++q) acc[q] = Cx<T>{T(0), T(0)};
  int64_t r = r0 + lane;
  for (; r + 64 * (kColsUnroll - 1) < r1; r += 64 * kColsUnroll) {
    Cx<T> xv[kColsUnroll], av[kColsUnroll][kColsPerWave];
#pragma unroll
    for (int u = 0; u < kColsUnroll; ++u) {
      const int64_t ri = r + 64 * u;
      xv[u] = ld_elem<T, CPLX>(v, ri * incv);
#pragma unroll
      for (int q = 0; q < kColsPerWave; ++q) av[u][q] = ld_elem<T, CPLX>(a, ri + col[q] * ld);
    }
#pragma unroll
    for (int u = 0; u < kColsUnroll; ++u)
#pragma unroll
      for (int q = 0; q < kColsPerWave; ++q)
        acc[q] = madd<T, CPLX>(acc[q], conj_if<T, CPLX>(av[u][q], conj), xv[u]);
  }
  for (; r < r1; r += 64) {
    const Cx<T> xv = ld_elem<T, CPLX>(v, r * incv);
#pragma unroll
    for (int q = 0; q < kColsPerWave; ++q)
      acc[q] = madd<T, CPLX>(acc[q], conj_if<T, CPLX>(ld_elem<T, CPLX>(a, r + col[q] * ld), conj),
                             xv);
  }
#pragma unroll
  for (int q = 0; q < kColsPerWave; ++q) {
    acc[q].re = wave_sum(acc[q].re);
    if constexpr (CPLX) acc[q].im = wave_sum(acc[q].im);
  }
  if (lane >= kColsPerWave) return;
  // Lane q stores column q (static indexing: a select chain, not an array
  // index by lane).
  Cx<T> mine = acc[0];
#pragma unroll
  for (int q = 1; q < kColsPerWave; ++q)
    if (lane == q) mine = acc[q];
  const int64_t c = cbase + lane;
  if (c >= C) return;
  if (part) {
    st_elem<T, CPLX>(part, split * C + c, mine);
  } else {
    const Cx<T> yv = ld_elem<T, CPLX>(y, c * incy);
    st_elem<T, CPLX>(y, c * incy, combine<T, CPLX>(yv, mine, s, true));
  }
}

// Sums the split partials of each output in split order and applies alpha
// and beta; with nsplit == 0 only scales y (alpha == 0).
template <typename T, bool CPLX>
__global__ __launch_bounds__(kThreads) void gemv_finish(int64_t leny, int nsplit, const T *part,
                                                        T *y, int64_t incy, Scalars<T> s) {
  for (int64_t o = blockIdx.x * (int64_t)kThreads + threadIdx.x; o < leny;
       o += (int64_t)gridDim.x * kThreads) {
    Cx<T> acc = nsplit > 0 ? ld_elem<T, CPLX>(part, o) : Cx<T>{T(0), T(0)};
    // Eight partials in flight per round trip, added in split order.
    constexpr int kBatch = 8;
    int q = 1;
    for (; q + kBatch <= nsplit; q += kBatch) {
      Cx<T> pv[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) pv[u] = ld_elem<T, CPLX>(part, (q + u) * leny + o);
#pragma unroll
      for (int u = 0; u < kBatch; ++u) acc = Cx<T>{acc.re + pv[u].re, acc.im + pv[u].im};
    }
    for (; q < nsplit; ++q) {
      const Cx<T> p = ld_elem<T, CPLX>(part, q * leny + o);
      acc = Cx<T>{acc.re + p.re, acc.im + p.im};
    }
    const Cx<T> yv = ld_elem<T, CPLX>(y, o * incy);
    st_elem<T, CPLX>(y, o * incy, combine<T, CPLX>(yv, acc, s, nsplit > 0));
  }
}

// Rank-1 over the storage view: thread = one r, 8 columns. Logical (i, j)
// is (r, c) for ColMajor and (c, r) for RowMajor; x is indexed by i, y by j.
//   real:    a = fma(alpha x_i, y_j, a)                 (level2.rs:81-87)
//   complex: a = a + (alpha x_i) y_j', y_j' = conj(y_j) for gerc (complex2.rs:130-141)
template <typename T, bool CPLX, bool ROWMAJOR>
__global__ __launch_bounds__(kThreads) void rank1_kernel(int64_t R, int64_t C, T *a, int64_t ld,
                                                         const T *x, int64_t incx, const T *y,
                                                         int64_t incy, int conj, T ar, T ai,
                                                         int64_t nrb, int64_t nblocks) {
  for (int64_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const int64_t rb = b % nrb, cb = b / nrb;
    const int64_t r = rb * kThreads + threadIdx.x;
    const int64_t rr = r < R ? r : R - 1;
    const int64_t c0 = cb * kRank1Cols;
    Cx<T> av[kRank1Cols], xv[kRank1Cols], yv[kRank1Cols];
#pragma unroll
    for (int u = 0; u < kRank1Cols; ++u) {
      const int64_t c = c0 + u < C ? c0 + u : C - 1;
      const int64_t i = ROWMAJOR ? c : rr, j = ROWMAJOR ? rr : c;
      av[u] = ld_elem<T, CPLX>(a, rr + c * ld);
      xv[u] = ld_elem<T, CPLX>(x, i * incx);
      yv[u] = ld_elem<T, CPLX>(y, j * incy);
    }
    if (r >= R) continue;
#pragma unroll
    for (int u = 0; u < kRank1Cols; ++u) {
      const int64_t c = c0 + u;
      if (c >= C) break;
      Cx<T> out;
      if constexpr (CPLX) {
        const Cx<T> axi = cmul(Cx<T>{ar, ai}, xv[u]);
        const Cx<T> p = cmul(axi, conj_if<T, CPLX>(yv[u], conj));
        out = Cx<T>{av[u].re + p.re, av[u].im + p.im};
      } else {
        out = Cx<T>{fma_t(ar * xv[u].re, yv[u].re, av[u].re), T(0)};
      }
      st_elem<T, CPLX>(a, r + c * ld, out);
    }
  }
}

int64_t cdiv64(int64_t a, int64_t b) { return (a + b - 1) / b; }

template <typename T>
struct GemvPlan {
  bool rows;        // rows shape
  int64_t leny;     // outputs
  int64_t nb;       // row blocks (rows) or column blocks (cols)
  int64_t splits;
  int64_t per;      // columns (rows shape) or rows (cols shape) per split
};

template <typename T>
GemvPlan<T> plan_gemv(const GemvProblem<T> &p) {
  GemvPlan<T> g;
  g.rows = !p.cols_shape;
  if (g.rows) {
    g.leny = p.R;
    g.nb = cdiv64(p.R, kThreads);
    int64_t sp = g.nb > 0 ? cdiv64(kTargetBlocks, g.nb) : 1;
    const int64_t cap = cdiv64(p.C, 64);
    if (sp > cap) sp = cap;
    g.splits = sp < 1 ? 1 : sp;
    g.per = cdiv64(p.C, g.splits);
  } else {
    constexpr int kColsPerBlock = kColsPerWave * (kThreads / 64);
    g.leny = p.C;
    g.nb = cdiv64(p.C, kColsPerBlock);
    int64_t sp = g.nb > 0 ? cdiv64(kTargetBlocks, g.nb) : 1;
    const int64_t cap = cdiv64(p.R, 256);
    if (sp > cap) sp = cap;
    g.splits = sp < 1 ? 1 : sp;
    g.per = cdiv64(cdiv64(p.R, g.splits), 64) * 64;
  }
  return g;
}

template <typename T, bool CPLX>
hipError_t gemv_as(const GemvProblem<T> &p, void *ws, hipStream_t st) {
  const Scalars<T> s{p.alpha_re, p.alpha_im, p.beta_re, p.beta_im};
  const GemvPlan<T> g = plan_gemv(p);
  if (g.leny == 0) return hipSuccess;
  const bool alpha_zero = p.alpha_re == T(0) && p.alpha_im == T(0);
  if (alpha_zero) {
    if (!CPLX && p.beta_re == T(1)) return hipSuccess;
    int64_t blocks = cdiv64(g.leny, kThreads);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL((gemv_finish<T, CPLX>), dim3((unsigned)blocks), dim3(kThreads), 0, st,
                       g.leny, 0, (const T *)nullptr, p.y, p.incy, s);
    return hipGetLastError();
  }
  T *part = g.splits > 1 ? (T *)ws : nullptr;
  const unsigned grid = (unsigned)(g.nb * g.splits);
  if (g.rows)
    hipLaunchKernelGGL((gemv_rows<T, CPLX>), dim3(grid), dim3(kThreads), 0, st, p.R, p.C, p.a,
                       p.ld, (int)p.conj, p.v, p.incv, g.nb, g.per, part, p.y, p.incy, s);
  else
    hipLaunchKernelGGL((gemv_cols<T, CPLX>), dim3(grid), dim3(kThreads), 0, st, p.R, p.C, p.a,
                       p.ld, (int)p.conj, p.v, p.incv, g.nb, g.per, part, p.y, p.incy, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !part) return e;
  int64_t blocks = cdiv64(g.leny, kThreads);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL((gemv_finish<T, CPLX>), dim3((unsigned)blocks), dim3(kThreads), 0, st,
                     g.leny, (int)g.splits, (const T *)part, p.y, p.incy, s);
  return hipGetLastError();
}

template <typename T, bool CPLX, bool ROWMAJOR>
hipError_t rank1_as(const Rank1Problem<T> &p, hipStream_t st) {
  const int64_t nrb = cdiv64(p.R, kThreads), ncb = cdiv64(p.C, kRank1Cols);
  const int64_t nblocks = nrb * ncb;
  const int64_t grid = nblocks < (int64_t(1) << 20) ? nblocks : (int64_t(1) << 20);
  hipLaunchKernelGGL((rank1_kernel<T, CPLX, ROWMAJOR>), dim3((unsigned)grid), dim3(kThreads), 0,
                     st, p.R, p.C, p.a, p.ld, p.x, p.incx, p.y, p.incy, (int)p.conj, p.alpha_re,
                     p.alpha_im, nrb, nblocks);
  return hipGetLastError();
}

}  // namespace

template <typename T>
size_t gemv_workspace_bytes(const GemvProblem<T> &p) {
  const GemvPlan<T> g = plan_gemv(p);
  if (g.leny == 0 || g.splits <= 1) return 0;
  return (size_t)(g.splits * g.leny) * (p.cplx ? 2 : 1) * sizeof(T);
}

template <typename T>
hipError_t launch_gemv(const GemvProblem<T> &p, void *ws, hipStream_t st) {
  return p.cplx ? gemv_as<T, true>(p, ws, st) : gemv_as<T, false>(p, ws, st);
}

template <typename T>
hipError_t launch_rank1(const Rank1Problem<T> &p, hipStream_t st) {
  if (p.R == 0 || p.C == 0 || (p.alpha_re == T(0) && p.alpha_im == T(0))) return hipSuccess;
  if (p.cplx) return p.rowmajor ? rank1_as<T, true, true>(p, st) : rank1_as<T, true, false>(p, st);
  return p.rowmajor ? rank1_as<T, false, true>(p, st) : rank1_as<T, false, false>(p, st);
}

template size_t gemv_workspace_bytes<float>(const GemvProblem<float> &);
template size_t gemv_workspace_bytes<double>(const GemvProblem<double> &);
template hipError_t launch_gemv<float>(const GemvProblem<float> &, void *, hipStream_t);
template hipError_t launch_gemv<double>(const GemvProblem<double> &, void *, hipStream_t);
template hipError_t launch_rank1<float>(const Rank1Problem<float> &, hipStream_t);
template hipError_t launch_rank1<double>(const Rank1Problem<double> &, hipStream_t);

}  // namespace rblas
