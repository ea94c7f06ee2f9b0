// MFMA tile GEMM for gfx950 (MI355X): C := alpha * op(A) * op(B) + beta * C,
// column-major, f32 (v_mfma_f32_16x16x4f32) and f64 (v_mfma_f64_16x16x4f64).
//
// This replaces rblas's Goto/BLIS pipeline (kernel/gemm.rs:185-474 and the
// AVX2 microkernels gemm_avx2.rs:26-155):
//   - pack_a_panel / pack_b_panel  -> global -> register -> LDS tile loads,
//     double-buffered in LDS, one barrier per k-tile;
//   - the MR x NR register microtile -> a 64 x 64 per-wave tile of 4 x 4
//     16 x 16 MFMA accumulators;
//   - the jc/pc/ic loops -> one workgroup per 128 x 128 C tile, the whole k
//     range accumulated in registers (no KC re-blocking of C);
//   - materialize (level3.rs:179-222) -> none: each operand has a loader
//     for its stored orientation (M/N-contiguous or K-contiguous).
//
// The MFMA is fed op(B)^T as its first operand and op(A)^T as its second, so
// the accumulator's lane index runs down a column of C: 16 consecutive lanes
// hold 16 consecutive rows of one C column and the epilogue writes 128-byte
// (f64) runs of column-major C.
#include "kernels.hpp"

namespace rblas {
namespace {

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // Accumulator element r of lane l is row (l/16) + 4r of the 16 x 16 block
  // (measured on gfx950: the f64 form interleaves rows).
  static __device__ __forceinline__ int row(int lane, int r) {
    return (lane >> 4) + 4 * r;
  }
};

template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // Accumulator element r of lane l is row 4(l/16) + r of the 16 x 16 block.
  static __device__ __forceinline__ int row(int lane, int r) {
    return 4 * (lane >> 4) + r;
  }
};

constexpr int kThreads = 256;  // 4 waves of 64, 2 x 2 over the C tile
constexpr int kBM = 128;       // C tile rows
constexpr int kBN = 128;       // C tile columns
constexpr int kBK = 16;        // k per LDS stage (4 MFMA k-steps)
constexpr int kWaveM = 64;     // per-wave C tile
constexpr int kWaveN = 64;
constexpr int kFragM = kWaveM / 16;  // 4 MFMA blocks down
constexpr int kFragN = kWaveN / 16;  // 4 MFMA blocks across
constexpr int kLoads = kBM * kBK / kThreads;  // 8 elements per thread per operand

template <typename T>
struct Pad;
template <>
struct Pad<double> { static constexpr int v = 1; };
template <>
struct Pad<float> { static constexpr int v = 4; };

// One operand's k-tile: kBK x 128 elements staged as lds[k][mn].
// MN_MAJOR: stored with the m (or n) index contiguous (A NoTrans, B Trans).
// Otherwise stored with k contiguous (A Trans, B NoTrans).
template <typename T, bool MN_MAJOR>
struct TileLoader {
  const T *base;  // element (mn0, 0) of this tile row/column panel
  int64_t ld;
  int mn_lim;     // valid mn extent inside this tile (<= 128)
  int idx_mn, idx_k;  // this thread's first (mn, k) in the tile

  __device__ __forceinline__ void init(const T *x, int64_t ld_, int mn0, int mn_total,
                                       int tid) {
    ld = ld_;
    mn_lim = mn_total - mn0;
    if (MN_MAJOR) {
      idx_mn = tid % kBM;
      idx_k = tid / kBM;  // 0..1, stepped by 2
      base = x + mn0;
    } else {
      idx_k = tid % kBK;
      idx_mn = tid / kBK;  // 0..15, stepped by 16
      base = x + (int64_t)mn0 * ld;
    }
  }

  // Load k-tile starting at k0 into registers; out-of-range elements are 0.
  __device__ __forceinline__ void load(T (&r)[kLoads], int k0, int k_total) const {
#pragma unroll
    for (int s = 0; s < kLoads; ++s) {
      int mn, k;
      if (MN_MAJOR) {
        mn = idx_mn;
        k = idx_k + 2 * s;
      } else {
        mn = idx_mn + 16 * s;
        k = idx_k;
      }
      const int kg = k0 + k;
      const bool ok = (mn < mn_lim) && (kg < k_total);
      const T *src = MN_MAJOR ? base + mn + (int64_t)kg * ld : base + kg + (int64_t)mn * ld;
      r[s] = ok ? *src : T(0);
    }
  }

  __device__ __forceinline__ void store(const T (&r)[kLoads], T (*lds)[kBM + Pad<T>::v]) const {
#pragma unroll
    for (int s = 0; s < kLoads; ++s) {
      if (MN_MAJOR)
        lds[idx_k + 2 * s][idx_mn] = r[s];
      else
        lds[idx_k][idx_mn + 16 * s] = r[s];
    }
  }
};

// Map the linear workgroup id to a (tile_m, tile_n) pair. Workgroups are
// dispatched round-robin over the 8 XCDs, so consecutive ids land on
// different L2s; we first regroup ids so that each XCD owns a contiguous run
// of tiles, then walk that run in column groups of kGroupM tile rows so the
// workgroups resident on one XCD share A and B panels in its L2.
__device__ __forceinline__ void tile_coords(int tiles_m, int tiles_n, int &tm, int &tn) {
  constexpr int kXcds = 8;
  constexpr int kGroupM = 8;
  const int total = tiles_m * tiles_n;
  int t = blockIdx.x;
  if (total % kXcds == 0) t = (t % kXcds) * (total / kXcds) + t / kXcds;
  const int per_group = kGroupM * tiles_n;
  const int group = t / per_group;
  const int first_m = group * kGroupM;
  const int gm = min(tiles_m - first_m, kGroupM);
  const int in_group = t % per_group;
  tm = first_m + in_group % gm;
  tn = in_group / gm;
}

template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(kThreads) void gemm_tile_kernel(GemmProblem<T> p) {
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int P = Pad<T>::v;
  __shared__ T As[2][kBK][kBM + P];
  __shared__ T Bs[2][kBK][kBN + P];

  const int M = (int)p.m, N = (int)p.n, K = (int)p.k;
  const int tiles_m = (M + kBM - 1) / kBM;
  const int tiles_n = (N + kBN - 1) / kBN;
  int tm, tn;
  tile_coords(tiles_m, tiles_n, tm, tn);
  const int m0 = tm * kBM, n0 = tn * kBN;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave & 1) * kWaveM;
  const int wn = (wave >> 1) * kWaveN;

  // A NoTrans is M-contiguous; B Trans is N-contiguous.
  TileLoader<T, !TA> la;
  TileLoader<T, TB> lb;
  la.init(p.a, p.lda, m0, M, tid);
  lb.init(p.b, p.ldb, n0, N, tid);

  acc_t acc[kFragM][kFragN];
#pragma unroll
  for (int i = 0; i < kFragM; ++i)
#pragma unroll
    for (int j = 0; j < kFragN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  T ra[kLoads], rb[kLoads];
  const int ktiles = (K + kBK - 1) / kBK;
  la.load(ra, 0, K);
  lb.load(rb, 0, K);
  la.store(ra, As[0]);
  lb.store(rb, Bs[0]);
  __syncthreads();

  const int fk = lane >> 4;   // k row of this lane's fragment within a k-step
  const int fmn = lane & 15;  // m/n column of this lane's fragment

  for (int kt = 0; kt < ktiles; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < ktiles;
    if (more) {
      la.load(ra, (kt + 1) * kBK, K);
      lb.load(rb, (kt + 1) * kBK, K);
    }
#pragma unroll
    for (int ks = 0; ks < kBK; ks += 4) {
      T fa[kFragM], fb[kFragN];
#pragma unroll
      for (int i = 0; i < kFragM; ++i) fa[i] = As[cur][ks + fk][wm + 16 * i + fmn];
#pragma unroll
      for (int j = 0; j < kFragN; ++j) fb[j] = Bs[cur][ks + fk][wn + 16 * j + fmn];
#pragma unroll
      for (int i = 0; i < kFragM; ++i)
#pragma unroll
        for (int j = 0; j < kFragN; ++j) acc[i][j] = Mfma<T>::mma(fb[j], fa[i], acc[i][j]);
    }
    if (more) {
      la.store(ra, As[cur ^ 1]);
      lb.store(rb, Bs[cur ^ 1]);
    }
    __syncthreads();
  }

  // Epilogue: element (i, j) of C gets alpha*acc + beta*C; beta == 0 never
  // reads C. Lane l of block (fi, fj) owns row i = m0+wm+16fi+l%16 and
  // columns j = n0+wn+16fj+row(l, r).
  const T alpha = p.alpha, beta = p.beta;
  const bool read_c = beta != T(0);
#pragma unroll
  for (int i = 0; i < kFragM; ++i) {
    const int gi = m0 + wm + 16 * i + fmn;
    if (gi >= M) continue;
#pragma unroll
    for (int j = 0; j < kFragN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn + 16 * j + Mfma<T>::row(lane, r);
        if (gj >= N) continue;
        T *cp = p.c + gi + (int64_t)gj * p.ldc;
        const T v = acc[i][j][r];
        *cp = read_c ? __builtin_fma(alpha, v, beta * *cp) : alpha * v;
      }
    }
  }
}

template <typename T>
__global__ void scale_kernel(int64_t m, int64_t n, T beta, T *c, int64_t ldc) {
  const int64_t total = m * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx / m, i = idx - j * m;
    T *cp = c + i + j * ldc;
    *cp = beta == T(0) ? T(0) : beta * *cp;
  }
}

}  // namespace

template <typename T>
hipError_t launch_gemm(const GemmProblem<T> &p, hipStream_t stream) {
  const int64_t tiles = ((p.m + kBM - 1) / kBM) * ((p.n + kBN - 1) / kBN);
  dim3 grid((unsigned)tiles), block(kThreads);
  if (!p.ta && !p.tb)
    hipLaunchKernelGGL((gemm_tile_kernel<T, false, false>), grid, block, 0, stream, p);
  else if (!p.ta && p.tb)
    hipLaunchKernelGGL((gemm_tile_kernel<T, false, true>), grid, block, 0, stream, p);
  else if (p.ta && !p.tb)
    hipLaunchKernelGGL((gemm_tile_kernel<T, true, false>), grid, block, 0, stream, p);
  else
    hipLaunchKernelGGL((gemm_tile_kernel<T, true, true>), grid, block, 0, stream, p);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale(int64_t m, int64_t n, T beta, T *c, int64_t ldc, hipStream_t stream) {
  const int64_t total = m * n;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((scale_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, stream, m, n, beta,
                     c, ldc);
  return hipGetLastError();
}

template hipError_t launch_gemm<float>(const GemmProblem<float> &, hipStream_t);
template hipError_t launch_gemm<double>(const GemmProblem<double> &, hipStream_t);
template hipError_t launch_scale<float>(int64_t, int64_t, float, float *, int64_t, hipStream_t);
template hipError_t launch_scale<double>(int64_t, int64_t, double, double *, int64_t,
                                         hipStream_t);

}  // namespace rblas
