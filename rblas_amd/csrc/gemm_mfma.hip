// Single-problem MFMA GEMM for gfx950 (MI355X): C := alpha * op(A) * op(B) +
// beta * C, column-major, f32 (v_mfma_f32_16x16x4f32) and f64
// (v_mfma_f64_16x16x4f64). The tile machinery is in tile_gemm.hpp.
//
// This replaces rblas's Goto/BLIS pipeline (kernel/gemm.rs:185-474 and the
// AVX2 microkernels gemm_avx2.rs:26-155):
//   - pack_a_panel / pack_b_panel -> the buffer unit's direct global->LDS
//     loads (`buffer_load_dwordx4 ... lds`), three LDS stages in flight, one
//     barrier per k-tile, no staging registers;
//   - the MR x NR register microtile -> a per-wave block of 16 x 16 MFMA
//     accumulators (64 x 32 per wave in a 128 x 128 tile, 8 waves);
//   - the jc/pc/ic loops -> one workgroup per C tile (and per k split when
//     the tile grid alone cannot fill the 256 CUs), the k range accumulated
//     in registers;
//   - materialize (level3.rs:179-222) -> none: each operand is staged from
//     its stored orientation (M/N-contiguous or K-contiguous).
#include <algorithm>

// Three LDS stages: measured faster than four for the 128 x 128 tile
// (8192^3 DGEMM 66.6 against 65.9 TFLOP/s, same bits).
#define RBLAS_TILE_STAGES 3
#include "tile_gemm.hpp"

namespace rblas {
namespace {

// Tiles of the launch: the full tiles_m x tiles_n grid, or its triangle.
__device__ __forceinline__ int tile_count(int tiles_m, int tiles_n, int tri) {
  return tri ? tiles_m * (tiles_m + 1) / 2 : tiles_m * tiles_n;
}

// Workgroup u handles tile u % tiles of k split u / tiles. Splits cut the
// k-tiles into `splits` contiguous runs of `per` k-tiles; see split_handoff
// for how the partial sums meet.
template <typename T, bool TA, bool TB, int BM, int BN>
__global__ __launch_bounds__(kThreads) void gemm_tile_kernel(GemmProblem<T> p, int splits,
                                                             T *ws, int *counters) {
  typedef Geo<T> G;
  TileGemm<T, TA, TB, BM, BN> tg;
  const int M = (int)p.m, N = (int)p.n, K = (int)p.k;
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  const int tiles = tile_count(tiles_m, tiles_n, p.tri);
  const int split = blockIdx.x / tiles, t = blockIdx.x % tiles;
  int tm, tn;
  if (p.tri == 0)
    tile_coords(t, tiles_m, tiles_n, tm, tn);
  else
    tri_coords(t, p.tri, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int ktiles = (K + G::BK - 1) / G::BK;
  const int per = (ktiles + splits - 1) / splits;
  const int kb = min(K, split * per * G::BK), ke = min(K, (split + 1) * per * G::BK);
  tg.init(p, m0, n0, kb, ke);
  tg.mainloop();
  if (splits > 1 && !split_handoff(tg, ws, (int64_t)t * splits, splits, split, counters + t))
    return;
  tg.epilogue(p, m0, n0);
}

__global__ void zero_ints(int *x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0;
}

template <typename T>
__global__ void scale_tri_kernel(int64_t n, int tri, T beta, T *c, int64_t ldc) {
  const int64_t total = n * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx / n, i = idx - j * n;
    if ((tri == 1 && i > j) || (tri == 2 && i < j)) continue;
    T *cp = c + i + j * ldc;
    *cp = beta == T(0) ? T(0) : beta * *cp;
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

template <typename T, int BM, int BN>
hipError_t launch_tiles(const GemmProblem<T> &p, int splits, T *ws, int *counters,
                        hipStream_t st) {
  int64_t tiles = ((p.m + BM - 1) / BM) * ((p.n + BN - 1) / BN);
  if (p.tri != 0) {
    const int64_t t = (p.m + BM - 1) / BM;
    tiles = t * (t + 1) / 2;
  }
  dim3 grid((unsigned)(tiles * splits)), block(kThreads);
  if (!p.ta && !p.tb)
    hipLaunchKernelGGL((gemm_tile_kernel<T, false, false, BM, BN>), grid, block, 0, st, p, splits,
                       ws, counters);
  else if (!p.ta && p.tb)
    hipLaunchKernelGGL((gemm_tile_kernel<T, false, true, BM, BN>), grid, block, 0, st, p, splits,
                       ws, counters);
  else if (p.ta && !p.tb)
    hipLaunchKernelGGL((gemm_tile_kernel<T, true, false, BM, BN>), grid, block, 0, st, p, splits,
                       ws, counters);
  else
    hipLaunchKernelGGL((gemm_tile_kernel<T, true, true, BM, BN>), grid, block, 0, st, p, splits,
                       ws, counters);
  return hipGetLastError();
}

}  // namespace

template <typename T>
bool gemm_fits(const GemmProblem<T> &p) {
  // Largest in-tile byte offset for either operand orientation or for C.
  int64_t ld = p.lda > p.ldb ? p.lda : p.ldb;
  if (p.ldc > ld) ld = p.ldc;
  const int64_t span = (int64_t)(kMaxTile - 1) * ld + kMaxTile + Geo<T>::BK;
  return span * (int64_t)sizeof(T) < 0x7fffffffll - 16;
}

template <typename T>
GemmConfig gemm_config(const GemmProblem<T> &p) {
  GemmConfig cfg{128, 1};
  if (p.tri != 0) return cfg;
  if (gemm_wide_ok(p)) return GemmConfig{256, 1};
  const int64_t t128 = ((p.m + 127) / 128) * ((p.n + 127) / 128);
  const int64_t t64 = ((p.m + 63) / 64) * ((p.n + 63) / 64);
  const int64_t ktiles = (p.k + Geo<T>::BK - 1) / Geo<T>::BK;
  if (t128 >= kCUs) return cfg;
  // Under one 128-tile per CU: long k is better served by 128-tiles with k
  // split until the grid fills the CUs (each split keeping >= 64 k-tiles, so
  // the partial traffic stays small); otherwise 64-tiles, split as well when
  // even they leave most CUs idle.
  int64_t s = std::min<int64_t>((kCUs + t128 - 1) / t128, ktiles / 64);
  s = std::min<int64_t>(s, 16);
  if (s >= 2 && t128 * s >= kCUs / 2) {
    cfg.splits = (int)s;
    return cfg;
  }
  cfg.tile = 64;
  if (t64 < kCUs / 2) {
    s = std::min<int64_t>((kCUs + t64 - 1) / t64, ktiles / 64);
    s = std::min<int64_t>(s, 16);
    if (s >= 2) cfg.splits = (int)s;
  }
  return cfg;
}

template <typename T>
size_t gemm_workspace_bytes(const GemmProblem<T> &p, GemmConfig cfg) {
  if (cfg.splits <= 1) return 0;
  const int64_t tiles = ((p.m + cfg.tile - 1) / cfg.tile) * ((p.n + cfg.tile - 1) / cfg.tile);
  const size_t partial = (size_t)tiles * cfg.splits * cfg.tile * cfg.tile * sizeof(T);
  return ((partial + 255) & ~(size_t)255) + (size_t)tiles * sizeof(int);
}

template <typename T>
hipError_t launch_gemm_cfg(const GemmProblem<T> &p, GemmConfig cfg, void *ws, hipStream_t st) {
  if (cfg.tile == 256) return cfg.splits == 1 ? launch_gemm_wide(p, st) : hipErrorInvalidValue;
  if (!gemm_fits(p)) return hipErrorInvalidValue;
  if (p.tri != 0 && (p.m != p.n || cfg.splits != 1)) return hipErrorInvalidValue;
  if (cfg.tile != 64 && cfg.tile != 128) return hipErrorInvalidValue;
  T *part = nullptr;
  int *counters = nullptr;
  if (cfg.splits > 1) {
    if (ws == nullptr) return hipErrorInvalidValue;
    const int64_t tiles =
        ((p.m + cfg.tile - 1) / cfg.tile) * ((p.n + cfg.tile - 1) / cfg.tile);
    const size_t partial = (size_t)tiles * cfg.splits * cfg.tile * cfg.tile * sizeof(T);
    part = (T *)ws;
    counters = (int *)((char *)ws + ((partial + 255) & ~(size_t)255));
    hipLaunchKernelGGL(zero_ints, dim3((unsigned)((tiles + 255) / 256)), dim3(256), 0, st,
                       counters, (int)tiles);
  }
  if (cfg.tile == 128) return launch_tiles<T, 128, 128>(p, cfg.splits, part, counters, st);
  return launch_tiles<T, 64, 64>(p, cfg.splits, part, counters, st);
}

template <typename T>
hipError_t launch_gemm(const GemmProblem<T> &p, hipStream_t st) {
  GemmConfig cfg = gemm_config(p);
  // Without a workspace there is no k split: where the config wanted one
  // (fewer than one 128-tile per CU), 64-tiles give 4x the workgroups.
  if (cfg.splits > 1) cfg.tile = 64;
  cfg.splits = 1;
  return launch_gemm_cfg(p, cfg, nullptr, st);
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

template <typename T>
hipError_t launch_scale_tri(int64_t n, int tri, T beta, T *c, int64_t ldc, hipStream_t stream) {
  int64_t blocks = (n * n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((scale_tri_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, stream, n, tri,
                     beta, c, ldc);
  return hipGetLastError();
}

#define RBLAS_GEMM_INST(T)                                                                     \
  template bool gemm_fits<T>(const GemmProblem<T> &);                                          \
  template GemmConfig gemm_config<T>(const GemmProblem<T> &);                                  \
  template size_t gemm_workspace_bytes<T>(const GemmProblem<T> &, GemmConfig);                 \
  template hipError_t launch_gemm_cfg<T>(const GemmProblem<T> &, GemmConfig, void *,           \
                                         hipStream_t);                                         \
  template hipError_t launch_gemm<T>(const GemmProblem<T> &, hipStream_t);                     \
  template hipError_t launch_scale<T>(int64_t, int64_t, T, T *, int64_t, hipStream_t);         \
  template hipError_t launch_scale_tri<T>(int64_t, int, T, T *, int64_t, hipStream_t);
RBLAS_GEMM_INST(float)
RBLAS_GEMM_INST(double)
#undef RBLAS_GEMM_INST

}  // namespace rblas
