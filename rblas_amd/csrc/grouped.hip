// Grouped GEMM: hundreds of independent problems C_i := alpha_i op(A_i)
// op(B_i) + beta_i C_i (column-major, any transposes, uneven sizes) in ONE
// persistent launch.
//
// The host planner (grouped_api.cpp) turns every problem into work units --
// one per (C tile, k split), tile side 64 or 128 per problem -- and orders
// them longest first. The launch keeps about as many workgroups as fit on
// the 256 CUs; each pulls the next unit from a device-scope counter until
// the list is exhausted, so long and short units balance across the chip.
// The last workgroup to leave resets the counters, so the same launch can be
// replayed (hipGraph) without a memset. Split units meet through
// split_handoff (tile_gemm.hpp): ordered, so results do not depend on which
// workgroup ran what.
#include "grouped.hpp"
#include "tile_gemm.hpp"

namespace rblas {
namespace {

// Descriptor fields are loaded with vector loads, so the compiler cannot see
// that they are the same in every lane; readfirstlane makes them scalar again,
// so buffer resources and LDS addresses built from them need no per-lane loop.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t uni(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <typename P>
__device__ __forceinline__ P *uni(P *p) {
  return (P *)uni((int64_t)p);
}

// The descriptor, loaded once and made wave-uniform field by field.
template <typename T>
__device__ __forceinline__ GroupDesc<T> load_desc(const GroupDesc<T> *dp) {
  GroupDesc<T> d = *dp;
  d.a = uni(d.a);
  d.b = uni(d.b);
  d.c = uni(d.c);
  d.lda = uni(d.lda);
  d.ldb = uni(d.ldb);
  d.ldc = uni(d.ldc);
  d.m = uni(d.m);
  d.n = uni(d.n);
  d.k = uni(d.k);
  d.splits = uni(d.splits);
  d.tiles_m = uni(d.tiles_m);
  d.slot0 = uni(d.slot0);
  d.cnt0 = uni(d.cnt0);
  return d;
}

// One unit of one variant (transposes, tile side). TileGemm::init derives
// its lane addresses from an opaque thread id, so they are recomputed per
// unit rather than hoisted out of the persistent loop for all eight variants
// at once (which would need more registers than a wave has).
template <typename T, bool TA, bool TB, int BM>
__device__ __forceinline__ void run_product(const GroupDesc<T> *dp, int tile, int split, T *ws,
                                            int *counters) {
  typedef Geo<T> G;
  const GroupDesc<T> d = load_desc(dp);
  GemmProblem<T> p{TA, TB, d.m, d.n, d.k, d.alpha, d.beta, d.a, d.lda, d.b, d.ldb, d.c, d.ldc};
  const int tm = tile % d.tiles_m, tn = tile / d.tiles_m;
  const int m0 = tm * BM, n0 = tn * BM;
  const int ktiles = (d.k + G::BK - 1) / G::BK;
  const int per = (ktiles + d.splits - 1) / d.splits;
  const int kb = min(d.k, split * per * G::BK), ke = min(d.k, (split + 1) * per * G::BK);
  TileGemm<T, TA, TB, BM, BM> tg;
  tg.init(p, m0, n0, kb, ke);
  tg.mainloop();
  if (d.splits > 1 &&
      !split_handoff(tg, ws + d.slot0, (int64_t)tile * d.splits, d.splits, split,
                     counters + d.cnt0 + tile))
    return;
  tg.epilogue(p, m0, n0);
}

template <typename T, int BM>
__device__ __forceinline__ void run_tile(const GroupDesc<T> *dp, int ta, int tb, int tile,
                                         int split, T *ws, int *counters) {
  switch (ta * 2 + tb) {
    case 0: run_product<T, false, false, BM>(dp, tile, split, ws, counters); break;
    case 1: run_product<T, false, true, BM>(dp, tile, split, ws, counters); break;
    case 2: run_product<T, true, false, BM>(dp, tile, split, ws, counters); break;
    default: run_product<T, true, true, BM>(dp, tile, split, ws, counters); break;
  }
}

// k == 0 or alpha == 0: C := beta C on one 128 x 128 tile (beta == 0 stores
// exact zeros, kernel/gemm.rs:322-338).
template <typename T>
__device__ __forceinline__ void run_scale(const GroupDesc<T> &d, const WorkUnit &w) {
  const int tm = w.tile % d.tiles_m, tn = w.tile / d.tiles_m;
  const int m0 = tm * 128, n0 = tn * 128;
  for (int e = threadIdx.x; e < 128 * 128; e += kThreads) {
    const int i = m0 + e % 128, j = n0 + e / 128;
    if (i >= d.m || j >= d.n) continue;
    T *cp = d.c + i + (int64_t)j * d.ldc;
    *cp = d.beta == T(0) ? T(0) : d.beta * *cp;
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void grouped_gemm_kernel(const GroupDesc<T> *descs,
                                                                const WorkUnit *units, int nunits,
                                                                int *sched, T *ws,
                                                                int *counters) {
  __shared__ int s_unit;
  for (;;) {
    if (threadIdx.x == 0) s_unit = atomicAdd(&sched[0], 1);
    __syncthreads();
    const int u = __builtin_amdgcn_readfirstlane(s_unit);
    __syncthreads();  // every wave has read s_unit before it is overwritten
    if (u >= nunits) break;
    const WorkUnit w = units[u];
    const GroupDesc<T> *dp = descs + w.prob;
    const int mode = dp->mode, tile_side = dp->tile, ta = dp->ta, tb = dp->tb;
    if (mode == 1)
      run_scale(*dp, w);
    else if (tile_side == 128)
      run_tile<T, 128>(dp, ta, tb, w.tile, w.split, ws, counters);
    else
      run_tile<T, 64>(dp, ta, tb, w.tile, w.split, ws, counters);
  }
  // Every workgroup has drawn past the end before it counts itself out, so
  // the last one out can reset the schedule for the next launch.
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&sched[1], 1) == (int)gridDim.x - 1) {
      atomicExch(&sched[0], 0);
      atomicExch(&sched[1], 0);
    }
  }
}

__global__ void fill_ints(int *x, int64_t n, int v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

}  // namespace

template <typename T>
int grouped_gemm_grid() {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grouped_gemm_kernel<T>, kThreads, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return per_cu * kCUs;
}

template <typename T>
hipError_t launch_grouped_gemm(const GroupDesc<T> *descs, const WorkUnit *units, int nunits,
                               int grid, int *sched, T *ws, int *counters, hipStream_t st) {
  if (nunits <= 0) return hipSuccess;
  if (grid > nunits) grid = nunits;
  hipLaunchKernelGGL((grouped_gemm_kernel<T>), dim3((unsigned)grid), dim3(kThreads), 0, st, descs,
                     units, nunits, sched, ws, counters);
  return hipGetLastError();
}

hipError_t launch_fill_ints(int *x, int64_t n, int v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_ints, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n, v);
  return hipGetLastError();
}

template int grouped_gemm_grid<float>();
template int grouped_gemm_grid<double>();
template hipError_t launch_grouped_gemm<float>(const GroupDesc<float> *, const WorkUnit *, int, int,
                                               int *, float *, int *, hipStream_t);
template hipError_t launch_grouped_gemm<double>(const GroupDesc<double> *, const WorkUnit *, int,
                                                int, int *, double *, int *, hipStream_t);

}  // namespace rblas
