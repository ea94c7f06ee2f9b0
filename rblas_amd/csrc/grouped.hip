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
// Three LDS stages: measured faster than four for the 128 x 128 tile
// (8192^3 DGEMM 66.6 against 65.9 TFLOP/s, same bits).
#define RBLAS_TILE_STAGES 3
#include "tile_gemm.hpp"

namespace rblas {
namespace {

// Descriptors and work units are never written by the launch, so they are
// read through the constant address space: uniform addresses then load with
// scalar loads straight into SGPRs.
template <typename X>
__device__ __forceinline__ X const_load(const X *p) {
  static_assert(sizeof(X) % 4 == 0, "whole words");
  typedef const __attribute__((address_space(4))) uint32_t cword;
  const cword *q = (const cword *)p;
  X x;
  uint32_t *w = (uint32_t *)&x;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(X) / 4); ++i) w[i] = q[i];
  return x;
}

// One unit of one variant (transposes, tile side). TileGemm::init derives
// its lane addresses from an opaque thread id, so they are recomputed per
// unit rather than hoisted out of the persistent loop for all eight variants
// at once (which would need more registers than a wave has). Once the k loop
// is done (its last barrier drained every load, the claim's atomic with
// them), thread 0 publishes the next unit's index in *s_next.
template <typename T, bool TA, bool TB, int BM>
__device__ __forceinline__ void run_product(const GroupDesc<T> &d, const WorkUnit &w, T *ws,
                                            int *counters, int next, int *s_next) {
  typedef Geo<T> G;
  GemmProblem<T> p{TA, TB, d.m, d.n, d.k, d.alpha, d.beta, d.a, d.lda, d.b, d.ldb, d.c, d.ldc};
  const int tm = w.tile % d.tiles_m, tn = w.tile / d.tiles_m;
  const int m0 = tm * BM, n0 = tn * BM;
  const int ktiles = (d.k + G::BK - 1) / G::BK;
  const int per = (ktiles + d.splits - 1) / d.splits;
  const int kb = min(d.k, w.split * per * G::BK), ke = min(d.k, (w.split + 1) * per * G::BK);
  TileGemm<T, TA, TB, BM, BM> tg;
  tg.init(p, m0, n0, kb, ke);
  tg.mainloop();
  if (threadIdx.x == 0) *s_next = next;
  if (d.splits > 1 &&
      !split_handoff(tg, ws + d.slot0, (int64_t)w.tile * d.splits, d.splits, w.split,
                     counters + d.cnt0 + w.tile))
    return;
  tg.epilogue(p, m0, n0);
}

template <typename T, int BM>
__device__ __forceinline__ void run_tile(const GroupDesc<T> &d, const WorkUnit &w, T *ws,
                                         int *counters, int next, int *s_next) {
  switch (d.ta * 2 + d.tb) {
    case 0: run_product<T, false, false, BM>(d, w, ws, counters, next, s_next); break;
    case 1: run_product<T, false, true, BM>(d, w, ws, counters, next, s_next); break;
    case 2: run_product<T, true, false, BM>(d, w, ws, counters, next, s_next); break;
    default: run_product<T, true, true, BM>(d, w, ws, counters, next, s_next); break;
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

// The workgroup claims its next unit as it starts the current one (one
// atomic in flight under the unit's k loop), so the claim's latency is off
// the path between units. The index goes through two LDS slots used in
// turn, and units are separated by a barrier that waits for LDS only: a
// unit's C stores drain while the next unit loads its descriptor and its
// first stages.
template <typename T>
__global__ __launch_bounds__(kThreads) void grouped_gemm_kernel(const GroupDesc<T> *descs,
                                                                const WorkUnit *units, int nunits,
                                                                int *sched, T *ws,
                                                                int *counters) {
  __shared__ int s_unit[2];
  if (threadIdx.x == 0) s_unit[0] = atomicAdd(&sched[0], 1);
  __syncthreads();
  int u = __builtin_amdgcn_readfirstlane(s_unit[0]);
  int slot = 0;
  while (u < nunits) {
    int next = 0;
    if (threadIdx.x == 0) next = atomicAdd(&sched[0], 1);
    int *s_next = &s_unit[slot ^ 1];
    const WorkUnit w = const_load(units + u);
    const GroupDesc<T> d = const_load(descs + w.prob);
    if (d.mode == 1) {
      run_scale(d, w);
      if (threadIdx.x == 0) *s_next = next;
    } else if (d.tile == 128) {
      run_tile<T, 128>(d, w, ws, counters, next, s_next);
    } else {
      run_tile<T, 64>(d, w, ws, counters, next, s_next);
    }
    stage_barrier<63>();
    slot ^= 1;
    u = __builtin_amdgcn_readfirstlane(s_unit[slot]);
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
