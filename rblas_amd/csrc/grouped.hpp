// Device-side descriptors of the grouped (persistent) launches and their
// host launch functions.
#pragma once
#include "kernels.hpp"

namespace rblas {

// One gemm problem of a group, column-major after rblas's RowMajor swap.
template <typename T>
struct GroupDesc {
  const T *a;
  const T *b;
  T *c;
  int64_t lda, ldb, ldc;
  int m, n, k;
  int ta, tb;     // operand stored transposed (0/1)
  int tile;       // C tile side: 64 or 128
  int splits;     // k splits per tile (1 = none)
  int tiles_m;    // tiles along m (tile index t -> (t % tiles_m, t / tiles_m))
  int mode;       // 0: product, 1: C := beta C only (k == 0 or alpha == 0)
  T alpha, beta;
  int64_t slot0;  // element offset of this problem's split partials in the workspace
  int cnt0;       // first per-tile counter of this problem (split problems)
};

// One unit of work: C tile `tile` (and k split `split`) of problem `prob`.
struct WorkUnit {
  int prob, tile, split, pad;
};

// Workgroups the persistent launch keeps resident (occupancy x CUs).
template <typename T>
int grouped_gemm_grid();

// One persistent launch over `units` (nunits of them). sched[0..1] must be 0
// before the first launch (the launch leaves them 0); `counters` holds one
// zero-initialised int per split tile, `ws` the split partials
// (tile * tile values per (tile, split) slot).
template <typename T>
hipError_t launch_grouped_gemm(const GroupDesc<T> *descs, const WorkUnit *units, int nunits,
                               int grid, int *sched, T *ws, int *counters, hipStream_t stream);

hipError_t launch_fill_ints(int *x, int64_t n, int v, hipStream_t stream);

}  // namespace rblas
