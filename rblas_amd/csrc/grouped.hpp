// Device-side descriptors of the grouped (persistent) launches and their
// host launch functions.
#pragma once
#include <algorithm>

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

// One trsm problem of a group, in Left form: op(A) X = alpha B on `nrhs`
// right-hand sides of length q. A Right problem (X op(A) = alpha B) is stored
// as its transpose: trans and upper flipped, right = 1 (right-hand side c is
// row c of B).
template <typename T>
struct TrsmDesc {
  const T *a;
  T *b;
  int64_t lda, ldb;
  int q, nrhs;
  int trans, upper, unit, right;
  int forward;  // substitution from row 0 up (lower triangle in Left form)
  int width;    // right-hand sides per work unit (trsm_unit_width)
  T alpha;
  int64_t inv0;  // element offset of this problem's diagonal-block inverses
};

// q * width must fit the unit's LDS (f64).
constexpr int kTrsmLdsDoubles = 16384;

// Work-unit queues of a grouped trsm launch, one per XCD (workgroup i runs on
// XCD i % 8). The launch's sched block (kTrsmSchedBytes) holds the claim
// counter of queue x at int x * kTrsmLine and the exit counter at int
// kTrsmQueues * kTrsmLine -- each on a 128-byte line of its own, so claims on
// different queues do not queue up behind one another -- and, from int
// kTrsmQueueStart, the queues' unit offsets (kTrsmQueues + 1 ints, staged by
// the host).
constexpr int kTrsmQueues = 8;
constexpr int kTrsmLine = 32;
constexpr int kTrsmQueueStart = (kTrsmQueues + 1) * kTrsmLine;
constexpr int kTrsmSchedBytes = 1280;
static_assert((kTrsmQueueStart + kTrsmQueues + 1) * 4 <= kTrsmSchedBytes, "sched block");

// Right-hand sides per trsm work unit. Held in registers, orders up to 1024:
// 16, or 32 (each op(A) row block then feeds twice the MFMAs; fewer units);
// both need 32-bit buffer offsets up to (63 lda + 15) elements for their
// op(A) row loads. Held in LDS: 8, orders up to 2048. trsm_unit_width gives
// 16 or 8; trsm_wide_ok says whether 32 is possible, which a caller takes
// only while the units still fill about three quarters of the CUs
// (kTrsmFillUnits).
constexpr int kTrsmWideMaxOrder = 1024;
// A unit {problem, first right-hand side} of a grouped trsm stores its
// problem as ~problem (negative) when it runs 16 wide in a 32-wide problem
// (the problem's last unit, at most 16 right-hand sides).
__host__ __device__ inline int trsm_unit_problem(int x) { return x < 0 ? ~x : x; }
constexpr int kTrsmFillUnits = 3 * kCUs / 4;
template <typename T>
inline int trsm_unit_width(int64_t q, int64_t lda) {
  const bool fits = lda * 63 + 15 < ((int64_t)1 << 31) / (int64_t)sizeof(T);
  return q <= kTrsmLdsDoubles / 16 && fits ? 16 : 8;
}
template <typename T>
inline bool trsm_wide_ok(int64_t q, int64_t lda, int64_t nrhs) {
  return trsm_unit_width<T>(q, lda) == 16 && q <= kTrsmWideMaxOrder && nrhs > 16;
}

// A single call's register-resident solve is split over several workgroups
// per unit (consecutive runs of the diagonal blocks, in substitution order)
// when its units would leave CUs idle and it has at least
// kTrsmSplitMinBlocks diagonal blocks: the number of parts (1: no split),
// at most kTrsmSplitMaxParts and as many as the CUs hold. width 16 or 32
// only (the register-resident solve), orders up to kTrsmWideMaxOrder. Each
// step's chain is bound by the latency of its op(A) row loads, so fewer rows
// per workgroup shorten it only a little, and small orders, with few steps,
// lose to the hand-off.
constexpr int kTrsmSplitMinBlocks = 12;
constexpr int kTrsmSplitMaxParts = 4;
inline int trsm_split_parts(int64_t q, int64_t nrhs, int width) {
  if (width != 16 && width != 32) return 1;
  if (q < (kTrsmSplitMinBlocks - 1) * 64 + 1 || q > kTrsmWideMaxOrder) return 1;
  const int64_t units = (nrhs + width - 1) / width;
  const int64_t p = std::min<int64_t>(kTrsmSplitMaxParts, kCUs / units);
  return p >= 2 ? (int)p : 1;
}
// Doubles of the published X of a split solve (every unit, every block),
// for any width: at most (nrhs + 31) right-hand sides of the padded order.
inline int64_t trsm_split_x_doubles(int64_t q, int64_t nrhs) {
  return (nrhs + 31) * ((q + 63) / 64 * 64);
}

template <typename T>
int grouped_trsm_grid();

// One problem (d.inv0 = 0) without tables: one launch inverting its
// diagonal blocks into `inv`, one launch of one workgroup per `width`
// right-hand sides. Requires d.q * d.width <= kTrsmLdsDoubles.
template <typename T>
hipError_t launch_trsm_single(const TrsmDesc<T> &d, double *inv, hipStream_t stream);

// Inverts every 64 x 64 diagonal block of op(A) (d.a, d.lda, d.q, d.trans,
// d.upper, d.unit; other fields unused) into inv + d.inv0, 4096 f64 each.
template <typename T>
hipError_t launch_trsm_invert(const TrsmDesc<T> &d, double *inv, hipStream_t stream);

// Inverts the diagonal blocks listed in `blocks` ({problem, block}), then
// runs one persistent launch over `units` ({problem, first right-hand
// side}) in kTrsmQueues queues (sched: counters zero, queue offsets staged at
// kTrsmQueueStart; reset by the launch itself). `inv` holds 4096 f64 per
// diagonal block.
template <typename T>
hipError_t launch_grouped_trsm(const TrsmDesc<T> *descs, const int2 *blocks, int nblocks,
                               const int2 *units, int nunits, int grid, int *sched, double *inv,
                               hipStream_t stream);

}  // namespace rblas
