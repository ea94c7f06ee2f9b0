// Host side of the grouped launches: the planner's cost model, and the
// tables of one grouped gemm / trsm launch built problem by problem. Used by
// the grouped-plan C-ABI (persistent plans) and by the serving loop (one
// group per batch, uploaded asynchronously).
//
// Device memory of a group is one block laid out as
//   descriptors | work units | [blocks (trsm)] | sched + counters | workspace
// where everything before the workspace is staged on the host (zeros for
// sched and counters) and uploaded with one copy.
#pragma once
#include <vector>

#include "grouped.hpp"
#include "../../include/rblas_hip.h"

namespace rblas {

// Per-unit time model of a grouped gemm unit: c0 + c1 * (its k-tiles)
// microseconds of one CU, per (dtype f32/f64, tile side 64/128), and the
// longest unit left unsplit.
struct Tuning {
  double c0[2][2];
  double c1[2][2];
  double unit_cap_us[2];
};
Tuning tuning_snapshot();
void tuning_set(int dtype, const double *coeffs);  // {c0_64, c1_64, c0_128, c1_128, cap}
void tuning_get(int dtype, double *coeffs);

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename T>
struct GemmGroup {
  std::vector<GroupDesc<T>> descs;
  std::vector<WorkUnit> units;
  std::vector<double> cost;
  int64_t ws_elems = 0, counters = 0;
  double total_us = 0;
  rblas_grouped_info_t info{};

  // One normalized column-major problem with m, n > 0 (k == 0 or
  // alpha == 0: C := beta C; the caller drops beta == 1 no-ops).
  void add(const GemmProblem<T> &p, const Tuning &tu);
  // Orders the units longest first (stable).
  void finish();

  size_t off_units() const { return align256(descs.size() * sizeof(GroupDesc<T>)); }
  size_t off_sched() const { return off_units() + align256(units.size() * sizeof(WorkUnit)); }
  size_t off_ws() const { return off_sched() + align256(256 + (size_t)counters * sizeof(int)); }
  size_t staged_bytes() const { return off_ws(); }
  size_t bytes() const { return off_ws() + (size_t)ws_elems * sizeof(T); }
  size_t ws_bytes() const { return (size_t)ws_elems * sizeof(T); }
  void stage(char *host) const;  // writes staged_bytes() bytes
  // `dev` holds the staged tables; the workspace is at `ws`, or right after
  // the tables when ws is null.
  hipError_t launch(char *dev, int grid, hipStream_t st, char *ws = nullptr) const;
};

template <typename T>
struct TrsmGroup {
  std::vector<TrsmDesc<T>> descs;
  std::vector<int2> blocks, units;
  std::vector<double> cost;
  int64_t inv_elems = 0;
  int qstart[kTrsmQueues + 1] = {};  // units of queue x: [qstart[x], qstart[x + 1])
  rblas_grouped_info_t info{};

  // One problem with m, n > 0 and the order of A at most 2048, as rblas's
  // trsm arguments (stored A, column-major B).
  void add(bool left, bool trans, bool eff_upper, bool unit, int64_t m, int64_t n, T alpha,
           const T *a, int64_t lda, T *b, int64_t ldb);
  // Splits the units into one queue per XCD: a system's units stay together
  // (up to 32, one XCD's CUs, per chunk) so that they share op(A) through
  // that XCD's L2; chunks go longest unit first to the least loaded queue.
  void finish();

  size_t off_units() const { return align256(descs.size() * sizeof(TrsmDesc<T>)); }
  size_t off_blocks() const { return off_units() + align256(units.size() * sizeof(int2)); }
  size_t off_sched() const { return off_blocks() + align256(blocks.size() * sizeof(int2)); }
  size_t off_ws() const { return off_sched() + kTrsmSchedBytes; }
  size_t staged_bytes() const { return off_ws(); }
  size_t bytes() const { return off_ws() + (size_t)inv_elems * sizeof(double); }
  size_t ws_bytes() const { return (size_t)inv_elems * sizeof(double); }
  void stage(char *host) const;
  hipError_t launch(char *dev, int grid, hipStream_t st, char *ws = nullptr) const;
};

}  // namespace rblas
