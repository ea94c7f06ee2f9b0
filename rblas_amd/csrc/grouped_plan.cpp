// Grouped planner: per-problem launch shape from the measured cost model,
// work units ordered longest first, and the staged device tables.
#include "grouped_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace rblas {
namespace {

// Defaults: the autotuner's measurement on MI355X (rblas_amd/autotune.py,
// saved in rblas_amd/tuning_gfx950.json). The unit cap sits where splitting
// stopped paying on the batching workload (for f64 no cap timed 3% under a
// cap of 320 us; 320 is kept so that very long-k units stay splittable).
Tuning g_tuning = {{{4.189, 8.065}, {3.941, 10.431}},
                   {{0.779, 2.188}, {0.751, 2.141}},
                   {320.0, 320.0}};
std::mutex g_mu;

constexpr int kMinSplitKtiles = 16;
constexpr int kMaxSplits = 16;

int k_tile(int dtype) { return dtype == 0 ? 32 : 16; }

struct Shape {
  int tile, splits, tiles_m, tiles;
  double unit_us;
};

// k-loop time of a tile reaching rem_m x rem_n into the matrix, as a share of
// a full tile's: wave w holds row blocks 2i + w / 4 and column blocks
// 4j + w % 4 and skips those past the edge, so the tile takes about as long
// as its busiest SIMD's share of the blocks allows, above a floor (fitted
// to edge tiles timed alone on MI355X: 128-tiles 0.215 + 0.76 share,
// 64-tiles 0.46 + 0.52 share).
double edge_scale(int64_t rem_m, int64_t rem_n, int tile) {
  const int fm = tile / 32, fn = tile / 64;
  const int rb = (int)((std::min<int64_t>(rem_m, tile) + 15) / 16);
  const int cb = (int)((std::min<int64_t>(rem_n, tile) + 15) / 16);
  const int nbi = std::max(0, std::min(fm, (rb + 1) >> 1)) + std::max(0, std::min(fm, rb >> 1));
  int busiest = 0;
  for (int simd = 0; simd < 4; ++simd)
    busiest = std::max(busiest, std::max(0, std::min(fn, (cb - simd + 3) >> 2)) * nbi);
  if (busiest >= 2 * fm * fn) return 1.0;
  const double share = (double)busiest / (2 * fm * fn);
  return tile == 128 ? 0.215 + 0.76 * share : 0.46 + 0.52 * share;
}

// Sum of the k-loop scales of a tm x tn grid of tiles over m x n.
double grid_scale(int64_t m, int64_t n, int tile, int64_t tm, int64_t tn) {
  const int64_t em = m - (tm - 1) * tile, en = n - (tn - 1) * tile;  // edge extents
  return (double)((tm - 1) * (tn - 1)) + (double)(tn - 1) * edge_scale(em, tile, tile) +
         (double)(tm - 1) * edge_scale(tile, en, tile) + edge_scale(em, en, tile);
}

// The launch shape of one product problem: a function of (dtype, m, n, k)
// and the tuning table only, so a problem gives the same bits in any group.
Shape plan_shape(const Tuning &tu, int dtype, int64_t m, int64_t n, int64_t k) {
  const int64_t kt = (k + k_tile(dtype) - 1) / k_tile(dtype);
  Shape best{};
  double best_total = 0;
  for (int ti = 1; ti >= 0; --ti) {  // 128 first: ties keep the bigger tile
    const int tile = ti ? 128 : 64;
    const int64_t tm = (m + tile - 1) / tile, tn = (n + tile - 1) / tile;
    const double unit = tu.c0[dtype][ti] + tu.c1[dtype][ti] * (double)kt;
    const double total = (double)(tm * tn) * tu.c0[dtype][ti] +
                         tu.c1[dtype][ti] * (double)kt * grid_scale(m, n, tile, tm, tn);
    if (ti == 1 || total < best_total) {
      best = Shape{tile, 1, (int)tm, (int)(tm * tn), unit};
      best_total = total;
    }
  }
  const int ti = best.tile == 128 ? 1 : 0;
  if (best.unit_us > tu.unit_cap_us[dtype]) {
    int64_t s = (int64_t)std::ceil(best.unit_us / tu.unit_cap_us[dtype]);
    s = std::min<int64_t>(s, kt / kMinSplitKtiles);
    s = std::min<int64_t>(s, kMaxSplits);
    if (s >= 2) {
      const int64_t per = (kt + s - 1) / s;
      best.splits = (int)s;
      best.unit_us = tu.c0[dtype][ti] + tu.c1[dtype][ti] * (double)per;
    }
  }
  return best;
}

template <typename U>
void order_longest_first(std::vector<U> &units, const std::vector<double> &cost) {
  std::vector<int> order(units.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  std::vector<U> sorted(units.size());
  for (size_t i = 0; i < order.size(); ++i) sorted[i] = units[order[i]];
  units.swap(sorted);
}

}  // namespace

Tuning tuning_snapshot() {
  std::lock_guard<std::mutex> g(g_mu);
  return g_tuning;
}

void tuning_set(int dtype, const double *c) {
  std::lock_guard<std::mutex> g(g_mu);
  g_tuning.c0[dtype][0] = c[0];
  g_tuning.c1[dtype][0] = c[1];
  g_tuning.c0[dtype][1] = c[2];
  g_tuning.c1[dtype][1] = c[3];
  g_tuning.unit_cap_us[dtype] = c[4];
}

void tuning_get(int dtype, double *c) {
  std::lock_guard<std::mutex> g(g_mu);
  c[0] = g_tuning.c0[dtype][0];
  c[1] = g_tuning.c1[dtype][0];
  c[2] = g_tuning.c0[dtype][1];
  c[3] = g_tuning.c1[dtype][1];
  c[4] = g_tuning.unit_cap_us[dtype];
}

template <typename T>
void GemmGroup<T>::add(const GemmProblem<T> &p, const Tuning &tu) {
  const int dtype = sizeof(T) == 4 ? 0 : 1;
  GroupDesc<T> d{};
  d.a = p.a;
  d.b = p.b;
  d.c = p.c;
  d.lda = p.lda;
  d.ldb = p.ldb;
  d.ldc = p.ldc;
  d.m = (int)p.m;
  d.n = (int)p.n;
  d.k = (int)p.k;
  d.ta = p.ta;
  d.tb = p.tb;
  d.alpha = p.alpha;
  d.beta = p.beta;
  d.splits = 1;
  const int prob = (int)descs.size();
  if (p.k == 0 || p.alpha == T(0)) {
    d.mode = 1;
    d.tile = 128;
    d.tiles_m = (int)((p.m + 127) / 128);
    const int tiles = d.tiles_m * (int)((p.n + 127) / 128);
    for (int t = 0; t < tiles; ++t) {
      units.push_back(WorkUnit{prob, t, 0, 0});
      cost.push_back(tu.c0[dtype][1]);
      total_us += tu.c0[dtype][1];
    }
    ++info.scale_problems;
    descs.push_back(d);
    return;
  }
  const Shape sh = plan_shape(tu, dtype, p.m, p.n, p.k);
  d.mode = 0;
  d.tile = sh.tile;
  d.splits = sh.splits;
  d.tiles_m = sh.tiles_m;
  if (sh.splits > 1) {
    d.slot0 = ws_elems;
    d.cnt0 = (int)counters;
    ws_elems += (int64_t)sh.tiles * sh.splits * sh.tile * sh.tile;
    counters += sh.tiles;
    ++info.split_problems;
  }
  (sh.tile == 128 ? info.tiles128 : info.tiles64) += sh.tiles;
  const int ti = sh.tile == 128 ? 1 : 0;
  const double fixed = tu.c0[dtype][ti], kloop = sh.unit_us - fixed;
  for (int t = 0; t < sh.tiles; ++t) {
    // An edge tile's units are priced by the share of the k loop it runs.
    const int64_t rem_m = p.m - (int64_t)(t % sh.tiles_m) * sh.tile;
    const int64_t rem_n = p.n - (int64_t)(t / sh.tiles_m) * sh.tile;
    const double u = fixed + kloop * edge_scale(rem_m, rem_n, sh.tile);
    for (int s = 0; s < sh.splits; ++s) {
      units.push_back(WorkUnit{prob, t, s, 0});
      cost.push_back(u);
      total_us += u;
    }
  }
  info.flops += 2.0 * (double)p.m * (double)p.n * (double)p.k;
  descs.push_back(d);
}

template <typename T>
void GemmGroup<T>::finish() {
  order_longest_first(units, cost);
  info.units = units.size();
}

template <typename T>
void GemmGroup<T>::stage(char *h) const {
  std::memset(h, 0, staged_bytes());
  std::memcpy(h, descs.data(), descs.size() * sizeof(GroupDesc<T>));
  std::memcpy(h + off_units(), units.data(), units.size() * sizeof(WorkUnit));
}

template <typename T>
hipError_t GemmGroup<T>::launch(char *dev, int grid, hipStream_t st, char *ws) const {
  return launch_grouped_gemm<T>((const GroupDesc<T> *)dev, (const WorkUnit *)(dev + off_units()),
                                (int)units.size(), grid, (int *)(dev + off_sched()),
                                (T *)(ws ? ws : dev + off_ws()), (int *)(dev + off_sched() + 256),
                                st);
}

template <typename T>
void TrsmGroup<T>::add(bool left, bool trans, bool eff_upper, bool unit, int64_t m, int64_t n,
                       T alpha, const T *a, int64_t lda, T *b, int64_t ldb) {
  TrsmDesc<T> d{};
  d.a = a;
  d.b = b;
  d.lda = lda;
  d.ldb = ldb;
  d.q = (int)(left ? m : n);
  d.nrhs = (int)(left ? n : m);
  // Left form: a Right problem solves op(A)^T X^T = alpha B^T.
  d.trans = left ? trans : !trans;
  d.upper = left ? eff_upper : !eff_upper;
  d.unit = unit;
  d.right = !left;
  d.forward = !d.upper;
  d.width = trsm_unit_width<T>(d.q, d.lda);  // finish() may widen it to 32
  d.alpha = alpha;
  d.inv0 = inv_elems;
  const int prob = (int)descs.size();
  const int nblk = (d.q + 63) / 64;
  inv_elems += (int64_t)nblk * 64 * 64;
  for (int blk = 0; blk < nblk; ++blk) blocks.push_back(int2{prob, blk});
  info.flops += (double)d.q * (double)d.q * (double)d.nrhs;
  descs.push_back(d);
}

// Time model of one trsm unit, us of one CU (f64), fitted to the unit times
// of the batching workload with every CU busy: 4.8 us fixed, 4.8 us per
// 64-row diagonal block (its barrier-separated steps), and the update at
// about 0.19 TFLOP/s per CU.
static double trsm_unit_us(int q, int width) {
  return 4.8 + 4.8 * ((q + 63) / 64) + (double)q * (double)q * width / 1.87e5;
}

template <typename T>
void TrsmGroup<T>::finish() {
  // Widths: 32 for the largest eligible systems first, while the units still
  // fill kTrsmFillUnits CUs.
  int64_t total = 0;
  std::vector<int> wide;
  for (size_t p = 0; p < descs.size(); ++p) {
    const TrsmDesc<T> &d = descs[p];
    total += (d.nrhs + d.width - 1) / d.width;
    if (trsm_wide_ok<T>(d.q, d.lda, d.nrhs)) wide.push_back((int)p);
  }
  std::stable_sort(wide.begin(), wide.end(),
                   [&](int x, int y) { return descs[x].q > descs[y].q; });
  for (int p : wide) {
    TrsmDesc<T> &d = descs[p];
    const int64_t n16 = (d.nrhs + 15) / 16, n32 = (d.nrhs + 31) / 32;
    const int64_t after = total - n16 + n32;
    if (after < kTrsmFillUnits) continue;
    d.width = 32;
    total = after;
  }

  // Units, each problem's consecutively. A 32-wide problem whose last unit
  // holds at most 16 right-hand sides runs that unit 16 wide (its problem
  // index stored as ~p, see trsm_unit_problem): half the MFMA work for the
  // same result.
  units.clear();
  cost.clear();
  for (size_t p = 0; p < descs.size(); ++p) {
    const TrsmDesc<T> &d = descs[p];
    for (int c0 = 0; c0 < d.nrhs; c0 += d.width) {
      const bool narrow = d.width == 32 && d.nrhs - c0 <= 16;
      units.push_back(int2{narrow ? ~(int)p : (int)p, c0});
      cost.push_back(trsm_unit_us(d.q, narrow ? 16 : d.width));
    }
  }
  // Each problem's units are consecutive; a chunk is priced by its first
  // (widest) unit when ordered and by all of its units when dealt.
  struct Chunk {
    size_t u0, n;
    double unit_cost, total;
  };
  constexpr size_t kChunk = kCUs / kTrsmQueues;
  std::vector<Chunk> chunks;
  for (size_t u = 0; u < units.size();) {
    size_t e = u;
    while (e < units.size() && trsm_unit_problem(units[e].x) == trsm_unit_problem(units[u].x)) ++e;
    for (size_t c = u; c < e; c += kChunk) {
      const size_t n = std::min(kChunk, e - c);
      double total = 0;
      for (size_t i = c; i < c + n; ++i) total += cost[i];
      chunks.push_back(Chunk{c, n, cost[c], total});
    }
    u = e;
  }
  std::stable_sort(chunks.begin(), chunks.end(),
                   [](const Chunk &x, const Chunk &y) { return x.unit_cost > y.unit_cost; });
  std::vector<std::vector<int2>> q(kTrsmQueues);
  double load[kTrsmQueues] = {};
  for (const Chunk &c : chunks) {
    int best = 0;
    for (int x = 1; x < kTrsmQueues; ++x)
      if (load[x] < load[best]) best = x;
    load[best] += c.total;
    q[best].insert(q[best].end(), units.begin() + c.u0, units.begin() + c.u0 + c.n);
  }
  std::vector<int2> sorted;
  sorted.reserve(units.size());
  for (int x = 0; x < kTrsmQueues; ++x) {
    qstart[x] = (int)sorted.size();
    sorted.insert(sorted.end(), q[x].begin(), q[x].end());
  }
  qstart[kTrsmQueues] = (int)sorted.size();
  units.swap(sorted);
  info.units = units.size();
}

template <typename T>
void TrsmGroup<T>::stage(char *h) const {
  std::memset(h, 0, staged_bytes());
  std::memcpy(h, descs.data(), descs.size() * sizeof(TrsmDesc<T>));
  std::memcpy(h + off_units(), units.data(), units.size() * sizeof(int2));
  std::memcpy(h + off_blocks(), blocks.data(), blocks.size() * sizeof(int2));
  std::memcpy(h + off_sched() + kTrsmQueueStart * sizeof(int), qstart, sizeof(qstart));
}

template <typename T>
hipError_t TrsmGroup<T>::launch(char *dev, int grid, hipStream_t st, char *ws) const {
  return launch_grouped_trsm<T>((const TrsmDesc<T> *)dev, (const int2 *)(dev + off_blocks()),
                                (int)blocks.size(), (const int2 *)(dev + off_units()),
                                (int)units.size(), grid, (int *)(dev + off_sched()),
                                (double *)(ws ? ws : dev + off_ws()), st);
}

template struct GemmGroup<float>;
template struct GemmGroup<double>;
template struct TrsmGroup<float>;
template struct TrsmGroup<double>;

}  // namespace rblas
