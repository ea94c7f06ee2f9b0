// Grouped GEMM plans: validation of every problem exactly as rblas's gemm
// (level3.rs:32-62, 143-222), the shape-driven planner (tile side and k
// splits per problem from a measured cost model), and the C-ABI around one
// persistent launch, plain or replayed from a captured hipGraph.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "gemm_host.hpp"
#include "grouped.hpp"
#include "runtime.hpp"
#include "trsm_host.hpp"

struct rblas_grouped {
  int kind = 0;   // 0 gemm, 1 trsm
  int dtype = 0;  // 0 f32, 1 f64
  int device = 0;
  void *mem = nullptr;  // descriptors | units | sched | counters (| blocks) | workspace
  void *descs = nullptr, *units = nullptr, *ws = nullptr;
  int *sched = nullptr, *counters = nullptr;
  int nunits = 0, grid = 0;
  void *blocks = nullptr;  // trsm: diagonal blocks to invert
  int nblocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  rblas_grouped_info_t info{};
};

namespace rblas {
namespace {

// Per-unit time model: c0 + c1 * (k-tiles of the unit), in microseconds of
// one CU, per (dtype, tile side), plus the longest unit the planner leaves
// unsplit. The defaults are the autotuner's measurement on MI355X
// (rblas_amd/autotune.py, saved in rblas_amd/tuning_gfx950.json); the unit
// cap is set just above where splitting stopped paying on the batching
// workload, so only very long-k units are split. The table can be replaced
// at run time (rblas_grouped_set_tuning).
struct Tuning {
  double c0[2][2];  // [dtype f32, f64][tile 64, 128]
  double c1[2][2];
  double unit_cap_us[2];
};
Tuning g_tuning = {{{6.582, 16.869}, {6.517, 17.415}},
                   {{0.791, 2.306}, {0.737, 2.136}},
                   {320.0, 320.0}};
std::mutex g_tuning_mu;

constexpr int kMinSplitKtiles = 16;
constexpr int kMaxSplits = 16;

int k_tile(int dtype) { return dtype == 0 ? 32 : 16; }

// The launch shape the planner gives one product problem: a function of its
// shape (and the tuning table) only.
struct Shape {
  int tile, splits, tiles_m, tiles;
  double unit_us;
};

Shape plan_shape(const Tuning &tu, int dtype, int64_t m, int64_t n, int64_t k) {
  const int64_t kt = (k + k_tile(dtype) - 1) / k_tile(dtype);
  Shape best{};
  double best_total = 0;
  for (int ti = 1; ti >= 0; --ti) {  // 128 first: ties keep the bigger tile
    const int tile = ti ? 128 : 64;
    const int64_t tm = (m + tile - 1) / tile, tn = (n + tile - 1) / tile;
    const double unit = tu.c0[dtype][ti] + tu.c1[dtype][ti] * (double)kt;
    const double total = (double)(tm * tn) * unit;
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

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename T, typename P>
int32_t grouped_gemm_create(rblas_grouped_t **out, const P *probs, size_t count,
                            size_t *bad_index) {
  if (out) *out = nullptr;
  if (bad_index) *bad_index = 0;
  if (!out || (count && !probs)) {
    Check c;
    c.status = RBLAS_BAD_ARG;
    c.err = {RBLAS_BAD_ARG, 0, 0, 0};
    c.msg = "grouped gemm: null plan or problem pointer";
    return fail(c, nullptr);
  }
  if (count > (size_t)INT_MAX) {
    Check c;
    c.status = RBLAS_BAD_ARG;
    c.err = {RBLAS_BAD_ARG, 0, 0, 0};
    c.msg = "grouped gemm: too many problems";
    return fail(c, nullptr);
  }
  const int dtype = sizeof(T) == 4 ? 0 : 1;
  Tuning tu;
  {
    std::lock_guard<std::mutex> g(g_tuning_mu);
    tu = g_tuning;
  }
  std::vector<GroupDesc<T>> descs;
  std::vector<WorkUnit> units;
  std::vector<double> cost;
  descs.reserve(count);
  int64_t ws_elems = 0, counters = 0;
  rblas_grouped_info_t info{};
  info.problems = count;
  for (size_t i = 0; i < count; ++i) {
    const P &q = probs[i];
    ColMajorCall<T> cc;
    Check chk = normalize<T>(q.layout, q.transa, q.transb, q.m, q.n, q.k, q.alpha, q.a, q.a_len,
                             q.lda, q.b, q.b_len, q.ldb, q.beta, q.c, q.c_len, q.ldc, cc);
    if (!chk.ok()) {
      if (bad_index) *bad_index = i;
      chk.msg = "problem " + std::to_string(i) + ": " + chk.msg;
      return fail(chk, nullptr);
    }
    if (cc.m == 0 || cc.n == 0) continue;
    GroupDesc<T> d{};
    d.a = cc.a;
    d.b = cc.b;
    d.c = cc.c;
    d.lda = (int64_t)cc.lda;
    d.ldb = (int64_t)cc.ldb;
    d.ldc = (int64_t)cc.ldc;
    d.m = (int)cc.m;
    d.n = (int)cc.n;
    d.k = (int)cc.k;
    d.ta = cc.at;
    d.tb = cc.bt;
    d.alpha = cc.alpha;
    d.beta = cc.beta;
    d.splits = 1;
    const int prob = (int)descs.size();
    if (cc.k == 0 || cc.alpha == T(0)) {
      if (cc.beta == T(1)) continue;  // rblas leaves C untouched
      d.mode = 1;
      d.tile = 128;
      d.tiles_m = (int)((cc.m + 127) / 128);
      const int tiles = d.tiles_m * (int)((cc.n + 127) / 128);
      for (int t = 0; t < tiles; ++t) {
        units.push_back(WorkUnit{prob, t, 0, 0});
        cost.push_back(tu.c0[dtype][1]);
      }
      ++info.scale_problems;
      descs.push_back(d);
      continue;
    }
    GemmProblem<T> gp{cc.at, cc.bt, (int64_t)cc.m, (int64_t)cc.n, (int64_t)cc.k, cc.alpha, cc.beta,
                      cc.a, (int64_t)cc.lda, cc.b, (int64_t)cc.ldb, cc.c, (int64_t)cc.ldc};
    if (!gemm_fits(gp)) {
      if (bad_index) *bad_index = i;
      Check c;
      c.status = RBLAS_BAD_ARG;
      c.err = {RBLAS_BAD_ARG, 0, 0, 0};
      c.msg = "problem " + std::to_string(i) + ": leading dimension too large for the device";
      return fail(c, nullptr);
    }
    const Shape sh = plan_shape(tu, dtype, (int64_t)cc.m, (int64_t)cc.n, (int64_t)cc.k);
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
    for (int t = 0; t < sh.tiles; ++t)
      for (int s = 0; s < sh.splits; ++s) {
        units.push_back(WorkUnit{prob, t, s, 0});
        cost.push_back(sh.unit_us);
      }
    info.flops += 2.0 * (double)cc.m * (double)cc.n * (double)cc.k;
    descs.push_back(d);
  }
  // Longest units first (a stable order: ties keep problem / tile order).
  std::vector<int> order(units.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  std::vector<WorkUnit> sorted(units.size());
  double total_us = 0;
  for (size_t i = 0; i < order.size(); ++i) {
    sorted[i] = units[order[i]];
    total_us += cost[order[i]];
  }
  if (sorted.size() > (size_t)INT_MAX) {
    Check c;
    c.status = RBLAS_BAD_ARG;
    c.err = {RBLAS_BAD_ARG, 0, 0, 0};
    c.msg = "grouped gemm: too many work units";
    return fail(c, nullptr);
  }

  int device = 0;
  RBLAS_HIP_TRY(hipGetDevice(&device), "grouped gemm device");
  rblas_grouped_t *g = new rblas_grouped_t();
  g->dtype = dtype;
  g->device = device;
  g->nunits = (int)sorted.size();
  g->grid = grouped_gemm_grid<T>();
  const size_t b_desc = align256(descs.size() * sizeof(GroupDesc<T>));
  const size_t b_units = align256(sorted.size() * sizeof(WorkUnit));
  const size_t b_sched = 256;
  const size_t b_cnt = align256((size_t)counters * sizeof(int));
  const size_t b_ws = (size_t)ws_elems * sizeof(T);
  const size_t total = b_desc + b_units + b_sched + b_cnt + b_ws;
  hipError_t e = hipMalloc(&g->mem, total);
  if (e != hipSuccess) {
    delete g;
    return hip_fail(e, "grouped gemm plan memory");
  }
  char *base = (char *)g->mem;
  g->descs = base;
  g->units = base + b_desc;
  g->sched = (int *)(base + b_desc + b_units);
  g->counters = (int *)(base + b_desc + b_units + b_sched);
  g->ws = base + b_desc + b_units + b_sched + b_cnt;
  e = hipMemcpy(g->descs, descs.data(), descs.size() * sizeof(GroupDesc<T>),
                hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(g->units, sorted.data(), sorted.size() * sizeof(WorkUnit),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(g->sched, 0, b_sched + b_cnt);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    hipFree(g->mem);
    delete g;
    return hip_fail(e, "grouped gemm plan upload");
  }
  info.units = sorted.size();
  info.grid = std::min(g->grid, g->nunits);
  info.est_us = total_us / std::max(1, info.grid);
  info.workspace_bytes = total;
  g->info = info;
  *out = g;
  return RBLAS_OK;
}

template <typename T, typename P>
int32_t grouped_trsm_create(rblas_grouped_t **out, const P *probs, size_t count,
                            size_t *bad_index) {
  if (out) *out = nullptr;
  if (bad_index) *bad_index = 0;
  if (!out || (count && !probs)) return fail(bad_arg("grouped trsm: null plan or problem pointer"), nullptr);
  if (count > (size_t)INT_MAX) return fail(bad_arg("grouped trsm: too many problems"), nullptr);
  std::vector<TrsmDesc<T>> descs;
  std::vector<int2> blocks, units;
  std::vector<double> cost;
  int64_t inv_elems = 0;
  rblas_grouped_info_t info{};
  info.problems = count;
  for (size_t i = 0; i < count; ++i) {
    const P &q = probs[i];
    Check chk = trsm_validate<T>(q.side, q.uplo, q.trans, q.diag, q.m, q.n, q.a_len, q.lda,
                                 q.b_len, q.ldb);
    if (chk.ok() && q.m != 0 && q.n != 0 && (q.side == RBLAS_LEFT ? q.m : q.n) > 2048)
      chk = bad_arg("grouped trsm: the order of A must be at most 2048 (use rblas_?trsm_dev)");
    if (!chk.ok()) {
      if (bad_index) *bad_index = i;
      chk.msg = "problem " + std::to_string(i) + ": " + chk.msg;
      return fail(chk, nullptr);
    }
    if (q.m == 0 || q.n == 0) continue;
    const bool left = q.side == RBLAS_LEFT;
    const bool t = is_transposed(q.trans);
    const bool eff_upper = (q.uplo == RBLAS_UPPER) != t;
    TrsmDesc<T> d{};
    d.a = q.a;
    d.b = q.b;
    d.lda = (int64_t)q.lda;
    d.ldb = (int64_t)q.ldb;
    d.q = (int)(left ? q.m : q.n);
    d.nrhs = (int)(left ? q.n : q.m);
    // Left form: a Right problem solves op(A)^T X^T = alpha B^T.
    d.trans = left ? t : !t;
    d.upper = left ? eff_upper : !eff_upper;
    d.unit = q.diag == RBLAS_UNIT;
    d.right = !left;
    d.forward = !d.upper;
    d.width = d.q <= kTrsmLdsDoubles / 16 ? 16 : 8;
    d.alpha = q.alpha;
    d.inv0 = inv_elems;
    const int prob = (int)descs.size();
    const int nblk = (d.q + 63) / 64;
    inv_elems += (int64_t)nblk * 64 * 64;
    for (int b = 0; b < nblk; ++b) blocks.push_back(int2{prob, b});
    for (int c0 = 0; c0 < d.nrhs; c0 += d.width) {
      units.push_back(int2{prob, c0});
      cost.push_back((double)d.q * (double)d.q * d.width);
    }
    info.flops += (double)d.q * (double)d.q * (double)d.nrhs;
    descs.push_back(d);
  }
  std::vector<int> order(units.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  std::vector<int2> sorted(units.size());
  for (size_t i = 0; i < order.size(); ++i) sorted[i] = units[order[i]];

  int device = 0;
  RBLAS_HIP_TRY(hipGetDevice(&device), "grouped trsm device");
  rblas_grouped_t *g = new rblas_grouped_t();
  g->kind = 1;
  g->dtype = sizeof(T) == 4 ? 0 : 1;
  g->device = device;
  g->nunits = (int)sorted.size();
  g->nblocks = (int)blocks.size();
  g->grid = grouped_trsm_grid<T>();
  const size_t b_desc = align256(descs.size() * sizeof(TrsmDesc<T>));
  const size_t b_units = align256(sorted.size() * sizeof(int2));
  const size_t b_sched = 256;
  const size_t b_blocks = align256(blocks.size() * sizeof(int2));
  const size_t b_inv = (size_t)inv_elems * sizeof(double);
  const size_t total = b_desc + b_units + b_sched + b_blocks + b_inv;
  hipError_t e = hipMalloc(&g->mem, total ? total : 256);
  if (e != hipSuccess) {
    delete g;
    return hip_fail(e, "grouped trsm plan memory");
  }
  char *base = (char *)g->mem;
  g->descs = base;
  g->units = base + b_desc;
  g->sched = (int *)(base + b_desc + b_units);
  g->blocks = base + b_desc + b_units + b_sched;
  g->ws = base + b_desc + b_units + b_sched + b_blocks;
  e = hipMemcpy(g->descs, descs.data(), descs.size() * sizeof(TrsmDesc<T>), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(g->units, sorted.data(), sorted.size() * sizeof(int2), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(g->blocks, blocks.data(), blocks.size() * sizeof(int2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(g->sched, 0, b_sched);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    hipFree(g->mem);
    delete g;
    return hip_fail(e, "grouped trsm plan upload");
  }
  info.units = sorted.size();
  info.grid = std::min(g->grid, std::max(1, g->nunits));
  info.workspace_bytes = total;
  g->info = info;
  *out = g;
  return RBLAS_OK;
}

int32_t grouped_launch(rblas_grouped_t *g, hipStream_t st) {
  if (g->kind == 1) {
    if (g->dtype == 0)
      RBLAS_HIP_TRY(launch_grouped_trsm<float>((const TrsmDesc<float> *)g->descs,
                                               (const int2 *)g->blocks, g->nblocks,
                                               (const int2 *)g->units, g->nunits, g->grid,
                                               g->sched, (double *)g->ws, st),
                    "grouped trsm launch");
    else
      RBLAS_HIP_TRY(launch_grouped_trsm<double>((const TrsmDesc<double> *)g->descs,
                                                (const int2 *)g->blocks, g->nblocks,
                                                (const int2 *)g->units, g->nunits, g->grid,
                                                g->sched, (double *)g->ws, st),
                    "grouped trsm launch");
    return RBLAS_OK;
  }
  if (g->dtype == 0)
    RBLAS_HIP_TRY(launch_grouped_gemm<float>((const GroupDesc<float> *)g->descs,
                                             (const WorkUnit *)g->units, g->nunits, g->grid,
                                             g->sched, (float *)g->ws, g->counters, st),
                  "grouped gemm launch");
  else
    RBLAS_HIP_TRY(launch_grouped_gemm<double>((const GroupDesc<double> *)g->descs,
                                              (const WorkUnit *)g->units, g->nunits, g->grid,
                                              g->sched, (double *)g->ws, g->counters, st),
                  "grouped gemm launch");
  return RBLAS_OK;
}

Check null_plan() {
  Check c;
  c.status = RBLAS_BAD_ARG;
  c.err = {RBLAS_BAD_ARG, 0, 0, 0};
  c.msg = "grouped: null plan";
  return c;
}

}  // namespace
}  // namespace rblas

using namespace rblas;

extern "C" {

int32_t rblas_sgemm_grouped_create(rblas_grouped_t **plan, const rblas_sgemm_problem_t *problems,
                                   size_t count, size_t *bad_index) {
  return grouped_gemm_create<float>(plan, problems, count, bad_index);
}

int32_t rblas_dgemm_grouped_create(rblas_grouped_t **plan, const rblas_dgemm_problem_t *problems,
                                   size_t count, size_t *bad_index) {
  return grouped_gemm_create<double>(plan, problems, count, bad_index);
}

int32_t rblas_strsm_grouped_create(rblas_grouped_t **plan, const rblas_strsm_problem_t *problems,
                                   size_t count, size_t *bad_index) {
  return grouped_trsm_create<float>(plan, problems, count, bad_index);
}

int32_t rblas_dtrsm_grouped_create(rblas_grouped_t **plan, const rblas_dtrsm_problem_t *problems,
                                   size_t count, size_t *bad_index) {
  return grouped_trsm_create<double>(plan, problems, count, bad_index);
}

int32_t rblas_grouped_run(rblas_grouped_t *plan, void *stream) {
  if (!plan) return fail(null_plan(), nullptr);
  return grouped_launch(plan, (hipStream_t)stream);
}

int32_t rblas_grouped_run_graph(rblas_grouped_t *plan, void *stream) {
  if (!plan) return fail(null_plan(), nullptr);
  hipStream_t st = (hipStream_t)stream;
  if (st == nullptr) {
    Check c;
    c.status = RBLAS_BAD_ARG;
    c.err = {RBLAS_BAD_ARG, 0, 0, 0};
    c.msg = "grouped: graph replay needs a non-null stream";
    return fail(c, nullptr);
  }
  if (!plan->exec) {
    RBLAS_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "graph capture");
    const int32_t s = grouped_launch(plan, st);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (s != RBLAS_OK) {
      if (graph) hipGraphDestroy(graph);
      return s;
    }
    RBLAS_HIP_TRY(e, "graph capture end");
    plan->graph = graph;
    RBLAS_HIP_TRY(hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0), "graph instantiate");
  }
  RBLAS_HIP_TRY(hipGraphLaunch(plan->exec, st), "graph launch");
  return RBLAS_OK;
}

int32_t rblas_grouped_info(const rblas_grouped_t *plan, rblas_grouped_info_t *info) {
  if (!plan || !info) return fail(null_plan(), nullptr);
  *info = plan->info;
  return RBLAS_OK;
}

int32_t rblas_grouped_destroy(rblas_grouped_t *plan) {
  if (!plan) return RBLAS_OK;
  if (plan->exec) hipGraphExecDestroy(plan->exec);
  if (plan->graph) hipGraphDestroy(plan->graph);
  hipError_t e = plan->mem ? hipFree(plan->mem) : hipSuccess;
  delete plan;
  RBLAS_HIP_TRY(e, "grouped destroy");
  return RBLAS_OK;
}

int32_t rblas_grouped_set_tuning(int32_t dtype, const double *coeffs) {
  if ((dtype != 0 && dtype != 1) || !coeffs) {
    Check c;
    c.status = RBLAS_BAD_ARG;
    c.err = {RBLAS_BAD_ARG, 0, 0, 0};
    c.msg = "grouped tuning: dtype must be 0 (f32) or 1 (f64)";
    return fail(c, nullptr);
  }
  for (int i = 0; i < 5; ++i)
    if (!(coeffs[i] > 0) || !std::isfinite(coeffs[i])) {
      Check c;
      c.status = RBLAS_BAD_ARG;
      c.err = {RBLAS_BAD_ARG, 0, 0, 0};
      c.msg = "grouped tuning: coefficients must be positive and finite";
      return fail(c, nullptr);
    }
  std::lock_guard<std::mutex> g(g_tuning_mu);
  g_tuning.c0[dtype][0] = coeffs[0];
  g_tuning.c1[dtype][0] = coeffs[1];
  g_tuning.c0[dtype][1] = coeffs[2];
  g_tuning.c1[dtype][1] = coeffs[3];
  g_tuning.unit_cap_us[dtype] = coeffs[4];
  return RBLAS_OK;
}

int32_t rblas_grouped_get_tuning(int32_t dtype, double *coeffs) {
  if ((dtype != 0 && dtype != 1) || !coeffs) {
    Check c;
    c.status = RBLAS_BAD_ARG;
    c.err = {RBLAS_BAD_ARG, 0, 0, 0};
    c.msg = "grouped tuning: dtype must be 0 (f32) or 1 (f64)";
    return fail(c, nullptr);
  }
  std::lock_guard<std::mutex> g(g_tuning_mu);
  coeffs[0] = g_tuning.c0[dtype][0];
  coeffs[1] = g_tuning.c1[dtype][0];
  coeffs[2] = g_tuning.c0[dtype][1];
  coeffs[3] = g_tuning.c1[dtype][1];
  coeffs[4] = g_tuning.unit_cap_us[dtype];
  return RBLAS_OK;
}

}  // extern "C"
