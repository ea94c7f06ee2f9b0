// C-ABI of grouped plans: validation of every problem exactly as rblas's
// gemm (level3.rs:32-62, 143-222) or trsm (level3.rs:390-520), the planner
// (grouped_plan.cpp), and one persistent launch per run, plain or replayed
// from a captured hipGraph.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gemm_host.hpp"
#include "grouped_plan.hpp"
#include "runtime.hpp"
#include "trsm_host.hpp"

struct rblas_grouped {
  int kind = 0;   // 0 gemm, 1 trsm
  int dtype = 0;  // 0 f32, 1 f64
  void *mem = nullptr;
  int grid = 0;
  rblas::GemmGroup<float> gs;
  rblas::GemmGroup<double> gd;
  rblas::TrsmGroup<float> ts;
  rblas::TrsmGroup<double> td;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  rblas_grouped_info_t info{};
};

namespace rblas {
namespace {

Check problem_error(size_t i, Check c) {
  c.msg = "problem " + std::to_string(i) + ": " + c.msg;
  return c;
}

// Upload a finished group's tables (sched and counters staged as zeros).
template <typename G>
int32_t upload(rblas_grouped_t *g, const G &grp, int grid) {
  std::vector<char> staged(grp.staged_bytes());
  grp.stage(staged.data());
  hipError_t e = hipMalloc(&g->mem, std::max<size_t>(grp.bytes(), 256));
  if (e != hipSuccess) return hip_fail(e, "grouped plan memory");
  e = hipMemcpy(g->mem, staged.data(), staged.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "grouped plan upload");
  g->grid = std::min(grid, std::max<int>(1, (int)grp.units.size()));
  g->info = grp.info;
  g->info.grid = g->grid;
  g->info.workspace_bytes = grp.bytes();
  return RBLAS_OK;
}

template <typename T>
GemmGroup<T> &gemm_of(rblas_grouped_t *g);
template <>
GemmGroup<float> &gemm_of<float>(rblas_grouped_t *g) { return g->gs; }
template <>
GemmGroup<double> &gemm_of<double>(rblas_grouped_t *g) { return g->gd; }
template <typename T>
TrsmGroup<T> &trsm_of(rblas_grouped_t *g);
template <>
TrsmGroup<float> &trsm_of<float>(rblas_grouped_t *g) { return g->ts; }
template <>
TrsmGroup<double> &trsm_of<double>(rblas_grouped_t *g) { return g->td; }

template <typename T, typename P>
int32_t grouped_gemm_create(rblas_grouped_t **out, const P *probs, size_t count,
                            size_t *bad_index) {
  if (out) *out = nullptr;
  if (bad_index) *bad_index = 0;
  if (!out || (count && !probs)) return fail(bad_arg("grouped gemm: null plan or problems"), nullptr);
  if (count > (size_t)INT_MAX) return fail(bad_arg("grouped gemm: too many problems"), nullptr);
  const Tuning tu = tuning_snapshot();
  GemmGroup<T> grp;
  for (size_t i = 0; i < count; ++i) {
    const P &q = probs[i];
    ColMajorCall<T> cc;
    Check chk = normalize<T>(q.layout, q.transa, q.transb, q.m, q.n, q.k, q.alpha, q.a, q.a_len,
                             q.lda, q.b, q.b_len, q.ldb, q.beta, q.c, q.c_len, q.ldc, cc);
    GemmProblem<T> p{cc.at, cc.bt, (int64_t)cc.m, (int64_t)cc.n, (int64_t)cc.k, cc.alpha,
                     cc.beta, cc.a, (int64_t)cc.lda, cc.b, (int64_t)cc.ldb, cc.c, (int64_t)cc.ldc};
    if (chk.ok() && cc.m && cc.n && cc.k && cc.alpha != T(0) && !gemm_fits(p))
      chk = bad_arg("leading dimension too large for the device");
    if (!chk.ok()) {
      if (bad_index) *bad_index = i;
      return fail(problem_error(i, chk), nullptr);
    }
    if (cc.m == 0 || cc.n == 0) continue;
    if ((cc.k == 0 || cc.alpha == T(0)) && cc.beta == T(1)) continue;  // C untouched
    grp.add(p, tu);
  }
  grp.finish();
  grp.info.problems = count;
  if (grp.units.size() > (size_t)INT_MAX) return fail(bad_arg("grouped gemm: too many units"), nullptr);
  rblas_grouped_t *g = new rblas_grouped_t();
  g->kind = 0;
  g->dtype = sizeof(T) == 4 ? 0 : 1;
  gemm_of<T>(g) = std::move(grp);
  const GemmGroup<T> &kept = gemm_of<T>(g);
  const int32_t s = upload(g, kept, grouped_gemm_grid<T>());
  if (s != RBLAS_OK) {
    // The upload's failure is the one reported; the free is best effort.
    if (g->mem) static_cast<void>(hipFree(g->mem));
    delete g;
    return s;
  }
  g->info.est_us = kept.total_us / std::max(1, g->grid);
  *out = g;
  return RBLAS_OK;
}

template <typename T, typename P>
int32_t grouped_trsm_create(rblas_grouped_t **out, const P *probs, size_t count,
                            size_t *bad_index) {
  if (out) *out = nullptr;
  if (bad_index) *bad_index = 0;
  if (!out || (count && !probs)) return fail(bad_arg("grouped trsm: null plan or problems"), nullptr);
  if (count > (size_t)INT_MAX) return fail(bad_arg("grouped trsm: too many problems"), nullptr);
  TrsmGroup<T> grp;
  for (size_t i = 0; i < count; ++i) {
    const P &q = probs[i];
    Check chk = trsm_validate<T>(q.side, q.uplo, q.trans, q.diag, q.m, q.n, q.a_len, q.lda,
                                 q.b_len, q.ldb);
    if (chk.ok() && q.m != 0 && q.n != 0 && (q.side == RBLAS_LEFT ? q.m : q.n) > 2048)
      chk = bad_arg("grouped trsm: the order of A must be at most 2048 (use rblas_?trsm_dev)");
    if (!chk.ok()) {
      if (bad_index) *bad_index = i;
      return fail(problem_error(i, chk), nullptr);
    }
    if (q.m == 0 || q.n == 0) continue;
    const bool t = is_transposed(q.trans);
    grp.add(q.side == RBLAS_LEFT, t, (q.uplo == RBLAS_UPPER) != t, q.diag == RBLAS_UNIT,
            (int64_t)q.m, (int64_t)q.n, q.alpha, q.a, (int64_t)q.lda, q.b, (int64_t)q.ldb);
  }
  grp.finish();
  grp.info.problems = count;
  rblas_grouped_t *g = new rblas_grouped_t();
  g->kind = 1;
  g->dtype = sizeof(T) == 4 ? 0 : 1;
  trsm_of<T>(g) = std::move(grp);
  const int32_t s = upload(g, trsm_of<T>(g), grouped_trsm_grid<T>());
  if (s != RBLAS_OK) {
    // The upload's failure is the one reported; the free is best effort.
    if (g->mem) static_cast<void>(hipFree(g->mem));
    delete g;
    return s;
  }
  *out = g;
  return RBLAS_OK;
}

int32_t grouped_launch(rblas_grouped_t *g, hipStream_t st) {
  char *d = (char *)g->mem;
  hipError_t e;
  if (g->kind == 0)
    e = g->dtype == 0 ? g->gs.launch(d, g->grid, st) : g->gd.launch(d, g->grid, st);
  else
    e = g->dtype == 0 ? g->ts.launch(d, g->grid, st) : g->td.launch(d, g->grid, st);
  RBLAS_HIP_TRY(e, "grouped launch");
  return RBLAS_OK;
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
  if (!plan) return fail(bad_arg("grouped: null plan"), nullptr);
  return grouped_launch(plan, (hipStream_t)stream);
}

int32_t rblas_grouped_run_graph(rblas_grouped_t *plan, void *stream) {
  if (!plan) return fail(bad_arg("grouped: null plan"), nullptr);
  hipStream_t st = (hipStream_t)stream;
  if (st == nullptr) return fail(bad_arg("grouped: graph replay needs a non-null stream"), nullptr);
  if (!plan->exec) {
    RBLAS_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "graph capture");
    const int32_t s = grouped_launch(plan, st);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (s != RBLAS_OK) {
      // The launch's failure is the one reported; the destroy is best effort.
      if (graph) static_cast<void>(hipGraphDestroy(graph));
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
  if (!plan || !info) return fail(bad_arg("grouped: null plan"), nullptr);
  *info = plan->info;
  return RBLAS_OK;
}

int32_t rblas_grouped_destroy(rblas_grouped_t *plan) {
  if (!plan) return RBLAS_OK;
  // Everything is released; the first failure is reported.
  hipError_t e = plan->exec ? hipGraphExecDestroy(plan->exec) : hipSuccess;
  const hipError_t eg = plan->graph ? hipGraphDestroy(plan->graph) : hipSuccess;
  const hipError_t em = plan->mem ? hipFree(plan->mem) : hipSuccess;
  if (e == hipSuccess) e = eg;
  if (e == hipSuccess) e = em;
  delete plan;
  RBLAS_HIP_TRY(e, "grouped destroy");
  return RBLAS_OK;
}

int32_t rblas_grouped_set_tuning(int32_t dtype, const double *coeffs) {
  if ((dtype != 0 && dtype != 1) || !coeffs)
    return fail(bad_arg("grouped tuning: dtype must be 0 (f32) or 1 (f64)"), nullptr);
  for (int i = 0; i < 5; ++i)
    if (!(coeffs[i] > 0) || !std::isfinite(coeffs[i]))
      return fail(bad_arg("grouped tuning: coefficients must be positive and finite"), nullptr);
  tuning_set(dtype, coeffs);
  return RBLAS_OK;
}

int32_t rblas_grouped_get_tuning(int32_t dtype, double *coeffs) {
  if ((dtype != 0 && dtype != 1) || !coeffs)
    return fail(bad_arg("grouped tuning: dtype must be 0 (f32) or 1 (f64)"), nullptr);
  tuning_get(dtype, coeffs);
  return RBLAS_OK;
}

}  // extern "C"
