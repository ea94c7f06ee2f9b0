// Serving loop: a stream of gemm / trsm / complex-gemm requests replayed
// against the wall clock, with admission by HBM footprint and continuous
// batching (no rblas counterpart: rblas is a library of calls).
//
// The read-only A and B of gemm and complex requests are slices of an operand
// pool generated in HBM before the first arrival (resident data, as a
// server's weights would be); each request's written operands are generated
// in HBM at launch (serve.hip) and freed when its batch completes; its
// result is hashed on the device. The scheduler (the caller's thread) loops:
//   1. requests whose arrival time has passed join a FIFO queue;
//   2. finished batches (their event has completed) retire: their requests
//      complete, their footprint is released;
//   3. while a stream slot is free and the queue is not empty, the queue's
//      head requests are admitted in order as long as the admitted footprint
//      stays under the cap (and the batch under max_batch), and launched as
//      one batch on that slot: one generation launch, one launch splitting
//      every complex request into its 4M real planes, one grouped gemm
//      launch per dtype (the complex requests' two real products join the
//      f64 one), one grouped trsm launch per dtype, one launch combining the
//      complex results, one hash launch, an event.
// Requests arriving while batches run are therefore admitted as slots and
// memory free up, and their batch overlaps the running ones on the device.
// The plain mode (the per-call comparator) serves one request at a time in
// FIFO order with the single-call kernels.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <thread>
#include <vector>

#include "grouped_plan.hpp"
#include "runtime.hpp"
#include "serve.hpp"
#include "trsm_host.hpp"

namespace rblas {
namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

struct Req {
  rblas_serve_request_t spec;
  double footprint = 0, flops = 0;
  size_t bytes[4] = {0, 0, 0, 0};  // A, B, C, workspace (device bytes of the request's own)
  size_t off[4] = {0, 0, 0, 0};
  const char *pooled[2] = {nullptr, nullptr};  // A, B as operand-pool slices (gemm, complex)
  int64_t pool_len[2] = {0, 0};                // their lengths in pool elements
  int esize = 8;                               // real element size of the operands
};

// The read-only operand pools (one per element type) that gemm and complex
// requests take A and B from: element i is draw i of RBLAS_SERVE_POOL_SEED.
struct Pools {
  char *base[2] = {nullptr, nullptr};  // f32, f64
  int64_t elems[2] = {0, 0};
};

// Element offset of operand j (1: A, 2: B) of length len in a pool of
// `elems` elements, for a request drawn from `seed`.
int64_t pool_offset(uint64_t seed, int j, int64_t len, int64_t elems) {
  return (int64_t)(mix64(seed ^ (uint64_t)j) % (uint64_t)(elems - len + 1));
}

// A stream slot: one batch in flight at a time. Its device arena holds the
// batch's operands, tables and workspaces; it is sized before the run starts
// and reused batch after batch, so the timed loop allocates nothing (growing
// a stream-ordered pool while other batches run stalls the host for as long
// as the device is busy).
struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  bool busy = false;
  std::vector<int> reqs;
  char *pinned = nullptr;
  size_t pinned_cap = 0;
  char *arena = nullptr;
  size_t arena_cap = 0;
  size_t hash_off = 0;  // offset of the hash results in the staged block
};

// Make `cap` at least `need` (geometric growth); the slot must be idle.
hipError_t ensure_arena(Slot &s, size_t need) {
  if (need <= s.arena_cap) return hipSuccess;
  if (s.arena) (void)hipFree(s.arena);
  s.arena = nullptr;
  const size_t cap = std::max(need + need / 4, 2 * s.arena_cap);
  s.arena_cap = 0;
  hipError_t e = hipMalloc((void **)&s.arena, cap);
  if (e == hipSuccess) s.arena_cap = cap;
  return e;
}

hipError_t ensure_pinned(Slot &s, size_t need) {
  if (need <= s.pinned_cap) return hipSuccess;
  if (s.pinned) (void)hipHostFree(s.pinned);
  s.pinned = nullptr;
  const size_t cap = std::max(need, std::max<size_t>(2 * s.pinned_cap, 1 << 20));
  s.pinned_cap = 0;
  hipError_t e = hipHostMalloc((void **)&s.pinned, cap, hipHostMallocDefault);
  if (e == hipSuccess) s.pinned_cap = cap;
  return e;
}

size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }

// Operand sizes, footprint and algorithmic flops of one request.
Req describe(const rblas_serve_request_t &s) {
  Req r;
  r.spec = s;
  const double m = s.m, n = s.n, k = s.k;
  if (s.kind == RBLAS_SERVE_GEMM) {
    r.esize = s.dtype == 0 ? 4 : 8;
    r.pool_len[0] = (int64_t)s.m * s.k;
    r.pool_len[1] = (int64_t)s.k * s.n;
    r.bytes[2] = (size_t)s.m * s.n * r.esize;
    r.flops = 2 * m * n * k;
  } else if (s.kind == RBLAS_SERVE_TRSM) {
    r.esize = s.dtype == 0 ? 4 : 8;
    r.bytes[0] = (size_t)s.m * s.m * r.esize;
    r.bytes[2] = (size_t)s.m * s.n * r.esize;  // B, solved in place
    r.flops = m * m * n;
  } else {
    r.esize = 8;  // complex f64 as interleaved doubles
    r.pool_len[0] = 2 * (int64_t)s.m * s.k;
    r.pool_len[1] = 2 * (int64_t)s.k * s.n;
    r.bytes[2] = (size_t)s.m * s.n * 16;
    r.bytes[3] = cgemm_workspace_elems(s.m, s.n, s.k) * sizeof(double);
    r.flops = 8 * m * n * k;
  }
  r.footprint = (double)(r.bytes[0] + r.bytes[1] + r.bytes[2] + r.bytes[3]);
  return r;
}

// Seeds: the written operands of a request with seed s draw from offsets
// in the order A, B, C (workloads.py's draw order), each from offset =
// elements before it (pooled A and B counted).
struct Planner {
  const rblas_serve_config_t &cfg;
  std::vector<Req> &reqs;
  bool plain;

  // Device work of one batch on `slot`. Returns a HIP error or success.
  hipError_t launch(Slot &slot, const std::vector<int> &ids) {
    hipStream_t st = slot.st;
    // Operand memory of the batch.
    size_t total = 0;
    for (int id : ids) {
      Req &r = reqs[id];
      for (int j = 0; j < 4; ++j) {
        r.off[j] = total;
        total += round256(r.bytes[j]);
      }
    }
    // Per-call workspaces (plain mode) come after the operands.
    std::vector<size_t> call_ws(ids.size(), 0);
    if (plain) {
      for (size_t i = 0; i < ids.size(); ++i) {
        const Req &r = reqs[ids[i]];
        if (r.spec.kind == RBLAS_SERVE_GEMM) {
          const GemmConfig c = r.spec.dtype == 0 ? gemm_config(gemm_shape<float>(r))
                                                 : gemm_config(gemm_shape<double>(r));
          const size_t b = r.spec.dtype == 0 ? gemm_workspace_bytes(gemm_shape<float>(r), c)
                                             : gemm_workspace_bytes(gemm_shape<double>(r), c);
          call_ws[i] = total;
          total += round256(b);
        } else if (r.spec.kind == RBLAS_SERVE_TRSM) {
          call_ws[i] = total;
          total += round256(trsm_workspace_bytes(r.spec.m, r.spec.n, true));
        }
      }
    }
    const size_t ops_bytes = round256(total);
    hipError_t e = hipSuccess;
    // Operand addresses are arena offsets until the arena is sized (its size
    // depends on the groups' workspaces); rebase() then adds its address.
    char *ops = nullptr;

    // Host tables: generation, groups (batching mode), hash.
    std::vector<GenDesc> gens;
    std::vector<int2> gen_units;
    std::vector<HashDesc> hashes;
    std::vector<int2> hash_units;
    std::vector<CplxDesc> cplx;  // batched complex requests
    std::vector<int2> cplx_split_units, cplx_comb_units;
    const Tuning tu = tuning_snapshot();
    GemmGroup<float> g32;
    GemmGroup<double> g64;
    TrsmGroup<float> t32;
    TrsmGroup<double> t64;
    auto add_gen = [&](void *p, int64_t count, int dtype, int mode, int64_t n, uint64_t seed,
                       int64_t offset) {
      const int d = (int)gens.size();
      gens.push_back(GenDesc{p, count, n, seed, offset, dtype, mode});
      for (int64_t c = 0; c * kGenChunk < count; ++c) gen_units.push_back(int2{d, (int)c});
    };
    const int col_step = std::max(1, cfg.hash_col_step);
    auto add_hash = [&](const void *p, int64_t rows, int64_t cols, int eb) {
      const int d = (int)hashes.size();
      hashes.push_back(HashDesc{p, rows, cols, rows, eb, col_step});
      for (int64_t c = 0; c * kGenChunk < hash_elems(hashes.back()); ++c)
        hash_units.push_back(int2{d, (int)c});
    };
    // Group descriptors whose A and B are pool slices (absolute addresses,
    // not rebased onto the arena).
    std::vector<char> g32_abs, g64_abs;
    for (int id : ids) {
      const Req &r = reqs[id];
      const rblas_serve_request_t &s = r.spec;
      char *a = ops + r.off[0], *b = ops + r.off[1], *c = ops + r.off[2];
      if (s.kind == RBLAS_SERVE_GEMM) {
        const int64_t la = (int64_t)s.m * s.k, lb = (int64_t)s.k * s.n, lc = (int64_t)s.m * s.n;
        add_gen(c, lc, s.dtype, 0, 0, s.seed, la + lb);
        add_hash(c, s.m, s.n, r.esize);
        if (!plain) {
          if (s.dtype == 0) {
            g32.add(GemmProblem<float>{false, false, s.m, s.n, s.k, 1.0f, 0.5f,
                                       (const float *)r.pooled[0], s.m,
                                       (const float *)r.pooled[1], s.k, (float *)c, s.m},
                    tu);
            g32_abs.push_back(1);
          } else {
            g64.add(GemmProblem<double>{false, false, s.m, s.n, s.k, 1.0, 0.5,
                                        (const double *)r.pooled[0], s.m,
                                        (const double *)r.pooled[1], s.k, (double *)c, s.m},
                    tu);
            g64_abs.push_back(1);
          }
        }
      } else if (s.kind == RBLAS_SERVE_TRSM) {
        const int64_t la = (int64_t)s.m * s.m, lb = (int64_t)s.m * s.n;
        add_gen(a, la, s.dtype, 1, s.m, s.seed, 0);
        add_gen(c, lb, s.dtype, 0, 0, s.seed, la);
        add_hash(c, s.m, s.n, r.esize);
        if (!plain) {
          if (s.dtype == 0)
            t32.add(true, false, false, false, s.m, s.n, 1.0f, (const float *)a, s.m, (float *)c,
                    s.m);
          else
            t64.add(true, false, false, false, s.m, s.n, 1.0, (const double *)a, s.m,
                    (double *)c, s.m);
        }
      } else {
        const int64_t la = 2 * (int64_t)s.m * s.k, lb = 2 * (int64_t)s.k * s.n;
        const int64_t lc = 2 * (int64_t)s.m * s.n;
        add_gen(c, lc, 1, 0, 0, s.seed, la + lb);
        add_hash(c, 2 * (int64_t)s.m, s.n, 8);
        if (!plain) {
          // 4M: the planes in one batched split launch, the two real
          // products m x n x 2k in the batch's grouped f64 launch, C in one
          // batched combine launch.
          const int64_t m = s.m, n = s.n, k = s.k;
          double *ws = (double *)(ops + r.off[3]);
          double *ap = ws, *b1 = ap + 2 * m * k, *b2 = b1 + 2 * k * n, *pr = b2 + 2 * k * n;
          double *pi = pr + m * n;
          const int d = (int)cplx.size();
          cplx.push_back(CplxDesc{(const double *)r.pooled[0], (const double *)r.pooled[1],
                                  (double *)c, ws, m, n, k, 1.0, 0.0, 0.5, 0.0});
          const int64_t chunks = (m * k + kGenChunk - 1) / kGenChunk +
                                 (k * n + kGenChunk - 1) / kGenChunk;
          for (int64_t ch = 0; ch < chunks; ++ch) cplx_split_units.push_back(int2{d, (int)ch});
          for (int64_t ch = 0; ch * kGenChunk < m * n; ++ch)
            cplx_comb_units.push_back(int2{d, (int)ch});
          g64.add(GemmProblem<double>{false, false, m, n, 2 * k, 1.0, 0.0, ap, m, b1, 2 * k, pr,
                                      m},
                  tu);
          g64.add(GemmProblem<double>{false, false, m, n, 2 * k, 1.0, 0.0, ap, m, b2, 2 * k, pi,
                                      m},
                  tu);
          g64_abs.push_back(0);
          g64_abs.push_back(0);
        }
      }
    }
    g32.finish();
    g64.finish();
    t32.finish();
    t64.finish();
    // Device block: gen descs | gen units | hash descs | hash units | hash
    // out | each group's staged tables | each group's workspace. Everything
    // before the first workspace is staged in pinned memory and uploaded with
    // one copy (sched / counters / hash out staged as zeros).
    size_t off = 0;
    auto take = [&](size_t bytes) {
      const size_t o = off;
      off += round256(bytes);
      return o;
    };
    const size_t o_gd = take(gens.size() * sizeof(GenDesc));
    const size_t o_gu = take(gen_units.size() * sizeof(int2));
    const size_t o_hd = take(hashes.size() * sizeof(HashDesc));
    const size_t o_hu = take(hash_units.size() * sizeof(int2));
    const size_t o_ho = take(hashes.size() * sizeof(uint64_t));
    const size_t o_cd = take(cplx.size() * sizeof(CplxDesc));
    const size_t o_cs = take(cplx_split_units.size() * sizeof(int2));
    const size_t o_cc = take(cplx_comb_units.size() * sizeof(int2));
    const size_t o_g32 = take(g32.units.empty() ? 0 : g32.staged_bytes());
    const size_t o_g64 = take(g64.units.empty() ? 0 : g64.staged_bytes());
    const size_t o_t32 = take(t32.units.empty() ? 0 : t32.staged_bytes());
    const size_t o_t64 = take(t64.units.empty() ? 0 : t64.staged_bytes());
    const size_t staged = off;
    const size_t w_g32 = take(g32.units.empty() ? 0 : g32.ws_bytes());
    const size_t w_g64 = take(g64.units.empty() ? 0 : g64.ws_bytes());
    const size_t w_t32 = take(t32.units.empty() ? 0 : t32.ws_bytes());
    const size_t w_t64 = take(t64.units.empty() ? 0 : t64.ws_bytes());
    const size_t dev_bytes = off;
    if ((e = ensure_pinned(slot, staged)) != hipSuccess) return e;
    if ((e = ensure_arena(slot, ops_bytes + dev_bytes)) != hipSuccess) return e;
    ops = slot.arena;
    auto rebase = [&](auto *p) { return (decltype(p))((char *)p + (uintptr_t)ops); };
    for (size_t i = 0; i < g32.descs.size(); ++i) {
      auto &x = g32.descs[i];
      if (!g32_abs[i]) x.a = rebase(x.a), x.b = rebase(x.b);
      x.c = rebase(x.c);
    }
    for (size_t i = 0; i < g64.descs.size(); ++i) {
      auto &x = g64.descs[i];
      if (!g64_abs[i]) x.a = rebase(x.a), x.b = rebase(x.b);
      x.c = rebase(x.c);
    }
    for (auto &x : t32.descs) x.a = rebase(x.a), x.b = rebase(x.b);
    for (auto &x : t64.descs) x.a = rebase(x.a), x.b = rebase(x.b);
    for (auto &x : gens) x.ptr = rebase((char *)x.ptr);
    for (auto &x : hashes) x.ptr = rebase((const char *)x.ptr);
    for (auto &x : cplx) x.c = rebase(x.c), x.ws = rebase(x.ws);  // a, b: pool slices
    char *h = slot.pinned;
    std::memcpy(h + o_gd, gens.data(), gens.size() * sizeof(GenDesc));
    std::memcpy(h + o_gu, gen_units.data(), gen_units.size() * sizeof(int2));
    std::memcpy(h + o_hd, hashes.data(), hashes.size() * sizeof(HashDesc));
    std::memcpy(h + o_hu, hash_units.data(), hash_units.size() * sizeof(int2));
    std::memset(h + o_ho, 0, hashes.size() * sizeof(uint64_t));
    std::memcpy(h + o_cd, cplx.data(), cplx.size() * sizeof(CplxDesc));
    std::memcpy(h + o_cs, cplx_split_units.data(), cplx_split_units.size() * sizeof(int2));
    std::memcpy(h + o_cc, cplx_comb_units.data(), cplx_comb_units.size() * sizeof(int2));
    if (!g32.units.empty()) g32.stage(h + o_g32);
    if (!g64.units.empty()) g64.stage(h + o_g64);
    if (!t32.units.empty()) t32.stage(h + o_t32);
    if (!t64.units.empty()) t64.stage(h + o_t64);
    slot.hash_off = o_ho;
    char *d = slot.arena + ops_bytes;
    e = hipMemcpyAsync(d, h, staged, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    e = launch_gen((const GenDesc *)(d + o_gd), (const int2 *)(d + o_gu), (int)gen_units.size(),
                   st);
    if (e != hipSuccess) return e;
    e = launch_cplx_split((const CplxDesc *)(d + o_cd), (const int2 *)(d + o_cs),
                          (int)cplx_split_units.size(), st);
    if (e != hipSuccess) return e;
    if (!plain) {
      if (!g32.units.empty() &&
          (e = g32.launch(d + o_g32, grouped_gemm_grid<float>(), st, d + w_g32)) != hipSuccess)
        return e;
      if (!g64.units.empty() &&
          (e = g64.launch(d + o_g64, grouped_gemm_grid<double>(), st, d + w_g64)) != hipSuccess)
        return e;
      if (!t32.units.empty() &&
          (e = t32.launch(d + o_t32, grouped_trsm_grid<float>(), st, d + w_t32)) != hipSuccess)
        return e;
      if (!t64.units.empty() &&
          (e = t64.launch(d + o_t64, grouped_trsm_grid<double>(), st, d + w_t64)) != hipSuccess)
        return e;
    }
    for (size_t i = 0; i < ids.size(); ++i) {
      const Req &r = reqs[ids[i]];
      const rblas_serve_request_t &s = r.spec;
      const char *a = r.pooled[0] ? r.pooled[0] : ops + r.off[0];
      const char *b = r.pooled[1] ? r.pooled[1] : ops + r.off[1];
      char *c = ops + r.off[2];
      if (plain && s.kind == RBLAS_SERVE_ZGEMM) {
        CgemmProblem<double> p{0, 0, s.m, s.n, s.k, 1.0, 0.0, 0.5, 0.0, (const double *)a, s.m,
                               (const double *)b, s.k, (double *)c, s.m};
        e = launch_cgemm<double>(p, (double *)(ops + r.off[3]), st);
      } else if (plain && s.kind == RBLAS_SERVE_GEMM) {
        if (s.dtype == 0) {
          GemmProblem<float> p{false, false, s.m, s.n, s.k, 1.0f, 0.5f, (const float *)a, s.m,
                               (const float *)b, s.k, (float *)c, s.m};
          e = launch_gemm_cfg<float>(p, gemm_config(p), ops + call_ws[i], st);
        } else {
          GemmProblem<double> p{false, false, s.m, s.n, s.k, 1.0, 0.5, (const double *)a, s.m,
                                (const double *)b, s.k, (double *)c, s.m};
          e = launch_gemm_cfg<double>(p, gemm_config(p), ops + call_ws[i], st);
        }
      } else if (plain && s.kind == RBLAS_SERVE_TRSM) {
        if (s.dtype == 0)
          e = launch_trsm<float>(TrsmProblem<float>{true, false, false, false, s.m, s.n, 1.0f,
                                                    (const float *)a, s.m, (float *)c, s.m},
                                 ops + call_ws[i], st);
        else
          e = launch_trsm<double>(TrsmProblem<double>{true, false, false, false, s.m, s.n, 1.0,
                                                      (const double *)a, s.m, (double *)c, s.m},
                                  ops + call_ws[i], st);
      }
      if (e != hipSuccess) return e;
    }
    e = launch_cplx_combine((const CplxDesc *)(d + o_cd), (const int2 *)(d + o_cc),
                            (int)cplx_comb_units.size(), st);
    if (e != hipSuccess) return e;
    e = launch_hash((const HashDesc *)(d + o_hd), (const int2 *)(d + o_hu), (int)hash_units.size(),
                    (uint64_t *)(d + o_ho), st);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(h + o_ho, d + o_ho, hashes.size() * sizeof(uint64_t),
                       hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    // Captured results (tests): dense copies into the caller's host buffers.
    for (size_t i = 0; i < ids.size(); ++i) {
      for (int32_t ci = 0; ci < cfg.capture_count; ++ci) {
        if (cfg.capture_ids[ci] != ids[i] || !cfg.capture_out || !cfg.capture_out[ci]) continue;
        const Req &r = reqs[ids[i]];
        e = hipMemcpyAsync(cfg.capture_out[ci], ops + r.off[2], r.bytes[2], hipMemcpyDeviceToHost,
                           st);
        if (e != hipSuccess) return e;
      }
    }
    return hipEventRecord(slot.ev, st);
  }

  template <typename T>
  GemmProblem<T> gemm_shape(const Req &r) const {
    const rblas_serve_request_t &s = r.spec;
    return GemmProblem<T>{false, false, s.m, s.n, s.k, T(1), T(0.5), nullptr, s.m, nullptr, s.k,
                          nullptr, s.m};
  }
};

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0;
  const size_t i = std::min(v.size() - 1, (size_t)std::floor(p * (double)(v.size() - 1) + 0.5));
  std::nth_element(v.begin(), v.begin() + i, v.end());
  return v[i];
}

}  // namespace
}  // namespace rblas

using namespace rblas;

extern "C" int32_t rblas_serve_run(const rblas_serve_request_t *reqs_in, size_t count,
                                   const rblas_serve_config_t *cfg, rblas_serve_result_t *results,
                                   rblas_serve_stats_t *stats) {
  if ((count && (!reqs_in || !results)) || !cfg)
    return fail(bad_arg("serve: null requests, results or config"), nullptr);
  const bool plain = cfg->mode == RBLAS_SERVE_PLAIN;
  if (cfg->mode != RBLAS_SERVE_BATCHED && !plain) return fail(bad_arg("serve: unknown mode"), nullptr);
  const int slots_n = plain ? 1 : std::max(1, std::min(cfg->max_inflight, 8));
  const int max_batch = plain ? 1 : std::max(1, cfg->max_batch);
  if (cfg->capture_count < 0 || (cfg->capture_count > 0 && !cfg->capture_ids))
    return fail(bad_arg("serve: bad capture list"), nullptr);
  if (cfg->hash_col_step < 0) return fail(bad_arg("serve: negative hash_col_step"), nullptr);
  std::vector<Req> reqs(count);
  for (size_t i = 0; i < count; ++i) {
    const rblas_serve_request_t &s = reqs_in[i];
    const bool kind_ok = s.kind == RBLAS_SERVE_GEMM || s.kind == RBLAS_SERVE_TRSM ||
                         s.kind == RBLAS_SERVE_ZGEMM;
    if (!kind_ok || s.m <= 0 || s.n <= 0 || (s.kind != RBLAS_SERVE_TRSM && s.k <= 0) ||
        (s.dtype != 0 && s.dtype != 1) || (s.kind == RBLAS_SERVE_TRSM && s.m > 2048) ||
        !(s.arrival_s >= 0) || (i > 0 && s.arrival_s < reqs_in[i - 1].arrival_s))
      return fail(bad_arg("serve: request " + std::to_string(i) +
                          " is malformed (kind, sizes, dtype, trsm order <= 2048, arrival order)"),
                  nullptr);
    reqs[i] = describe(s);
    if (reqs[i].footprint > cfg->hbm_cap_bytes)
      return fail(bad_arg("serve: request " + std::to_string(i) + " exceeds the HBM cap alone"),
                  nullptr);
  }
  // The largest batch the FIFO policy can form: at most max_batch
  // consecutive requests whose footprints fit the cap together.
  double window = 0;
  for (size_t i = 0; i < count; ++i) {
    double sum = 0;
    for (size_t j = i; j < count && j < i + (size_t)max_batch; ++j) {
      if (sum + reqs[j].footprint > cfg->hbm_cap_bytes) break;
      sum += reqs[j].footprint;
    }
    window = std::max(window, sum);
  }
  // Operand pools, generated before the first arrival.
  struct PoolGuard {
    Pools p;
    ~PoolGuard() {
      for (char *b : p.base)
        if (b) (void)hipFree(b);
    }
  } pool;
  {
    int64_t need[2] = {0, 0};
    for (const Req &r : reqs) {
      const int dt = r.spec.kind == RBLAS_SERVE_ZGEMM ? 1 : r.spec.dtype;
      need[dt] = std::max(need[dt], std::max(r.pool_len[0], r.pool_len[1]));
    }
    std::vector<GenDesc> gd;
    std::vector<int2> gu;
    for (int dt = 0; dt < 2; ++dt) {
      if (need[dt] == 0) continue;
      const int64_t elems = std::max<int64_t>(int64_t(1) << 25, need[dt]);
      RBLAS_HIP_TRY(hipMalloc((void **)&pool.p.base[dt], (size_t)elems * (dt ? 8 : 4)),
                    "serve operand pool");
      pool.p.elems[dt] = elems;
      const int d = (int)gd.size();
      gd.push_back(GenDesc{pool.p.base[dt], elems, 0, RBLAS_SERVE_POOL_SEED, 0, dt, 0});
      for (int64_t c = 0; c * kGenChunk < elems; ++c) gu.push_back(int2{d, (int)c});
    }
    if (!gd.empty()) {
      const size_t bd = round256(gd.size() * sizeof(GenDesc)), bu = gu.size() * sizeof(int2);
      char *tab = nullptr;
      RBLAS_HIP_TRY(hipMalloc((void **)&tab, bd + bu), "serve pool tables");
      hipError_t e = hipMemcpy(tab, gd.data(), gd.size() * sizeof(GenDesc), hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemcpy(tab + bd, gu.data(), bu, hipMemcpyHostToDevice);
      if (e == hipSuccess) e = launch_gen((const GenDesc *)tab, (const int2 *)(tab + bd),
                                          (int)gu.size(), nullptr);
      if (e == hipSuccess) e = hipDeviceSynchronize();
      const hipError_t f = hipFree(tab);
      if (e == hipSuccess) e = f;
      RBLAS_HIP_TRY(e, "serve pool generation");
    }
    for (Req &r : reqs) {
      if (r.pool_len[0] == 0) continue;
      const int dt = r.spec.kind == RBLAS_SERVE_ZGEMM ? 1 : r.spec.dtype;
      const size_t es = dt ? 8 : 4;
      for (int j = 0; j < 2; ++j)
        r.pooled[j] = pool.p.base[dt] +
                      es * (size_t)pool_offset(r.spec.seed, j + 1, r.pool_len[j], pool.p.elems[dt]);
    }
  }
  std::vector<Slot> slots(slots_n);
  for (Slot &s : slots) {
    RBLAS_HIP_TRY(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking), "serve stream");
    RBLAS_HIP_TRY(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming), "serve event");
    // Operands plus room for tables and workspaces (the arena grows if a
    // batch needs more).
    RBLAS_HIP_TRY(ensure_arena(s, (size_t)(window * 1.05) + (64u << 20)), "serve arena");
    RBLAS_HIP_TRY(ensure_pinned(s, 8u << 20), "serve staging");
  }
  Planner planner{*cfg, reqs, plain};
  std::deque<int> queue;
  size_t next = 0, done = 0;
  double inflight_fp = 0, peak_fp = 0;
  int batches = 0, max_busy = 0;
  double flops_total = 0;
  int32_t status = RBLAS_OK;
  const auto t0 = Clock::now();
  while (done < count && status == RBLAS_OK) {
    bool progress = false;
    const double now = seconds_since(t0);
    while (next < count && reqs_in[next].arrival_s <= now) {
      queue.push_back((int)next);
      results[next] = rblas_serve_result_t{};
      results[next].arrival_s = reqs_in[next].arrival_s;
      ++next;
      progress = true;
    }
    for (Slot &s : slots) {
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.ev);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) {
        status = hip_fail(q, "serve batch");
        break;
      }
      const double t = seconds_since(t0);
      const uint64_t *h = (const uint64_t *)(s.pinned + s.hash_off);
      for (size_t i = 0; i < s.reqs.size(); ++i) {
        const int id = s.reqs[i];
        results[id].done_s = t;
        results[id].hash = h[i];
        inflight_fp -= reqs[id].footprint;
        ++done;
      }
      s.busy = false;
      s.reqs.clear();
      progress = true;
    }
    if (status != RBLAS_OK) break;
    for (Slot &s : slots) {
      if (s.busy || queue.empty()) continue;
      std::vector<int> ids;
      while (!queue.empty() && (int)ids.size() < max_batch &&
             inflight_fp + reqs[queue.front()].footprint <= cfg->hbm_cap_bytes) {
        const int id = queue.front();
        queue.pop_front();
        inflight_fp += reqs[id].footprint;
        results[id].admit_s = seconds_since(t0);
        results[id].flops = reqs[id].flops;
        results[id].footprint_bytes = reqs[id].footprint;
        results[id].batch = batches;
        flops_total += reqs[id].flops;
        ids.push_back(id);
      }
      if (ids.empty()) break;  // the head waits for memory
      peak_fp = std::max(peak_fp, inflight_fp);
      const hipError_t e = planner.launch(s, ids);
      if (e != hipSuccess) {
        status = hip_fail(e, "serve launch");
        break;
      }
      const double enq = seconds_since(t0);
      for (int id : ids) results[id].enqueued_s = enq;
      s.reqs = ids;
      s.busy = true;
      ++batches;
      progress = true;
    }
    int busy = 0;
    for (const Slot &s : slots) busy += s.busy;
    max_busy = std::max(max_busy, busy);
    if (!progress) std::this_thread::yield();
  }
  for (Slot &s : slots) {
    if (s.st) (void)hipStreamSynchronize(s.st);
    if (s.arena) (void)hipFree(s.arena);
    if (s.pinned) (void)hipHostFree(s.pinned);
    if (s.ev) (void)hipEventDestroy(s.ev);
    if (s.st) (void)hipStreamDestroy(s.st);
  }
  if (status != RBLAS_OK) return status;
  if (stats) {
    std::vector<double> lat(count);
    double wall = 0;
    for (size_t i = 0; i < count; ++i) {
      lat[i] = results[i].done_s - results[i].arrival_s;
      wall = std::max(wall, results[i].done_s);
    }
    *stats = rblas_serve_stats_t{};
    stats->wall_s = wall;
    stats->flops = flops_total;
    stats->p50_s = percentile(lat, 0.50);
    stats->p99_s = percentile(lat, 0.99);
    stats->batches = batches;
    stats->max_inflight_seen = max_busy;
    stats->peak_footprint_bytes = peak_fp;
  }
  return RBLAS_OK;
}
