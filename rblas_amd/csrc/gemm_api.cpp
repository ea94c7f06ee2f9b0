// C-ABI gemm entry points: level3::gemm / level3::try_gemm (level3.rs:32-103)
// for host slices, plus the HBM-resident `_dev` form.
#include <climits>
#include <mutex>
#include <string>

#include "gemm_host.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace rblas {
namespace {

// Whether a normalized call reaches the product kernel (else it is a no-op
// or a beta-only update).
template <typename T>
bool runs_product(const ColMajorCall<T> &cc) {
  return cc.m != 0 && cc.n != 0 && cc.k != 0 && cc.alpha != T(0);
}

// Split planes for an SGEMM compute mode (0 = plain f32).
inline int split_planes(int32_t compute) {
  return compute == RBLAS_COMPUTE_BF16X3 ? 2 : compute == RBLAS_COMPUTE_BF16X6 ? 3 : 0;
}

// The shape-only problem of a normalized call (pointers unset).
template <typename T>
GemmProblem<T> shape_of(const ColMajorCall<T> &cc) {
  return GemmProblem<T>{cc.at, cc.bt, (int64_t)cc.m, (int64_t)cc.n, (int64_t)cc.k, cc.alpha,
                        cc.beta, nullptr, (int64_t)cc.lda, nullptr, (int64_t)cc.ldb, nullptr,
                        (int64_t)cc.ldc};
}

// Device workspace of a call: the bf16 planes of a split SGEMM, or the k-split
// partials of the MFMA gemm's launch shape.
template <typename T>
size_t workspace_bytes(const ColMajorCall<T> &cc, int planes) {
  if (!runs_product(cc)) return 0;
  if (planes) return sgemm_split_workspace((int64_t)cc.m, (int64_t)cc.n, (int64_t)cc.k, planes);
  const GemmProblem<T> p = shape_of(cc);
  return gemm_workspace_bytes<T>(p, gemm_config<T>(p));
}

// f64: no split compute modes (`planes` is used by the f32 form only).
template <typename T>
hipError_t launch_product(const GemmProblem<T> &p, int /*planes*/, void *ws, hipStream_t st) {
  return launch_gemm_cfg<T>(p, gemm_config<T>(p), ws, st);
}
template <>
hipError_t launch_product<float>(const GemmProblem<float> &p, int planes, void *ws,
                                 hipStream_t st) {
  if (planes) return launch_sgemm_split(p, planes, ws, st);
  return launch_gemm_cfg<float>(p, gemm_config<float>(p), ws, st);
}

// Enqueue the device work of a normalized call whose pointers are in HBM.
// Mirrors gemm_contig's control flow (kernel/gemm.rs:198-205): nothing when
// m or n is 0; beta-only update when k is 0 or alpha is 0. `ws` holds
// workspace_bytes(cc, planes).
template <typename T>
hipError_t enqueue(const ColMajorCall<T> &cc, const T *a, size_t lda, const T *b, size_t ldb, T *c,
                   size_t ldc, int planes, void *ws, hipStream_t st) {
  if (cc.m == 0 || cc.n == 0) return hipSuccess;
  if (cc.k == 0 || cc.alpha == T(0)) {
    if (cc.beta == T(1)) return hipSuccess;
    return launch_scale<T>((int64_t)cc.m, (int64_t)cc.n, cc.beta, c, (int64_t)ldc, st);
  }
  GemmProblem<T> p{cc.at, cc.bt, (int64_t)cc.m, (int64_t)cc.n, (int64_t)cc.k, cc.alpha, cc.beta,
                   a, (int64_t)lda, b, (int64_t)ldb, c, (int64_t)ldc};
  return launch_product<T>(p, planes, ws, st);
}

template <typename T>
int32_t gemm_host(const ColMajorCall<T> &cc, int planes = 0) {
  if (cc.m == 0 || cc.n == 0) return RBLAS_OK;
  // beta 1 with no product leaves C untouched (kernel/gemm.rs:202-205).
  if ((cc.k == 0 || cc.alpha == T(0)) && cc.beta == T(1)) return RBLAS_OK;
  CtxLease lease;
  if (!lease.get()) return hip_fail(lease.status(), "stream create");
  Ctx &ctx = *lease.get();
  const bool need_ab = cc.k != 0 && cc.alpha != T(0);
  const size_t total = (need_ab ? cc.m * cc.k + cc.k * cc.n : 0) + cc.m * cc.n;
  HostXfer x(ctx, sizeof(T), total * sizeof(T));
  Staged sa, sb, sc;
  if (need_ab) {
    sa.plan(cc.at ? cc.k : cc.m, cc.at ? cc.m : cc.k, cc.lda, x.pinned());
    sb.plan(cc.bt ? cc.n : cc.k, cc.bt ? cc.k : cc.n, cc.ldb, x.pinned());
  }
  sc.plan(cc.m, cc.n, cc.ldc, x.pinned());
  RBLAS_HIP_TRY(ctx.ensure(0, sa.elems * sizeof(T)), "device buffer A");
  RBLAS_HIP_TRY(ctx.ensure(1, sb.elems * sizeof(T)), "device buffer B");
  RBLAS_HIP_TRY(ctx.ensure(2, sc.elems * sizeof(T)), "device buffer C");
  RBLAS_HIP_TRY(ctx.ensure(3, workspace_bytes(cc, planes)), "device workspace");
  RBLAS_HIP_TRY(x.reserve(sa, sb, sc), "pinned staging buffer");
  T *da = (T *)ctx.buf[0], *db = (T *)ctx.buf[1], *dc = (T *)ctx.buf[2];
  hipStream_t st = ctx.stream;
  if (need_ab) {
    RBLAS_HIP_TRY(x.in(sa, cc.a, da), "copy A to device");
    RBLAS_HIP_TRY(x.in(sb, cc.b, db), "copy B to device");
  }
  RBLAS_HIP_TRY(x.in(sc, cc.c, dc), "copy C to device");
  RBLAS_HIP_TRY(enqueue(cc, da, sa.dev_ld, db, sb.dev_ld, dc, sc.dev_ld, planes, ctx.buf[3], st),
                "gemm launch");
  RBLAS_HIP_TRY(x.out(sc, dc, cc.c), "copy C to host");
  RBLAS_HIP_TRY(x.finish(), "gemm");
  return RBLAS_OK;
}

template <typename T>
int32_t gemm_entry(int32_t layout, int32_t ta, int32_t tb, size_t m, size_t n, size_t k, T alpha,
                   const T *a, size_t a_len, size_t lda, const T *b, size_t b_len, size_t ldb,
                   T beta, T *c, size_t c_len, size_t ldc) {
  ColMajorCall<T> cc;
  Check chk = normalize(layout, ta, tb, m, n, k, alpha, a, a_len, lda, b, b_len, ldb, beta, c,
                        c_len, ldc, cc);
  if (!chk.ok()) return fail(chk, nullptr);
  return gemm_host(cc);
}

template <typename T>
int32_t gemm_try_entry(int32_t layout, int32_t ta, int32_t tb, size_t m, size_t n, size_t k,
                       T alpha, const T *a, size_t a_len, size_t lda, const T *b, size_t b_len,
                       size_t ldb, T beta, T *c, size_t c_len, size_t ldc, rblas_error_t *err) {
  if (err) *err = {RBLAS_OK, 0, 0, 0};
  if (!valid_layout(layout) || !valid_trans(ta) || !valid_trans(tb)) {
    Check bad;
    bad.status = RBLAS_BAD_ENUM;
    bad.err = {RBLAS_BAD_ENUM, 0, 0, 0};
    bad.msg = "try_gemm: enum argument out of range";
    return fail(bad, err);
  }
  Check chk = try_gemm_check(layout, ta, tb, m, n, k, a_len, lda, b_len, ldb, c_len, ldc);
  if (!chk.ok()) return fail(chk, err);
  ColMajorCall<T> cc;
  chk = normalize(layout, ta, tb, m, n, k, alpha, a, a_len, lda, b, b_len, ldb, beta, c, c_len,
                  ldc, cc);
  if (!chk.ok()) return fail(chk, err);
  return gemm_host(cc);
}

template <typename T>
int32_t gemm_dev_entry(int32_t layout, int32_t ta, int32_t tb, size_t m, size_t n, size_t k,
                       T alpha, const T *a, size_t a_len, size_t lda, const T *b, size_t b_len,
                       size_t ldb, T beta, T *c, size_t c_len, size_t ldc, void *stream,
                       int planes = 0) {
  ColMajorCall<T> cc;
  Check chk = normalize(layout, ta, tb, m, n, k, alpha, a, a_len, lda, b, b_len, ldb, beta, c,
                        c_len, ldc, cc);
  if (!chk.ok()) return fail(chk, nullptr);
  hipStream_t st = (hipStream_t)stream;
  const size_t wsb = workspace_bytes(cc, planes);
  void *ws = nullptr;
  if (wsb) RBLAS_HIP_TRY(pool_alloc(&ws, wsb, st), "gemm workspace");
  hipError_t e = enqueue(cc, cc.a, cc.lda, cc.b, cc.ldb, cc.c, cc.ldc, planes, ws, st);
  e = pool_release(ws, st, e);
  RBLAS_HIP_TRY(e, "gemm launch");
  return RBLAS_OK;
}

Check bad_compute() {
  Check bad;
  bad.status = RBLAS_BAD_ENUM;
  bad.err = {RBLAS_BAD_ENUM, 0, 0, 0};
  bad.msg = "sgemm_ex: compute mode out of range";
  return bad;
}

inline bool valid_compute(int32_t c) {
  return c == RBLAS_COMPUTE_F32 || c == RBLAS_COMPUTE_BF16X3 || c == RBLAS_COMPUTE_BF16X6;
}

// ---------------------------------------------------------------- symm
// level3::symm (level3.rs:301-381): rblas builds the dense symmetric A from
// its stored `uplo` triangle (full_symmetric: every read is a stored entry,
// the largest index (q-1)(lda+1)), then calls gemm(ColMajor, N, N) with it
// as the left (Side Left) or right (Side Right) operand. The same checks in
// the same order: A's largest index, then that gemm's.
template <typename T>
Check symm_validate(int32_t side, int32_t uplo, size_t m, size_t n, T alpha, size_t a_len,
                    size_t lda, const T *b, size_t b_len, size_t ldb, T beta, T *c, size_t c_len,
                    size_t ldc, ColMajorCall<T> &cc) {
  if ((side != RBLAS_LEFT && side != RBLAS_RIGHT) || (uplo != RBLAS_UPPER && uplo != RBLAS_LOWER)) {
    Check bad;
    bad.status = RBLAS_BAD_ENUM;
    bad.err = {RBLAS_BAD_ENUM, 0, 0, 0};
    bad.msg = "symm: enum argument out of range";
    return bad;
  }
  const size_t q = side == RBLAS_LEFT ? m : n;
  if (q > 0) {
    size_t hi;
    const bool fits = max_index(q, q, 1, lda, hi);
    if (!fits || hi >= a_len)
      return panic("index out of bounds: the len is " + std::to_string(a_len) +
                   " but the index is " + (fits ? std::to_string(hi) : std::string("past usize")));
  }
  if (side == RBLAS_LEFT)
    return normalize<T>(RBLAS_COL_MAJOR, 0, 0, m, n, q, alpha, nullptr, q * q, q, b, b_len, ldb,
                        beta, c, c_len, ldc, cc);
  return normalize<T>(RBLAS_COL_MAJOR, 0, 0, m, n, q, alpha, b, b_len, ldb, nullptr, q * q, q,
                      beta, c, c_len, ldc, cc);
}

// Dense-A bytes at the front of a symm workspace (256-byte aligned).
inline size_t sym_bytes(size_t q, size_t elem) { return (q * q * elem + 255) / 256 * 256; }

// Enqueue a validated symm whose operands are in HBM: mirror A into `full`,
// then the gemm on it.
template <typename T>
hipError_t symm_enqueue(const ColMajorCall<T> &cc, bool left, bool upper, const T *a, size_t lda,
                        const T *b, size_t ldb, T *c, size_t ldc, T *full, void *gemm_ws,
                        hipStream_t st) {
  const size_t q = left ? cc.m : cc.n;
  if (runs_product(cc)) {
    hipError_t e = launch_sym_full<T>(a, (int64_t)lda, (int64_t)q, upper, false, false, full, st);
    if (e != hipSuccess) return e;
  }
  if (left) return enqueue(cc, (const T *)full, q, b, ldb, c, ldc, 0, gemm_ws, st);
  return enqueue(cc, b, ldb, (const T *)full, q, c, ldc, 0, gemm_ws, st);
}

template <typename T>
int32_t symm_host(int32_t side, int32_t uplo, size_t m, size_t n, T alpha, const T *a,
                  size_t a_len, size_t lda, const T *b, size_t b_len, size_t ldb, T beta, T *c,
                  size_t c_len, size_t ldc) {
  ColMajorCall<T> cc;
  Check chk = symm_validate<T>(side, uplo, m, n, alpha, a_len, lda, b, b_len, ldb, beta, c, c_len,
                               ldc, cc);
  if (!chk.ok()) return fail(chk, nullptr);
  if (m == 0 || n == 0) return RBLAS_OK;
  const bool left = side == RBLAS_LEFT;
  const size_t q = left ? m : n;
  const bool prod = runs_product(cc);
  CtxLease lease;
  if (!lease.get()) return hip_fail(lease.status(), "stream create");
  Ctx &ctx = *lease.get();
  HostXfer x(ctx, sizeof(T), ((prod ? q * q + m * n : 0) + m * n) * sizeof(T));
  Staged sa, sb, sc;
  if (prod) {
    sa.plan(q, q, lda, x.pinned());
    sb.plan(m, n, ldb, x.pinned());
  }
  sc.plan(m, n, ldc, x.pinned());
  const size_t full_b = prod ? sym_bytes(q, sizeof(T)) : 0;
  RBLAS_HIP_TRY(ctx.ensure(0, sa.elems * sizeof(T)), "device buffer A");
  RBLAS_HIP_TRY(ctx.ensure(1, sb.elems * sizeof(T)), "device buffer B");
  RBLAS_HIP_TRY(ctx.ensure(2, sc.elems * sizeof(T)), "device buffer C");
  RBLAS_HIP_TRY(ctx.ensure(3, full_b + workspace_bytes(cc, 0)), "device workspace");
  RBLAS_HIP_TRY(x.reserve(sa, sb, sc), "pinned staging buffer");
  T *da = (T *)ctx.buf[0], *db = (T *)ctx.buf[1], *dc = (T *)ctx.buf[2];
  if (prod) {
    RBLAS_HIP_TRY(x.in(sa, a, da), "copy A to device");
    RBLAS_HIP_TRY(x.in(sb, b, db), "copy B to device");
  }
  RBLAS_HIP_TRY(x.in(sc, c, dc), "copy C to device");
  RBLAS_HIP_TRY(symm_enqueue<T>(cc, left, uplo == RBLAS_UPPER, da, sa.dev_ld, db, sb.dev_ld, dc,
                                sc.dev_ld, (T *)ctx.buf[3], (char *)ctx.buf[3] + full_b,
                                ctx.stream),
                "symm launch");
  RBLAS_HIP_TRY(x.out(sc, dc, c), "copy C to host");
  RBLAS_HIP_TRY(x.finish(), "symm");
  return RBLAS_OK;
}

template <typename T>
int32_t symm_dev(int32_t side, int32_t uplo, size_t m, size_t n, T alpha, const T *a, size_t a_len,
                 size_t lda, const T *b, size_t b_len, size_t ldb, T beta, T *c, size_t c_len,
                 size_t ldc, void *stream) {
  ColMajorCall<T> cc;
  Check chk = symm_validate<T>(side, uplo, m, n, alpha, a_len, lda, b, b_len, ldb, beta, c, c_len,
                               ldc, cc);
  if (!chk.ok()) return fail(chk, nullptr);
  if (m == 0 || n == 0) return RBLAS_OK;
  const bool left = side == RBLAS_LEFT;
  const size_t q = left ? m : n;
  hipStream_t st = (hipStream_t)stream;
  const size_t full_b = runs_product(cc) ? sym_bytes(q, sizeof(T)) : 0;
  const size_t wsb = full_b + workspace_bytes(cc, 0);
  void *ws = nullptr;
  if (wsb) RBLAS_HIP_TRY(pool_alloc(&ws, wsb, st), "symm workspace");
  hipError_t e = symm_enqueue<T>(cc, left, uplo == RBLAS_UPPER, a, lda, b, ldb, c, ldc, (T *)ws,
                                 ws ? (char *)ws + full_b : nullptr, st);
  e = pool_release(ws, st, e);
  RBLAS_HIP_TRY(e, "symm launch");
  return RBLAS_OK;
}

// ------------------------------------------------- no-allocation DGEMM
// The k-split partial tiles and per-tile counters of rblas_dgemm_dev_noalloc
// live in this fixed device array instead of the stream-ordered pool. Its
// size covers every split configuration gemm_config gives f64: a split runs
// only with fewer than 256 tiles of 128^2 (or 128 of 64^2) and at most
// ceil(256 / tiles) splits, so tiles x splits <= 510 at 128^2 (255 x 2) and
// <= 383 at 64^2: 510 x 128 x 128 x 8 bytes = 63.75 MiB of partials, plus
// at most 255 counters. Every GPU the library runs on holds one copy until
// the process ends; calls on one GPU share it.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (the array's name is only a host-side stand-in).
hipError_t noalloc_scratch(void **out) {
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

int32_t dgemm_dev_noalloc(int32_t layout, int32_t ta, int32_t tb, size_t m, size_t n, size_t k,
                          double alpha, const double *a, size_t a_len, size_t lda, const double *b,
                          size_t b_len, size_t ldb, double beta, double *c, size_t c_len,
                          size_t ldc, void *stream) {
  ColMajorCall<double> cc;
  Check chk = normalize(layout, ta, tb, m, n, k, alpha, a, a_len, lda, b, b_len, ldb, beta, c,
                        c_len, ldc, cc);
  if (!chk.ok()) return fail(chk, nullptr);
  const size_t wsb = workspace_bytes(cc, 0);
  if (wsb > kNoAllocScratchBytes) {
    Check bad;
    bad.status = RBLAS_BAD_ARG;
    bad.err = {RBLAS_BAD_ARG, 0, 0, 0};
    bad.msg = "dgemm_dev_noalloc: the k-split workspace (" + std::to_string(wsb) +
              " bytes) exceeds the fixed device scratch (" +
              std::to_string(kNoAllocScratchBytes) + " bytes)";
    return fail(bad, nullptr);
  }
  void *ws = nullptr;
  if (wsb) RBLAS_HIP_TRY(noalloc_scratch(&ws), "dgemm_dev_noalloc scratch address");
  // The counters are cleared by a fill kernel inside the launch, and the
  // partials are written before they are read, so a previous call's
  // contents never reach this one.
  RBLAS_HIP_TRY(enqueue(cc, cc.a, cc.lda, cc.b, cc.ldb, cc.c, cc.ldc, 0, ws, (hipStream_t)stream),
                "dgemm_dev_noalloc launch");
  return RBLAS_OK;
}

}  // namespace
}  // namespace rblas

using namespace rblas;

extern "C" {

int32_t rblas_dgemm_dev_noalloc(int32_t layout, int32_t transa, int32_t transb, size_t m,
                                size_t n, size_t k, double alpha, const double *a, size_t a_len,
                                size_t lda, const double *b, size_t b_len, size_t ldb,
                                double beta, double *c, size_t c_len, size_t ldc, void *stream) {
  return dgemm_dev_noalloc(layout, transa, transb, m, n, k, alpha, a, a_len, lda, b, b_len, ldb,
                           beta, c, c_len, ldc, stream);
}

int32_t rblas_sgemm(int32_t layout, int32_t transa, int32_t transb, size_t m, size_t n, size_t k,
                    float alpha, const float *a, size_t a_len, size_t lda, const float *b,
                    size_t b_len, size_t ldb, float beta, float *c, size_t c_len, size_t ldc) {
  return gemm_entry<float>(layout, transa, transb, m, n, k, alpha, a, a_len, lda, b, b_len, ldb,
                           beta, c, c_len, ldc);
}

int32_t rblas_dgemm(int32_t layout, int32_t transa, int32_t transb, size_t m, size_t n, size_t k,
                    double alpha, const double *a, size_t a_len, size_t lda, const double *b,
                    size_t b_len, size_t ldb, double beta, double *c, size_t c_len, size_t ldc) {
  return gemm_entry<double>(layout, transa, transb, m, n, k, alpha, a, a_len, lda, b, b_len, ldb,
                            beta, c, c_len, ldc);
}

int32_t rblas_sgemm_try(int32_t layout, int32_t transa, int32_t transb, size_t m, size_t n,
                        size_t k, float alpha, const float *a, size_t a_len, size_t lda,
                        const float *b, size_t b_len, size_t ldb, float beta, float *c,
                        size_t c_len, size_t ldc, rblas_error_t *err) {
  return gemm_try_entry<float>(layout, transa, transb, m, n, k, alpha, a, a_len, lda, b, b_len,
                               ldb, beta, c, c_len, ldc, err);
}

int32_t rblas_dgemm_try(int32_t layout, int32_t transa, int32_t transb, size_t m, size_t n,
                        size_t k, double alpha, const double *a, size_t a_len, size_t lda,
                        const double *b, size_t b_len, size_t ldb, double beta, double *c,
                        size_t c_len, size_t ldc, rblas_error_t *err) {
  return gemm_try_entry<double>(layout, transa, transb, m, n, k, alpha, a, a_len, lda, b, b_len,
                                ldb, beta, c, c_len, ldc, err);
}

int32_t rblas_sgemm_ex(int32_t layout, int32_t transa, int32_t transb, size_t m, size_t n,
                       size_t k, float alpha, const float *a, size_t a_len, size_t lda,
                       const float *b, size_t b_len, size_t ldb, float beta, float *c,
                       size_t c_len, size_t ldc, int32_t compute) {
  if (!valid_compute(compute)) return fail(bad_compute(), nullptr);
  ColMajorCall<float> cc;
  Check chk = normalize(layout, transa, transb, m, n, k, alpha, a, a_len, lda, b, b_len, ldb,
                        beta, c, c_len, ldc, cc);
  if (!chk.ok()) return fail(chk, nullptr);
  return gemm_host(cc, split_planes(compute));
}

int32_t rblas_sgemm_ex_dev(int32_t layout, int32_t transa, int32_t transb, size_t m, size_t n,
                           size_t k, float alpha, const float *a, size_t a_len, size_t lda,
                           const float *b, size_t b_len, size_t ldb, float beta, float *c,
                           size_t c_len, size_t ldc, int32_t compute, void *stream) {
  if (!valid_compute(compute)) return fail(bad_compute(), nullptr);
  return gemm_dev_entry<float>(layout, transa, transb, m, n, k, alpha, a, a_len, lda, b, b_len,
                               ldb, beta, c, c_len, ldc, stream, split_planes(compute));
}

int32_t rblas_sgemm_dev(int32_t layout, int32_t transa, int32_t transb, size_t m, size_t n,
                        size_t k, float alpha, const float *a, size_t a_len, size_t lda,
                        const float *b, size_t b_len, size_t ldb, float beta, float *c,
                        size_t c_len, size_t ldc, void *stream) {
  return gemm_dev_entry<float>(layout, transa, transb, m, n, k, alpha, a, a_len, lda, b, b_len,
                               ldb, beta, c, c_len, ldc, stream);
}

int32_t rblas_dgemm_dev(int32_t layout, int32_t transa, int32_t transb, size_t m, size_t n,
                        size_t k, double alpha, const double *a, size_t a_len, size_t lda,
                        const double *b, size_t b_len, size_t ldb, double beta, double *c,
                        size_t c_len, size_t ldc, void *stream) {
  return gemm_dev_entry<double>(layout, transa, transb, m, n, k, alpha, a, a_len, lda, b, b_len,
                                ldb, beta, c, c_len, ldc, stream);
}

int32_t rblas_ssymm(int32_t side, int32_t uplo, size_t m, size_t n, float alpha, const float *a,
                    size_t a_len, size_t lda, const float *b, size_t b_len, size_t ldb, float beta,
                    float *c, size_t c_len, size_t ldc) {
  return symm_host<float>(side, uplo, m, n, alpha, a, a_len, lda, b, b_len, ldb, beta, c, c_len,
                          ldc);
}

int32_t rblas_dsymm(int32_t side, int32_t uplo, size_t m, size_t n, double alpha, const double *a,
                    size_t a_len, size_t lda, const double *b, size_t b_len, size_t ldb,
                    double beta, double *c, size_t c_len, size_t ldc) {
  return symm_host<double>(side, uplo, m, n, alpha, a, a_len, lda, b, b_len, ldb, beta, c, c_len,
                           ldc);
}

int32_t rblas_ssymm_dev(int32_t side, int32_t uplo, size_t m, size_t n, float alpha,
                        const float *a, size_t a_len, size_t lda, const float *b, size_t b_len,
                        size_t ldb, float beta, float *c, size_t c_len, size_t ldc, void *stream) {
  return symm_dev<float>(side, uplo, m, n, alpha, a, a_len, lda, b, b_len, ldb, beta, c, c_len,
                         ldc, stream);
}

int32_t rblas_dsymm_dev(int32_t side, int32_t uplo, size_t m, size_t n, double alpha,
                        const double *a, size_t a_len, size_t lda, const double *b, size_t b_len,
                        size_t ldb, double beta, double *c, size_t c_len, size_t ldc,
                        void *stream) {
  return symm_dev<double>(side, uplo, m, n, alpha, a, a_len, lda, b, b_len, ldb, beta, c, c_len,
                          ldc, stream);
}

}  // extern "C"
