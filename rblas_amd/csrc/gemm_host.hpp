// Host-side normalisation of one rblas gemm call into its column-major form,
// with rblas's panics and the device path's limits checked. Shared by the
// single-call entry points and the grouped planner.
#pragma once
#include <climits>

#include "kernels.hpp"
#include "validate.hpp"

namespace rblas {

// Column-major form of one gemm call, after rblas's RowMajor swap
// (level3.rs:56-60: A<->B, m<->n, transposes swapped).
template <typename T>
struct ColMajorCall {
  bool at, bt;
  size_t m, n, k;
  T alpha, beta;
  const T *a;
  size_t a_len, lda;
  const T *b;
  size_t b_len, ldb;
  T *c;
  size_t c_len, ldc;
};

template <typename T>
inline Check normalize(int32_t layout, int32_t ta, int32_t tb, size_t m, size_t n, size_t k, T alpha,
                const T *a, size_t a_len, size_t lda, const T *b, size_t b_len, size_t ldb,
                T beta, T *c, size_t c_len, size_t ldc, ColMajorCall<T> &out) {
  Check bad;
  if (!valid_layout(layout) || !valid_trans(ta) || !valid_trans(tb)) {
    bad.status = RBLAS_BAD_ENUM;
    bad.err = {RBLAS_BAD_ENUM, 0, 0, 0};
    bad.msg = "gemm: enum argument out of range";
    return bad;
  }
  if (layout == RBLAS_COL_MAJOR)
    out = {is_transposed(ta), is_transposed(tb), m, n, k, alpha, beta, a, a_len, lda,
           b, b_len, ldb, c, c_len, ldc};
  else
    out = {is_transposed(tb), is_transposed(ta), n, m, k, alpha, beta, b, b_len, ldb,
           a, a_len, lda, c, c_len, ldc};
  Check pc = gemm_panic_check(out.at, out.bt, out.m, out.n, out.k, out.a_len, out.lda, out.b_len,
                              out.ldb, out.c_len, out.ldc, out.alpha == T(0), out.beta == T(1));
  if (!pc.ok()) return pc;
  // The kernels index tiles with 32-bit m, n, k.
  if (out.m > INT_MAX || out.n > INT_MAX || out.k > INT_MAX) {
    bad.status = RBLAS_BAD_ARG;
    bad.err = {RBLAS_BAD_ARG, 0, 0, 0};
    bad.msg = "gemm: m, n and k must each be below 2^31 on the device path";
    return bad;
  }
  // rblas walks C's columns one after another, so columns that overlap
  // (ldc < m) see each other's updates in that order; device tiles update
  // them at the same time. Refused whenever C is written.
  const bool touches_c = out.m && out.n && (out.beta != T(1) || (out.k && out.alpha != T(0)));
  if (touches_c && out.n > 1 && out.ldc < out.m) {
    bad.status = RBLAS_BAD_ARG;
    bad.err = {RBLAS_BAD_ARG, 0, 0, 0};
    bad.msg = "gemm: ldc < m (overlapping columns of C) is not supported on the device";
    return bad;
  }
  // A 128-wide tile of either operand, and of C, must span under 2 GiB (the
  // tile loaders and the epilogue address it with 31-bit byte offsets).
  size_t ldmax = out.lda > out.ldb ? out.lda : out.ldb;
  if (out.ldc > ldmax) ldmax = out.ldc;
  const bool runs_kernel = out.m && out.n && out.k && out.alpha != T(0);
  if (runs_kernel && (ldmax * 128 + 160) * sizeof(T) >= 0x7fffffe0ull) {
    bad.status = RBLAS_BAD_ARG;
    bad.err = {RBLAS_BAD_ARG, 0, 0, 0};
    bad.msg = "gemm: lda/ldb/ldc above 2^31/(128*sizeof(T)) elements is not supported on the device";
    return bad;
  }
  return Check{};
}

}  // namespace rblas
