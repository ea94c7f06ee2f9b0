// Host-side argument checks that reproduce rblas's error behaviour exactly.
// Pure C++ (no HIP calls): every check runs before the device is touched.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/rblas_hip.h"

namespace rblas {

inline bool is_transposed(int32_t t) { return t == RBLAS_TRANS || t == RBLAS_CONJ_TRANS; }
inline bool valid_layout(int32_t v) { return v == RBLAS_ROW_MAJOR || v == RBLAS_COL_MAJOR; }
inline bool valid_trans(int32_t v) { return v >= RBLAS_NO_TRANS && v <= RBLAS_CONJ_TRANS; }
inline bool valid_uplo(int32_t v) { return v == RBLAS_UPPER || v == RBLAS_LOWER; }
inline bool valid_side(int32_t v) { return v == RBLAS_LEFT || v == RBLAS_RIGHT; }
inline bool valid_diag(int32_t v) { return v == RBLAS_NON_UNIT || v == RBLAS_UNIT; }

struct Check {
  int32_t status = RBLAS_OK;
  rblas_error_t err{RBLAS_OK, 0, 0, 0};
  std::string msg;
  bool ok() const { return status == RBLAS_OK; }
};

// `check_matrix` (level3.rs:107-140): one stored rows x cols operand against
// its leading dimension and slice length. rows or cols == 0 passes; then the
// leading-dimension test comes before the length test. usize arithmetic is
// unsigned and wraps, as rblas's release build does.
inline Check check_matrix(int32_t layout, char which, size_t rows, size_t cols, size_t lead,
                          size_t len) {
  Check c;
  if (rows == 0 || cols == 0) return c;
  size_t min_lead, need;
  if (layout == RBLAS_ROW_MAJOR) {
    min_lead = cols;
    need = (rows - 1) * lead + cols;
  } else {
    min_lead = rows;
    need = (cols - 1) * lead + rows;
  }
  if (lead < min_lead) {
    c.status = RBLAS_INVALID_LEADING_DIM;
    c.err = {RBLAS_INVALID_LEADING_DIM, which, (uint64_t)lead, (uint64_t)min_lead};
    c.msg = std::string("invalid leading dimension for `") + which + "`: got " +
            std::to_string(lead) + ", need >= " + std::to_string(min_lead);
    return c;
  }
  if (len < need) {
    c.status = RBLAS_BUFFER_TOO_SMALL;
    c.err = {RBLAS_BUFFER_TOO_SMALL, which, (uint64_t)len, (uint64_t)need};
    c.msg = std::string("buffer `") + which + "` too small: got " + std::to_string(len) +
            " elements, need >= " + std::to_string(need);
  }
  return c;
}

// try_gemm's validation order (level3.rs:84-98): a, then b, then c, in the
// caller's layout, with stored shapes swapped under transposition.
inline Check try_gemm_check(int32_t layout, int32_t ta, int32_t tb, size_t m, size_t n, size_t k,
                            size_t a_len, size_t lda, size_t b_len, size_t ldb, size_t c_len,
                            size_t ldc) {
  const bool at = is_transposed(ta), bt = is_transposed(tb);
  Check c = check_matrix(layout, 'a', at ? k : m, at ? m : k, lda, a_len);
  if (!c.ok()) return c;
  c = check_matrix(layout, 'b', bt ? n : k, bt ? k : n, ldb, b_len);
  if (!c.ok()) return c;
  return check_matrix(layout, 'c', m, n, ldc, c_len);
}

inline Check panic(const std::string &m) {
  Check c;
  c.status = RBLAS_PANIC;
  c.err = {RBLAS_PANIC, 0, 0, 0};
  c.msg = m;
  return c;
}

// Largest index touched by a k-loop/i-loop gather x[i*si + p*sp] for i < r,
// p < q (r, q > 0), with overflow reported as "out of bounds".
inline bool max_index(size_t r, size_t q, size_t si, size_t sp, size_t &out) {
  size_t a, b;
  if (__builtin_mul_overflow(r - 1, si, &a)) return false;
  if (__builtin_mul_overflow(q - 1, sp, &b)) return false;
  return !__builtin_add_overflow(a, b, &out);
}

// Where rblas's non-checked `gemm` panics, in its order, after the RowMajor
// -> ColMajor swap (level3.rs:52-61). Arguments are column-major here: op(A)
// is m x k, op(B) is k x n.
//   NN: three asserts (level3.rs:164-166), evaluated with wrapping usize
//       arithmetic: `len >= (k-1)*ld + m || k == 0`.
//   otherwise: materialize's gathers index a then b (level3.rs:193-220) and
//       panic on the first out-of-range index; then gemm_contig returns when
//       m or n is 0 (kernel/gemm.rs:198), apply_beta slices each C column
//       when beta != 1 (kernel/gemm.rs:326-337).
//   In both paths a C that is too short for the microkernel's raw-pointer
//   writes (reached when k != 0 and alpha != 0) is undefined behaviour in
//   rblas; we report it as a panic too.
inline Check gemm_panic_check(bool at, bool bt, size_t m, size_t n, size_t k, size_t a_len,
                              size_t lda, size_t b_len, size_t ldb, size_t c_len, size_t ldc,
                              bool alpha_zero, bool beta_one) {
  if (!at && !bt) {
    if (!(a_len >= (k - 1) * lda + m || k == 0)) return panic("gemm: A too small");
    if (!(b_len >= (n - 1) * ldb + k || n == 0)) return panic("gemm: B too small");
    if (!(c_len >= (n - 1) * ldc + m || n == 0)) return panic("gemm: C too small");
    return Check{};
  }
  if (m > 0 && k > 0) {
    size_t hi;
    bool fits = at ? max_index(m, k, lda, 1, hi) : max_index(m, k, 1, lda, hi);
    if (!fits || hi >= a_len)
      return panic("index out of bounds: the len is " + std::to_string(a_len) +
                   " but the index is " + (fits ? std::to_string(hi) : std::string("past usize")));
  }
  if (n > 0 && k > 0) {
    size_t hi;
    bool fits = bt ? max_index(n, k, 1, ldb, hi) : max_index(n, k, ldb, 1, hi);
    if (!fits || hi >= b_len)
      return panic("index out of bounds: the len is " + std::to_string(b_len) +
                   " but the index is " + (fits ? std::to_string(hi) : std::string("past usize")));
  }
  if (m == 0 || n == 0) return Check{};
  const bool touches_c = !beta_one || (k != 0 && !alpha_zero);
  if (touches_c) {
    size_t hi;
    bool fits = max_index(m, n, 1, ldc, hi);
    if (!fits || hi >= c_len)
      return panic("range end index " + (fits ? std::to_string(hi + 1) : std::string("past usize")) +
                   " out of range for slice of length " + std::to_string(c_len));
  }
  return Check{};
}

}  // namespace rblas
