// Host-side argument checks of rblas's syrk / trsm shared by the single-call
// entry points and the grouped planner. Pure C++ (no HIP calls).
#pragma once
#include <climits>
#include <string>

#include "validate.hpp"

namespace rblas {

inline Check bad_arg(const std::string &m) {
  Check c;
  c.status = RBLAS_BAD_ARG;
  c.err = {RBLAS_BAD_ARG, 0, 0, 0};
  c.msg = m;
  return c;
}

inline Check bad_enum(const char *m) {
  Check c;
  c.status = RBLAS_BAD_ENUM;
  c.err = {RBLAS_BAD_ENUM, 0, 0, 0};
  c.msg = m;
  return c;
}

inline Check oob(size_t len, bool fits, size_t idx) {
  return panic("index out of bounds: the len is " + std::to_string(len) + " but the index is " +
               (fits ? std::to_string(idx) : std::string("past usize")));
}

inline bool ld_too_large(size_t ld, size_t elem) { return (ld * 128 + 160) * elem >= 0x7fffffe0ull; }

// Where rblas's trsm panics (level3.rs:403-520). Nothing is indexed when m
// or n is 0. Otherwise the alpha pre-scale (alpha != 1) and the
// substitutions index every B entry up to (m-1) + (n-1)*ldb, and every
// substitution step reads the diagonal of op(A) (finish's a_ii argument is
// evaluated for Unit too), whose largest index is (q-1)*(lda+1), q = m (Left)
// or n (Right); the off-diagonal reads stay below it.
template <typename T>
inline Check trsm_validate(int32_t side, int32_t uplo, int32_t trans, int32_t diag, size_t m, size_t n,
                    size_t a_len, size_t lda, size_t b_len, size_t ldb) {
  if (!valid_side(side) || !valid_uplo(uplo) || !valid_trans(trans) || !valid_diag(diag))
    return bad_enum("trsm: enum argument out of range");
  if (m == 0 || n == 0) return Check{};
  size_t hb;
  const bool fb = max_index(m, n, 1, ldb, hb);
  if (!fb || hb >= b_len) return oob(b_len, fb, hb);
  const size_t q = side == RBLAS_LEFT ? m : n;
  size_t ha;
  const bool fa = max_index(q, q, 1, lda, ha);  // (q-1) + (q-1)*lda
  if (!fa || ha >= a_len) return oob(a_len, fa, ha);
  if (m > INT_MAX || n > INT_MAX) return bad_arg("trsm: m and n must be below 2^31 on the device");
  if (n > 1 && ldb < m)
    return bad_arg("trsm: ldb < m (overlapping columns of B) is not supported on the device");
  if (ld_too_large(lda, sizeof(T)) || ld_too_large(ldb, sizeof(T)))
    return bad_arg("trsm: lda/ldb above 2^31/(128*sizeof(T)) elements is not supported on the device");
  return Check{};
}


}  // namespace rblas
