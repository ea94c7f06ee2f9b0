// ORACLE — test infrastructure only. A CPU restatement of rblas's algorithms
// for the Level-3 hot path, used by tests/ as the checker and by bench.py as
// the timed CPU baseline. It is never linked into or called by the product
// library (rblas_amd/librblas_hip.so).
//
// Each function names the rblas source it restates (paths relative to the
// rblas repository). Arithmetic follows rblas step for step: the same
// MC/KC/NC blocking and MR x NR tiles, the same FMA chains in the AVX2
// microkernels, fused write-back on full tiles and unfused `+=` on edge
// tiles, beta applied first, materialized transposes, unfused substitution
// in trsm, the fma-based complex multiply, and the 2-accumulator AVX2 dot.
// Build with -ffp-contract=off so that only the FMAs written here are fused.
//
// Pinning (what checks each function besides this file):
//   gemm / gemm_contig  : reference unit tests restated in tests/ (17x13x19
//                         sin/cos, 8^3 alpha/beta spot value, NN/TN/row-major
//                         KATs, the 200x180x150 and 64x5000x48 LCG cases),
//                         exact small-integer products, numpy float64 matmul.
//   try_gemm            : the exact BlasError values of level3.rs:593-639.
//   syrk                : level3.rs:739-824 KATs + numpy.
//   trsm                : x = (1,2,3) KAT, round-trip residuals, numpy solve.
//   complex gemm (4M)   : complex3.rs:373-437 cases + numpy complex matmul.
//   dot / axpy / nrm2   : level1.rs:170-228 KATs (32, [12,24,36,48], 5,
//                         sqrt(3)*1e200), numpy.
//   complex trsm        : complex3.rs:439-478 round trip + numpy solve.
//   gemv / ger          : level2.rs:110-187 KATs, numpy (all layouts,
//                         transposes, strides), exact small integers.
//   complex gemv / geru / gerc : complex2.rs:150-182 KATs, numpy.
#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// kernel/gemm.rs:173-175
constexpr size_t MC = 384, KC = 512, NC = 4096;

template <typename T>
struct Tile;
// kernel/gemm.rs:139-146 (x86_64): f32 16 x 6, f64 8 x 6.
template <>
struct Tile<float> { static constexpr size_t MR = 16, NR = 6; };
template <>
struct Tile<double> { static constexpr size_t MR = 8, NR = 6; };

bool has_avx2_fma() {
  static const bool v = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
  return v;
}

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// gemm_avx2.rs:26-85 — 16 x 6 f32 tile, two 8-lane accumulators per column.
__attribute__((target("avx2,fma"))) void micro_16x6_f32(size_t mr, size_t nr, size_t kc,
                                                        float alpha, const float *ap,
                                                        const float *bp, float *c, size_t ldc,
                                                        size_t ci, size_t cj) {
  __m256 lo[6], hi[6];
#pragma GCC unroll 6
  for (int s = 0; s < 6; ++s) lo[s] = hi[s] = _mm256_setzero_ps();
  for (size_t p = 0; p < kc; ++p) {
    const float *a = ap + p * 16, *b = bp + p * 6;
    __m256 alo = _mm256_loadu_ps(a), ahi = _mm256_loadu_ps(a + 8);
#pragma GCC unroll 6
    for (int s = 0; s < 6; ++s) {
      __m256 bs = _mm256_broadcast_ss(b + s);
      lo[s] = _mm256_fmadd_ps(alo, bs, lo[s]);
      hi[s] = _mm256_fmadd_ps(ahi, bs, hi[s]);
    }
  }
  __m256 va = _mm256_set1_ps(alpha);
  if (mr == 16 && nr == 6) {
    for (int s = 0; s < 6; ++s) {
      float *cp = c + ci + (cj + s) * ldc;
      _mm256_storeu_ps(cp, _mm256_fmadd_ps(va, lo[s], _mm256_loadu_ps(cp)));
      _mm256_storeu_ps(cp + 8, _mm256_fmadd_ps(va, hi[s], _mm256_loadu_ps(cp + 8)));
    }
  } else {
    float buf[16];
    for (size_t s = 0; s < nr; ++s) {
      _mm256_storeu_ps(buf, lo[s]);
      _mm256_storeu_ps(buf + 8, hi[s]);
      float *col = c + ci + (cj + s) * ldc;
      for (size_t r = 0; r < mr; ++r) {
        float t = alpha * buf[r];  // two roundings: rblas's unfused `+=`
        col[r] = col[r] + t;
      }
    }
  }
}

// gemm_avx2.rs:97-155 — 8 x 6 f64 tile, two 4-lane accumulators per column.
__attribute__((target("avx2,fma"))) void micro_8x6_f64(size_t mr, size_t nr, size_t kc,
                                                       double alpha, const double *ap,
                                                       const double *bp, double *c, size_t ldc,
                                                       size_t ci, size_t cj) {
  __m256d lo[6], hi[6];
#pragma GCC unroll 6
  for (int s = 0; s < 6; ++s) lo[s] = hi[s] = _mm256_setzero_pd();
  for (size_t p = 0; p < kc; ++p) {
    const double *a = ap + p * 8, *b = bp + p * 6;
    __m256d alo = _mm256_loadu_pd(a), ahi = _mm256_loadu_pd(a + 4);
#pragma GCC unroll 6
    for (int s = 0; s < 6; ++s) {
      __m256d bs = _mm256_broadcast_sd(b + s);
      lo[s] = _mm256_fmadd_pd(alo, bs, lo[s]);
      hi[s] = _mm256_fmadd_pd(ahi, bs, hi[s]);
    }
  }
  __m256d va = _mm256_set1_pd(alpha);
  if (mr == 8 && nr == 6) {
    for (int s = 0; s < 6; ++s) {
      double *cp = c + ci + (cj + s) * ldc;
      _mm256_storeu_pd(cp, _mm256_fmadd_pd(va, lo[s], _mm256_loadu_pd(cp)));
      _mm256_storeu_pd(cp + 4, _mm256_fmadd_pd(va, hi[s], _mm256_loadu_pd(cp + 4)));
    }
  } else {
    double buf[8];
    for (size_t s = 0; s < nr; ++s) {
      _mm256_storeu_pd(buf, lo[s]);
      _mm256_storeu_pd(buf + 4, hi[s]);
      double *col = c + ci + (cj + s) * ldc;
      for (size_t r = 0; r < mr; ++r) {
        double t = alpha * buf[r];
        col[r] = col[r] + t;
      }
    }
  }
}

// kernel/gemm.rs:641-670 — the scalar tile used when AVX2/FMA is absent.
template <typename T>
void micro_raw_scalar(size_t mr, size_t nr, size_t kc, T alpha, const T *ap, const T *bp, T *c,
                      size_t ldc, size_t ci, size_t cj) {
  constexpr size_t MR = Tile<T>::MR, NR = Tile<T>::NR;
  T acc[MR][NR] = {};
  for (size_t p = 0; p < kc; ++p)
    for (size_t r = 0; r < MR; ++r)
      for (size_t s = 0; s < NR; ++s) acc[r][s] = std::fma(ap[p * MR + r], bp[p * NR + s], acc[r][s]);
  for (size_t s = 0; s < nr; ++s)
    for (size_t r = 0; r < mr; ++r) {
      T *cp = c + ci + r + (cj + s) * ldc;
      *cp = std::fma(alpha, acc[r][s], *cp);
    }
}

void micro(size_t mr, size_t nr, size_t kc, float alpha, const float *ap, const float *bp,
           float *c, size_t ldc, size_t ci, size_t cj) {
  if (has_avx2_fma()) return micro_16x6_f32(mr, nr, kc, alpha, ap, bp, c, ldc, ci, cj);
  micro_raw_scalar<float>(mr, nr, kc, alpha, ap, bp, c, ldc, ci, cj);
}
void micro(size_t mr, size_t nr, size_t kc, double alpha, const double *ap, const double *bp,
           double *c, size_t ldc, size_t ci, size_t cj) {
  if (has_avx2_fma()) return micro_8x6_f64(mr, nr, kc, alpha, ap, bp, c, ldc, ci, cj);
  micro_raw_scalar<double>(mr, nr, kc, alpha, ap, bp, c, ldc, ci, cj);
}

// kernel/gemm.rs:343-371
template <typename T>
void pack_a_panel(size_t mc, size_t kc, const T *a, size_t lda, size_t ic, size_t pc, T *pack) {
  constexpr size_t MR = Tile<T>::MR;
  size_t dst = 0;
  for (size_t i = 0; i < mc; i += MR) {
    size_t mr = std::min(MR, mc - i);
    for (size_t p = 0; p < kc; ++p) {
      size_t acol = pc + p;
      for (size_t r = 0; r < mr; ++r) pack[dst++] = a[(ic + i + r) + acol * lda];
      for (size_t r = mr; r < MR; ++r) pack[dst++] = T(0);
    }
  }
}

// kernel/gemm.rs:374-400
template <typename T>
void pack_b_panel(size_t kc, size_t nc, const T *b, size_t ldb, size_t pc, size_t jc, T *pack) {
  constexpr size_t NR = Tile<T>::NR;
  size_t dst = 0;
  for (size_t j = 0; j < nc; j += NR) {
    size_t nr = std::min(NR, nc - j);
    for (size_t p = 0; p < kc; ++p) {
      size_t brow = pc + p;
      for (size_t s = 0; s < nr; ++s) pack[dst++] = b[brow + (jc + j + s) * ldb];
      for (size_t s = nr; s < NR; ++s) pack[dst++] = T(0);
    }
  }
}

// kernel/gemm.rs:404-429
template <typename T>
void macrokernel(size_t mc, size_t nc, size_t kc, T alpha, const T *pa, const T *pb, T *c,
                 size_t ldc, size_t ic, size_t jc) {
  constexpr size_t MR = Tile<T>::MR, NR = Tile<T>::NR;
  for (size_t j = 0; j < nc; j += NR) {
    size_t nr = std::min(NR, nc - j);
    const T *bp = pb + (j / NR) * (kc * NR);
    for (size_t i = 0; i < mc; i += MR) {
      size_t mr = std::min(MR, mc - i);
      const T *ap = pa + (i / MR) * (kc * MR);
      micro(mr, nr, kc, alpha, ap, bp, c, ldc, ic + i, jc + j);
    }
  }
}

// kernel/gemm.rs:238-272 (scratch zero-filled like grow_zeroed, :95-103)
template <typename T>
void gemm_band(size_t m, size_t nc, size_t k, T alpha, const T *a, size_t lda, const T *b,
               size_t ldb, size_t j0, T *cband, size_t ldc) {
  constexpr size_t MR = Tile<T>::MR, NR = Tile<T>::NR;
  std::vector<T> pa(round_up(std::min(MC, m), MR) * std::min(KC, k), T(0));
  std::vector<T> pb(std::min(KC, k) * round_up(nc, NR), T(0));
  for (size_t pc = 0; pc < k; pc += KC) {
    size_t kc = std::min(KC, k - pc);
    pack_b_panel(kc, nc, b, ldb, pc, j0, pb.data());
    for (size_t ic = 0; ic < m; ic += MC) {
      size_t mc = std::min(MC, m - ic);
      pack_a_panel(mc, kc, a, lda, ic, pc, pa.data());
      macrokernel(mc, nc, kc, alpha, pa.data(), pb.data(), cband, ldc, ic, 0);
    }
  }
}

// kernel/gemm.rs:322-338
template <typename T>
void apply_beta(size_t m, size_t n, T beta, T *c, size_t ldc) {
  if (beta == T(1)) return;
  for (size_t j = 0; j < n; ++j) {
    T *col = c + j * ldc;
    if (beta == T(0))
      for (size_t i = 0; i < m; ++i) col[i] = T(0);
    else
      for (size_t i = 0; i < m; ++i) col[i] = beta * col[i];
  }
}

// kernel/gemm.rs:185-230 and the column-band parallel driver :277-313.
// threads <= 1 is the serial build; threads > 1 restates the `threads`
// feature with `threads` workers (rayon's pool size) pulling NR-aligned
// column chunks of ~4 tasks per worker from a shared counter.
template <typename T>
void gemm_contig(size_t m, size_t n, size_t k, T alpha, const T *a, size_t lda, const T *b,
                 size_t ldb, T beta, T *c, size_t ldc, int threads) {
  constexpr size_t NR = Tile<T>::NR;
  if (m == 0 || n == 0) return;
  apply_beta(m, n, beta, c, ldc);
  if (k == 0 || alpha == T(0)) return;
  const bool big_enough = n >= 2 * NR && (m * n * k) >= (size_t(1) << 18);
  if (threads > 1 && big_enough) {
    size_t target_tasks = std::max<size_t>(size_t(threads) * 4, 1);
    size_t chunk_cols = (n + target_tasks - 1) / target_tasks;
    chunk_cols = std::min(round_up(std::max(chunk_cols, NR), NR), NC);
    size_t nchunks = (n + chunk_cols - 1) / chunk_cols;
    std::atomic<size_t> next{0};
    auto worker = [&]() {
      for (;;) {
        size_t ci = next.fetch_add(1);
        if (ci >= nchunks) break;
        size_t j0 = ci * chunk_cols;
        size_t nc = std::min(chunk_cols, n - j0);
        gemm_band(m, nc, k, alpha, a, lda, b, ldb, j0, c + j0 * ldc, ldc);
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    return;
  }
  for (size_t jc = 0; jc < n; jc += NC) {
    size_t nc = std::min(NC, n - jc);
    gemm_band(m, nc, k, alpha, a, lda, b, ldb, jc, c + jc * ldc, ldc);
  }
}

// level3.rs:179-222
template <typename T>
void materialize(bool at, bool bt, size_t m, size_t n, size_t k, const T *a, size_t lda,
                 const T *b, size_t ldb, std::vector<T> &am, std::vector<T> &bm) {
  am.assign(m * k, T(0));
  bm.assign(k * n, T(0));
  for (size_t p = 0; p < k; ++p)
    for (size_t i = 0; i < m; ++i) am[i + p * m] = at ? a[p + i * lda] : a[i + p * lda];
  for (size_t j = 0; j < n; ++j)
    for (size_t p = 0; p < k; ++p) bm[p + j * k] = bt ? b[j + p * ldb] : b[p + j * ldb];
}

// level3.rs:143-175
template <typename T>
void gemm_colmajor(bool at, bool bt, size_t m, size_t n, size_t k, T alpha, const T *a,
                   size_t lda, const T *b, size_t ldb, T beta, T *c, size_t ldc, int threads) {
  if (!at && !bt) {
    gemm_contig(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, threads);
    return;
  }
  std::vector<T> am, bm;
  materialize(at, bt, m, n, k, a, lda, b, ldb, am, bm);
  gemm_contig(m, n, k, alpha, am.data(), m, bm.data(), k, beta, c, ldc, threads);
}

// level3.rs:32-62 (layout 0 = RowMajor, 1 = ColMajor; trans 1/2 transposed)
template <typename T>
void gemm(int layout, int ta, int tb, size_t m, size_t n, size_t k, T alpha, const T *a,
          size_t lda, const T *b, size_t ldb, T beta, T *c, size_t ldc, int threads) {
  bool at = ta != 0, bt = tb != 0;
  if (layout == 1)
    gemm_colmajor(at, bt, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, threads);
  else
    gemm_colmajor(bt, at, n, m, k, alpha, b, ldb, a, lda, beta, c, ldc, threads);
}

// level3.rs:236-294
template <typename T>
void syrk(int uplo, int trans, size_t n, size_t k, T alpha, const T *a, size_t lda, T beta, T *c,
          size_t ldc, int threads) {
  if (n == 0) return;
  int ta = trans != 0 ? 1 : 0, tb = trans != 0 ? 0 : 1;
  std::vector<T> p(n * n, T(0));
  gemm<T>(1, ta, tb, n, n, k, alpha, a, lda, a, lda, T(0), p.data(), n, threads);
  for (size_t j = 0; j < n; ++j) {
    size_t lo = uplo == 0 ? 0 : j, hi = uplo == 0 ? j + 1 : n;
    for (size_t i = lo; i < hi; ++i) {
      size_t cij = i + j * ldc;
      T scaled = beta == T(0) ? T(0) : beta * c[cij];
      c[cij] = p[i + j * n] + scaled;
    }
  }
}

// level3.rs:390-520 (neg_mul_add unfused: acc - a*x with two roundings)
template <typename T>
void trsm(int side, int uplo, int trans, int diag, size_t m, size_t n, T alpha, const T *a,
          size_t lda, T *b, size_t ldb) {
  if (alpha != T(1))
    for (size_t j = 0; j < n; ++j)
      for (size_t i = 0; i < m; ++i) b[i + j * ldb] = alpha * b[i + j * ldb];
  const bool tr = trans != 0;
  auto aop = [&](size_t r, size_t c) { return tr ? a[c + r * lda] : a[r + c * lda]; };
  const bool upper = (uplo == 0) != tr;
  const bool unit = diag == 1;
  auto finish = [&](T s, T aii) { return unit ? s : s / aii; };
  if (side == 0) {
    for (size_t j = 0; j < n; ++j) {
      T *col = b + j * ldb;
      if (upper) {
        for (size_t ii = m; ii-- > 0;) {
          T s = col[ii];
          for (size_t kk = ii + 1; kk < m; ++kk) {
            T prod = aop(ii, kk) * col[kk];
            s = s - prod;
          }
          col[ii] = finish(s, aop(ii, ii));
        }
      } else {
        for (size_t ii = 0; ii < m; ++ii) {
          T s = col[ii];
          for (size_t kk = 0; kk < ii; ++kk) {
            T prod = aop(ii, kk) * col[kk];
            s = s - prod;
          }
          col[ii] = finish(s, aop(ii, ii));
        }
      }
    }
  } else {
    for (size_t i = 0; i < m; ++i) {
      for (size_t t = 0; t < n; ++t) {
        size_t j = upper ? t : n - 1 - t;
        T s = b[i + j * ldb];
        if (upper) {
          for (size_t kk = 0; kk < j; ++kk) {
            T prod = b[i + kk * ldb] * aop(kk, j);
            s = s - prod;
          }
        } else {
          for (size_t kk = j + 1; kk < n; ++kk) {
            T prod = b[i + kk * ldb] * aop(kk, j);
            s = s - prod;
          }
        }
        b[i + j * ldb] = finish(s, aop(j, j));
      }
    }
  }
}

// complex.rs:101-111 — (a+bi)(c+di) with rblas's fma pattern.
template <typename T>
void cmul(T ar, T ai, T br, T bi, T &re, T &im) {
  T t = ai * bi;
  re = std::fma(ar, br, T(0) - t);
  T u = ai * br;
  im = std::fma(ar, bi, u);
}

// complex3.rs:252-349 — complex trsm by substitution: alpha pre-scale when
// alpha != 1 (complex multiply), s = s - aop*x per component, finish =
// cdiv(s, a_ii) = s * conj(a_ii) / norm_sqr(a_ii) with norm_sqr =
// fma(re, re, im*im) (complex.rs:68-70, complex3.rs:344-349); ConjTrans
// conjugates op(A) (conj: im = 0 - im). Interleaved {re, im} storage.
template <typename T>
void ctrsm(int side, int uplo, int trans, int diag, size_t m, size_t n, T alr, T ali, const T *a,
           size_t lda, T *b, size_t ldb) {
  if (!(alr == T(1) && ali == T(0)))
    for (size_t j = 0; j < n; ++j)
      for (size_t i = 0; i < m; ++i) {
        T *x = b + 2 * (i + j * ldb);
        T re, im;
        cmul(alr, ali, x[0], x[1], re, im);
        x[0] = re;
        x[1] = im;
      }
  const bool tr = trans != 0, cj = trans == 2;
  auto aop = [&](size_t r, size_t c, T &re, T &im) {
    const size_t idx = tr ? c + r * lda : r + c * lda;
    re = a[2 * idx];
    im = cj ? T(0) - a[2 * idx + 1] : a[2 * idx + 1];
  };
  const bool upper = (uplo == 0) != tr;
  const bool unit = diag == 1;
  auto finish = [&](T sr, T si, T ar, T ai, T *x) {
    if (unit) {
      x[0] = sr;
      x[1] = si;
      return;
    }
    const T d = std::fma(ar, ar, ai * ai);
    T nr, ni;
    cmul(sr, si, ar, T(0) - ai, nr, ni);
    x[0] = nr / d;
    x[1] = ni / d;
  };
  auto step = [&](T &sr, T &si, T pr_a, T pi_a, T pr_b, T pi_b) {
    T re, im;
    cmul(pr_a, pi_a, pr_b, pi_b, re, im);
    sr = sr - re;
    si = si - im;
  };
  if (side == 0) {
    for (size_t j = 0; j < n; ++j) {
      T *col = b + 2 * j * ldb;
      for (size_t t = 0; t < m; ++t) {
        const size_t ii = upper ? m - 1 - t : t;
        T sr = col[2 * ii], si = col[2 * ii + 1];
        const size_t k0 = upper ? ii + 1 : 0, k1 = upper ? m : ii;
        for (size_t kk = k0; kk < k1; ++kk) {
          T ar, ai;
          aop(ii, kk, ar, ai);
          step(sr, si, ar, ai, col[2 * kk], col[2 * kk + 1]);
        }
        T dr, di;
        aop(ii, ii, dr, di);
        finish(sr, si, dr, di, col + 2 * ii);
      }
    }
  } else {
    for (size_t i = 0; i < m; ++i) {
      for (size_t t = 0; t < n; ++t) {
        const size_t j = upper ? t : n - 1 - t;
        T *x = b + 2 * (i + j * ldb);
        T sr = x[0], si = x[1];
        const size_t k0 = upper ? 0 : j + 1, k1 = upper ? j : n;
        for (size_t kk = k0; kk < k1; ++kk) {
          T ar, ai;
          aop(kk, j, ar, ai);
          const T *y = b + 2 * (i + kk * ldb);
          step(sr, si, y[0], y[1], ar, ai);
        }
        T dr, di;
        aop(j, j, dr, di);
        finish(sr, si, dr, di, x);
      }
    }
  }
}

// level2.rs:11-61 — gemv. (rs, cs) = layout.strides(lda): RowMajor (lda, 1),
// ColMajor (1, lda). y := beta y unless beta == 1 (0 stores 0); nothing more
// when alpha == 0; NoTrans: per row an fma chain over j, then
// y_i = fma(alpha, acc, y_i); Trans: xi = alpha * x_i, then
// y_j = fma(a_ij, xi, y_j) for i outer, j inner.
template <typename T>
void gemv(int layout, int trans, size_t m, size_t n, T alpha, const T *a, size_t lda, const T *x,
          size_t incx, T beta, T *y, size_t incy) {
  const size_t rs = layout == 0 ? lda : 1, cs = layout == 0 ? 1 : lda;
  const bool tr = trans != 0;
  const size_t leny = tr ? n : m;
  if (beta != T(1))  // level2.rs:91-104
    for (size_t t = 0; t < leny; ++t) y[t * incy] = beta == T(0) ? T(0) : beta * y[t * incy];
  if (alpha == T(0)) return;
  if (!tr) {
    for (size_t i = 0; i < m; ++i) {
      T acc = T(0);
      for (size_t j = 0; j < n; ++j) acc = std::fma(a[i * rs + j * cs], x[j * incx], acc);
      y[i * incy] = std::fma(alpha, acc, y[i * incy]);
    }
  } else {
    for (size_t i = 0; i < m; ++i) {
      const T xi = alpha * x[i * incx];
      for (size_t j = 0; j < n; ++j) y[j * incy] = std::fma(a[i * rs + j * cs], xi, y[j * incy]);
    }
  }
}

// level2.rs:65-88 — ger: nothing when alpha == 0; a_ij = fma(alpha * x_i,
// y_j, a_ij), i outer, j inner.
template <typename T>
void ger(int layout, size_t m, size_t n, T alpha, const T *x, size_t incx, const T *y,
         size_t incy, T *a, size_t lda) {
  if (alpha == T(0)) return;
  const size_t rs = layout == 0 ? lda : 1, cs = layout == 0 ? 1 : lda;
  for (size_t i = 0; i < m; ++i) {
    const T axi = alpha * x[i * incx];
    for (size_t j = 0; j < n; ++j) a[i * rs + j * cs] = std::fma(axi, y[j * incy], a[i * rs + j * cs]);
  }
}

// complex2.rs:11-76 — complex gemv. y := beta y always (its skip condition
// never holds; beta == 0 stores 0); nothing more when alpha == 0; per output
// acc = sum of aval * x (complex multiply, then add), y = y + alpha * acc.
// NoTrans: rows i over j; Trans / ConjTrans: columns j over i, ConjTrans
// conjugating A.
template <typename T>
void cgemv(int layout, int trans, size_t m, size_t n, T alr, T ali, const T *a, size_t lda,
           const T *x, size_t incx, T ber, T bei, T *y, size_t incy) {
  const size_t rs = layout == 0 ? lda : 1, cs = layout == 0 ? 1 : lda;
  const bool tr = trans != 0, conj = trans == 2;
  const size_t leny = tr ? n : m;
  const bool beta_zero = ber == T(0) && bei == T(0);
  for (size_t t = 0; t < leny; ++t) {
    T *yy = y + 2 * t * incy;
    if (beta_zero) {
      yy[0] = T(0);
      yy[1] = T(0);
    } else {
      T re, im;
      cmul(ber, bei, yy[0], yy[1], re, im);
      yy[0] = re;
      yy[1] = im;
    }
  }
  if (alr == T(0) && ali == T(0)) return;
  auto term = [&](size_t i, size_t j, size_t xi, T &sr, T &si) {
    const T *av = a + 2 * (i * rs + j * cs);
    const T ar = av[0], ai = conj ? T(0) - av[1] : av[1];
    T re, im;
    cmul(ar, ai, x[2 * xi], x[2 * xi + 1], re, im);
    sr = sr + re;
    si = si + im;
  };
  auto finish = [&](size_t o, T sr, T si) {
    T *yy = y + 2 * o * incy;
    T re, im;
    cmul(alr, ali, sr, si, re, im);
    yy[0] = yy[0] + re;
    yy[1] = yy[1] + im;
  };
  if (!tr) {
    for (size_t i = 0; i < m; ++i) {
      T sr = T(0), si = T(0);
      for (size_t j = 0; j < n; ++j) term(i, j, j * incx, sr, si);
      finish(i, sr, si);
    }
  } else {
    for (size_t j = 0; j < n; ++j) {
      T sr = T(0), si = T(0);
      for (size_t i = 0; i < m; ++i) term(i, j, i * incx, sr, si);
      finish(j, sr, si);
    }
  }
}

// complex2.rs:112-142 — geru / gerc: nothing when alpha == 0;
// axi = alpha * x_i, a_ij = a_ij + axi * y_j (y_j conjugated for gerc).
template <typename T>
void crank1(int layout, size_t m, size_t n, T alr, T ali, const T *x, size_t incx, const T *y,
            size_t incy, T *a, size_t lda, bool conj) {
  if (alr == T(0) && ali == T(0)) return;
  const size_t rs = layout == 0 ? lda : 1, cs = layout == 0 ? 1 : lda;
  for (size_t i = 0; i < m; ++i) {
    T xr, xi;
    cmul(alr, ali, x[2 * i * incx], x[2 * i * incx + 1], xr, xi);
    for (size_t j = 0; j < n; ++j) {
      const T yr = y[2 * j * incy], yi = conj ? T(0) - y[2 * j * incy + 1] : y[2 * j * incy + 1];
      T re, im;
      cmul(xr, xi, yr, yi, re, im);
      T *av = a + 2 * (i * rs + j * cs);
      av[0] = av[0] + re;
      av[1] = av[1] + im;
    }
  }
}

// complex3.rs:27-125 — 4M complex gemm on split planes (column-major).
// Interleaved {re, im} storage; trans 2 = ConjTrans conjugates.
template <typename T>
void cgemm(int ta, int tb, size_t m, size_t n, size_t k, T alr, T ali, const T *a, size_t lda,
           const T *b, size_t ldb, T ber, T bei, T *c, size_t ldc, int threads) {
  auto split = [](int trans, size_t rows, size_t cols, const T *x, size_t ld, std::vector<T> &re,
                  std::vector<T> &im) {
    re.assign(rows * cols, T(0));
    im.assign(rows * cols, T(0));
    const bool tr = trans != 0, conj = trans == 2;
    for (size_t j = 0; j < cols; ++j)
      for (size_t i = 0; i < rows; ++i) {
        size_t idx = tr ? (j + i * ld) : (i + j * ld);
        T vr = x[2 * idx], vi = x[2 * idx + 1];
        if (conj) vi = T(0) - vi;
        re[i + j * rows] = vr;
        im[i + j * rows] = vi;
      }
  };
  std::vector<T> ar, ai, br, bi;
  split(ta, m, k, a, lda, ar, ai);
  split(tb, k, n, b, ldb, br, bi);
  std::vector<T> pr(m * n, T(0)), pi(m * n, T(0));
  auto rg = [&](T alpha, const std::vector<T> &x, const std::vector<T> &y, T beta,
                std::vector<T> &z) {
    gemm<T>(1, 0, 0, m, n, k, alpha, x.data(), m, y.data(), k, beta, z.data(), m, threads);
  };
  rg(T(1), ar, br, T(0), pr);
  rg(T(0) - T(1), ai, bi, T(1), pr);
  rg(T(1), ar, bi, T(0), pi);
  rg(T(1), ai, br, T(1), pi);
  const bool beta_zero = ber == T(0) && bei == T(0);
  for (size_t j = 0; j < n; ++j)
    for (size_t i = 0; i < m; ++i) {
      size_t cij = i + j * ldc;
      T sr = 0, si = 0;
      if (!beta_zero) cmul(ber, bei, c[2 * cij], c[2 * cij + 1], sr, si);
      T qr, qi;
      cmul(alr, ali, pr[i + j * m], pi[i + j * m], qr, qi);
      c[2 * cij] = qr + sr;
      c[2 * cij + 1] = qi + si;
    }
}

// avx2.rs:30-56 / :105-131 and the scalar tail; level1.rs:45-66.
__attribute__((target("avx2,fma"))) float dot_avx2(size_t n, const float *x, const float *y) {
  __m256 acc0 = _mm256_setzero_ps(), acc1 = _mm256_setzero_ps();
  size_t i = 0;
  for (; i + 16 <= n; i += 16) {
    acc0 = _mm256_fmadd_ps(_mm256_loadu_ps(x + i), _mm256_loadu_ps(y + i), acc0);
    acc1 = _mm256_fmadd_ps(_mm256_loadu_ps(x + i + 8), _mm256_loadu_ps(y + i + 8), acc1);
  }
  for (; i + 8 <= n; i += 8)
    acc0 = _mm256_fmadd_ps(_mm256_loadu_ps(x + i), _mm256_loadu_ps(y + i), acc0);
  __m256 v = _mm256_add_ps(acc0, acc1);
  __m128 s = _mm_add_ps(_mm256_castps256_ps128(v), _mm256_extractf128_ps(v, 1));
  s = _mm_add_ps(s, _mm_movehl_ps(s, s));
  s = _mm_add_ss(s, _mm_shuffle_ps(s, s, 0x55));
  float acc = _mm_cvtss_f32(s);
  for (; i < n; ++i) acc = std::fma(x[i], y[i], acc);
  return acc;
}
__attribute__((target("avx2,fma"))) double dot_avx2(size_t n, const double *x, const double *y) {
  __m256d acc0 = _mm256_setzero_pd(), acc1 = _mm256_setzero_pd();
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    acc0 = _mm256_fmadd_pd(_mm256_loadu_pd(x + i), _mm256_loadu_pd(y + i), acc0);
    acc1 = _mm256_fmadd_pd(_mm256_loadu_pd(x + i + 4), _mm256_loadu_pd(y + i + 4), acc1);
  }
  for (; i + 4 <= n; i += 4)
    acc0 = _mm256_fmadd_pd(_mm256_loadu_pd(x + i), _mm256_loadu_pd(y + i), acc0);
  __m256d v = _mm256_add_pd(acc0, acc1);
  __m128d s = _mm_add_pd(_mm256_castpd256_pd128(v), _mm256_extractf128_pd(v, 1));
  s = _mm_add_sd(s, _mm_unpackhi_pd(s, s));
  double acc = _mm_cvtsd_f64(s);
  for (; i < n; ++i) acc = std::fma(x[i], y[i], acc);
  return acc;
}

// scalar.rs:15-34 (no AVX2)
template <typename T>
T dot_scalar(size_t n, const T *x, const T *y) {
  T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  size_t chunks = n / 4;
  for (size_t q = 0; q < chunks; ++q) {
    size_t i = q * 4;
    a0 = std::fma(x[i], y[i], a0);
    a1 = std::fma(x[i + 1], y[i + 1], a1);
    a2 = std::fma(x[i + 2], y[i + 2], a2);
    a3 = std::fma(x[i + 3], y[i + 3], a3);
  }
  T acc = (a0 + a1) + (a2 + a3);
  for (size_t i = chunks * 4; i < n; ++i) acc = std::fma(x[i], y[i], acc);
  return acc;
}

template <typename T>
T dot(size_t n, const T *x, size_t incx, const T *y, size_t incy) {
  if (n == 0) return T(0);
  if (incx == 1 && incy == 1) return has_avx2_fma() ? dot_avx2(n, x, y) : dot_scalar(n, x, y);
  T acc = 0;
  for (size_t t = 0, ix = 0, iy = 0; t < n; ++t, ix += incx, iy += incy)
    acc = std::fma(x[ix], y[iy], acc);
  return acc;
}

// level1.rs:17-41 (every path is y = fma(alpha, x, y) per element)
template <typename T>
void axpy(size_t n, T alpha, const T *x, size_t incx, T *y, size_t incy) {
  if (n == 0 || alpha == T(0)) return;
  for (size_t t = 0, ix = 0, iy = 0; t < n; ++t, ix += incx, iy += incy)
    y[iy] = std::fma(alpha, x[ix], y[iy]);
}

// level1.rs:96-123
template <typename T>
T nrm2(size_t n, const T *x, size_t incx) {
  if (n == 0) return T(0);
  T scale = 0, ssq = 1;
  for (size_t t = 0, ix = 0; t < n; ++t, ix += incx) {
    T xi = x[ix];
    if (xi != T(0)) {
      T a = std::fabs(xi);
      if (scale < a) {
        T r = scale / a;
        ssq = T(1) + ssq * (r * r);
        scale = a;
      } else {
        T r = a / scale;
        ssq = ssq + r * r;
      }
    }
  }
  return scale * std::sqrt(ssq);
}

}  // namespace

extern "C" {

int oracle_has_avx2(void) { return has_avx2_fma() ? 1 : 0; }

void oracle_sgemm(int layout, int ta, int tb, size_t m, size_t n, size_t k, float alpha,
                  const float *a, size_t lda, const float *b, size_t ldb, float beta, float *c,
                  size_t ldc, int threads) {
  gemm<float>(layout, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, threads);
}
void oracle_dgemm(int layout, int ta, int tb, size_t m, size_t n, size_t k, double alpha,
                  const double *a, size_t lda, const double *b, size_t ldb, double beta,
                  double *c, size_t ldc, int threads) {
  gemm<double>(layout, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, threads);
}
void oracle_ssyrk(int uplo, int trans, size_t n, size_t k, float alpha, const float *a,
                  size_t lda, float beta, float *c, size_t ldc, int threads) {
  syrk<float>(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, threads);
}
void oracle_dsyrk(int uplo, int trans, size_t n, size_t k, double alpha, const double *a,
                  size_t lda, double beta, double *c, size_t ldc, int threads) {
  syrk<double>(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, threads);
}
void oracle_strsm(int side, int uplo, int trans, int diag, size_t m, size_t n, float alpha,
                  const float *a, size_t lda, float *b, size_t ldb) {
  trsm<float>(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb);
}
void oracle_dtrsm(int side, int uplo, int trans, int diag, size_t m, size_t n, double alpha,
                  const double *a, size_t lda, double *b, size_t ldb) {
  trsm<double>(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb);
}
void oracle_cgemm(int ta, int tb, size_t m, size_t n, size_t k, float alr, float ali,
                  const float *a, size_t lda, const float *b, size_t ldb, float ber, float bei,
                  float *c, size_t ldc, int threads) {
  cgemm<float>(ta, tb, m, n, k, alr, ali, a, lda, b, ldb, ber, bei, c, ldc, threads);
}
void oracle_zgemm(int ta, int tb, size_t m, size_t n, size_t k, double alr, double ali,
                  const double *a, size_t lda, const double *b, size_t ldb, double ber,
                  double bei, double *c, size_t ldc, int threads) {
  cgemm<double>(ta, tb, m, n, k, alr, ali, a, lda, b, ldb, ber, bei, c, ldc, threads);
}
void oracle_sgemv(int layout, int trans, size_t m, size_t n, float alpha, const float *a,
                  size_t lda, const float *x, size_t incx, float beta, float *y, size_t incy) {
  gemv<float>(layout, trans, m, n, alpha, a, lda, x, incx, beta, y, incy);
}
void oracle_dgemv(int layout, int trans, size_t m, size_t n, double alpha, const double *a,
                  size_t lda, const double *x, size_t incx, double beta, double *y, size_t incy) {
  gemv<double>(layout, trans, m, n, alpha, a, lda, x, incx, beta, y, incy);
}
void oracle_sger(int layout, size_t m, size_t n, float alpha, const float *x, size_t incx,
                 const float *y, size_t incy, float *a, size_t lda) {
  ger<float>(layout, m, n, alpha, x, incx, y, incy, a, lda);
}
void oracle_dger(int layout, size_t m, size_t n, double alpha, const double *x, size_t incx,
                 const double *y, size_t incy, double *a, size_t lda) {
  ger<double>(layout, m, n, alpha, x, incx, y, incy, a, lda);
}
void oracle_cgemv(int layout, int trans, size_t m, size_t n, float alr, float ali, const float *a,
                  size_t lda, const float *x, size_t incx, float ber, float bei, float *y,
                  size_t incy) {
  cgemv<float>(layout, trans, m, n, alr, ali, a, lda, x, incx, ber, bei, y, incy);
}
void oracle_zgemv(int layout, int trans, size_t m, size_t n, double alr, double ali,
                  const double *a, size_t lda, const double *x, size_t incx, double ber,
                  double bei, double *y, size_t incy) {
  cgemv<double>(layout, trans, m, n, alr, ali, a, lda, x, incx, ber, bei, y, incy);
}
void oracle_crank1(int layout, size_t m, size_t n, float alr, float ali, const float *x,
                   size_t incx, const float *y, size_t incy, float *a, size_t lda, int conj) {
  crank1<float>(layout, m, n, alr, ali, x, incx, y, incy, a, lda, conj != 0);
}
void oracle_zrank1(int layout, size_t m, size_t n, double alr, double ali, const double *x,
                   size_t incx, const double *y, size_t incy, double *a, size_t lda, int conj) {
  crank1<double>(layout, m, n, alr, ali, x, incx, y, incy, a, lda, conj != 0);
}
void oracle_ctrsm(int side, int uplo, int trans, int diag, size_t m, size_t n, float alr,
                  float ali, const float *a, size_t lda, float *b, size_t ldb) {
  ctrsm<float>(side, uplo, trans, diag, m, n, alr, ali, a, lda, b, ldb);
}
void oracle_ztrsm(int side, int uplo, int trans, int diag, size_t m, size_t n, double alr,
                  double ali, const double *a, size_t lda, double *b, size_t ldb) {
  ctrsm<double>(side, uplo, trans, diag, m, n, alr, ali, a, lda, b, ldb);
}
float oracle_sdot(size_t n, const float *x, size_t incx, const float *y, size_t incy) {
  return dot<float>(n, x, incx, y, incy);
}
double oracle_ddot(size_t n, const double *x, size_t incx, const double *y, size_t incy) {
  return dot<double>(n, x, incx, y, incy);
}
void oracle_saxpy(size_t n, float alpha, const float *x, size_t incx, float *y, size_t incy) {
  axpy<float>(n, alpha, x, incx, y, incy);
}
void oracle_daxpy(size_t n, double alpha, const double *x, size_t incx, double *y, size_t incy) {
  axpy<double>(n, alpha, x, incx, y, incy);
}
float oracle_snrm2(size_t n, const float *x, size_t incx) { return nrm2<float>(n, x, incx); }
double oracle_dnrm2(size_t n, const double *x, size_t incx) { return nrm2<double>(n, x, incx); }

}  // extern "C"
