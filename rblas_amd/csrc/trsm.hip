// Blocked triangular solve on gfx950 (level3::trsm, level3.rs:390-520):
// op(A) X = alpha B (Left) or X op(A) = alpha B (Right), in place in B.
//
// rblas substitutes element by element. Here, for an order q = m (Left) or
// n (Right) up to 2048, the solve is the LDS-resident one of grouped_trsm.hip
// (one launch inverting the diagonal blocks, one launch of one workgroup per
// 16 or 8 right-hand sides, which stay in LDS through the whole
// substitution). For larger orders:
//   1. B := alpha B when alpha != 1 (level3.rs:403-409; every entry, so a
//      zero alpha keeps NaN).
//   2. launch_trsm_invert inverts every 64 x 64 diagonal block of op(A) at once,
//      in f64 for f32 and f64 operands alike (Unit: the stored diagonal is
//      not read).
//   3. The diagonal range is solved by recursive halving: solve the half
//      that comes first in substitution order, subtract its contribution from
//      the other half with one MFMA gemm (k = the first half's length, so the
//      bulk of the flops runs at large k), solve the other half. A 64-row
//      leaf is solved by trsm_apply_* in place: X_b = inv_b * B_b (Left) or
//      B_b * inv_b (Right) in f64, summing only over the triangle, so a
//      non-finite B entry reaches exactly the entries that depend on it in
//      rblas's substitution (no 0 * NaN from structural zeros).
#include "grouped.hpp"

namespace rblas {
namespace {

constexpr int kNB = 64;
constexpr int64_t kDirectMaxOrder = 2048;  // q * 8 fits the LDS block

// op(A)(r, c) of the stored column-major A.
template <typename T>
__device__ __forceinline__ T op_a(const T *a, int64_t lda, bool trans, int64_t r, int64_t c) {
  return trans ? a[c + r * lda] : a[r + c * lda];
}

// Left leaf: B[r0 + i, col] := sum_{j in tri(i)} inv[i][j] * B[r0 + j, col]
// for i, j < nb and the kLeafCols columns of this workgroup. 256 threads:
// lane row i = t % 64 (coalesced on column-major B), columns t / 64 + 4 r.
// 32 columns per workgroup put 256 workgroups on a 64 x 8192 leaf (64 columns
// gave 128, half the CUs, and a 16-column FMA chain per lane; 16 columns
// measured 1% slower in the leaf pair, whose workgroups each re-read two
// inverses and a block of op(A)).
constexpr int kLeafCols = 32;
template <typename T>
__global__ __launch_bounds__(256) void trsm_apply_left(const double *inv, bool upper, int nb,
                                                       int64_t n, T *b, int64_t ldb) {
  constexpr int R = kLeafCols / 4;  // columns per lane
  __shared__ double invT[kNB][kNB];     // invT[j][i] = inv(i, j)
  __shared__ double Bs[kLeafCols][kNB];  // Bs[c][j] = B(r0 + j, col0 + c)
  const int t = threadIdx.x, i = t % kNB, c0 = t / kNB;
  const int64_t col0 = (int64_t)blockIdx.x * kLeafCols;
  const int ncol = (int)((n - col0) < kLeafCols ? (n - col0) : kLeafCols);
  // Every load issued before any is used, without branches (an entry outside
  // the block reads b[0] and is replaced by 0).
  double iv[16];
  T raw[R];
#pragma unroll
  for (int r = 0; r < 16; ++r) iv[r] = inv[t + 256 * r];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = c0 + 4 * r;
    const bool ok = c < ncol && i < nb;
    raw[r] = *(ok ? b + i + (col0 + c) * ldb : b);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int idx = t + 256 * r;
    invT[idx / kNB][idx % kNB] = iv[r];
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = c0 + 4 * r;
    Bs[c][i] = (c < ncol && i < nb) ? (double)raw[r] : 0.0;
  }
  __syncthreads();
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  // Row i sums over its triangle only: j = 0..i (lower) or nb-1 down to i
  // (upper). j is the same in every lane still running, so Bs reads
  // broadcast; lanes leave the loop when their triangle ends.
  const int cnt = i < nb ? (upper ? nb - i : i + 1) : 0;
  for (int s = 0; s < cnt; ++s) {
    const int j = upper ? nb - 1 - s : s;
    const double w = invT[j][i];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = __builtin_fma(w, Bs[c0 + 4 * r][j], acc[r]);
  }
  if (i < nb) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int c = c0 + 4 * r;
      if (c < ncol) b[i + (col0 + c) * ldb] = (T)acc[r];
    }
  }
}

// Left leaf pair: the recursion's last level in one launch. A diagonal
// range of 65..128 rows splits as the solver splits it, first f_len rows at
// f_off then s_len rows at s_off (one of them a whole 64-row block); per
// kLeafCols columns: X_f = inv_f B_f (triangle only, as trsm_apply_left);
// B_s -= op(A)[s, f] X_f (a_sf points at op(A)(s_lo, f_lo)); X_s = inv_s B_s.
// It replaces a leaf, a k = 64 gemm update of 64 rows and a second leaf.
template <typename T>
__global__ __launch_bounds__(256) void trsm_apply_left_pair(const double *inv_f,
                                                            const double *inv_s, const T *a_sf,
                                                            int64_t lda, bool trans, bool upper,
                                                            int f_off, int f_len, int s_off,
                                                            int s_len, int64_t n, T *b,
                                                            int64_t ldb) {
  constexpr int R = kLeafCols / 4;  // columns per lane
  __shared__ double W[kNB][kNB];              // W[j][i]: inv_f(i, j), op(A)(s+i, f+j), inv_s(i, j)
  __shared__ double Bs[kLeafCols][2 * kNB];   // Bs[c][r] = B(lo + r, col0 + c)
  const int t = threadIdx.x, i = t % kNB, c0 = t / kNB;
  const int len = f_len + s_len;
  const int64_t col0 = (int64_t)blockIdx.x * kLeafCols;
  const int ncol = (int)((n - col0) < kLeafCols ? (n - col0) : kLeafCols);
  // Every global load issued before any is used; an entry outside its block
  // reads element 0 and is replaced by 0.
  double ivf[16], ivs[16];
  T av[16], raw[2][R];
#pragma unroll
  for (int r = 0; r < 16; ++r) ivf[r] = inv_f[t + 256 * r];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = h * kNB + i, c = c0 + 4 * r;
      raw[h][r] = *(row < len && c < ncol ? b + row + (col0 + c) * ldb : b);
    }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int idx = t + 256 * r, ai = idx % kNB, aj = idx / kNB;
    const bool ok = ai < s_len && aj < f_len;
    av[r] = *(ok ? (trans ? a_sf + aj + ai * lda : a_sf + ai + aj * lda) : a_sf);
    ivs[r] = inv_s[t + 256 * r];
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int idx = t + 256 * r;
    W[idx / kNB][idx % kNB] = ivf[r];
  }
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = h * kNB + i, c = c0 + 4 * r;
      Bs[c][row] = (row < len && c < ncol) ? (double)raw[h][r] : 0.0;
    }
  __syncthreads();
  double acc[R];
  // X_f: row i of the f block sums over its triangle (as trsm_apply_left).
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  int cnt = i < f_len ? (upper ? f_len - i : i + 1) : 0;
  for (int s = 0; s < cnt; ++s) {
    const int j = upper ? f_len - 1 - s : s;
    const double w = W[j][i];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = __builtin_fma(w, Bs[c0 + 4 * r][f_off + j], acc[r]);
  }
  __syncthreads();  // every lane has read W and the f rows
  if (i < f_len) {
#pragma unroll
    for (int r = 0; r < R; ++r) Bs[c0 + 4 * r][f_off + i] = acc[r];
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int idx = t + 256 * r, ai = idx % kNB, aj = idx / kNB;
    W[aj][ai] = (ai < s_len && aj < f_len) ? (double)av[r] : 0.0;
  }
  __syncthreads();
  // B_s -= op(A)[s, f] X_f: lanes read only f rows and write only their own
  // s row.
  if (i < s_len) {
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    for (int j = 0; j < f_len; ++j) {
      const double w = W[j][i];
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] = __builtin_fma(w, Bs[c0 + 4 * r][f_off + j], acc[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) Bs[c0 + 4 * r][s_off + i] -= acc[r];
  }
  __syncthreads();  // every lane has read W; the s rows are updated
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int idx = t + 256 * r;
    W[idx / kNB][idx % kNB] = ivs[r];
  }
  __syncthreads();
  // X_s over the s block's triangle.
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  cnt = i < s_len ? (upper ? s_len - i : i + 1) : 0;
  for (int s = 0; s < cnt; ++s) {
    const int j = upper ? s_len - 1 - s : s;
    const double w = W[j][i];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = __builtin_fma(w, Bs[c0 + 4 * r][s_off + j], acc[r]);
  }
  if (i < s_len) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int c = c0 + 4 * r;
      if (c < ncol) b[s_off + i + (col0 + c) * ldb] = (T)acc[r];
    }
  }
  if (i < f_len) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int c = c0 + 4 * r;
      if (c < ncol) b[f_off + i + (col0 + c) * ldb] = (T)Bs[c][f_off + i];
    }
  }
}

// Right leaf: B[row, r0 + j] := sum_{p in tri(j)} B[row, r0 + p] * inv[p][j]
// for the kLeafRows rows of this workgroup. Loads and stores use lane = row
// t % kLeafRows (runs of 32 rows of column-major B); the sums use lane =
// column j = t % 64 and rows t / 64 + 4 r, so each lane's triangle (p = 0..j
// upper, nb-1 down to j lower) is its own loop trip count, p is the same in
// every lane still running and the row reads broadcast. 32 rows per
// workgroup put 256 workgroups on an 8192 x 64 leaf (64 rows gave 128, half
// the CUs; 16 measured 1.7% slower in the leaf pair, whose workgroups each
// re-read two inverses and a block of op(A)).
constexpr int kLeafRows = 32;
template <typename T>
__global__ __launch_bounds__(256) void trsm_apply_right(const double *inv, bool upper, int nb,
                                                        int64_t m, T *b, int64_t ldb) {
  constexpr int R = kLeafRows / 4;       // rows per lane in the sums
  constexpr int L = kNB * kLeafRows / 256;  // B entries per lane in loads / stores
  __shared__ double inv_t[kNB][kNB + 1];        // inv_t[p][j] = inv(p, j)
  __shared__ double Bs[kNB][kLeafRows + 1];     // Bs[p][i] = B(row0 + i, r0 + p)
  const int t = threadIdx.x, i = t % kLeafRows, c0 = t / kLeafRows;
  const int64_t row0 = (int64_t)blockIdx.x * kLeafRows;
  const bool row_ok = row0 + i < m;
  double iv[16];
  T raw[L];
#pragma unroll
  for (int r = 0; r < 16; ++r) iv[r] = inv[t + 256 * r];
#pragma unroll
  for (int r = 0; r < L; ++r) {
    const int p = c0 + (256 / kLeafRows) * r;
    const bool ok = row_ok && p < nb;
    raw[r] = *(ok ? b + row0 + i + p * ldb : b);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int idx = t + 256 * r;
    inv_t[idx % kNB][idx / kNB] = iv[r];
  }
#pragma unroll
  for (int r = 0; r < L; ++r) {
    const int p = c0 + (256 / kLeafRows) * r;
    Bs[p][i] = (row_ok && p < nb) ? (double)raw[r] : 0.0;
  }
  __syncthreads();
  const int j = t % kNB, rw = t / kNB;  // column j; rows rw + 4 r
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  const int cnt = j < nb ? (upper ? j + 1 : nb - j) : 0;
  for (int s = 0; s < cnt; ++s) {
    const int p = upper ? s : nb - 1 - s;
    const double w = inv_t[p][j];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = __builtin_fma(Bs[p][rw + 4 * r], w, acc[r]);
  }
  __syncthreads();  // every lane has read Bs
#pragma unroll
  for (int r = 0; r < R; ++r) Bs[j][rw + 4 * r] = acc[r];
  __syncthreads();
  if (row_ok) {
#pragma unroll
    for (int r = 0; r < L; ++r) {
      const int p = c0 + (256 / kLeafRows) * r;
      if (p < nb) b[row0 + i + p * ldb] = (T)Bs[p][i];
    }
  }
}

// Right leaf pair: the recursion's last level on the Right in one launch.
// The diagonal range of 65..128 columns splits as the solver splits it
// (first f_len columns at f_off, then s_len at s_off); per kLeafRows rows of
// B: X_f = B_f inv_f (as trsm_apply_right), B_s -= X_f op(A)[f, s] (a_fs
// points at op(A)(f_lo, s_lo)), X_s = B_s inv_s.
template <typename T>
__global__ __launch_bounds__(256) void trsm_apply_right_pair(const double *inv_f,
                                                             const double *inv_s, const T *a_fs,
                                                             int64_t lda, bool trans, bool upper,
                                                             int f_off, int f_len, int s_off,
                                                             int s_len, int64_t m, T *b,
                                                             int64_t ldb) {
  constexpr int R = kLeafRows / 4;               // rows per lane in the sums
  constexpr int L = 2 * kNB * kLeafRows / 256;   // B entries per lane in loads / stores
  __shared__ double W[kNB][kNB + 1];             // W[p][j]: inv_f(p, j), op(A)(f+p, s+j), inv_s(p, j)
  __shared__ double Bs[2 * kNB][kLeafRows + 1];  // Bs[p][i] = B(row0 + i, lo + p)
  const int t = threadIdx.x, i = t % kLeafRows, c0 = t / kLeafRows;
  const int len = f_len + s_len;
  const int64_t row0 = (int64_t)blockIdx.x * kLeafRows;
  const bool row_ok = row0 + i < m;
  // Every global load issued before any is used; an entry outside its block
  // reads element 0 and is replaced by 0.
  double ivf[16], ivs[16];
  T av[16], raw[L];
#pragma unroll
  for (int r = 0; r < 16; ++r) ivf[r] = inv_f[t + 256 * r];
#pragma unroll
  for (int r = 0; r < L; ++r) {
    const int p = c0 + (256 / kLeafRows) * r;
    raw[r] = *(row_ok && p < len ? b + row0 + i + p * ldb : b);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int idx = t + 256 * r, ap = idx % kNB, aj = idx / kNB;
    const bool ok = ap < f_len && aj < s_len;
    av[r] = *(ok ? (trans ? a_fs + aj + ap * lda : a_fs + ap + aj * lda) : a_fs);
    ivs[r] = inv_s[t + 256 * r];
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int idx = t + 256 * r;
    W[idx % kNB][idx / kNB] = ivf[r];
  }
#pragma unroll
  for (int r = 0; r < L; ++r) {
    const int p = c0 + (256 / kLeafRows) * r;
    Bs[p][i] = (row_ok && p < len) ? (double)raw[r] : 0.0;
  }
  __syncthreads();
  const int j = t % kNB, rw = t / kNB;  // column j; rows rw + 4 r
  double acc[R];
  // X_f: column j of the f block sums over its triangle (as trsm_apply_right).
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  int cnt = j < f_len ? (upper ? j + 1 : f_len - j) : 0;
  for (int s = 0; s < cnt; ++s) {
    const int q = upper ? s : f_len - 1 - s;
    const double w = W[q][j];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = __builtin_fma(Bs[f_off + q][rw + 4 * r], w, acc[r]);
  }
  __syncthreads();  // every lane has read W and the f columns
  if (j < f_len) {
#pragma unroll
    for (int r = 0; r < R; ++r) Bs[f_off + j][rw + 4 * r] = acc[r];
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int idx = t + 256 * r, ap = idx % kNB, aj = idx / kNB;
    W[ap][aj] = (ap < f_len && aj < s_len) ? (double)av[r] : 0.0;
  }
  __syncthreads();
  // B_s -= X_f op(A)[f, s]: lanes read only f columns and write only their
  // own entries of s column j.
  if (j < s_len) {
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    for (int q = 0; q < f_len; ++q) {
      const double w = W[q][j];
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] = __builtin_fma(Bs[f_off + q][rw + 4 * r], w, acc[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) Bs[s_off + j][rw + 4 * r] -= acc[r];
  }
  __syncthreads();  // every lane has read W; the s columns are updated
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int idx = t + 256 * r;
    W[idx % kNB][idx / kNB] = ivs[r];
  }
  __syncthreads();
  // X_s over the s block's triangle.
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  cnt = j < s_len ? (upper ? j + 1 : s_len - j) : 0;
  for (int s = 0; s < cnt; ++s) {
    const int q = upper ? s : s_len - 1 - s;
    const double w = W[q][j];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = __builtin_fma(Bs[s_off + q][rw + 4 * r], w, acc[r]);
  }
  __syncthreads();  // every lane has read the s columns
  if (j < s_len) {
#pragma unroll
    for (int r = 0; r < R; ++r) Bs[s_off + j][rw + 4 * r] = acc[r];
  }
  __syncthreads();
  if (row_ok) {
#pragma unroll
    for (int r = 0; r < L; ++r) {
      const int p = c0 + (256 / kLeafRows) * r;
      if (p < len) b[row0 + i + p * ldb] = (T)Bs[p][i];
    }
  }
}

// B := alpha * B over m x n.
template <typename T>
__global__ void mul_kernel(int64_t m, int64_t n, T alpha, T *b, int64_t ldb) {
  const int64_t total = m * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / m, i = e - j * m;
    b[i + j * ldb] = alpha * b[i + j * ldb];
  }
}

unsigned grid_for(int64_t total) {
  int64_t b = (total + 255) / 256;
  if (b > 16384) b = 16384;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Pointer to op(A)(r0, c0) as a gemm operand (transposed storage when trans).
template <typename T>
const T *sub(const T *a, int64_t lda, bool trans, int64_t r0, int64_t c0) {
  return trans ? a + c0 + r0 * lda : a + r0 + c0 * lda;
}

template <typename T>
struct Solver {
  const TrsmProblem<T> &p;
  const double *inv;
  hipStream_t st;
  bool forward;  // substitution runs from index 0 up

  // Solve the diagonal range [lo, lo + len) (lo a multiple of 64), every
  // contribution from outside it having been subtracted already.
  hipError_t range(int64_t lo, int64_t len) {
    if (len <= kNB) {
      const double *ib = inv + (lo / kNB) * kNB * kNB;
      if (p.left)
        hipLaunchKernelGGL((trsm_apply_left<T>),
                           dim3((unsigned)((p.n + kLeafCols - 1) / kLeafCols)), dim3(256), 0,
                           st, ib, p.upper, (int)len, p.n, p.b + lo, p.ldb);
      else
        hipLaunchKernelGGL((trsm_apply_right<T>),
                           dim3((unsigned)((p.m + kLeafRows - 1) / kLeafRows)), dim3(256), 0,
                           st, ib, p.upper, (int)len, p.m, p.b + lo * p.ldb, p.ldb);
      return hipGetLastError();
    }
    const int64_t h = ((len / 2 + kNB - 1) / kNB) * kNB;
    int64_t f_lo = lo, f_len = h, s_lo = lo + h, s_len = len - h;
    if (!forward) {
      f_lo = lo + h;
      f_len = len - h;
      s_lo = lo;
      s_len = h;
    }
    if (p.left && len <= 2 * kNB) {
      hipLaunchKernelGGL((trsm_apply_left_pair<T>),
                         dim3((unsigned)((p.n + kLeafCols - 1) / kLeafCols)), dim3(256), 0, st,
                         inv + (f_lo / kNB) * kNB * kNB, inv + (s_lo / kNB) * kNB * kNB,
                         sub(p.a, p.lda, p.trans, s_lo, f_lo), p.lda, p.trans, p.upper,
                         (int)(f_lo - lo), (int)f_len, (int)(s_lo - lo), (int)s_len, p.n,
                         p.b + lo, p.ldb);
      return hipGetLastError();
    }
    if (!p.left && len <= 2 * kNB) {
      hipLaunchKernelGGL((trsm_apply_right_pair<T>),
                         dim3((unsigned)((p.m + kLeafRows - 1) / kLeafRows)), dim3(256), 0, st,
                         inv + (f_lo / kNB) * kNB * kNB, inv + (s_lo / kNB) * kNB * kNB,
                         sub(p.a, p.lda, p.trans, f_lo, s_lo), p.lda, p.trans, p.upper,
                         (int)(f_lo - lo), (int)f_len, (int)(s_lo - lo), (int)s_len, p.m,
                         p.b + lo * p.ldb, p.ldb);
      return hipGetLastError();
    }
    hipError_t e = range(f_lo, f_len);
    if (e != hipSuccess) return e;
    if (p.left) {
      // B[second rows] -= op(A)[second, first] * X[first rows]
      GemmProblem<T> u{p.trans, false, s_len, p.n, f_len, T(-1), T(1),
                       sub(p.a, p.lda, p.trans, s_lo, f_lo), p.lda, p.b + f_lo, p.ldb,
                       p.b + s_lo, p.ldb};
      e = launch_gemm<T>(u, st);
    } else {
      // B[:, second] -= X[:, first] * op(A)[first, second]
      GemmProblem<T> u{false, p.trans, p.m, s_len, f_len, T(-1), T(1), p.b + f_lo * p.ldb, p.ldb,
                       sub(p.a, p.lda, p.trans, f_lo, s_lo), p.lda, p.b + s_lo * p.ldb, p.ldb};
      e = launch_gemm<T>(u, st);
    }
    if (e != hipSuccess) return e;
    return range(s_lo, s_len);
  }
};

}  // namespace

size_t trsm_workspace_bytes(int64_t m, int64_t n, bool left) {
  const int64_t q = left ? m : n, nrhs = left ? n : m;
  const int64_t nblk = (q + kNB - 1) / kNB;
  size_t bytes = (size_t)(nblk * kNB * kNB) * sizeof(double);
  // A split register-resident solve (orders up to kTrsmWideMaxOrder): its
  // published X and one flag per unit (at most nrhs / 16 + 1 units).
  if (q > (kTrsmSplitMinBlocks - 1) * kNB && q <= kTrsmWideMaxOrder)
    bytes += (size_t)trsm_split_x_doubles(q, nrhs) * sizeof(double) +
             (size_t)(nrhs / 16 + 1) * sizeof(int);
  return bytes;
}

template <typename T>
hipError_t launch_trsm(const TrsmProblem<T> &p, void *ws, hipStream_t st) {
  const int64_t m = p.m, n = p.n;
  if (m == 0 || n == 0) return hipSuccess;
  const int64_t q = p.left ? m : n;
  if (q <= kDirectMaxOrder && (p.left ? n : m) <= (int64_t)INT32_MAX - 64) {
    // Orders up to 2048: the right-hand sides of each workgroup stay in LDS
    // and the whole solve is two launches (grouped_trsm.hip), in Left form
    // (a Right problem is solved on B^T with op(A)^T).
    TrsmDesc<T> d{};
    d.a = p.a;
    d.b = p.b;
    d.lda = p.lda;
    d.ldb = p.ldb;
    d.q = (int)q;
    d.nrhs = (int)(p.left ? n : m);
    d.trans = p.left ? p.trans : !p.trans;
    d.upper = p.left ? p.upper : !p.upper;
    d.unit = p.unit;
    d.right = !p.left;
    d.forward = !d.upper;
    d.width = trsm_unit_width<T>(q, p.lda);
    if (trsm_wide_ok<T>(q, p.lda, d.nrhs) && (d.nrhs + 31) / 32 >= kTrsmFillUnits) d.width = 32;
    d.alpha = p.alpha;
    d.inv0 = 0;
    return launch_trsm_single<T>(d, (double *)ws, st);
  }
  double *inv = (double *)ws;
  if (p.alpha != T(1))
    hipLaunchKernelGGL((mul_kernel<T>), dim3(grid_for(m * n)), dim3(256), 0, st, m, n, p.alpha,
                       p.b, p.ldb);
  TrsmDesc<T> dd{};
  dd.a = p.a;
  dd.lda = p.lda;
  dd.q = (int)q;
  dd.trans = p.trans;
  dd.upper = p.upper;
  dd.unit = p.unit;
  hipError_t e = launch_trsm_invert<T>(dd, inv, st);
  if (e != hipSuccess) return e;
  // Left+lower and Right+upper substitute from index 0 up, the others down.
  Solver<T> s{p, inv, st, p.left ? !p.upper : p.upper};
  return s.range(0, q);
}

template hipError_t launch_trsm<float>(const TrsmProblem<float> &, void *, hipStream_t);
template hipError_t launch_trsm<double>(const TrsmProblem<double> &, void *, hipStream_t);

}  // namespace rblas
