// Grouped triangular solve: many independent rblas trsm problems
// (level3.rs:390-520) in one persistent launch, after one batched launch that
// inverts every 64 x 64 diagonal block of every op(A).
//
// Every problem is solved as a Left problem: X op(A) = alpha B (Right) is
// op(A)^T X^T = alpha B^T, so a Right problem becomes Left with op(A)
// transposed (trans flipped, triangle flipped) on B^T. A work unit is one
// system and W consecutive right-hand sides (columns of B for Left, rows of
// B for Right). The unit keeps its W right-hand sides as f64 -- 16 of them
// for orders up to 1024 in the waves' MFMA accumulators (solve_unit_reg), 8
// for orders up to 2048 in LDS (solve_unit) -- then walks the diagonal blocks
// in substitution order:
//   X_b = inv_b * B_b   (f64 MFMA over the 16 x 16 sub-blocks of inv_b; a
//                        diagonal sub-block whose B rows hold a non-finite
//                        entry is summed over its triangle only, scalar f64,
//                        so no structural zero ever multiplies it: a
//                        non-finite entry reaches exactly what rblas's
//                        substitution reaches)
//   B_rest -= op(A)[rest, b] * X_b   (f64 MFMA, op(A) streamed from HBM/L2)
// and writes the right-hand sides back. All arithmetic is f64 (f32 operands
// are widened on load and rounded once on store).
#include "grouped.hpp"

namespace rblas {
namespace {

#ifndef RBLAS_TRSM_WAVES
#define RBLAS_TRSM_WAVES 8
#endif
constexpr int kWaves = RBLAS_TRSM_WAVES;  // solve workgroup: 8 waves
constexpr int kThreads = 64 * kWaves;
// Inversion workgroups: 8 waves for the batched launch (four share a CU),
// 16 for a single call's few blocks (latency-bound: more lanes to stage and
// store each block).
constexpr int kInvThreads = 512;
constexpr int kInvThreadsSingle = 1024;
constexpr int kNB = 64;             // diagonal block
constexpr int kLdsDoubles = kTrsmLdsDoubles;  // q * W limit (128 KiB)
constexpr int kSub = 256;           // doubles in a 16 x 16 sub-block
constexpr int kInvLds = 10 * kSub;  // the 10 sub-blocks of a triangular 64 x 64 block
constexpr int kInvPerThread = (kInvLds + kThreads - 1) / kThreads;
constexpr int kLoadBatch = 8;       // right-hand-side loads in flight per thread

typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f64x4 mma64(double p, double q, f64x4 acc) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(p, q, acc, 0, 0, 0);
}

template <typename T>
__device__ __forceinline__ double opa(const TrsmDesc<T> &d, int64_t r, int64_t c) {
  return (double)(d.trans ? d.a[c + r * d.lda] : d.a[r + c * d.lda]);
}

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Value of v in lane l (l wave-uniform), in every lane.
__device__ __forceinline__ double lane_value(double v, int l) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Lane S's value of each row of 16 lanes, in every lane of the row (DPP
// row broadcast); S must be a constant once the caller's loop is unrolled.
template <int S>
__device__ __forceinline__ int row_bcast_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x150 + S, 0xf, 0xf, false);
}
template <int S>
__device__ __forceinline__ double row_bcast_d(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)row_bcast_i<S>((int)(unsigned)u);
  const unsigned hi = (unsigned)row_bcast_i<S>((int)(unsigned)(u >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double row_bcast(double v, int s) {
  switch (s) {
    case 0: return row_bcast_d<0>(v);
    case 1: return row_bcast_d<1>(v);
    case 2: return row_bcast_d<2>(v);
    case 3: return row_bcast_d<3>(v);
    case 4: return row_bcast_d<4>(v);
    case 5: return row_bcast_d<5>(v);
    case 6: return row_bcast_d<6>(v);
    case 7: return row_bcast_d<7>(v);
    case 8: return row_bcast_d<8>(v);
    case 9: return row_bcast_d<9>(v);
    case 10: return row_bcast_d<10>(v);
    case 11: return row_bcast_d<11>(v);
    case 12: return row_bcast_d<12>(v);
    case 13: return row_bcast_d<13>(v);
    case 14: return row_bcast_d<14>(v);
    default: return row_bcast_d<15>(v);
  }
}

// LDS of one diagonal-block inversion: the lower triangle of a 64 x 64 block
// as its 16 x 16 sub-blocks (I, J), I >= J, 256 doubles each, element (r, c)
// of a sub-block at c * 16 + (r ^ c): the 16 rows of a column are contiguous
// (a permutation of them), and the 16 columns of one row fall on distinct
// banks. Ls holds the six strictly lower sub-blocks of L; Iv the ten of the
// inverse. The four diagonal sub-blocks of L live in Iv's slots 1, 3, 4 and 6
// (off-diagonal slots of the inverse, written only after the diagonal
// inverses have been formed from them). 32 KiB in all, so four inversions
// (8 waves each) share a CU. (With 4 waves, five fit the 160 KiB and the
// batching workload's 1256 blocks run in one round, but each block took
// longer and the launch no less: 25.6 us.)
constexpr int kInvOff = 6;   // strictly lower 16 x 16 sub-blocks
constexpr int kInvTri = 10;  // lower sub-blocks including the diagonal
__device__ __forceinline__ int sub_elem(int r, int c) { return c * 16 + (r ^ c); }
// Slot of sub-block (I, J), I > J, in Ls; of (I, J), I >= J, in Iv; of the
// diagonal sub-block I of L (in Iv).
__device__ __forceinline__ int ls_slot(int I, int J) { return I * (I - 1) / 2 + J; }
__device__ __forceinline__ int iv_slot(int I, int J) { return I * (I + 1) / 2 + J; }
__device__ __forceinline__ int ld_slot(int I) { return I == 0 ? 1 : I == 1 ? 3 : I == 2 ? 4 : 6; }

struct InvLds {
  double ls[kInvOff * kSub];
  double iv[kInvTri * kSub];
  // L(row, col) of the padded block, lower form (0 above the diagonal).
  __device__ __forceinline__ double l(int row, int col) const {
    const int I = row >> 4, J = col >> 4;
    if (I == J) return iv[ld_slot(I) * kSub + sub_elem(row & 15, col & 15)];
    if (I > J) return ls[ls_slot(I, J) * kSub + sub_elem(row & 15, col & 15)];
    return 0.0;
  }
  // Element (row, col) of the inverse, row >> 4 >= col >> 4.
  __device__ __forceinline__ double &inv(int row, int col) {
    return iv[iv_slot(row >> 4, col >> 4) * kSub + sub_elem(row & 15, col & 15)];
  }
};

// acc += L[bi, bk] * Inv[bk, bj] over one 16 x 16 block product (f64 MFMA,
// k in steps of 4), bi > bk >= bj. Output layout: acc[r] holds row
// (lane >> 4) + 4 r, column lane & 15 -- also the B-operand layout of the
// next product's k step r.
__device__ __forceinline__ f64x4 mm16(InvLds &S, int bi, int bk, int bj, f64x4 acc, int lane) {
  const double *X = S.ls + ls_slot(bi, bk) * kSub;
  const double *Y = S.iv + iv_slot(bk, bj) * kSub;
#pragma unroll
  for (int t4 = 0; t4 < 4; ++t4) {
    const int m = 4 * t4 + (lane >> 4);
    acc = mma64(X[sub_elem(lane & 15, m)], Y[sub_elem(m, lane & 15)], acc);
  }
  return acc;
}

// Inverse of diagonal block `blk` of op(A) (as the problem is solved: Left
// form), f64, column-major 64 x 64 at inv + d.inv0 + blk * 4096, zero outside
// the block's nb x nb corner. An upper block is inverted as the lower one
// P U P (P reverses the nb indices). The padded block is L below nb and the
// identity past it.
//
// When every entry of the block's triangle is finite, blockwise with MFMA:
// the four 16 x 16 diagonal inverses D_b = L_bb^-1 by substitution (wave b;
// lane (i, g) holds row i of columns g, g + 4, g + 8, g + 12; step k scales
// row k by 1 / L(k, k), broadcasts it across the 16-lane row and subtracts
// L(i, k) times it from every later row i), then the off-diagonal blocks by
// distance d = i - j = 1, 2, 3 (i > j):
//   inv_ij = -D_i (sum_{k=j}^{i-1} L_ik inv_kj),
// each a handful of 16 x 16 f64 MFMA products, one wave per block.
//
// Otherwise (an Inf or NaN in the triangle) the whole block by
// column-oriented substitution against the identity, in registers: lane i
// holds row i, wave w columns 8w..8w+7; step k scales row k by 1 / L(k, k)
// and subtracts L(i, k) times it from every later row i, so each entry gets
// exactly the operations of a row-by-row substitution of its column and a
// non-finite entry reaches no more of the inverse than it does there.
template <int NT, typename T>
__device__ __forceinline__ void invert_block(const TrsmDesc<T> &d, int blk, double *inv) {
  static_assert(NT % 64 == 0 && kNB % (NT / 64) == 0, "inversion workgroup shape");
  __shared__ InvLds S;
  const int64_t r0 = (int64_t)blk * kNB;
  const int nb = (int)min((int64_t)kNB, (int64_t)d.q - r0);
  const int t = threadIdx.x;
  // Staged: the 6 strictly lower sub-blocks of L into S.ls, the 4 diagonal
  // ones into their Iv slots; kFill elements per thread.
  constexpr int kFill = (kInvTri * kSub + NT - 1) / NT;
  static_assert((kInvTri * kSub) % 64 == 0, "staging rounds are whole waves");
  T raw[kFill];
#pragma unroll
  for (int u = 0; u < kFill; ++u) {  // all loads issued first, no branches
    const int idx = min(t + u * NT, kInvTri * kSub - 1);
    const int q = idx / kSub, e = idx % kSub;
    const int I = q < kInvOff ? (q < 1 ? 1 : q < 3 ? 2 : 3) : q - kInvOff;
    const int J = q < kInvOff ? q - I * (I - 1) / 2 : I;
    const int i = 16 * I + e % 16, k = 16 * J + e / 16;
    const bool in = i < nb && k < nb && i > k;
    const int64_t r = r0 + (d.upper ? nb - 1 - i : i), c = r0 + (d.upper ? nb - 1 - k : k);
    const bool diag = i == k && i < nb && !d.unit;
    const T *p = d.trans ? d.a + c + r * d.lda : d.a + r + c * d.lda;
    raw[u] = *((in || diag) ? p : d.a);
  }
  bool fin = true;
#pragma unroll
  for (int u = 0; u < kFill; ++u) {
    const int idx = t + u * NT;
    if (idx >= kInvTri * kSub) break;  // wave-uniform: whole waves past the end
    const int q = idx / kSub, e = idx % kSub;
    const int I = q < kInvOff ? (q < 1 ? 1 : q < 3 ? 2 : 3) : q - kInvOff;
    const int J = q < kInvOff ? q - I * (I - 1) / 2 : I;
    const int i = 16 * I + e % 16, k = 16 * J + e / 16;
    const bool in = i < nb && k < nb && i > k;
    const bool diag = i == k && i < nb && !d.unit;
    const double v = (in || diag) ? (double)raw[u] : (i == k ? 1.0 : 0.0);  // padding: identity
    fin = fin && __builtin_isfinite(v);
    const int o = sub_elem(e % 16, e / 16);
    if (q < kInvOff)
      S.ls[q * kSub + o] = v;
    else
      S.iv[ld_slot(I) * kSub + o] = v;
  }
  // The barrier below also publishes the staged L. For the batched launch
  // each wave's finiteness vote goes to Iv slot 8 (the block (3, 2) of the
  // inverse, written only after a later barrier), so it needs no LDS beyond
  // InvLds (four per CU); a single call's 16-wave workgroups (two per CU
  // whatever their LDS) take the barrier's own reduction, one LDS round
  // trip shorter.
  const int lane = t & 63, w = uni(t >> 6);
  bool all_fin = true;
  if constexpr (NT == kInvThreadsSingle) {
    all_fin = __syncthreads_and(fin);
  } else {
    const bool wave_fin = __all(fin);
    if (lane == 0) S.iv[iv_slot(3, 2) * kSub + w] = wave_fin ? 1.0 : 0.0;
    __syncthreads();
#pragma unroll
    for (int v = 0; v < NT / 64; ++v) all_fin = all_fin && S.iv[iv_slot(3, 2) * kSub + v] != 0.0;
  }
  double *out = inv + d.inv0 + (int64_t)blk * kNB * kNB;
  if (all_fin) {
    if (w < 4) {  // D_w by substitution, from the L diagonal slot to Iv's diagonal slot
      const int i = lane & 15, g = lane >> 4, o = 16 * w;
      const double *DL = S.iv + ld_slot(w) * kSub;  // L's diagonal sub-block w
      const double rl = 1.0 / DL[sub_elem(i, i)];
      double x[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) x[r] = i == g + 4 * r ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const double lik = DL[sub_elem(i, k)];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double y = x[r] * rl;
          const double xk = row_bcast(y, k);
          const double upd = __builtin_fma(-lik, xk, x[r]);
          x[r] = i == k ? y : (i > k ? upd : x[r]);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) S.inv(o + i, o + g + 4 * r) = x[r];
    }
    // Every wave is done with the L diagonal slots before the products below
    // overwrite them.
    __syncthreads();
#pragma unroll
    for (int dist = 1; dist < 4; ++dist) {
      if (w < 4 - dist) {  // block (i, j) = (w + dist, w)
        const int bi = w + dist, bj = w;
        f64x4 tacc = {0, 0, 0, 0};
        for (int bk = bj; bk < bi; ++bk) tacc = mm16(S, bi, bk, bj, tacc, lane);
        f64x4 pacc = {0, 0, 0, 0};
        const double *D = S.iv + iv_slot(bi, bi) * kSub;
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4) {
          const int m = 4 * t4 + (lane >> 4);
          pacc = mma64(D[sub_elem(lane & 15, m)], tacc[t4], pacc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
          S.inv(16 * bi + (lane >> 4) + 4 * r, 16 * bj + (lane & 15)) = -pacc[r];
      }
      __syncthreads();
    }
    constexpr int kOut = kNB * kNB / NT;
#pragma unroll
    for (int u = 0; u < kOut; ++u) {
      const int idx = t + u * NT;
      const int li = idx % kNB, lj = idx / kNB;
      const bool in = li < nb && lj < nb;
      const int i = (d.upper && in) ? nb - 1 - li : li;
      const int j = (d.upper && in) ? nb - 1 - lj : lj;
      out[i + j * kNB] = (in && li / 16 >= lj / 16) ? S.inv(li, lj) : 0.0;
    }
    return;
  }
  constexpr int kCols = kNB / (NT / 64);  // columns per wave
  const int j0 = kCols * w;
  const double rl = 1.0 / S.l(lane, lane);
  double x[kCols];
#pragma unroll
  for (int c = 0; c < kCols; ++c) x[c] = lane == j0 + c ? 1.0 : 0.0;
  for (int k = j0; k < kNB; ++k) {
    const double lik = S.l(lane, k);
#pragma unroll
    for (int c = 0; c < kCols; ++c) {
      if (k >= j0 + c) {  // wave-uniform
        // Every lane forms x / L(lane, lane); lane k's is x_k / L(k, k), the
        // value broadcast (the same product the row-wise substitution forms).
        const double y = x[c] * rl;
        const double xk = lane_value(y, k);
        const double upd = __builtin_fma(-lik, xk, x[c]);
        x[c] = lane == k ? y : (lane > k ? upd : x[c]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < kCols; ++c) {
    const int jl = j0 + c;
    const bool in = lane < nb && jl < nb;
    const int i = (d.upper && lane < nb) ? nb - 1 - lane : lane;
    const int j = (d.upper && jl < nb) ? nb - 1 - jl : jl;
    out[i + j * kNB] = in ? x[c] : 0.0;
  }
}

template <typename T>
__global__ __launch_bounds__(kInvThreads) void grouped_trsm_inv(const TrsmDesc<T> *descs,
                                                                const int2 *blocks, double *inv) {
  const int2 u = blocks[blockIdx.x];
  invert_block<kInvThreads>(descs[u.x], u.y, inv);
}

// One problem passed by value: block blockIdx.x. Workgroup 0 also clears
// the nflags flags of a split solve.
template <typename T>
__global__ __launch_bounds__(kInvThreadsSingle) void single_trsm_inv(TrsmDesc<T> d, double *inv,
                                                               int *flags, int nflags) {
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < nflags; i += kInvThreadsSingle) flags[i] = 0;
  invert_block<kInvThreadsSingle>(d, blockIdx.x, inv);
}

// Address of right-hand side c, row r of the unit's B (Left: B(r, c0 + c);
// Right: B(c0 + c, r)).
template <typename T>
__device__ __forceinline__ int64_t b_index(const TrsmDesc<T> &d, int64_t r, int64_t c) {
  return d.right ? c + r * d.ldb : r + c * d.ldb;
}

// LDS copy of a diagonal-block inverse: its 10 nonzero 16 x 16 sub-blocks,
// each column-major. Sub-block (bi, bj) of a lower block (bi >= bj) sits at
// slot bi (bi + 1) / 2 + bj; an upper block uses the slot of (3 - bi, 3 - bj).
__device__ __forceinline__ int sub_slot(bool upper, int bi, int bj) {
  if (upper) {
    bi = 3 - bi;
    bj = 3 - bj;
  }
  return (bi * (bi + 1) / 2 + bj) * kSub;
}

// Element e (< kInvLds) of the LDS copy: its offset in the 64 x 64 inverse.
__device__ __forceinline__ int inv_src(bool upper, int e) {
  const int slot = e / kSub, w = e % kSub;
  // slot -> (bi, bj) of the lower form: 0:(0,0) 1:(1,0) 2:(1,1) 3:(2,0) ...
  const int bi = slot >= 6 ? 3 : slot >= 3 ? 2 : slot >= 1 ? 1 : 0;
  int bj = slot - bi * (bi + 1) / 2;
  int ri = bi;
  if (upper) {
    ri = 3 - bi;
    bj = 3 - bj;
  }
  return (16 * ri + w % 16) + (16 * bj + w / 16) * kNB;
}

// Raw loads of op(A)(rb + l%16, r0 + 4t + l/16), t = 0..15, for lane l. An
// entry outside rows [.., hi) or block columns [0, nb) loads op(A)(0, 0)
// instead (a valid address), so every load is issued without a branch; the
// out-of-range entries are zeroed where they are used (update_rows), which
// leaves the loads in flight until then.
template <typename T>
__device__ __forceinline__ void load_rows(const TrsmDesc<T> &d, int rb, int hi, int r0, int nb,
                                          int64_t si, int64_t sj, int lane, T (&v)[16]) {
  const int kq = lane >> 4;
  const int64_t ar = rb + (lane & 15);
  const bool row_ok = ar < hi;
  const T *ap = d.a + ar * si + (int64_t)(r0 + kq) * sj;
#pragma unroll
  for (int t4 = 0; t4 < 16; ++t4) {
    const bool ok = row_ok && 4 * t4 + kq < nb;
    v[t4] = *(ok ? ap + (int64_t)(4 * t4) * sj : d.a);
  }
}

// Bs[i][col] -= (op(A)[rb .. rb+15, block] * X_b)[i][col] for rows i < hi,
// from the raw loads of load_rows.
template <typename T>
__device__ __forceinline__ void update_rows(double *Bs, const T (&pv)[16], const double (&xq)[16],
                                            int rb, int hi, int nb, int W, int lane) {
  const int col = lane & 15, kq = lane >> 4;
  const bool row_ok = rb + (lane & 15) < hi;
  f64x4 u = {0, 0, 0, 0};
#pragma unroll
  for (int t4 = 0; t4 < 16; ++t4) {
    const bool ok = row_ok && 4 * t4 + kq < nb;
    u = mma64(ok ? (double)pv[t4] : 0.0, xq[t4], u);
  }
  if (col < W) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = rb + kq + 4 * r;
      if (i < hi) Bs[i * W + col] -= u[r];
    }
  }
}

// Bs: the kernel's LDS block of kLdsDoubles doubles, used as Bs[r * W + c];
// Is: kInvLds doubles for the current diagonal-block inverse.
//
// Per diagonal block b, in substitution order:
//   1. every wave issues the op(A) loads of its first 16-row block of the
//      update (they do not depend on X_b);
//   2. waves 0-3 form X_b = inv_b * B_b from LDS (rows 16w .. 16w + 15);
//   3. X_b replaces B_b in LDS;
//   4. every wave issues the loads of inv_{b+1}, then runs the update
//      B_rest -= op(A)[rest, b] X_b over its 16-row blocks (stride 16 x
//      kWaves rows), loading the next row block before the MFMAs of this one;
//   5. inv_{b+1} goes to LDS.
template <typename T, int W>
__device__ __forceinline__ void solve_unit(const TrsmDesc<T> &d, int c0, const double *inv, double *Bs,
                           double *Is) {
  // Opaque copy of the thread id: in the persistent loop every per-lane
  // offset is recomputed per unit instead of hoisted out of it (and spilled).
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  const int lane = t & 63, wave = uni(t >> 6);
  const int q = d.q;
  const int nc = min(W, d.nrhs - c0);
  const int qpad = (q + kNB - 1) / kNB * kNB;
  const bool upper = d.upper;
  // Load (alpha applied to every entry, as rblas's pre-scale: 0 * NaN = NaN),
  // 8 loads per thread in flight at a time, without branches.
  const int total = qpad * W;
  for (int e0 = t; e0 < total; e0 += kLoadBatch * kThreads) {
    T raw[kLoadBatch];
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int e = e0 + u * kThreads;
      const int r = e / W, c = e % W;
      const bool ok = e < total && r < q && c < nc;
      raw[u] = *(ok ? d.b + b_index(d, r, c0 + c) : d.b);
    }
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int e = e0 + u * kThreads;
      const int r = e / W, c = e % W;
      const T v = d.alpha != T(1) ? d.alpha * raw[u] : raw[u];
      if (e < total) Bs[e] = (r < q && c < nc) ? (double)v : 0.0;
    }
  }
  const int nblk = qpad / kNB;
  {
    const double *ib = inv + d.inv0 + (int64_t)(d.forward ? 0 : nblk - 1) * kNB * kNB;
    double iv[kInvPerThread];
#pragma unroll
    for (int u = 0; u < kInvPerThread; ++u) {
      const int e = t + u * kThreads;
      iv[u] = ib[e < kInvLds ? inv_src(upper, e) : 0];
    }
#pragma unroll
    for (int u = 0; u < kInvPerThread; ++u) {
      const int e = t + u * kThreads;
      if (e < kInvLds) Is[e] = iv[u];
    }
  }
  __syncthreads();
  const int col = lane & 15;  // MFMA output column / Q-operand column
  const int kq = lane >> 4;   // k within a 4-step
  const int64_t si = d.trans ? d.lda : 1;  // op(A) row step
  const int64_t sj = d.trans ? 1 : d.lda;  // op(A) column step
  for (int s = 0; s < nblk; ++s) {
    const int b = d.forward ? s : nblk - 1 - s;
    const int r0 = b * kNB;
    const int nb = min(kNB, q - r0);
    const int lo = upper ? 0 : r0 + nb, hi = upper ? r0 : q;
    // 1. First update row block of this wave.
    int rb = lo + 16 * wave;
    T pa[16], pb[16];
    if (rb < hi) load_rows(d, rb, hi, r0, nb, si, sj, lane, pa);
    // 2. X_b = inv_b * B_b: wave w < 4 owns rows 16w .. 16w+15 of the block.
    f64x4 acc = {0, 0, 0, 0};
    const int iw = wave;
    const bool xw = iw < 4 && 16 * iw < nb;
    if (xw) {
#pragma unroll
      for (int jb = 0; jb < 4; ++jb) {
        const bool use = (upper ? jb > iw : jb < iw) && 16 * jb < nb;  // wave-uniform
        if (use) {
          const double *sb = Is + sub_slot(upper, iw, jb);
          double pa[4], qa[4];
#pragma unroll
          for (int t4 = 0; t4 < 4; ++t4) {
            const int j = 4 * t4 + kq;
            pa[t4] = sb[(lane & 15) + 16 * j];
            qa[t4] = (col < W) ? Bs[(r0 + 16 * jb + j) * W + col] : 0.0;
          }
#pragma unroll
          for (int t4 = 0; t4 < 4; ++t4) acc = mma64(pa[t4], qa[t4], acc);
        }
      }
      // Diagonal sub-block. Its inverse holds exact zeros outside the
      // triangle, so when this sub-block's B rows are all finite the full
      // 16 x 16 product (4 MFMAs) adds only exact zeros beyond the
      // triangle's terms. A non-finite entry must not reach entries that do
      // not depend on it (0 * Inf is NaN), so then the triangle is summed
      // alone, scalar f64 (predicated: every lane runs the same 16 steps),
      // and each X entry gets exactly the terms rblas's substitution gives it.
      const double *sd = Is + sub_slot(upper, iw, iw);
      double pd[4], qd[4];
      bool fin = true;
#pragma unroll
      for (int t4 = 0; t4 < 4; ++t4) {
        const int j = 4 * t4 + kq;
        pd[t4] = sd[(lane & 15) + 16 * j];
        qd[t4] = (col < W) ? Bs[(r0 + 16 * iw + j) * W + col] : 0.0;
        fin = fin && __builtin_isfinite(qd[t4]);
      }
      if (__all(fin)) {
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4) acc = mma64(pd[t4], qd[t4], acc);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int il = kq + 4 * r;  // row within the sub-block (MFMA output layout)
          double sum = acc[r];
#pragma unroll
          for (int jl = 0; jl < 16; ++jl) {
            const int j = 16 * iw + jl;
            const bool on = (upper ? jl >= il : jl <= il) && j < nb && col < W;
            const double w = on ? sd[il + 16 * jl] : 0.0;
            const double x = on ? Bs[(r0 + j) * W + col] : 0.0;
            sum = on ? __builtin_fma(w, x, sum) : sum;
          }
          acc[r] = sum;
        }
      }
    }
    __syncthreads();  // every wave has read B_b and inv_b
    // 3.
    if (xw && col < W) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * iw + kq + 4 * r;
        if (i < nb) Bs[(r0 + i) * W + col] = acc[r];
      }
    }
    // 4. Next inverse into registers.
    const bool more = s + 1 < nblk;
    double ipre[kInvPerThread];
    if (more) {
      const int bn = d.forward ? b + 1 : b - 1;
      const double *ib = inv + d.inv0 + (int64_t)bn * kNB * kNB;
#pragma unroll
      for (int u = 0; u < kInvPerThread; ++u) {
        const int e = t + u * kThreads;
        ipre[u] = ib[e < kInvLds ? inv_src(upper, e) : 0];
      }
    }
    __syncthreads();  // X_b visible
    double xq[16];
#pragma unroll
    for (int t4 = 0; t4 < 16; ++t4) {
      const int j = 4 * t4 + kq;
      xq[t4] = (col < W && j < nb) ? Bs[(r0 + j) * W + col] : 0.0;
    }
    // Two register sets, each loaded one row block ahead of its MFMAs.
    while (rb < hi) {
      load_rows(d, rb + 16 * kWaves, hi, r0, nb, si, sj, lane, pb);
      update_rows(Bs, pa, xq, rb, hi, nb, W, lane);
      rb += 16 * kWaves;
      if (rb >= hi) break;
      load_rows(d, rb + 16 * kWaves, hi, r0, nb, si, sj, lane, pa);
      update_rows(Bs, pb, xq, rb, hi, nb, W, lane);
      rb += 16 * kWaves;
    }
    // 5.
    if (more) {
#pragma unroll
      for (int u = 0; u < kInvPerThread; ++u) {
        const int e = t + u * kThreads;
        if (e < kInvLds) Is[e] = ipre[u];
      }
    }
    __syncthreads();
  }
  for (int e = t; e < q * W; e += kThreads) {
    const int r = e / W, c = e % W;
    if (c < nc) d.b[b_index(d, r, c0 + c)] = (T)Bs[e];
  }
}

// Buffer loads for solve_unit_reg's op(A) rows: one 32-bit lane offset and a
// scalar step per k, instead of a 64-bit address per load. An offset of
// kPast lies beyond the resource's num_records (clamped to 2^31 - 1), which
// the hardware reads as 0. In-block offsets stay below 2^31 for the leading
// dimensions trsm_unit_width gives width 16.
constexpr uint32_t kPast = 0x80000000u;
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                   uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                 uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// op(A)(rb + l%16, r0 + 4t + l/16), t = 0..15, for lane l; entries outside
// rows [.., hi) or block columns [0, nb) read as 0.
template <typename T>
__device__ __forceinline__ void load_rows_buf(const TrsmDesc<T> &d, int rb, int hi, int r0, int nb,
                                              int64_t si, int64_t sj, int lane, T (&v)[16]) {
  const int kq = lane >> 4;
  const T *o = d.a + (int64_t)rb * si + (int64_t)r0 * sj;
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(o), (short)0, 0x7fffffff, 0x00020000);
  const bool row_ok = rb + (lane & 15) < hi;
  const uint32_t base = row_ok ? (uint32_t)(((lane & 15) * si + kq * sj) * (int64_t)sizeof(T)) : kPast;
  const uint32_t step = (uint32_t)(4 * sj * (int64_t)sizeof(T));
#pragma unroll
  for (int t4 = 0; t4 < 16; ++t4)
    v[t4] = buf_load<T>(rs, 4 * t4 + kq < nb ? base : kPast, (uint32_t)t4 * step);
}

typedef __attribute__((address_space(3))) void *lds_void_ptr;

// Workgroup barrier that leaves this wave's last N vector-memory loads in
// flight and publishes its LDS stores: vmcnt(N) lgkmcnt(0), then s_barrier.
// Loads retire in order, so everything issued before those N (an LDS DMA
// included) has landed. __syncthreads() would drain every load instead.
template <int N>
__device__ __forceinline__ void barrier_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt((N & 0xf) | ((N >> 4) << 14) | 0x0070);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// acc += (-X_b)^T op(A)[rb .. rb+15, block]^T (transposed accumulator layout
// of solve_unit_reg) from the loads of load_rows_buf and X fragments xq.
template <typename T>
__device__ __forceinline__ f64x4 update_acc(f64x4 acc, const T (&pv)[16],
                                            const double (&xq)[16]) {
#pragma unroll
  for (int t4 = 0; t4 < 16; ++t4) acc = mma64(xq[t4], (double)pv[t4], acc);
  return acc;
}

// The same over half a block's columns (k in [32 half, 32 half + 32)) for two
// column blocks, X fragments read from Xs (row stride kS) per product.
template <typename T>
__device__ __forceinline__ void update_half(f64x4 (&acc)[2], const T (&pv)[8], const double *Xs,
                                            int kS, int half, int lane) {
  const int kq = lane >> 4, li = lane & 15;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const double a = (double)pv[t];
    const double *xr = Xs + (32 * half + 4 * t + kq) * kS + li;
    acc[0] = mma64(xr[0], a, acc[0]);
    acc[1] = mma64(xr[16], a, acc[1]);
  }
}

// op(A)(rb + l%16, r0 + 32 half + 4t + l/16), t = 0..7, as load_rows_buf.
template <typename T>
__device__ __forceinline__ void load_half_buf(const TrsmDesc<T> &d, int rb, int hi, int r0, int nb,
                                              int64_t si, int64_t sj, int half, int lane,
                                              T (&v)[8]) {
  const int kq = lane >> 4;
  const T *o = d.a + (int64_t)rb * si + (int64_t)(r0 + 32 * half) * sj;
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(o), (short)0, 0x7fffffff, 0x00020000);
  const bool row_ok = rb + (lane & 15) < hi;
  const uint32_t base = row_ok ? (uint32_t)(((lane & 15) * si + kq * sj) * (int64_t)sizeof(T)) : kPast;
  const uint32_t step = (uint32_t)(4 * sj * (int64_t)sizeof(T));
#pragma unroll
  for (int t = 0; t < 8; ++t)
    v[t] = buf_load<T>(rs, 32 * half + 4 * t + kq < nb ? base : kPast, (uint32_t)t * step);
}

// A diagonal-block inverse (64 x 64 f64 at ib) into Is in the sub-block
// order of inv_src, by LDS DMA: 16 bytes a lane, two consecutive entries of
// one sub-block column. Completes with the wave's vector-memory count.
__device__ __forceinline__ void dma_inverse(const double *ib, bool upper, double *Is, int wave,
                                            int lane) {
  const __amdgpu_buffer_rsrc_t ir = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<double *>(ib), (short)0, kNB * kNB * (int)sizeof(double), 0x00020000);
#pragma unroll
  for (int u = 0; u < (kInvLds / 2 + kThreads - 1) / kThreads; ++u) {
    const int p0 = (u * kWaves + wave) * 64;  // first entry pair of this wave's share
    if (p0 < kInvLds / 2)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          ir, (lds_void_ptr)(Is + 2 * p0), 16,
          (uint32_t)inv_src(upper, 2 * (p0 + lane)) * (uint32_t)sizeof(double), 0, 0, 0);
  }
}

// Row block of a unit's accumulator j in wave w (substitution order h =
// w + kWaves j, mirrored for a backward solve).
__device__ __forceinline__ int row_block(bool upper, int nrb, int w, int j) {
  return upper ? nrb - 1 - (w + kWaves * j) : w + kWaves * j;
}

// Rows 16 oi .. 16 oi + 15 of X = inv B for one diagonal block (B staged in
// Bb, its inverse in Is), transposed accumulator layout: f64 MFMA over the
// sub-blocks; the diagonal sub-block as a full 16 x 16 product when its B
// rows are all finite (the inverse's exact zeros outside the triangle add
// exact zeros), else its triangle alone, scalar f64, so a non-finite entry
// reaches exactly the X entries rblas's substitution reaches.
template <int CB>
__device__ __forceinline__ void form_x(const double *Bb, int kS, const double *Is, bool upper,
                                       int oi, int nb, int lane, f64x4 (&x)[CB]) {
  const int li = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int cb = 0; cb < CB; ++cb) x[cb] = f64x4{0, 0, 0, 0};
  if (16 * oi >= nb) return;
#pragma unroll
  for (int jb = 0; jb < 4; ++jb) {
    const bool use = (upper ? jb > oi : jb < oi) && 16 * jb < nb;  // wave-uniform
    if (use) {
      const double *sb = Is + sub_slot(upper, oi, jb);
      double iq[4];
#pragma unroll
      for (int t4 = 0; t4 < 4; ++t4) iq[t4] = sb[li + 16 * (4 * t4 + kq)];
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4)
          x[cb] = mma64(Bb[(16 * jb + 4 * t4 + kq) * kS + 16 * cb + li], iq[t4], x[cb]);
    }
  }
  const double *sd = Is + sub_slot(upper, oi, oi);
  double pd[4], qd[CB][4];
  bool fin = true;
#pragma unroll
  for (int t4 = 0; t4 < 4; ++t4) {
    const int j = 4 * t4 + kq;
    pd[t4] = sd[li + 16 * j];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      qd[cb][t4] = Bb[(16 * oi + j) * kS + 16 * cb + li];
      fin = fin && __builtin_isfinite(qd[cb][t4]);
    }
  }
  if (__all(fin)) {
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
      for (int t4 = 0; t4 < 4; ++t4) x[cb] = mma64(qd[cb][t4], pd[t4], x[cb]);
  } else {
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int c = 16 * cb + kq + 4 * r;
        double sum = x[cb][r];
#pragma unroll
        for (int jl = 0; jl < 16; ++jl) {
          const int j = 16 * oi + jl;
          const bool on = (upper ? jl >= li : jl <= li) && j < nb;
          const double w = on ? sd[li + 16 * jl] : 0.0;
          const double xv = on ? Bb[j * kS + c] : 0.0;
          sum = on ? __builtin_fma(w, xv, sum) : sum;
        }
        x[cb][r] = sum;
      }
  }
}

// Diagonal block solved at step st.
__device__ __forceinline__ int diag_block(bool forward, int nblk, int st) {
  return forward ? st : nblk - 1 - st;
}

// Applies X_s to accumulator a (16 rows at row block start rb) from the
// op(A) loads in flight in `cur` (CB = 1: the whole row block; CB = 2: its
// first half), loading the next row block (rn, when `more`) into `nxt`.
template <typename T, int CB>
__device__ __forceinline__ void apply_x(const TrsmDesc<T> &d, f64x4 (&a)[CB],
                                        T (&cur)[CB == 1 ? 16 : 8], T (&nxt)[CB == 1 ? 16 : 8],
                                        const double (&xq)[16], const double *Xs, int kS, int rb,
                                        int rn, bool more, int hi, int r0, int nb, int64_t si,
                                        int64_t sj, int lane) {
  if constexpr (CB == 1) {
    if (more) load_rows_buf(d, rn, hi, r0, nb, si, sj, lane, nxt);
    a[0] = update_acc(a[0], cur, xq);
  } else {
    load_half_buf(d, rb, hi, r0, nb, si, sj, 1, lane, nxt);
    update_half(a, cur, Xs, kS, 0, lane);
    if (more) load_half_buf(d, rn, hi, r0, nb, si, sj, 0, lane, cur);
    update_half(a, nxt, Xs, kS, 1, lane);
  }
}

// A unit's solve split over several workgroups (single calls whose units
// would leave most CUs idle), one per part: part p holds the row blocks of
// diagonal blocks [bbeg, bend) (in substitution order). For the blocks
// before its own it waits for *flag and reads -X_b from xst instead of
// forming it; then it solves its own blocks and, when later parts exist,
// publishes each -X_b it forms, f64, at xst + b * kNB * W (row-major, W
// doubles a row), and raises *flag to b + 1. *flag counts the leading blocks
// published: a part publishes only after it has read every earlier block,
// so the parts write it in turn. Every row gets the same f64 operations in
// the same order as in one workgroup.
// Data and flag go through relaxed agent-scope atomics, which are performed
// at the level coherent across XCDs (stores written through to it, loads
// served from it): the owner waves wait for their data stores to complete
// (a workgroup-scope release) before the barrier after which thread 0
// stores the flag, and the reader issues its data loads only after its poll
// has returned the new flag. Agent-scope release / acquire fences instead
// would write back or invalidate the XCD's whole L2 on every block (and
// with it the op(A) rows the units share there): with a release on the flag
// store alone, q = nrhs = 1024 took 0.119 ms against 0.112.
// flag null: no split.
struct Split {
  int bbeg, bend;
  int *flag;
  double *xst;
};
constexpr Split kNoSplit{0, 1 << 24, nullptr, nullptr};  // bend: past any order

// Bound on the polls of a workgroup for one block (each after a short
// sleep), so that every wave ends even if the block is never published; -X
// then reads as NaN, which the result carries.
constexpr int kSplitPolls = 1 << 24;

// -X_b of an earlier part: waits for the flag, then reads the 16 rows of
// sub-block oi into Xn in the layout the owners write.
template <int CB>
__device__ __forceinline__ void fetch_x(const Split &sp, int b, int oi, int lane, double *Xn,
                                        int kS) {
  constexpr int W = 16 * CB;
  bool ok = true;
  if (lane == 0) {
    int polls = 0;
    while (__hip_atomic_load(sp.flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < b + 1) {
      __builtin_amdgcn_s_sleep(2);
      if (++polls == kSplitPolls) {
        ok = false;
        break;
      }
    }
  }
  ok = __shfl(ok ? 1 : 0, 0) != 0;
  const int li = lane & 15, kq = lane >> 4;
  const int row = 16 * oi + li;
  double *src = sp.xst + (int64_t)b * kNB * W + (int64_t)row * W;
  double v[CB][4];
#pragma unroll
  for (int cb = 0; cb < CB; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      v[cb][r] = __hip_atomic_load(src + 16 * cb + kq + 4 * r, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
  for (int cb = 0; cb < CB; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      Xn[row * kS + 16 * cb + kq + 4 * r] = ok ? v[cb][r] : __builtin_nan("");
}

// Units of 16 CB right-hand sides (CB = 1 or 2; orders up to 16 kWaves RB)
// held in registers instead of LDS, as f64 MFMA accumulators. Row blocks (16
// rows) are numbered h = 0, 1, ... in substitution order -- h = g for a
// forward solve, nrb - 1 - g for a backward one (g the row block, nrb the
// padded count) -- so that step s solves h in [4s, 4s + 4) (diagonal block
// D_s) and updates h in [4s + 4, nrb). Row block h belongs to wave
// h % kWaves; its j-th row block (h = wave + kWaves j) is held transposed in
// CB accumulators: element r of lane l is B(16 g + (l & 15),
// 16 cb + (l >> 4) + 4 r). An update is acc += (-X_s)^T op(A)[rows, block]^T
// -- the X fragments are the product's first operand, the streamed op(A)
// rows its second -- so B passes through LDS only for the 64 rows of a
// diagonal block. The accumulators rotate: acc[0] is always the wave's first
// row block not yet solved, and a solved one moves to the back, so every
// index below is a constant.
//
// Look-ahead by one block: the four waves owning D_{s+1} apply X_s to those
// rows first and stage them, so that X_{s+1} is formed while the other
// waves are still applying X_s to the rest. Iteration s = -1 .. nblk - 2:
//   A. every wave applies X_s to acc[0] (for the owners of D_{s+1}, their
//      D_{s+1} rows), loading acc[1]'s op(A) rows; the owners stage acc[0]
//      in Bb (rows at or past q as zeros);
//   -- barrier (Bb staged; inv_{s+1} landed);
//   C. inv_{s+2} starts for its Is buffer (LDS DMA); the owners form
//      X_{s+1} = inv_{s+1} B_{s+1} (f64 MFMA; the scalar triangle for a
//      diagonal sub-block with a non-finite B entry), keep it in acc[0] and
//      publish -X_{s+1} in the other Xs buffer; every wave applies X_s to
//      acc[1 ..], loading each row block's op(A) one block ahead (X
//      fragments in registers for CB = 1, read from Xs per product for
//      CB = 2), then starts the loads for its acc[0] of iteration s + 1;
//   -- barrier (X_{s+1} visible; X_s, Bb and inv_s free);
//   the owners rotate their solved acc[0] to the back.
// Neither barrier waits for the op(A) loads in flight.
template <typename T, int RB, int CB>
__device__ __forceinline__ void solve_unit_reg(const TrsmDesc<T> &d, int c0, const double *inv,
                                               double *Bs, double *Is0,
                                               const Split sp = kNoSplit) {
  static_assert(kWaves == 8, "owners of a diagonal block are waves 0-3 or 4-7");
  constexpr int W = 16 * CB;
  constexpr int kS = W + 1;                 // Bb / Xs row stride in doubles
  constexpr int kLoads = CB == 1 ? 16 : 8;  // op(A) loads per row block (or half)
  // LDS: Bb, Xs of even and odd steps, the second inverse buffer.
  double *const Bb = Bs;
  double *const Is1 = Bs + 3 * kNB * kS;
  // Opaque copy of the thread id: in the persistent loop every per-lane
  // offset is recomputed per unit instead of hoisted out of it (and spilled).
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  const int lane = t & 63, wave = uni(t >> 6);
  const int q = d.q;
  const int nc = min(W, d.nrhs - c0);
  const int qpad = (q + kNB - 1) / kNB * kNB;
  const int nrb = qpad / 16;
  const bool upper = d.upper, fwd = d.forward;
  const int li = lane & 15, kq = lane >> 4;
  const int nblk = qpad / kNB;
  // This workgroup's row blocks h in [hbeg, hend) (all of them unless split),
  // and the last diagonal block it solves plus one.
  const int hbeg = 4 * sp.bbeg;
  const int hend = min(nrb, 4 * sp.bend);
  const int send = min(nblk, sp.bend);
  const bool publish = sp.flag != nullptr && sp.bend < nblk;  // later parts wait for it
  // Accumulators of this wave with h >= 4 st: h = hbeg + wave + kWaves j < hend.
  auto live_from = [hbeg, hend](int st, int w, int) {
    const int first = hbeg + w;
    if (first >= hend) return 0;
    const int lo = 4 * st > first ? (4 * st - first + kWaves - 1) / kWaves : 0;
    return (hend - first + kWaves - 1) / kWaves - lo;
  };
  const int nown = live_from(0, wave, nrb);  // this wave's row blocks
  if (hbeg == 0)
    dma_inverse(inv + d.inv0 + (int64_t)diag_block(fwd, nblk, 0) * kNB * kNB, upper, Is0, wave,
                lane);
  // B, alpha applied to every entry on load (rblas's pre-scale); row blocks
  // past the order are neither loaded nor stored.
  const int64_t rstep = d.right ? d.ldb : 1, cstep = d.right ? 1 : d.ldb;
  T *const bl = d.b + (int64_t)li * rstep + (int64_t)(c0 + kq) * cstep;
  f64x4 acc[RB][CB];
  {
    T raw[RB][CB][4];
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      const int g = row_block(upper, nrb, hbeg + wave, j);
      const bool live = j < nown && 16 * g < q;
      const bool rok = 16 * g + li < q;
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          raw[j][cb][r] = T(0);
          if (live && rok && 16 * cb + kq + 4 * r < nc)
            raw[j][cb][r] = bl[(int64_t)(16 * g) * rstep + (int64_t)(16 * cb + 4 * r) * cstep];
        }
    }
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      const int g = row_block(upper, nrb, hbeg + wave, j);
      const bool rok = j < nown && 16 * g + li < q;
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const T v = d.alpha != T(1) ? d.alpha * raw[j][cb][r] : raw[j][cb][r];
          acc[j][cb][r] = (rok && 16 * cb + kq + 4 * r < nc) ? (double)v : 0.0;
        }
    }
  }
  const int64_t si = d.trans ? d.lda : 1;  // op(A) row step
  const int64_t sj = d.trans ? 1 : d.lda;  // op(A) column step
  T pa[kLoads], pb[kLoads];
  double xq[16];
  int done = 0;  // rotations so far: acc[k] holds the wave's row block j = k + done
  for (int s = -1; s + 1 < send; ++s) {  // the last block's step has nothing left to update
    const bool upd = s >= 0;
    const int r0 = upd ? diag_block(fwd, nblk, s) * kNB : 0;
    const int nb = upd ? min(kNB, q - r0) : 0;
    const int hi = upper ? r0 : q;
    const int nact = upd ? live_from(s + 1, wave, nrb) : 0;  // accumulators X_s reaches
    // Owners of D_{s+1}: h = 4s + 4 + i4, this wave's acc[0]. For a block of
    // an earlier part, waves 0-3 read its published X instead (one sub-block
    // each).
    const bool next = s + 1 < nblk;
    const bool own_next = next && 4 * s + 4 >= hbeg;
    const int i4 = own_next ? wave - ((4 * s + 4 - hbeg) % kWaves) : wave;
    const bool owner = own_next && i4 >= 0 && i4 < 4;
    const bool fetcher = next && !own_next && wave < 4;
    const int oi = upper ? 3 - i4 : i4;  // sub-block within D_{s+1}
    const int r0n = next ? diag_block(fwd, nblk, s + 1) * kNB : 0;
    const int nbn = next ? min(kNB, q - r0n) : 0;
    const double *Xs = Bs + (1 + (s & 1)) * kNB * kS;
    if (CB == 1 && upd) {
#pragma unroll
      for (int t4 = 0; t4 < 16; ++t4) xq[t4] = Xs[(4 * t4 + kq) * kS + li];
    }
    // A.
    bool more = false;
    if (nact > 0) {
      more = nact > 1;
      apply_x<T, CB>(d, acc[0], pa, pb, xq, Xs, kS, 16 * row_block(upper, nrb, hbeg + wave, done),
                     16 * row_block(upper, nrb, hbeg + wave, done + 1), more, hi, r0, nb, si, sj, lane);
    }
    if (owner) {
      const int row = 16 * oi + li;
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          Bb[row * kS + 16 * cb + kq + 4 * r] = r0n + row < q ? acc[0][cb][r] : 0.0;
    }
    if (more)
      barrier_vm<kLoads>();  // B_{s+1} staged; inv_{s+1} landed
    else
      barrier_vm<0>();
    // C.
    if (s + 2 < send && 4 * (s + 2) >= hbeg)
      dma_inverse(inv + d.inv0 + (int64_t)diag_block(fwd, nblk, s + 2) * kNB * kNB, upper,
                  (s & 1) ? Is1 : Is0, wave, lane);
    double *Xn = Bs + (1 + ((s + 1) & 1)) * kNB * kS;
    if (owner) {
      f64x4 x[CB];
      form_x<CB>(Bb, kS, (s & 1) ? Is0 : Is1, upper, oi, nbn, lane, x);
      const int row = 16 * oi + li;
      double *pub = publish ? sp.xst + (int64_t)(s + 1) * kNB * W + (int64_t)row * W : nullptr;
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        acc[0][cb] = x[cb];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double v = row < nbn ? -x[cb][r] : 0.0;
          Xn[row * kS + 16 * cb + kq + 4 * r] = v;
          if (publish)
            __hip_atomic_store(pub + 16 * cb + kq + 4 * r, v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      // -X_{s+1} stored before this wave reaches the barrier ahead of the flag.
      if (publish) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
    if (fetcher) fetch_x<CB>(sp, s + 1, oi, lane, Xn, kS);
#pragma unroll
    for (int k = 1; k < RB; ++k) {
      if (k < nact) {
        const int rb = 16 * row_block(upper, nrb, hbeg + wave, done + k);
        const int rn = 16 * row_block(upper, nrb, hbeg + wave, done + k + 1);
        // CB = 1 alternates the two sets; CB = 2 always starts from pa.
        if (CB == 1 && (k & 1))
          apply_x<T, CB>(d, acc[k], pb, pa, xq, Xs, kS, rb, rn, k + 1 < nact, hi, r0, nb, si, sj,
                         lane);
        else
          apply_x<T, CB>(d, acc[k], pa, pb, xq, Xs, kS, rb, rn, k + 1 < nact, hi, r0, nb, si, sj,
                         lane);
      }
    }
    // Loads for acc[0] of iteration s + 1 (after this iteration's rotation).
    const int first = done + (owner ? 1 : 0);
    bool inflight = false;
    if (s + 1 < nblk && live_from(s + 2, wave, nrb) > 0) {
      const int r0x = diag_block(fwd, nblk, s + 1) * kNB;
      if constexpr (CB == 1)
        load_rows_buf(d, 16 * row_block(upper, nrb, hbeg + wave, first), upper ? r0x : q, r0x,
                      min(kNB, q - r0x), si, sj, lane, pa);
      else
        load_half_buf(d, 16 * row_block(upper, nrb, hbeg + wave, first), upper ? r0x : q, r0x,
                      min(kNB, q - r0x), si, sj, 0, lane, pa);
      inflight = true;
    }
    if (inflight)
      barrier_vm<kLoads>();  // X_{s+1} visible
    else
      barrier_vm<0>();
    if (publish && own_next && t == 0)  // the owners formed X_{s+1}; their stores completed
      __hip_atomic_store(sp.flag, s + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (owner) {  // acc[0] solved: rotate it to the back
      f64x4 tmp[CB];
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) tmp[cb] = acc[0][cb];
#pragma unroll
      for (int k = 0; k + 1 < RB; ++k)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) acc[k][cb] = acc[k + 1][cb];
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) acc[RB - 1][cb] = tmp[cb];
      ++done;
    }
  }
  // After `done` rotations acc[k] holds the wave's row block j = (k + done) % RB.
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const int j = (k + done) % RB;
    const int g = row_block(upper, nrb, hbeg + wave, j);
    const bool rok = j < nown && 16 * g + li < q;
    if (j < nown && 16 * g < q) {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (rok && 16 * cb + kq + 4 * r < nc)
            bl[(int64_t)(16 * g) * rstep + (int64_t)(16 * cb + 4 * r) * cstep] = (T)acc[k][cb][r];
    }
  }
}

// Widths 32 and 16 through the register-resident solve, width 8 through the
// LDS one (trsm_unit_width picks).
template <typename T>
__device__ __forceinline__ void solve_any(const TrsmDesc<T> &d, int c0, const double *inv,
                                          double *Bs, double *Is,
                                          const Split sp = kNoSplit) {
  constexpr int RB = kTrsmLdsDoubles / 16 / (16 * kWaves);
  if (d.width == 32)
    solve_unit_reg<T, kTrsmWideMaxOrder / (16 * kWaves), 2>(d, c0, inv, Bs, Is, sp);
  else if (d.width == 16)
    solve_unit_reg<T, RB, 1>(d, c0, inv, Bs, Is, sp);
  else
    solve_unit<T, 8>(d, c0, inv, Bs, Is);
}

// Persistent: each workgroup claims units from its XCD's queue (the
// planner keeps a system's units in one queue, so they meet op(A) in one
// L2), then helps a queue that has fallen behind. The last workgroup out
// resets the counters, so the launch can be replayed.
constexpr int kStealMin = 4;
template <typename T>
__global__ __launch_bounds__(kThreads) void grouped_trsm_kernel(const TrsmDesc<T> *descs,
                                                                const int2 *units, int *sched,
                                                                const double *inv) {
  __shared__ __align__(16) double Bs[kLdsDoubles];
  __shared__ __align__(16) double Is[kInvLds];
  __shared__ int s_unit;
  const int home = (int)(blockIdx.x % kTrsmQueues);
  // Queue offsets and whether this workgroup's own queue has run dry, in LDS
  // (read by thread 0 only; kept out of the solve's registers).
  __shared__ int qs[kTrsmQueues + 1];
  __shared__ int home_open;
  if (threadIdx.x == 0) {
    for (int x = 0; x <= kTrsmQueues; ++x) qs[x] = sched[kTrsmQueueStart + x];
    home_open = 1;
  }
  for (;;) {
    if (threadIdx.x == 0) {
      int u = -1;
      if (home_open) {
        const int c = atomicAdd(&sched[home * kTrsmLine], 1);
        if (c < qs[home + 1] - qs[home])
          u = qs[home] + c;
        else
          home_open = 0;
      }
      // Own queue drained: read every counter at once and help the queue
      // furthest behind, if it still holds kStealMin units (fewer are left to
      // its own workgroups -- every queue has some, the grid being at least
      // kTrsmQueues -- which keeps the drained tail from stealing units their
      // owners are about to claim). A lost race re-reads; each loss means
      // that queue fell below the threshold, so this ends.
      for (int round = 0; u < 0 && round < kTrsmQueues; ++round) {
        int left[kTrsmQueues];
#pragma unroll
        for (int x = 0; x < kTrsmQueues; ++x)
          left[x] = qs[x + 1] - qs[x] -
                    __hip_atomic_load(&sched[x * kTrsmLine], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
        int pick = 0;
#pragma unroll
        for (int x = 1; x < kTrsmQueues; ++x)
          if (left[x] > left[pick]) pick = x;
        if (left[pick] < kStealMin) break;
        const int c = atomicAdd(&sched[pick * kTrsmLine], 1);
        if (c < qs[pick + 1] - qs[pick]) u = qs[pick] + c;
      }
      s_unit = u;
    }
    __syncthreads();
    const int u = uni(s_unit);
    __syncthreads();
    if (u < 0) break;
    const int2 w = units[u];
    const int x = uni(w.x);
    TrsmDesc<T> d = descs[trsm_unit_problem(x)];
    if (x < 0) d.width = 16;
    solve_any(d, uni(w.y), inv, Bs, Is);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&sched[kTrsmQueues * kTrsmLine], 1) == (int)gridDim.x - 1)
      for (int x = 0; x <= kTrsmQueues; ++x) atomicExch(&sched[x * kTrsmLine], 0);
  }
}

// One problem passed by value: workgroup u solves right-hand sides
// u*width .. u*width + width - 1. Split (parts > 1): workgroup u + p units
// solves part p of unit u (diagonal blocks [p nblk / parts, (p + 1) nblk /
// parts)), with flags[u] and the published X of unit u at
// xst + u * nblk * kNB * width; earlier parts have the lower ids, so they are
// dispatched first.
template <typename T>
__global__ __launch_bounds__(kThreads) void single_trsm_solve(TrsmDesc<T> d, const double *inv,
                                                              int units, int parts, int *flags,
                                                              double *xst) {
  __shared__ __align__(16) double Bs[kLdsDoubles];
  __shared__ __align__(16) double Is[kInvLds];
  const int u = (int)blockIdx.x % units, part = (int)blockIdx.x / units;
  Split sp = kNoSplit;
  if (parts > 1) {
    const int nblk = (d.q + kNB - 1) / kNB;
    sp = Split{part * nblk / parts, (part + 1) * nblk / parts, flags + u,
               xst + (int64_t)u * nblk * kNB * d.width};
  }
  solve_any(d, u * d.width, inv, Bs, Is, sp);
}

}  // namespace

template <typename T>
hipError_t launch_trsm_single(const TrsmDesc<T> &d, double *inv, hipStream_t st) {
  if (d.q <= 0 || d.nrhs <= 0) return hipSuccess;
  const int nblk = (d.q + kNB - 1) / kNB;
  const int units = (d.nrhs + d.width - 1) / d.width;
  const int parts = trsm_split_parts(d.q, d.nrhs, d.width);
  int *flags = nullptr;
  double *xst = nullptr;
  if (parts > 1) {
    xst = inv + (int64_t)nblk * kNB * kNB;
    flags = (int *)(xst + trsm_split_x_doubles(d.q, d.nrhs));
  }
  hipLaunchKernelGGL((single_trsm_inv<T>), dim3((unsigned)nblk), dim3(kInvThreadsSingle), 0, st, d, inv,
                     flags, parts > 1 ? units : 0);
  hipLaunchKernelGGL((single_trsm_solve<T>), dim3((unsigned)(parts * units)), dim3(kThreads), 0,
                     st, d, (const double *)inv, units, parts, flags, xst);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_trsm_invert(const TrsmDesc<T> &d, double *inv, hipStream_t st) {
  if (d.q <= 0) return hipSuccess;
  const int nblk = (d.q + kNB - 1) / kNB;
  hipLaunchKernelGGL((single_trsm_inv<T>), dim3((unsigned)nblk), dim3(kInvThreadsSingle), 0, st, d, inv,
                     (int *)nullptr, 0);
  return hipGetLastError();
}

template hipError_t launch_trsm_invert<float>(const TrsmDesc<float> &, double *, hipStream_t);
template hipError_t launch_trsm_invert<double>(const TrsmDesc<double> &, double *, hipStream_t);
template hipError_t launch_trsm_single<float>(const TrsmDesc<float> &, double *, hipStream_t);
template hipError_t launch_trsm_single<double>(const TrsmDesc<double> &, double *, hipStream_t);

template <typename T>
int grouped_trsm_grid() {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grouped_trsm_kernel<T>, kThreads, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return per_cu * kCUs;
}

template <typename T>
hipError_t launch_grouped_trsm(const TrsmDesc<T> *descs, const int2 *blocks, int nblocks,
                               const int2 *units, int nunits, int grid, int *sched, double *inv,
                               hipStream_t st) {
  if (nblocks > 0)
    hipLaunchKernelGGL((grouped_trsm_inv<T>), dim3((unsigned)nblocks), dim3(kInvThreads), 0, st, descs,
                       blocks, inv);
  if (nunits > 0) {
    // Every queue needs a workgroup of its own (units below the stealing
    // threshold are left to it); surplus workgroups find nothing and leave.
    if (grid > nunits) grid = nunits;
    if (grid < kTrsmQueues) grid = kTrsmQueues;
    hipLaunchKernelGGL((grouped_trsm_kernel<T>), dim3((unsigned)grid), dim3(kThreads), 0, st,
                       descs, units, sched, (const double *)inv);
  }
  return hipGetLastError();
}

template int grouped_trsm_grid<float>();
template int grouped_trsm_grid<double>();
template hipError_t launch_grouped_trsm<float>(const TrsmDesc<float> *, const int2 *, int,
                                               const int2 *, int, int, int *, double *,
                                               hipStream_t);
template hipError_t launch_grouped_trsm<double>(const TrsmDesc<double> *, const int2 *, int,
                                                const int2 *, int, int, int *, double *,
                                                hipStream_t);

}  // namespace rblas
