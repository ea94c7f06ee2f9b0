// Grouped triangular solve: many independent rblas trsm problems
// (level3.rs:390-520) in one persistent launch, after one batched launch that
// inverts every 64 x 64 diagonal block of every op(A).
//
// Every problem is solved as a Left problem: X op(A) = alpha B (Right) is
// op(A)^T X^T = alpha B^T, so a Right problem becomes Left with op(A)
// transposed (trans flipped, triangle flipped) on B^T. A work unit is one
// system and W consecutive right-hand sides (columns of B for Left, rows of
// B for Right). The unit keeps its W right-hand sides in LDS as f64
// (q x W, q the order of A, q * W <= 16384), then walks the diagonal blocks
// in substitution order:
//   X_b = inv_b * B_b   (f64 MFMA over the strictly off-diagonal 16 x 16
//                        sub-blocks of inv_b, scalar f64 over the triangle of
//                        each diagonal sub-block, so no structural zero ever
//                        multiplies a B entry: a non-finite entry reaches
//                        exactly what rblas's substitution reaches)
//   B_rest -= op(A)[rest, b] * X_b   (f64 MFMA, op(A) streamed from HBM/L2)
// and writes the right-hand sides back. All arithmetic is f64 (f32 operands
// are widened on load and rounded once on store).
#include "grouped.hpp"

namespace rblas {
namespace {

constexpr int kThreads = 256;       // 4 waves
constexpr int kNB = 64;             // diagonal block
constexpr int kLdsDoubles = kTrsmLdsDoubles;  // q * W limit (128 KiB)

typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f64x4 mma64(double p, double q, f64x4 acc) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(p, q, acc, 0, 0, 0);
}

template <typename T>
__device__ __forceinline__ double opa(const TrsmDesc<T> &d, int64_t r, int64_t c) {
  return (double)(d.trans ? d.a[c + r * d.lda] : d.a[r + c * d.lda]);
}

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Inverse of diagonal block `blk` of op(A) (as the problem is solved: Left
// form), f64, column-major 64 x 64 at inv + d.inv0 + blk * 4096. Thread j
// builds column j by substitution against e_j.
template <typename T>
__global__ __launch_bounds__(kNB) void grouped_trsm_inv(const TrsmDesc<T> *descs,
                                                        const int2 *blocks, double *inv) {
  __shared__ double L[kNB][kNB + 1];
  __shared__ double X[kNB][kNB + 1];
  const int2 u = blocks[blockIdx.x];
  const TrsmDesc<T> d = descs[u.x];
  const int64_t r0 = (int64_t)u.y * kNB;
  const int nb = (int)min((int64_t)kNB, (int64_t)d.q - r0);
  const int t = threadIdx.x;
  for (int idx = t; idx < kNB * kNB; idx += kNB) {
    const int r = idx % kNB, c = idx / kNB;
    double v = 0.0;
    const bool in_tri = d.upper ? r <= c : r >= c;
    if (r < nb && c < nb && in_tri) v = (d.unit && r == c) ? 1.0 : opa(d, r0 + r, r0 + c);
    L[r][c] = v;
    X[r][c] = 0.0;
  }
  __syncthreads();
  const int j = t;
  if (j < nb) {
    X[j][j] = 1.0 / L[j][j];
    if (!d.upper) {
      for (int i = j + 1; i < nb; ++i) {
        double s = 0.0;
        for (int kk = j; kk < i; ++kk) s = __builtin_fma(L[i][kk], X[kk][j], s);
        X[i][j] = -s / L[i][i];
      }
    } else {
      for (int i = j - 1; i >= 0; --i) {
        double s = 0.0;
        for (int kk = i + 1; kk <= j; ++kk) s = __builtin_fma(L[i][kk], X[kk][j], s);
        X[i][j] = -s / L[i][i];
      }
    }
  }
  __syncthreads();
  double *out = inv + d.inv0 + (int64_t)u.y * kNB * kNB;
  for (int idx = t; idx < kNB * kNB; idx += kNB) out[idx] = X[idx % kNB][idx / kNB];
}

// Address of right-hand side c, row r of the unit's B (Left: B(r, c0 + c);
// Right: B(c0 + c, r)).
template <typename T>
__device__ __forceinline__ int64_t b_index(const TrsmDesc<T> &d, int64_t r, int64_t c) {
  return d.right ? c + r * d.ldb : r + c * d.ldb;
}

// Bs: the kernel's LDS block of kLdsDoubles doubles, used as Bs[r * W + c].
template <typename T, int W>
__device__ void solve_unit(const TrsmDesc<T> &d, int c0, const double *inv, double *Bs) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int q = d.q;
  const int nc = min(W, d.nrhs - c0);
  const int qpad = (q + kNB - 1) / kNB * kNB;
  // Load (alpha applied to every entry, as rblas's pre-scale: 0 * NaN = NaN).
  for (int e = t; e < qpad * W; e += kThreads) {
    const int r = e / W, c = e % W;
    double v = 0.0;
    if (r < q && c < nc) {
      v = (double)d.b[b_index(d, r, c0 + c)];
      if (d.alpha != T(1)) v = (double)(d.alpha * (T)v);
    }
    Bs[e] = v;
  }
  __syncthreads();
  const int nblk = qpad / kNB;
  const int col = lane & 15;          // MFMA output column / Q-operand column
  const int kq = lane >> 4;           // k within a 4-step
  for (int s = 0; s < nblk; ++s) {
    const int b = d.forward ? s : nblk - 1 - s;
    const int r0 = b * kNB;
    const int nb = min(kNB, q - r0);
    const double *ib = inv + d.inv0 + (int64_t)b * kNB * kNB;
    // X_b = inv_b * B_b: wave w owns rows 16w .. 16w+15 of the block.
    f64x4 acc = {0, 0, 0, 0};
    const int iw = wave;
    if (16 * iw < nb) {
      // Off-diagonal 16 x 16 sub-blocks: j-blocks before iw (lower) or after
      // (upper); unrolled, the loads of a sub-block are issued together.
#pragma unroll
      for (int jb = 0; jb < 4; ++jb) {
        const bool use = (d.upper ? jb > iw : jb < iw) && 16 * jb < nb;  // wave-uniform
        if (use) {
          double pv[4], qv[4];
#pragma unroll
          for (int t4 = 0; t4 < 4; ++t4) {
            const int j = 16 * jb + 4 * t4 + kq;
            pv[t4] = ib[(int64_t)j * kNB + 16 * iw + (lane & 15)];
            qv[t4] = (col < W) ? Bs[(r0 + j) * W + col] : 0.0;
          }
#pragma unroll
          for (int t4 = 0; t4 < 4; ++t4) acc = mma64(pv[t4], qv[t4], acc);
        }
      }
      // Diagonal sub-block: only the triangle, scalar f64 (predicated, so
      // every lane runs the same 16 steps).
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int il = kq + 4 * r;  // row within the sub-block (MFMA output layout)
        const int i = 16 * iw + il;
        double sum = acc[r];
#pragma unroll
        for (int jl = 0; jl < 16; ++jl) {
          const int j = 16 * iw + jl;
          const bool on = (d.upper ? jl >= il : jl <= il) && j < nb && col < W;
          const double w = on ? ib[(int64_t)j * kNB + i] : 0.0;
          const double x = on ? Bs[(r0 + j) * W + col] : 0.0;
          sum = on ? __builtin_fma(w, x, sum) : sum;
        }
        acc[r] = sum;
      }
    }
    __syncthreads();  // every wave has read B_b
    if (16 * iw < nb && col < W) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * iw + kq + 4 * r;
        if (i < nb) Bs[(r0 + i) * W + col] = acc[r];
      }
    }
    __syncthreads();  // X_b visible
    // B_rest -= op(A)[rest, b] * X_b, rest = rows after the block (lower) or
    // before it (upper), in 16-row blocks spread over the waves. Lane l reads
    // op(A)(rb + l%16, r0 + 4t + l/16) for t = 0..15: all 16 loads of a
    // row-block are issued before its 16 MFMAs.
    const int lo = d.upper ? 0 : r0 + nb, hi = d.upper ? r0 : q;
    const int64_t si = d.trans ? d.lda : 1;  // op(A) row step
    const int64_t sj = d.trans ? 1 : d.lda;  // op(A) column step
    double xq[16];
#pragma unroll
    for (int t4 = 0; t4 < 16; ++t4) {
      const int j = 4 * t4 + kq;
      xq[t4] = (col < W && j < nb) ? Bs[(r0 + j) * W + col] : 0.0;
    }
    for (int rb = lo + 16 * wave; rb < hi; rb += 64) {
      const int64_t ar = rb + (lane & 15);
      const bool row_ok = ar < hi;
      const T *ap = d.a + ar * si + (int64_t)(r0 + kq) * sj;
      double pv[16];
#pragma unroll
      for (int t4 = 0; t4 < 16; ++t4) {
        const int j = 4 * t4 + kq;
        pv[t4] = (row_ok && j < nb) ? (double)ap[(int64_t)(4 * t4) * sj] : 0.0;
      }
      f64x4 u = {0, 0, 0, 0};
#pragma unroll
      for (int t4 = 0; t4 < 16; ++t4) u = mma64(pv[t4], xq[t4], u);
      if (col < W) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = rb + kq + 4 * r;
          if (i < hi) Bs[i * W + col] -= u[r];
        }
      }
    }
    __syncthreads();
  }
  for (int e = t; e < q * W; e += kThreads) {
    const int r = e / W, c = e % W;
    if (c < nc) d.b[b_index(d, r, c0 + c)] = (T)Bs[e];
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void grouped_trsm_kernel(const TrsmDesc<T> *descs,
                                                                const int2 *units, int nunits,
                                                                int *sched, const double *inv) {
  __shared__ double Bs[kLdsDoubles];
  __shared__ int s_unit;
  for (;;) {
    if (threadIdx.x == 0) s_unit = atomicAdd(&sched[0], 1);
    __syncthreads();
    const int u = uni(s_unit);
    __syncthreads();
    if (u >= nunits) break;
    const int2 w = units[u];
    const TrsmDesc<T> d = descs[uni(w.x)];
    if (d.width == 16)
      solve_unit<T, 16>(d, uni(w.y), inv, Bs);
    else
      solve_unit<T, 8>(d, uni(w.y), inv, Bs);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&sched[1], 1) == (int)gridDim.x - 1) {
      atomicExch(&sched[0], 0);
      atomicExch(&sched[1], 0);
    }
  }
}

}  // namespace

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
    hipLaunchKernelGGL((grouped_trsm_inv<T>), dim3((unsigned)nblocks), dim3(kNB), 0, st, descs,
                       blocks, inv);
  if (nunits > 0) {
    if (grid > nunits) grid = nunits;
    hipLaunchKernelGGL((grouped_trsm_kernel<T>), dim3((unsigned)grid), dim3(kThreads), 0, st,
                       descs, units, nunits, sched, (const double *)inv);
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
