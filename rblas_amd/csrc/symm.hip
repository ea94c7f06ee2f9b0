// Full q x q matrix from one stored triangle, for symm / hemm
// (level3.rs:301-381 full_symmetric, complex3.rs:188-218 he_or_sy): rblas
// materializes the symmetric (or Hermitian) A densely and runs the product
// through gemm; here the dense copy is written in HBM by one launch and the
// product runs on the MFMA gemm (real) or the 4M complex gemm.
//
// out(i, j) = a(i, j) where (i, j) is in the stored `uplo` triangle, else
// a(j, i) -- conjugated for Hermitian A, whose diagonal keeps only its real
// part. Only stored entries are read. Each 32 x 32 output tile is one of:
//   - inside the stored triangle: a column-contiguous copy;
//   - inside the mirrored triangle: the transposed source tile is read
//     column-contiguously into LDS and written back transposed;
//   - on the diagonal: element by element.
#include "kernels.hpp"

namespace rblas {
namespace {

constexpr int kTile = 32;
constexpr int kRows = 8;  // block = 32 x 8 threads, 4 rows each

template <typename T>
struct Elem {  // one real, or one interleaved complex, element
  T v[2];
};

template <typename T, bool CPLX>
__device__ __forceinline__ Elem<T> load_elem(const T *a, int64_t idx) {
  Elem<T> e;
  if (CPLX) {
    e.v[0] = a[2 * idx];
    e.v[1] = a[2 * idx + 1];
  } else {
    e.v[0] = a[idx];
    e.v[1] = T(0);
  }
  return e;
}

template <typename T, bool CPLX>
__device__ __forceinline__ void store_elem(T *o, int64_t idx, Elem<T> e) {
  if (CPLX) {
    o[2 * idx] = e.v[0];
    o[2 * idx + 1] = e.v[1];
  } else {
    o[idx] = e.v[0];
  }
}

template <typename T, bool CPLX>
__global__ __launch_bounds__(kTile *kRows) void sym_full_kernel(const T *a, int64_t lda, int q,
                                                                int upper, int herm, T *out) {
  __shared__ Elem<T> s[kTile][kTile + 1];
  const int ti = blockIdx.x, tj = blockIdx.y;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = ti * kTile, j0 = tj * kTile;
  // Tile classes for the stored triangle i <= j (upper) or i >= j (lower).
  const bool stored_tile = upper ? ti < tj : ti > tj;
  const bool mirror_tile = upper ? ti > tj : ti < tj;
  if (stored_tile) {
    const int i = i0 + tx;
#pragma unroll
    for (int r = 0; r < kTile / kRows; ++r) {
      const int j = j0 + ty + kRows * r;
      if (i < q && j < q) store_elem<T, CPLX>(out, i + (int64_t)j * q,
                                              load_elem<T, CPLX>(a, i + (int64_t)j * lda));
    }
    return;
  }
  if (mirror_tile) {
    // Source tile: rows j0.., columns i0.. of A (stored), read along its rows.
#pragma unroll
    for (int r = 0; r < kTile / kRows; ++r) {
      const int sr = j0 + tx, sc = i0 + ty + kRows * r;
      if (sr < q && sc < q) s[tx][ty + kRows * r] = load_elem<T, CPLX>(a, sr + (int64_t)sc * lda);
    }
    __syncthreads();
    const int i = i0 + tx;
#pragma unroll
    for (int r = 0; r < kTile / kRows; ++r) {
      const int jl = ty + kRows * r, j = j0 + jl;
      if (i < q && j < q) {
        Elem<T> e = s[jl][tx];  // a(j, i)
        if (CPLX && herm) e.v[1] = -e.v[1];
        store_elem<T, CPLX>(out, i + (int64_t)j * q, e);
      }
    }
    return;
  }
  // Diagonal tile.
  const int i = i0 + tx;
#pragma unroll
  for (int r = 0; r < kTile / kRows; ++r) {
    const int j = j0 + ty + kRows * r;
    if (i >= q || j >= q) continue;
    const bool here = upper ? i <= j : i >= j;
    Elem<T> e = load_elem<T, CPLX>(a, here ? i + (int64_t)j * lda : j + (int64_t)i * lda);
    if (CPLX && herm) {
      if (i == j)
        e.v[1] = T(0);
      else if (!here)
        e.v[1] = -e.v[1];
    }
    store_elem<T, CPLX>(out, i + (int64_t)j * q, e);
  }
}

}  // namespace

template <typename T>
hipError_t launch_sym_full(const T *a, int64_t lda, int64_t q, bool upper, bool complex_elems,
                           bool herm, T *out, hipStream_t st) {
  if (q <= 0) return hipSuccess;
  const unsigned t = (unsigned)((q + kTile - 1) / kTile);
  const dim3 grid(t, t), block(kTile, kRows);
  if (complex_elems)
    hipLaunchKernelGGL((sym_full_kernel<T, true>), grid, block, 0, st, a, lda, (int)q,
                       (int)upper, (int)herm, out);
  else
    hipLaunchKernelGGL((sym_full_kernel<T, false>), grid, block, 0, st, a, lda, (int)q,
                       (int)upper, 0, out);
  return hipGetLastError();
}

template hipError_t launch_sym_full<float>(const float *, int64_t, int64_t, bool, bool, bool,
                                           float *, hipStream_t);
template hipError_t launch_sym_full<double>(const double *, int64_t, int64_t, bool, bool, bool,
                                            double *, hipStream_t);

}  // namespace rblas
