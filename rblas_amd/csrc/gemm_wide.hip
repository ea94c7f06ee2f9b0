// Wide MFMA tiles for large single GEMMs: 128 x 256 C tiles, eight waves of
// 64 x 64 (four by four 16 x 16 f64 / f32 MFMA blocks each), so every LDS
// barrier and every fragment read feeds twice the MFMA work of the 128 x 128
// tile. The stages are larger (A 128 x BK, B 256 x BK), so this translation
// unit keeps three of them (53 KiB each, 159 KiB of the 160 KiB LDS), and
// its waves compute every block of an edge tile rather than skipping those
// past the edge (the k-loop copies that skipping needs cost more registers
// than the tile has to spare). The k order of every C entry is the one of
// the 128 x 128 and 64 x 64 tiles (tile_gemm.hpp), so the results are the
// same bits.
//
// Used (after the RowMajor swap) for f64 NN, TN and NT and for f32 NN and NT;
// see gemm_wide_shape_ok. Measured at 8192^3 (bench, HBM-resident operands):
// DGEMM NN 65.9 -> 68.4 TFLOP/s, TN 65.6 -> 67.1, NT 65.7 -> 67.6; SGEMM NN
// 128 -> 131.8, NT 129.5 -> 134.9; ZGEMM 4096^3 (4M, its real products
// here) 62.8 -> 66.2.
#define RBLAS_TILE_STAGES 3
#define RBLAS_TILE_MAX_B 256
#include "tile_gemm.hpp"

namespace rblas {
namespace {

constexpr int kWideM = 128, kWideN = 256;

template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(kThreads) void gemm_wide_kernel(GemmProblem<T> p) {
  TileGemm<T, TA, TB, kWideM, kWideN, false> tg;
  const int tiles_m = (int)((p.m + kWideM - 1) / kWideM);
  const int tiles_n = (int)((p.n + kWideN - 1) / kWideN);
  int tm, tn;
  tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * kWideM, n0 = tn * kWideN;
  tg.init(p, m0, n0, 0, (int)p.k);
  tg.mainloop();
  tg.epilogue(p, m0, n0);
}

}  // namespace

template <typename T>
bool gemm_wide_ok(const GemmProblem<T> &p) {
  if (p.tri != 0) return false;
  // The layouts where it measured faster at 8192^3 (TFLOP/s, against 128 x
  // 128 tiles): f64 NN 68.4 / 65.9, TN 67.1 / 65.6, NT 67.6 / 65.7 (TT 62.6
  // / 65.7: its kernel spills); f32 NN 131.9 / 128, NT 134.9 / 129.5 (TN
  // 125.6 / 130.2). So op(A) stored K-contiguous takes it only in f64 with
  // op(B) stored K-contiguous too.
  if (p.ta && (sizeof(T) == 4 || p.tb)) return false;
  const int64_t tiles = ((p.m + kWideM - 1) / kWideM) * ((p.n + kWideN - 1) / kWideN);
  if (tiles < 2 * kCUs) return false;  // fills the CUs at least twice
  // Largest in-tile byte offsets (31-bit buffer addressing): op(A) over 128
  // rows, B over 256 columns, C over 256 columns.
  const int64_t span_a = (int64_t)(kWideM - 1) * p.lda + kWideM + Geo<T>::BK;
  const int64_t span_b = (int64_t)(kWideN - 1) * p.ldb + Geo<T>::BK;
  const int64_t span_c = (int64_t)(kWideN - 1) * p.ldc + kWideM;
  const int64_t span = span_a > span_b ? (span_a > span_c ? span_a : span_c)
                                       : (span_b > span_c ? span_b : span_c);
  return span * (int64_t)sizeof(T) < 0x7fffffffll - 16;
}

template <typename T>
hipError_t launch_gemm_wide(const GemmProblem<T> &p, hipStream_t st) {
  if (!gemm_wide_ok(p)) return hipErrorInvalidValue;
  const int64_t tiles = ((p.m + kWideM - 1) / kWideM) * ((p.n + kWideN - 1) / kWideN);
  const dim3 grid((unsigned)tiles), block(kThreads);
  if (!p.ta && !p.tb)
    hipLaunchKernelGGL((gemm_wide_kernel<T, false, false>), grid, block, 0, st, p);
  else if (!p.ta)
    hipLaunchKernelGGL((gemm_wide_kernel<T, false, true>), grid, block, 0, st, p);
  else if (!p.tb)
    hipLaunchKernelGGL((gemm_wide_kernel<T, true, false>), grid, block, 0, st, p);
  else
    hipLaunchKernelGGL((gemm_wide_kernel<T, true, true>), grid, block, 0, st, p);
  return hipGetLastError();
}

template bool gemm_wide_ok<float>(const GemmProblem<float> &);
template bool gemm_wide_ok<double>(const GemmProblem<double> &);
template hipError_t launch_gemm_wide<float>(const GemmProblem<float> &, hipStream_t);
template hipError_t launch_gemm_wide<double>(const GemmProblem<double> &, hipStream_t);

}  // namespace rblas
