// Device helpers of the serving loop: synthetic operand generation in HBM
// and an order-independent hash of each result.
//
// Generation is the counter-based splitmix64 of rblas_amd/workloads.py, bit
// for bit: element i of a draw (seed s, offset o) is
// mix(s + (o + i + 1) * 0x9E3779B97F4A7C15); uniform(-1, 1) takes the top 24
// (f32) or 53 (f64) bits, u * 2^-bits, then 2u - 1 in the element type. A
// trsm matrix (mode 1, order n, column-major) keeps the lower triangle:
// off-diagonal u / sqrt(n) (in f64, then rounded), diagonal 2 + |u|, zeros
// above.
//
// The hash of a rows x cols column-major block is the sum mod 2^64 over its
// elements of mix(bits(x_e) ^ (e * golden)), e the column-major index: a sum,
// so the order in which blocks add their parts does not matter.
#include "serve.hpp"

namespace rblas {
namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
__device__ __forceinline__ T uniform_at(uint64_t seed, int64_t idx);
template <>
__device__ __forceinline__ float uniform_at<float>(uint64_t seed, int64_t idx) {
  const uint64_t z = mix(seed + (uint64_t)(idx + 1) * kGolden);
  const float u = (float)(uint32_t)(z >> 40) * 0x1p-24f;
  return 2.0f * u - 1.0f;
}
template <>
__device__ __forceinline__ double uniform_at<double>(uint64_t seed, int64_t idx) {
  const uint64_t z = mix(seed + (uint64_t)(idx + 1) * kGolden);
  const double u = (double)(z >> 11) * 0x1p-53;
  return 2.0 * u - 1.0;
}

template <typename T>
__device__ __forceinline__ void gen_range(const GenDesc &d, int64_t lo, int64_t hi) {
  T *p = (T *)d.ptr;
  for (int64_t e = lo + threadIdx.x; e < hi; e += blockDim.x) {
    const T u = uniform_at<T>(d.seed, d.offset + e);
    T v = u;
    if (d.mode == 1) {
      const int64_t i = e % d.n, j = e / d.n;
      if (i > j)
        v = (T)((double)u / __builtin_sqrt((double)d.n));
      else if (i == j)
        v = (T)(2.0 + __builtin_fabs((double)u));
      else
        v = T(0);
    }
    p[e] = v;
  }
}

__global__ __launch_bounds__(256) void gen_kernel(const GenDesc *descs, const int2 *units) {
  const int2 u = units[blockIdx.x];
  const GenDesc d = descs[u.x];
  const int64_t lo = (int64_t)u.y * kGenChunk;
  const int64_t hi = min(d.count, lo + (int64_t)kGenChunk);
  if (d.dtype == 0)
    gen_range<float>(d, lo, hi);
  else
    gen_range<double>(d, lo, hi);
}

__global__ __launch_bounds__(256) void hash_kernel(const HashDesc *descs, const int2 *units,
                                                   unsigned long long *out) {
  __shared__ unsigned long long part[256];
  const int2 u = units[blockIdx.x];
  const HashDesc d = descs[u.x];
  const int64_t total = d.rows * d.cols;
  const int64_t lo = (int64_t)u.y * kGenChunk;
  const int64_t hi = min(total, lo + (int64_t)kGenChunk);
  unsigned long long acc = 0;
  for (int64_t e = lo + threadIdx.x; e < hi; e += blockDim.x) {
    const int64_t i = e % d.rows, j = e / d.rows;
    const int64_t at = i + j * d.ld;
    uint64_t bits;
    if (d.elem_bytes == 4)
      bits = ((const uint32_t *)d.ptr)[at];
    else
      bits = ((const uint64_t *)d.ptr)[at];
    acc += mix(bits ^ ((uint64_t)e * kGolden));
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&out[u.x], part[0]);
}

}  // namespace

hipError_t launch_gen(const GenDesc *descs, const int2 *units, int nunits, hipStream_t st) {
  if (nunits <= 0) return hipSuccess;
  hipLaunchKernelGGL(gen_kernel, dim3((unsigned)nunits), dim3(256), 0, st, descs, units);
  return hipGetLastError();
}

hipError_t launch_hash(const HashDesc *descs, const int2 *units, int nunits, uint64_t *out,
                       hipStream_t st) {
  if (nunits <= 0) return hipSuccess;
  hipLaunchKernelGGL(hash_kernel, dim3((unsigned)nunits), dim3(256), 0, st, descs, units,
                     (unsigned long long *)out);
  return hipGetLastError();
}

}  // namespace rblas
