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
// elements (or those of every col_step-th column) of
// mix(bits(x_e) ^ (e * golden)), e the column-major index: a sum, so the
// order in which blocks add their parts does not matter.
#include "serve.hpp"

namespace rblas {
namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t mix(uint64_t z) { return mix64(z); }

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

// Row i and column j of flat index e = i + j * rows, for the first index a
// thread visits; later indices advance (i, j) by the block stride without a
// division (a 64-bit division is a long instruction sequence on the GPU).
__device__ __forceinline__ void split_index(int64_t e, int64_t rows, int64_t &i, int64_t &j) {
  j = e / rows;
  i = e - j * rows;
}
__device__ __forceinline__ void advance(int64_t step, int64_t rows, int64_t &i, int64_t &j) {
  i += step;
  while (i >= rows) {
    i -= rows;
    ++j;
  }
}

template <typename T>
__device__ __forceinline__ void gen_range(const GenDesc &d, int64_t lo, int64_t hi) {
  T *p = (T *)d.ptr;
  const int64_t e0 = lo + threadIdx.x;
  if (d.mode == 0) {
    for (int64_t e = e0; e < hi; e += blockDim.x) p[e] = uniform_at<T>(d.seed, d.offset + e);
    return;
  }
  int64_t i, j;
  split_index(e0, d.n, i, j);
  for (int64_t e = e0; e < hi; e += blockDim.x) {
    const T u = uniform_at<T>(d.seed, d.offset + e);
    T v;
    if (i > j)
      v = (T)((double)u / __builtin_sqrt((double)d.n));
    else if (i == j)
      v = (T)(2.0 + __builtin_fabs((double)u));
    else
      v = T(0);
    p[e] = v;
    advance(blockDim.x, d.n, i, j);
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

// Each thread hashes kHashPer elements (indices lo + threadIdx.x + 256 t):
// all their loads are issued before any is used, so a thread waits for
// memory once, not once per element. Element e of the visited sequence is
// row e % rows of column col_step * (e / rows) (32-bit division: a hashed
// block holds fewer than 2^31 elements).
constexpr int kHashPer = kGenChunk / 256;

__global__ __launch_bounds__(256) void hash_kernel(const HashDesc *descs, const int2 *units,
                                                   unsigned long long *out) {
  __shared__ unsigned long long part[256];
  const int2 u = units[blockIdx.x];
  const HashDesc d = descs[u.x];
  const int64_t total = hash_elems(d);
  const int64_t lo = (int64_t)u.y * kGenChunk;
  const uint32_t rows = (uint32_t)d.rows;
  uint64_t bits[kHashPer];
  uint64_t idx[kHashPer];
#pragma unroll
  for (int t = 0; t < kHashPer; ++t) {
    const int64_t e = lo + threadIdx.x + 256 * t;
    const uint32_t jj = (uint32_t)e / rows, i = (uint32_t)e - jj * rows;
    const int64_t j = (int64_t)jj * d.col_step;
    const int64_t at = e < total ? i + j * d.ld : 0;
    idx[t] = e < total ? (uint64_t)(i + j * d.rows) : ~0ull;
    if (d.elem_bytes == 4)
      bits[t] = ((const uint32_t *)d.ptr)[at];
    else
      bits[t] = ((const uint64_t *)d.ptr)[at];
  }
  unsigned long long acc = 0;
#pragma unroll
  for (int t = 0; t < kHashPer; ++t)
    if (idx[t] != ~0ull) acc += mix(bits[t] ^ (idx[t] * kGolden));
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&out[u.x], part[0]);
}

// Complex requests (complex3.rs:27-74 as 4M, NoTrans / NoTrans, dense
// column-major operands): the real planes of every complex request of a
// batch in one launch, before the batch's grouped real gemms compute
// P_re = [Ar | Ai][Br; -Bi] and P_im = [Ar | Ai][Bi; Br]; then one launch
// applies C = alpha P + (beta == 0 ? 0 : beta C) with rblas's complex
// multiply (complex.rs:101-111) to every request. Units are {request, chunk
// of kGenChunk elements}; a split request's chunks cover op(A)'s m k
// entries first, then op(B)'s k n.
__device__ __forceinline__ int64_t a_chunks(const CplxDesc &d) {
  return (d.m * d.k + kGenChunk - 1) / kGenChunk;
}

// The two kernels below move kGenChunk elements per workgroup in batches of
// kCplxBatch per thread, all of a batch's loads issued before its stores: a
// load after a store to a buffer the compiler cannot tell apart would
// otherwise wait for memory once per element.
constexpr int kCplxBatch = 8;

__global__ __launch_bounds__(256) void cplx_split_kernel(const CplxDesc *descs,
                                                         const int2 *units) {
  const int2 u = units[blockIdx.x];
  const CplxDesc d = descs[u.x];
  const int64_t mk = d.m * d.k, kn = d.k * d.n;
  double *ap = d.ws, *b1 = ap + 2 * mk, *b2 = b1 + 2 * kn;
  const int64_t ca = a_chunks(d);
  if (u.y < ca) {  // [Ar | Ai]: entry e = i + p m of op(A)
    const int64_t lo = (int64_t)u.y * kGenChunk, hi = min(mk, lo + (int64_t)kGenChunk);
    for (int64_t b0 = lo + threadIdx.x; b0 < hi; b0 += 256 * kCplxBatch) {
      double re[kCplxBatch], im[kCplxBatch];
#pragma unroll
      for (int t = 0; t < kCplxBatch; ++t) {
        const int64_t e = min(b0 + 256 * t, hi - 1);
        re[t] = d.a[2 * e];
        im[t] = d.a[2 * e + 1];
      }
#pragma unroll
      for (int t = 0; t < kCplxBatch; ++t) {
        const int64_t e = b0 + 256 * t;
        if (e < hi) {
          ap[e] = re[t];
          ap[e + mk] = im[t];
        }
      }
    }
    return;
  }
  // [Br; -Bi] and [Bi; Br] (2k x n, ld 2k): entry e = p + j k of op(B)
  // (32-bit division: k n < 2^31).
  const int64_t lo = (u.y - ca) * kGenChunk, hi = min(kn, lo + (int64_t)kGenChunk);
  const uint32_t k = (uint32_t)d.k;
  for (int64_t b0 = lo + threadIdx.x; b0 < hi; b0 += 256 * kCplxBatch) {
    double re[kCplxBatch], im[kCplxBatch];
#pragma unroll
    for (int t = 0; t < kCplxBatch; ++t) {
      const int64_t e = min(b0 + 256 * t, hi - 1);
      re[t] = d.b[2 * e];
      im[t] = d.b[2 * e + 1];
    }
#pragma unroll
    for (int t = 0; t < kCplxBatch; ++t) {
      const int64_t e = b0 + 256 * t;
      if (e >= hi) continue;
      const uint32_t j = (uint32_t)e / k, p = (uint32_t)e - j * k;
      const int64_t col = (int64_t)j * 2 * d.k;
      b1[col + p] = re[t];
      b1[col + d.k + p] = -im[t];
      b2[col + p] = im[t];
      b2[col + d.k + p] = re[t];
    }
  }
}

__global__ __launch_bounds__(256) void cplx_combine_kernel(const CplxDesc *descs,
                                                           const int2 *units) {
  const int2 u = units[blockIdx.x];
  const CplxDesc d = descs[u.x];
  const int64_t mk = d.m * d.k, kn = d.k * d.n, mn = d.m * d.n;
  const double *pr = d.ws + 2 * mk + 4 * kn, *pi = pr + mn;
  const bool beta_zero = d.beta_re == 0.0 && d.beta_im == 0.0;
  const int64_t lo = (int64_t)u.y * kGenChunk, hi = min(mn, lo + (int64_t)kGenChunk);
  for (int64_t b0 = lo + threadIdx.x; b0 < hi; b0 += 256 * kCplxBatch) {
    double p_re[kCplxBatch], p_im[kCplxBatch], c_re[kCplxBatch], c_im[kCplxBatch];
#pragma unroll
    for (int t = 0; t < kCplxBatch; ++t) {
      const int64_t e = min(b0 + 256 * t, hi - 1);
      p_re[t] = pr[e];
      p_im[t] = pi[e];
      c_re[t] = beta_zero ? 0.0 : d.c[2 * e];
      c_im[t] = beta_zero ? 0.0 : d.c[2 * e + 1];
    }
#pragma unroll
    for (int t = 0; t < kCplxBatch; ++t) {
      const int64_t e = b0 + 256 * t;
      if (e >= hi) continue;
      double s_re = 0.0, s_im = 0.0;
      if (!beta_zero) {
        s_re = __builtin_fma(d.beta_re, c_re[t], -(d.beta_im * c_im[t]));
        s_im = __builtin_fma(d.beta_re, c_im[t], d.beta_im * c_re[t]);
      }
      const double q_re = __builtin_fma(d.alpha_re, p_re[t], -(d.alpha_im * p_im[t]));
      const double q_im = __builtin_fma(d.alpha_re, p_im[t], d.alpha_im * p_re[t]);
      d.c[2 * e] = q_re + s_re;
      d.c[2 * e + 1] = q_im + s_im;
    }
  }
}

}  // namespace

hipError_t launch_cplx_split(const CplxDesc *descs, const int2 *units, int nunits,
                             hipStream_t st) {
  if (nunits <= 0) return hipSuccess;
  hipLaunchKernelGGL(cplx_split_kernel, dim3((unsigned)nunits), dim3(256), 0, st, descs, units);
  return hipGetLastError();
}

hipError_t launch_cplx_combine(const CplxDesc *descs, const int2 *units, int nunits,
                               hipStream_t st) {
  if (nunits <= 0) return hipSuccess;
  hipLaunchKernelGGL(cplx_combine_kernel, dim3((unsigned)nunits), dim3(256), 0, st, descs, units);
  return hipGetLastError();
}

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
