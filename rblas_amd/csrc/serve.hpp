// Device helpers of the serving loop (serve.hip): operand generation and
// result hashing over lists of blocks, one launch each per batch.
#pragma once
#include "kernels.hpp"

namespace rblas {

constexpr int kGenChunk = 8192;  // elements per generation / hash / 4M work unit (32 per thread)

// Fill `count` elements at ptr: mode 0 uniform(-1, 1); mode 1 a trsm matrix
// of order n (lower triangle, see serve.hip).
struct GenDesc {
  void *ptr;
  int64_t count;
  int64_t n;
  uint64_t seed;
  int64_t offset;
  int dtype;  // 0 f32, 1 f64
  int mode;
};

// Hash of a rows x cols column-major block with leading dimension ld, over
// its columns 0, col_step, 2 col_step, ...
struct HashDesc {
  const void *ptr;
  int64_t rows, cols, ld;
  int elem_bytes;  // 4 or 8
  int col_step;    // >= 1
};

// Elements the hash of d visits.
__host__ __device__ inline int64_t hash_elems(const HashDesc &d) {
  return d.rows * ((d.cols + d.col_step - 1) / d.col_step);
}

// The splitmix64 finalizer (also the pool slice offsets of the serving loop).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// One complex request (NoTrans / NoTrans, dense column-major interleaved
// complex f64 operands): op(A) m x k at a, op(B) k x n at b, C m x n at c;
// ws holds the 4M planes [Ar | Ai] (m x 2k), [Br; -Bi], [Bi; Br] (2k x n
// each), P_re, P_im (m x n each), cgemm_workspace_elems(m, n, k) doubles.
struct CplxDesc {
  const double *a;
  const double *b;
  double *c;
  double *ws;
  int64_t m, n, k;
  double alpha_re, alpha_im, beta_re, beta_im;
};

// units: {descriptor, chunk of kGenChunk elements}. `out` (one uint64 per
// hash descriptor) must be zero before the launch.
hipError_t launch_gen(const GenDesc *descs, const int2 *units, int nunits, hipStream_t stream);
hipError_t launch_hash(const HashDesc *descs, const int2 *units, int nunits, uint64_t *out,
                       hipStream_t stream);
// The 4M planes of every listed complex request (split units cover each
// request's m k entries of op(A), then its k n of op(B)), and the final
// C = alpha P + beta C (combine units cover m n).
hipError_t launch_cplx_split(const CplxDesc *descs, const int2 *units, int nunits,
                             hipStream_t stream);
hipError_t launch_cplx_combine(const CplxDesc *descs, const int2 *units, int nunits,
                               hipStream_t stream);

}  // namespace rblas
