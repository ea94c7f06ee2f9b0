// Device helpers of the serving loop (serve.hip): operand generation and
// result hashing over lists of blocks, one launch each per batch.
#pragma once
#include "kernels.hpp"

namespace rblas {

constexpr int kGenChunk = 65536;  // elements per generation / hash work unit

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

// Hash of a rows x cols column-major block with leading dimension ld.
struct HashDesc {
  const void *ptr;
  int64_t rows, cols, ld;
  int elem_bytes;  // 4 or 8
  int pad;
};

// units: {descriptor, chunk of kGenChunk elements}. `out` (one uint64 per
// hash descriptor) must be zero before the launch.
hipError_t launch_gen(const GenDesc *descs, const int2 *units, int nunits, hipStream_t stream);
hipError_t launch_hash(const HashDesc *descs, const int2 *units, int nunits, uint64_t *out,
                       hipStream_t stream);

}  // namespace rblas
