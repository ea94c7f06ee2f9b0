// Fused multiply-add in the precision of its operands. `__builtin_fma` is the
// double builtin: on floats it computes in f64 and rounds twice (and costs
// four conversions), which is not rblas's single-rounding f32 `mul_add`.
#pragma once
#include <hip/hip_runtime.h>

namespace rblas {

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

}  // namespace rblas
