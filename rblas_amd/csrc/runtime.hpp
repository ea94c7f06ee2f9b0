// Device context shared by the C-ABI entry points: per-thread error text, a
// pool of (stream, grow-only device buffers) used by the host-slice entry
// points, and helpers that stage one stored operand between host and HBM.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "validate.hpp"

namespace rblas {

void set_error(const std::string &msg);

// Stream-ordered device allocation from the current device's default memory
// pool, which is configured once to keep freed memory cached (no release to
// the OS between calls). Free with hipFreeAsync on the same stream.
hipError_t pool_alloc(void **ptr, size_t bytes, hipStream_t stream);
int32_t fail(const Check &c, rblas_error_t *err);
int32_t hip_fail(hipError_t e, const char *what);

// One host-slice call's device workspace. Buffers only grow; a context is
// owned by one call at a time.
struct Ctx {
  hipStream_t stream = nullptr;
  void *buf[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t cap[4] = {0, 0, 0, 0};
  hipError_t ensure(int slot, size_t bytes);
};

// RAII checkout of a context from the process-wide pool.
class CtxLease {
 public:
  CtxLease();
  ~CtxLease();
  Ctx *get() const { return ctx_; }
  hipError_t status() const { return status_; }

 private:
  Ctx *ctx_ = nullptr;
  hipError_t status_ = hipSuccess;
};

// A stored column-major rows x cols operand with leading dimension ld, as
// seen by the host-slice entry points. When ld >= rows it is staged densely
// (device ld = rows) with a 2-D copy; otherwise (columns overlap, which the
// NN path of rblas allows) its linear extent is copied with the same ld.
struct Staged {
  size_t rows = 0, cols = 0, ld = 0;  // host view
  size_t dev_ld = 0;
  size_t elems = 0;  // device elements
  bool dense = true;
  void plan(size_t r, size_t c, size_t l) {
    rows = r;
    cols = c;
    ld = l;
    if (r == 0 || c == 0) {
      elems = 0;
      dev_ld = r > 0 ? r : 1;
      return;
    }
    dense = l >= r;
    dev_ld = dense ? r : l;
    elems = dense ? r * c : (c - 1) * l + r;
  }
};

template <typename T>
hipError_t stage_in(const Staged &s, const T *host, T *dev, hipStream_t st) {
  if (s.elems == 0) return hipSuccess;
  if (s.dense)
    return hipMemcpy2DAsync(dev, s.dev_ld * sizeof(T), host, s.ld * sizeof(T), s.rows * sizeof(T),
                            s.cols, hipMemcpyHostToDevice, st);
  return hipMemcpyAsync(dev, host, s.elems * sizeof(T), hipMemcpyHostToDevice, st);
}

template <typename T>
hipError_t stage_out(const Staged &s, const T *dev, T *host, hipStream_t st) {
  if (s.elems == 0) return hipSuccess;
  if (s.dense)
    return hipMemcpy2DAsync(host, s.ld * sizeof(T), dev, s.dev_ld * sizeof(T), s.rows * sizeof(T),
                            s.cols, hipMemcpyDeviceToHost, st);
  return hipMemcpyAsync(host, dev, s.elems * sizeof(T), hipMemcpyDeviceToHost, st);
}

}  // namespace rblas

#define RBLAS_HIP_TRY(expr, what)                      \
  do {                                                 \
    hipError_t e_ = (expr);                            \
    if (e_ != hipSuccess) return rblas::hip_fail(e_, what); \
  } while (0)
