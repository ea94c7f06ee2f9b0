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
// Returns `ptr` (from pool_alloc, or null) to the pool once the work enqueued
// on `stream` before it completes. The result is `prior` when that is an
// error (the launch's failure is the one to report), else the free's.
hipError_t pool_release(void *ptr, hipStream_t stream, hipError_t prior);
int32_t fail(const Check &c, rblas_error_t *err);
int32_t hip_fail(hipError_t e, const char *what);

// One host-slice call's device workspace and pinned host staging buffer.
// Buffers only grow; a context is owned by one call at a time.
struct Ctx {
  hipStream_t stream = nullptr;
  void *buf[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t cap[4] = {0, 0, 0, 0};
  char *pinned = nullptr;
  size_t pinned_cap = 0;
  hipError_t ensure(int slot, size_t bytes);
  hipError_t ensure_pinned(size_t bytes);
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

// Transfers of one host-slice call that together stay under this size go
// through the context's pinned buffer: one CPU copy (packing strided
// columns) and one DMA per operand, instead of pageable copies, which the
// runtime splits per column for strided layouts.
constexpr size_t kPinnedMax = size_t(64) << 20;

// A stored column-major rows x cols operand with leading dimension ld, as
// seen by the host-slice entry points, and how it is laid out on the device:
//   - packed (staged through the pinned buffer, or ld == rows): dense,
//     device ld = rows;
//   - linear (large, ld > rows): the host's linear extent (cols-1)*ld + rows
//     as one copy, device ld = ld;
//   - ld < rows (columns overlap, which the NN path of rblas allows): the
//     linear extent, device ld = ld.
struct Staged {
  size_t rows = 0, cols = 0, ld = 0;  // host view
  size_t dev_ld = 1;
  size_t elems = 0;       // device elements
  bool packed = true;     // device ld = rows, host columns copied one by one
  bool pinned = false;    // staged through the pinned buffer
  size_t pin_off = 0;     // byte offset in the pinned buffer
  void plan(size_t r, size_t c, size_t l, bool use_pinned) {
    rows = r;
    cols = c;
    ld = l;
    pinned = use_pinned;
    if (r == 0 || c == 0) {
      elems = 0;
      dev_ld = r > 0 ? r : 1;
      return;
    }
    packed = l >= r && (use_pinned || l == r);
    dev_ld = packed ? r : l;
    elems = packed ? r * c : (c - 1) * l + r;
  }
  size_t bytes(size_t elem) const { return elems * elem; }
};

// The staging of one host-slice call. Usage: plan every operand with
// pinned() as its mode, reserve(), then in()/out() in stream order, and
// finish() (waits for the stream, then unpacks pinned outputs).
class HostXfer {
 public:
  // elem: bytes per element; total_bytes: an upper bound of the bytes the
  // call stages (each operand counted once).
  HostXfer(Ctx &ctx, size_t elem, size_t total_bytes)
      : ctx_(ctx), pinned_(total_bytes > 0 && total_bytes <= kPinnedMax), elem_(elem) {}
  bool pinned() const { return pinned_; }
  // Allocates the pinned buffer for the planned operands (each at its own
  // offset, 256-byte aligned).
  template <typename... S>
  hipError_t reserve(S &...ops) {
    size_t off = 0;
    ((ops.pin_off = off, off += (ops.elems * elem_ + 255) / 256 * 256), ...);
    return pinned_ ? ctx_.ensure_pinned(off) : hipSuccess;
  }
  hipError_t in(const Staged &s, const void *host, void *dev);
  hipError_t out(const Staged &s, const void *dev, void *host);
  hipError_t finish();

 private:
  struct Pending {
    const Staged *s;
    void *host;
  };
  Ctx &ctx_;
  bool pinned_;
  size_t elem_;
  Pending outs_[2];
  int nouts_ = 0;
};

}  // namespace rblas

#define RBLAS_HIP_TRY(expr, what)                      \
  do {                                                 \
    hipError_t e_ = (expr);                            \
    if (e_ != hipSuccess) return rblas::hip_fail(e_, what); \
  } while (0)
