// Device context pool, error reporting and the device-memory/stream plumbing
// entry points of the C ABI.
#include "runtime.hpp"

#include <cstring>
#include <mutex>
#include <vector>

namespace rblas {

namespace {
thread_local std::string g_msg = "ok";
std::mutex g_pool_mu;
std::vector<Ctx *> g_pool;
}  // namespace

void set_error(const std::string &msg) { g_msg = msg; }

hipError_t pool_alloc(void **ptr, size_t bytes, hipStream_t stream) {
  static std::mutex mu;
  static std::vector<int> configured;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  {
    std::lock_guard<std::mutex> g(mu);
    bool done = false;
    for (int d : configured) done |= d == dev;
    if (!done) {
      hipMemPool_t pool;
      e = hipDeviceGetDefaultMemPool(&pool, dev);
      if (e != hipSuccess) return e;
      uint64_t keep = ~0ull;
      e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
      if (e != hipSuccess) return e;
      configured.push_back(dev);
    }
  }
  return hipMallocAsync(ptr, bytes, stream);
}

int32_t fail(const Check &c, rblas_error_t *err) {
  if (err) *err = c.err;
  if (c.status != RBLAS_OK) set_error(c.msg);
  return c.status;
}

int32_t hip_fail(hipError_t e, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? RBLAS_ALLOC_FAILED : RBLAS_HIP_ERROR;
}

hipError_t pool_release(void *ptr, hipStream_t stream, hipError_t prior) {
  const hipError_t e = ptr ? hipFreeAsync(ptr, stream) : hipSuccess;
  return prior != hipSuccess ? prior : e;
}

hipError_t Ctx::ensure(int slot, size_t bytes) {
  if (bytes <= cap[slot]) return hipSuccess;
  if (buf[slot]) {
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    e = hipFree(buf[slot]);
    buf[slot] = nullptr;
    cap[slot] = 0;
    if (e != hipSuccess) return e;
  }
  size_t want = bytes < 4096 ? 4096 : bytes;
  hipError_t e = hipMalloc(&buf[slot], want);
  if (e != hipSuccess) {
    buf[slot] = nullptr;
    return e;
  }
  cap[slot] = want;
  return hipSuccess;
}

hipError_t Ctx::ensure_pinned(size_t bytes) {
  if (bytes <= pinned_cap) return hipSuccess;
  if (pinned) {
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    e = hipHostFree(pinned);
    pinned = nullptr;
    pinned_cap = 0;
    if (e != hipSuccess) return e;
  }
  size_t want = bytes < (size_t(1) << 20) ? (size_t(1) << 20) : bytes;
  hipError_t e = hipHostMalloc((void **)&pinned, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    pinned = nullptr;
    return e;
  }
  pinned_cap = want;
  return hipSuccess;
}

namespace {

// Copies a rows x cols column-major operand (ld_src -> ld_dst, in elements
// of `elem` bytes) on the host.
template <typename E>
void copy_strided(E *dst, size_t ld_dst, const E *src, size_t ld_src, size_t cols) {
  for (size_t c = 0; c < cols; ++c) dst[c * ld_dst] = src[c * ld_src];
}

void copy_columns(char *dst, size_t ld_dst, const char *src, size_t ld_src, size_t rows,
                  size_t cols, size_t elem) {
  if (ld_dst == rows && ld_src == rows) {
    std::memcpy(dst, src, rows * cols * elem);
    return;
  }
  if (rows == 1 && elem == 4) {  // strided vectors: one element per column
    copy_strided((uint32_t *)dst, ld_dst, (const uint32_t *)src, ld_src, cols);
    return;
  }
  if (rows == 1 && elem == 8) {
    copy_strided((uint64_t *)dst, ld_dst, (const uint64_t *)src, ld_src, cols);
    return;
  }
  for (size_t c = 0; c < cols; ++c)
    std::memcpy(dst + c * ld_dst * elem, src + c * ld_src * elem, rows * elem);
}

}  // namespace

hipError_t HostXfer::in(const Staged &s, const void *host, void *dev) {
  if (s.elems == 0) return hipSuccess;
  const size_t bytes = s.elems * elem_;
  if (!s.pinned)  // contiguous (ld == rows) or the linear extent
    return hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx_.stream);
  char *pin = ctx_.pinned + s.pin_off;
  if (s.packed)
    copy_columns(pin, s.rows, (const char *)host, s.ld, s.rows, s.cols, elem_);
  else
    std::memcpy(pin, host, bytes);
  return hipMemcpyAsync(dev, pin, bytes, hipMemcpyHostToDevice, ctx_.stream);
}

hipError_t HostXfer::out(const Staged &s, const void *dev, void *host) {
  if (s.elems == 0) return hipSuccess;
  const size_t bytes = s.elems * elem_;
  if (!s.pinned) return hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx_.stream);
  if (nouts_ == 2) return hipErrorInvalidValue;
  outs_[nouts_++] = Pending{&s, host};
  return hipMemcpyAsync(ctx_.pinned + s.pin_off, dev, bytes, hipMemcpyDeviceToHost, ctx_.stream);
}

hipError_t HostXfer::finish() {
  hipError_t e = hipStreamSynchronize(ctx_.stream);
  if (e != hipSuccess) return e;
  for (int i = 0; i < nouts_; ++i) {
    const Staged &s = *outs_[i].s;
    const char *pin = ctx_.pinned + s.pin_off;
    if (s.packed)
      copy_columns((char *)outs_[i].host, s.ld, pin, s.rows, s.rows, s.cols, elem_);
    else
      std::memcpy(outs_[i].host, pin, s.elems * elem_);
  }
  nouts_ = 0;
  return hipSuccess;
}

CtxLease::CtxLease() {
  {
    std::lock_guard<std::mutex> g(g_pool_mu);
    if (!g_pool.empty()) {
      ctx_ = g_pool.back();
      g_pool.pop_back();
      return;
    }
  }
  ctx_ = new Ctx();
  status_ = hipStreamCreateWithFlags(&ctx_->stream, hipStreamNonBlocking);
  if (status_ != hipSuccess) {
    delete ctx_;
    ctx_ = nullptr;
  }
}

CtxLease::~CtxLease() {
  if (!ctx_) return;
  std::lock_guard<std::mutex> g(g_pool_mu);
  g_pool.push_back(ctx_);
}

}  // namespace rblas

using namespace rblas;

extern "C" {

const char *rblas_last_error_message(void) { return g_msg.c_str(); }

int32_t rblas_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int32_t rblas_malloc(void **ptr, size_t bytes) {
  if (!ptr) return fail(panic("rblas_malloc: null out-pointer"), nullptr);
  RBLAS_HIP_TRY(hipMalloc(ptr, bytes == 0 ? 1 : bytes), "hipMalloc");
  return RBLAS_OK;
}

int32_t rblas_free(void *ptr) {
  RBLAS_HIP_TRY(hipFree(ptr), "hipFree");
  return RBLAS_OK;
}

int32_t rblas_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream) {
  RBLAS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream),
                "hipMemcpyAsync(HtoD)");
  RBLAS_HIP_TRY(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  return RBLAS_OK;
}

int32_t rblas_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream) {
  RBLAS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream),
                "hipMemcpyAsync(DtoH)");
  RBLAS_HIP_TRY(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  return RBLAS_OK;
}

int32_t rblas_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream) {
  RBLAS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream),
                "hipMemcpyAsync(DtoD)");
  return RBLAS_OK;
}

int32_t rblas_stream_create(void **stream) {
  hipStream_t s;
  RBLAS_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  *stream = (void *)s;
  return RBLAS_OK;
}

int32_t rblas_stream_destroy(void *stream) {
  RBLAS_HIP_TRY(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
  return RBLAS_OK;
}

int32_t rblas_stream_synchronize(void *stream) {
  RBLAS_HIP_TRY(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  return RBLAS_OK;
}

int32_t rblas_event_create(void **event) {
  hipEvent_t e;
  RBLAS_HIP_TRY(hipEventCreate(&e), "hipEventCreate");
  *event = (void *)e;
  return RBLAS_OK;
}

int32_t rblas_event_destroy(void *event) {
  RBLAS_HIP_TRY(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
  return RBLAS_OK;
}

int32_t rblas_event_record(void *event, void *stream) {
  RBLAS_HIP_TRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
  return RBLAS_OK;
}

int32_t rblas_event_elapsed_ms(float *ms, void *start, void *stop) {
  RBLAS_HIP_TRY(hipEventSynchronize((hipEvent_t)stop), "hipEventSynchronize");
  RBLAS_HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop),
                "hipEventElapsedTime");
  return RBLAS_OK;
}

int32_t rblas_device_synchronize(void) {
  RBLAS_HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return RBLAS_OK;
}

}  // extern "C"
