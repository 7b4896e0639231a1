// mf_internal.hpp -- context, status handling, device workspace and timing
// shared by every translation unit of libmagforce.  Host-side only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "mf_capi.h"

namespace mf {

// Thread-local message of the last failure, returned by mf_last_error().
void set_thread_error(const std::string &msg);

// A status carrying its message.  Functions return Status; the C entry points
// turn it into an int and store the message in the context.
struct Status {
  int code = MF_OK;
  std::string msg;
  bool ok() const { return code == MF_OK; }
  static Status Ok() { return {}; }
  static Status Err(int c, std::string m) { return {c, std::move(m)}; }
};

#define MF_TRY(expr)                   \
  do {                                 \
    ::mf::Status _s = (expr);          \
    if (!_s.ok()) return _s;           \
  } while (0)

#define MF_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t _e = (call);                                                          \
    if (_e != hipSuccess)                                                            \
      return ::mf::Status::Err(MF_ERR_HIP, std::string(#call " failed: ") +         \
                                               hipGetErrorString(_e) + " at " +     \
                                               __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

// Checks the launch that was just enqueued.
#define MF_LAUNCH_CHECK() MF_HIP(hipGetLastError())

#define MF_ARG(cond, msg)                                              \
  do {                                                                 \
    if (!(cond)) return ::mf::Status::Err(MF_ERR_INVALID_ARG, (msg));  \
  } while (0)

// Owning device allocation.
// The allocation hook of mf_set_alloc_hook (fault injection); consulted by
// every DeviceBuffer growth.  Returns true when the hook refuses `bytes`.
bool alloc_refused(size_t bytes);

class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  DeviceBuffer(DeviceBuffer &&o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) {
    o.ptr_ = nullptr;
    o.bytes_ = 0;
  }
  DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
    if (this != &o) {
      release();
      ptr_ = o.ptr_;
      bytes_ = o.bytes_;
      o.ptr_ = nullptr;
      o.bytes_ = 0;
    }
    return *this;
  }
  ~DeviceBuffer() { release(); }

  // Grows to at least `bytes` (contents are not preserved on growth).
  Status reserve(size_t bytes) {
    if (bytes <= bytes_) return Status::Ok();
    release();
    size_t want = bytes < 256 ? 256 : bytes;
    if (alloc_refused(want))
      return Status::Err(MF_ERR_ALLOC, "allocation of " + std::to_string(want) +
                                           " bytes refused by the allocation hook");
    hipError_t e = hipMalloc(&ptr_, want);
    if (e != hipSuccess) {
      ptr_ = nullptr;
      return Status::Err(MF_ERR_ALLOC, "hipMalloc(" + std::to_string(want) +
                                           ") failed: " + hipGetErrorString(e));
    }
    bytes_ = want;
    return Status::Ok();
  }
  void release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
  }
  template <class T>
  T *as() const { return static_cast<T *>(ptr_); }
  void *get() const { return ptr_; }
  size_t bytes() const { return bytes_; }

 private:
  void *ptr_ = nullptr;
  size_t bytes_ = 0;
};

// Per-kernel-family timing with HIP events (only when enabled).
struct TimingRecord {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double ms = 0.0;
  int64_t launches = 0;
};

}  // namespace mf

struct mf_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  // Second stream (non-blocking, default priority) on which the FMM products
  // run their near field beside the far field; forked from and joined to
  // `stream` by these events within every product.
  hipStream_t aux_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::string last_error;
  bool timing = false;
  std::map<std::string, mf::TimingRecord> timers;
  std::vector<hipEvent_t> event_pool;
  // Named workspaces, grown on demand and kept for the context's lifetime.
  std::map<std::string, mf::DeviceBuffer> ws;
  // Long-lived per-context objects (operator tables), keyed by name.
  std::map<std::string, std::shared_ptr<void>> cache;
  // Pinned host staging for scalar read-backs.
  double *host_scalars = nullptr;
  int num_cu = 256;
};

namespace mf {

// Returns a device workspace of at least `bytes` for `name`.
Status workspace(mf_ctx *ctx, const std::string &name, size_t bytes, void **out);

// Records timing events around a kernel family when timing is enabled, on
// the stream the family is launched on (the context's stream by default).
class KernelTimer {
 public:
  KernelTimer(mf_ctx *ctx, const char *name);
  KernelTimer(mf_ctx *ctx, const char *name, hipStream_t stream);
  ~KernelTimer();
  void stop();

 private:
  mf_ctx *ctx_;
  const char *name_;
  hipStream_t stream_ = nullptr;
  hipEvent_t start_ = nullptr, stop_ = nullptr;
};

// Fork/join of the context's auxiliary stream around one product: fork()
// makes it wait for the work issued so far on the context stream, record()
// marks the end of the work issued on it, join() makes the context stream
// wait for that.  The destructor joins on paths that return early, so work
// on the auxiliary stream never overtakes the context stream's next use of
// the same buffers.  Inside a stream capture the two events pull the
// auxiliary stream into the graph and back.
class NearOverlap {
 public:
  explicit NearOverlap(mf_ctx *ctx) : ctx_(ctx) {}
  NearOverlap(const NearOverlap &) = delete;
  NearOverlap &operator=(const NearOverlap &) = delete;
  ~NearOverlap() {
    if (!forked_ || joined_) return;
    if (!recorded_) (void)hipEventRecord(ctx_->ev_join, ctx_->aux_stream);
    (void)hipStreamWaitEvent(ctx_->stream, ctx_->ev_join, 0);
  }
  Status fork() {
    MF_HIP(hipEventRecord(ctx_->ev_fork, ctx_->stream));
    MF_HIP(hipStreamWaitEvent(ctx_->aux_stream, ctx_->ev_fork, 0));
    forked_ = true;
    return Status::Ok();
  }
  Status record() {
    MF_HIP(hipEventRecord(ctx_->ev_join, ctx_->aux_stream));
    recorded_ = true;
    return Status::Ok();
  }
  Status join() {
    if (!forked_) return Status::Ok();
    MF_HIP(hipStreamWaitEvent(ctx_->stream, ctx_->ev_join, 0));
    joined_ = true;
    return Status::Ok();
  }

 private:
  mf_ctx *ctx_;
  bool forked_ = false, recorded_ = false, joined_ = false;
};

// Looks up, once per device, the device address of the scratch array of
// mf_field_allpairs_noalloc (mf_noalloc.hip); `device` must be current.
Status noalloc_setup(int device);

// Collects finished event pairs into the totals (synchronizes on them).
Status flush_timers(mf_ctx *ctx);

// Turns a Status into the C return code and records the message.
int finish(mf_ctx *ctx, const Status &s);

inline unsigned grid_for(int64_t n, int block) {
  return static_cast<unsigned>((n + block - 1) / block);
}

}  // namespace mf
