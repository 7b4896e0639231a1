// mf_capi.cpp -- extern "C" entry points (include/mf_capi.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <atomic>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "mf_allpairs.hpp"
#include "mf_capi.h"
#include "mf_internal.hpp"
#include "mf_plan.hpp"
#include "mf_vector.hpp"

namespace mf {

namespace {
thread_local std::string t_last_error;
}

void set_thread_error(const std::string &msg) { t_last_error = msg; }

int finish(mf_ctx *ctx, const Status &s) {
  if (!s.ok()) {
    if (ctx) ctx->last_error = s.msg;
    set_thread_error(s.msg);
  }
  return s.code;
}

Status workspace(mf_ctx *ctx, const std::string &name, size_t bytes, void **out) {
  DeviceBuffer &b = ctx->ws[name];
  if (b.bytes() < bytes) {
    // A workspace can be in use by work already enqueued on the stream.
    MF_HIP(hipStreamSynchronize(ctx->stream));
    MF_TRY(b.reserve(bytes + bytes / 8));
  }
  *out = b.get();
  return Status::Ok();
}

KernelTimer::KernelTimer(mf_ctx *ctx, const char *name) : KernelTimer(ctx, name, ctx->stream) {}

KernelTimer::KernelTimer(mf_ctx *ctx, const char *name, hipStream_t stream)
    : ctx_(ctx), name_(name), stream_(stream) {
  if (!ctx_->timing) return;
  auto take = [&]() {
    hipEvent_t e = nullptr;
    if (!ctx_->event_pool.empty()) {
      e = ctx_->event_pool.back();
      ctx_->event_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    return e;
  };
  start_ = take();
  stop_ = take();
  if (start_ && stop_) (void)hipEventRecord(start_, stream_);
}

void KernelTimer::stop() {
  if (!ctx_ || !start_ || !stop_) return;
  (void)hipEventRecord(stop_, stream_);
  ctx_->timers[name_].pending.emplace_back(start_, stop_);
  start_ = stop_ = nullptr;
}

KernelTimer::~KernelTimer() { stop(); }

Status flush_timers(mf_ctx *ctx) {
  for (auto &kv : ctx->timers) {
    for (auto &pr : kv.second.pending) {
      MF_HIP(hipEventSynchronize(pr.second));
      float ms = 0.f;
      MF_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
      kv.second.ms += ms;
      kv.second.launches += 1;
      ctx->event_pool.push_back(pr.first);
      ctx->event_pool.push_back(pr.second);
    }
    kv.second.pending.clear();
  }
  return Status::Ok();
}

// ---- plans -----------------------------------------------------------------

namespace {
__global__ void to_f64_kernel(int64_t n, const SrcRec *__restrict__ rec, double *__restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SrcRec r = rec[i];
  out[i] = r.x;
  out[n + i] = r.y;
  out[2 * n + i] = r.z;
}
}  // namespace

Status plan_set_positions(mf_plan *plan, int64_t n, const float *x, const float *y,
                          const float *z) {
  mf_ctx *ctx = plan->ctx;
  MF_ARG(n >= 0, "n must be >= 0");
  MF_ARG(n == 0 || (x && y && z), "positions must be non-null");
  MF_ARG(plan->nsys == 1 || n == plan->sys_offsets.back(),
         "a batch plan keeps its particle count (offsets[nsys])");
  plan->ready = false;
  plan->n = n;
  plan->pos64_valid = false;
  const int64_t np = padded_sources(n);
  MF_HIP(hipStreamSynchronize(ctx->stream));
  MF_TRY(plan->rec.reserve(sizeof(SrcRec) * static_cast<size_t>(np > 0 ? np : 1)));
  MF_TRY(plan->boxes.reserve(sizeof(float4) * 2 * static_cast<size_t>(np / kTile + 1)));
  MF_TRY(pack_sources(ctx, n, x, y, z, nullptr, nullptr, nullptr, 0.f, plan->rec.as<SrcRec>()));
  MF_TRY(tile_boxes(ctx, n, plan->rec.as<SrcRec>(), plan->boxes.as<float4>()));
  if (plan->opts.mode == MF_MODE_FMM) MF_TRY(fmm_build(plan));
  plan->ready = true;
  return Status::Ok();
}

Status plan_pos64(mf_plan *plan, const double **px, const double **py, const double **pz) {
  const int64_t n = plan->n;
  if (!plan->pos64_valid) {
    MF_HIP(hipStreamSynchronize(plan->ctx->stream));
    MF_TRY(plan->pos64.reserve(sizeof(double) * 3 * static_cast<size_t>(n > 0 ? n : 1)));
    if (n > 0) {
      hipLaunchKernelGGL(to_f64_kernel, dim3(grid_for(n, 256)), dim3(256), 0, plan->ctx->stream, n,
                         plan->rec.as<SrcRec>(), plan->pos64.as<double>());
      MF_LAUNCH_CHECK();
    }
    plan->pos64_valid = true;
  }
  double *p = plan->pos64.as<double>();
  *px = p;
  *py = p + n;
  *pz = p + 2 * n;
  return Status::Ok();
}

Status plan_apply_G(mf_plan *plan, float scale, const float *px, const float *py,
                    const float *pz, const float *hconst, float *ox, float *oy, float *oz) {
  if (plan->opts.mode == MF_MODE_FMM)
    return fmm_apply_G(plan, scale, px, py, pz, hconst, ox, oy, oz);
  mf_ctx *ctx = plan->ctx;
  SrcRec *rec = plan->rec.as<SrcRec>();
  MF_TRY(pack_moments(ctx, plan->n, px, py, pz, scale, rec));
  return field_allpairs_launch(ctx, reinterpret_cast<const float4 *>(rec), 2, plan->n, rec,
                               plan->boxes.as<float4>(), plan->n, plan->fc, hconst, ox, oy, oz);
}

Status plan_force(mf_plan *plan, const float *mx, const float *my, const float *mz,
                  const float *const *text9, float *fx, float *fy, float *fz) {
  if (plan->opts.mode == MF_MODE_FMM) return fmm_force(plan, mx, my, mz, text9, fx, fy, fz);
  mf_ctx *ctx = plan->ctx;
  SrcRec *rec = plan->rec.as<SrcRec>();
  MF_TRY(pack_moments(ctx, plan->n, mx, my, mz, 1.f, rec));
  return force_allpairs_launch(ctx, rec, plan->boxes.as<float4>(), plan->n, plan->hc, text9, fx,
                               fy, fz);
}

namespace {
// mf_set_alloc_hook's hook and its argument, swapped together under a lock
std::mutex g_hook_mu;
mf_alloc_hook g_hook = nullptr;
void *g_hook_user = nullptr;
std::atomic<bool> g_hook_set{false};
}  // namespace

// Options of the solve entry points: tolerances finite and not negative,
// precision codes in range.
Status check_solve_opts(const mf_solve_opts *o) {
  if (!o) return Status::Ok();
  MF_ARG(std::isfinite(o->rtol) && o->rtol >= 0.0, "rtol must be finite and >= 0");
  MF_ARG(std::isfinite(o->inner_rtol) && o->inner_rtol >= 0.0,
         "inner_rtol must be finite and >= 0");
  MF_ARG(std::isfinite(o->relax_rtol) && o->relax_rtol >= 0.0,
         "relax_rtol must be finite and >= 0");
  MF_ARG(o->inner_m2l_precision >= 0 && o->inner_m2l_precision <= MF_M2L_BF16X1_CUT + 1,
         "inner_m2l_precision out of range");
  MF_ARG(o->relax_m2l_precision >= 0 && o->relax_m2l_precision <= MF_M2L_BF16X1_CUT + 1,
         "relax_m2l_precision out of range");
  return Status::Ok();
}

bool alloc_refused(size_t bytes) {
  if (!g_hook_set.load(std::memory_order_acquire)) return false;
  std::lock_guard<std::mutex> lock(g_hook_mu);
  return g_hook != nullptr && g_hook(static_cast<uint64_t>(bytes), g_hook_user) != 0;
}

}  // namespace mf

using mf::Status;

extern "C" {

int mf_abi_version(void) { return MF_ABI_VERSION; }

int mf_set_alloc_hook(mf_alloc_hook hook, void *user) {
  std::lock_guard<std::mutex> lock(mf::g_hook_mu);
  mf::g_hook = hook;
  mf::g_hook_user = user;
  mf::g_hook_set.store(hook != nullptr, std::memory_order_release);
  return MF_OK;
}

const char *mf_status_string(int status) {
  switch (status) {
    case MF_OK: return "ok";
    case MF_ERR_INVALID_ARG: return "invalid argument";
    case MF_ERR_HIP: return "HIP runtime error";
    case MF_ERR_ALLOC: return "device allocation failed";
    case MF_ERR_NOT_CONVERGED: return "solver did not converge";
    case MF_ERR_UNSUPPORTED: return "unsupported";
    case MF_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

const char *mf_last_error(void) { return mf::t_last_error.c_str(); }

int mf_ctx_create(int device, mf_ctx **out) {
  if (!out) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "out is null"));
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0)
    return mf::finish(nullptr, Status::Err(MF_ERR_NO_DEVICE, "no HIP device available"));
  if (device < 0 || device >= count)
    return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "device index out of range"));
  auto run = [&]() -> Status {
    MF_HIP(hipSetDevice(device));
    // address of the no-alloc entry point's scratch on this device
    MF_TRY(mf::noalloc_setup(device));
    mf_ctx *ctx = new (std::nothrow) mf_ctx();
    if (!ctx) return Status::Err(MF_ERR_ALLOC, "out of host memory");
    ctx->device = device;
    hipError_t err = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (err != hipSuccess) {
      delete ctx;
      return Status::Err(MF_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(err));
    }
    ctx->stream = ctx->own_stream;
    err = hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming);
    if (err == hipSuccess)
      err = hipHostMalloc(reinterpret_cast<void **>(&ctx->host_scalars), sizeof(double) * 128);
    if (err != hipSuccess) {
      if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
      if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
      if (ctx->aux_stream) (void)hipStreamDestroy(ctx->aux_stream);
      (void)hipStreamDestroy(ctx->own_stream);
      delete ctx;
      return Status::Err(MF_ERR_ALLOC, std::string("context set-up: ") + hipGetErrorString(err));
    }
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cu > 0)
      ctx->num_cu = cu;
    *out = ctx;
    return Status::Ok();
  };
  return mf::finish(nullptr, run());
}

int mf_ctx_destroy(mf_ctx *ctx) {
  if (!ctx) return MF_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->timers)
    for (auto &pr : kv.second.pending) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  ctx->ws.clear();
  if (ctx->host_scalars) (void)hipHostFree(ctx->host_scalars);
  if (ctx->aux_stream) {
    (void)hipStreamSynchronize(ctx->aux_stream);
    (void)hipStreamDestroy(ctx->aux_stream);
  }
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return MF_OK;
}

const char *mf_ctx_last_error(const mf_ctx *ctx) { return ctx ? ctx->last_error.c_str() : ""; }

int mf_ctx_get_default(mf_ctx **out) {
  static std::mutex mu;
  static std::vector<mf_ctx *> per_device;
  if (!out) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "out is null"));
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return mf::finish(nullptr, Status::Err(MF_ERR_NO_DEVICE, "no current HIP device"));
  std::lock_guard<std::mutex> lock(mu);
  if (static_cast<int>(per_device.size()) <= dev) per_device.resize(dev + 1, nullptr);
  if (!per_device[dev]) {
    mf_ctx *ctx = nullptr;
    int rc = mf_ctx_create(dev, &ctx);
    if (rc != MF_OK) return rc;
    ctx->stream = nullptr;  // HIP null stream
    per_device[dev] = ctx;
  }
  *out = per_device[dev];
  return MF_OK;
}

int mf_ctx_set_stream(mf_ctx *ctx, void *hip_stream) {
  if (!ctx) return MF_ERR_INVALID_ARG;
  (void)hipStreamSynchronize(ctx->stream);  // workspaces may be in use on the old stream
  ctx->stream = static_cast<hipStream_t>(hip_stream);
  return MF_OK;
}

void *mf_ctx_get_stream(mf_ctx *ctx) { return ctx ? static_cast<void *>(ctx->stream) : nullptr; }

int mf_ctx_synchronize(mf_ctx *ctx) {
  if (!ctx) return MF_ERR_INVALID_ARG;
  auto run = [&]() -> Status {
    MF_HIP(hipStreamSynchronize(ctx->stream));
    return Status::Ok();
  };
  return mf::finish(ctx, run());
}

int mf_ctx_enable_timing(mf_ctx *ctx, int enable) {
  if (!ctx) return MF_ERR_INVALID_ARG;
  ctx->timing = enable != 0;
  return MF_OK;
}

int mf_ctx_kernel_time(mf_ctx *ctx, const char *name, double *ms_total, int64_t *launches) {
  if (!ctx || !name) return MF_ERR_INVALID_ARG;
  Status s = mf::flush_timers(ctx);
  if (!s.ok()) return mf::finish(ctx, s);
  auto it = ctx->timers.find(name);
  if (ms_total) *ms_total = it == ctx->timers.end() ? 0.0 : it->second.ms;
  if (launches) *launches = it == ctx->timers.end() ? 0 : it->second.launches;
  return MF_OK;
}

int mf_ctx_reset_timing(mf_ctx *ctx) {
  if (!ctx) return MF_ERR_INVALID_ARG;
  Status s = mf::flush_timers(ctx);
  for (auto &kv : ctx->timers) {
    kv.second.ms = 0.0;
    kv.second.launches = 0;
  }
  return mf::finish(ctx, s);
}

int mf_field_allpairs(mf_ctx *ctx, int64_t n_t, const float *tx, const float *ty,
                      const float *tz, int64_t n_s, const float *sx, const float *sy,
                      const float *sz, const float *mx, const float *my, const float *mz,
                      float h, const float *hconst, float *ox, float *oy, float *oz) {
  if (!ctx) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "ctx is null"));
  auto run = [&]() -> Status {
    MF_ARG(n_t >= 0 && n_s >= 0, "sizes must be >= 0");
    MF_ARG(std::isfinite(h) && h > 0.f, "h must be finite and > 0");
    MF_ARG(n_t == 0 || (tx && ty && tz && ox && oy && oz), "target/output pointers are null");
    MF_ARG(n_s == 0 || (sx && sy && sz && mx && my && mz), "source pointers are null");
    MF_HIP(hipSetDevice(ctx->device));
    const int64_t np = mf::padded_sources(n_s);
    void *p = nullptr;
    MF_TRY(mf::workspace(ctx, "ext_src", sizeof(mf::SrcRec) * static_cast<size_t>(np + 1), &p));
    mf::SrcRec *src = static_cast<mf::SrcRec *>(p);
    MF_TRY(mf::workspace(ctx, "ext_boxes", sizeof(float4) * 2 * static_cast<size_t>(np / mf::kTile + 1), &p));
    float4 *boxes = static_cast<float4 *>(p);
    // pack_sources writes whole tiles: size the target records to the padded count
    MF_TRY(mf::workspace(ctx, "ext_tgt",
                         sizeof(mf::SrcRec) * static_cast<size_t>(mf::padded_sources(n_t) + 1), &p));
    mf::SrcRec *tgt = static_cast<mf::SrcRec *>(p);
    MF_TRY(mf::pack_sources(ctx, n_s, sx, sy, sz, mx, my, mz, 1.f, src));
    MF_TRY(mf::tile_boxes(ctx, n_s, src, boxes));
    if (n_t > 0) {
      // exact copy of the targets (no padding needed: pack n_t of them)
      MF_TRY(mf::pack_sources(ctx, n_t, tx, ty, tz, nullptr, nullptr, nullptr, 0.f, tgt));
    }
    return mf::field_allpairs_launch(ctx, reinterpret_cast<const float4 *>(tgt), 2, n_t, src,
                                     boxes, n_s, mf::make_field_consts(h), hconst, ox, oy, oz);
  };
  return mf::finish(ctx, run());
}

int mf_field_allpairs_f64(mf_ctx *ctx, int64_t n_t, const double *tx, const double *ty,
                          const double *tz, int64_t n_s, const double *sx, const double *sy,
                          const double *sz, const double *mx, const double *my,
                          const double *mz, double h, const double *hconst, double *ox,
                          double *oy, double *oz) {
  if (!ctx) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "ctx is null"));
  auto run = [&]() -> Status {
    MF_ARG(n_t >= 0 && n_s >= 0, "sizes must be >= 0");
    MF_ARG(std::isfinite(h) && h > 0.0, "h must be finite and > 0");
    MF_HIP(hipSetDevice(ctx->device));
    return mf::field_allpairs_f64(ctx, n_t, tx, ty, tz, n_s, sx, sy, sz, mx, my, mz, h, hconst,
                                  ox, oy, oz);
  };
  return mf::finish(ctx, run());
}

namespace {
// Shared body of mf_hessian_allpairs and mf_hessian_allpairs9.
int hessian_entry(mf_ctx *ctx, int64_t n_t, const float *tx, const float *ty, const float *tz,
                  int64_t n_s, const float *sx, const float *sy, const float *sz, const float *mx,
                  const float *my, const float *mz, float h, float *const dst[9],
                  const int sym[9]) {
  if (!ctx) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "ctx is null"));
  auto run = [&]() -> Status {
    MF_ARG(n_t >= 0 && n_s >= 0, "sizes must be >= 0");
    MF_ARG(std::isfinite(h) && h > 0.f, "h must be finite and > 0");
    MF_ARG(n_t == 0 || (tx && ty && tz), "target pointers are null");
    for (int k = 0; k < 9; ++k) MF_ARG(n_t == 0 || sym[k] < 0 || dst[k], "output pointers are null");
    MF_ARG(n_s == 0 || (sx && sy && sz && mx && my && mz), "source pointers are null");
    MF_HIP(hipSetDevice(ctx->device));
    const int64_t np = mf::padded_sources(n_s);
    void *p = nullptr;
    MF_TRY(mf::workspace(ctx, "ext_src", sizeof(mf::SrcRec) * static_cast<size_t>(np + 1), &p));
    mf::SrcRec *src = static_cast<mf::SrcRec *>(p);
    MF_TRY(mf::workspace(ctx, "ext_boxes", sizeof(float4) * 2 * static_cast<size_t>(np / mf::kTile + 1), &p));
    float4 *boxes = static_cast<float4 *>(p);
    MF_TRY(mf::workspace(ctx, "ext_tgt",
                         sizeof(mf::SrcRec) * static_cast<size_t>(mf::padded_sources(n_t) + 1), &p));
    mf::SrcRec *tgt = static_cast<mf::SrcRec *>(p);
    MF_TRY(mf::pack_sources(ctx, n_s, sx, sy, sz, mx, my, mz, 1.f, src));
    MF_TRY(mf::tile_boxes(ctx, n_s, src, boxes));
    if (n_t > 0) MF_TRY(mf::pack_sources(ctx, n_t, tx, ty, tz, nullptr, nullptr, nullptr, 0.f, tgt));
    // the kernel reads the records' position halves with a stride of two float4
    MF_TRY(mf::workspace(ctx, "ext_tpos", sizeof(float4) * static_cast<size_t>(n_t + 1), &p));
    float4 *tpos = static_cast<float4 *>(p);
    if (n_t > 0)
      MF_HIP(hipMemcpy2DAsync(tpos, sizeof(float4), tgt, sizeof(mf::SrcRec), sizeof(float4), n_t,
                              hipMemcpyDeviceToDevice, ctx->stream));
    float *d[9];
    int sy9[9];
    for (int k = 0; k < 9; ++k) {
      d[k] = sym[k] < 0 ? nullptr : dst[k];
      sy9[k] = sym[k] < 0 ? 0 : sym[k];
    }
    return mf::hessian_allpairs_launch(ctx, tpos, n_t, src, boxes, n_s, mf::make_hess_consts(h), d,
                                       sy9);
  };
  return mf::finish(ctx, run());
}
}  // namespace

int mf_hessian_allpairs(mf_ctx *ctx, int64_t n_t, const float *tx, const float *ty,
                        const float *tz, int64_t n_s, const float *sx, const float *sy,
                        const float *sz, const float *mx, const float *my, const float *mz,
                        float h, float *out6) {
  float *dst[9] = {};
  const int sym[9] = {0, 1, 2, 3, 4, 5, -1, -1, -1};
  for (int k = 0; k < 6; ++k) dst[k] = out6 ? out6 + static_cast<int64_t>(k) * n_t : nullptr;
  return hessian_entry(ctx, n_t, tx, ty, tz, n_s, sx, sy, sz, mx, my, mz, h, dst, sym);
}

int mf_hessian_allpairs9(mf_ctx *ctx, int64_t n_t, const float *tx, const float *ty,
                         const float *tz, int64_t n_s, const float *sx, const float *sy,
                         const float *sz, const float *mx, const float *my, const float *mz,
                         float h, float *const *out9) {
  if (!out9) return mf::finish(ctx, Status::Err(MF_ERR_INVALID_ARG, "out9 is null"));
  // row a, column b of T -> symmetric component: xx xy xz / xy yy yz / xz yz zz
  const int sym[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
  float *dst[9];
  for (int k = 0; k < 9; ++k) dst[k] = out9[k];
  return hessian_entry(ctx, n_t, tx, ty, tz, n_s, sx, sy, sz, mx, my, mz, h, dst, sym);
}

int mf_force_allpairs(mf_ctx *ctx, int64_t n, const float *x, const float *y, const float *z,
                      const float *mx, const float *my, const float *mz, float h,
                      const float *const *text9, float *fx, float *fy, float *fz) {
  if (!ctx) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "ctx is null"));
  auto run = [&]() -> Status {
    MF_ARG(n >= 0, "n must be >= 0");
    MF_ARG(std::isfinite(h) && h > 0.f, "h must be finite and > 0");
    MF_ARG(n == 0 || (x && y && z && mx && my && mz && fx && fy && fz), "pointers are null");
    MF_HIP(hipSetDevice(ctx->device));
    const int64_t np = mf::padded_sources(n);
    void *p = nullptr;
    MF_TRY(mf::workspace(ctx, "ext_src", sizeof(mf::SrcRec) * static_cast<size_t>(np + 1), &p));
    mf::SrcRec *src = static_cast<mf::SrcRec *>(p);
    MF_TRY(mf::workspace(ctx, "ext_boxes", sizeof(float4) * 2 * static_cast<size_t>(np / mf::kTile + 1), &p));
    float4 *boxes = static_cast<float4 *>(p);
    MF_TRY(mf::pack_sources(ctx, n, x, y, z, mx, my, mz, 1.f, src));
    MF_TRY(mf::tile_boxes(ctx, n, src, boxes));
    return mf::force_allpairs_launch(ctx, src, boxes, n, mf::make_hess_consts(h), text9, fx, fy,
                                     fz);
  };
  return mf::finish(ctx, run());
}

int mf_force_allpairs_f64(mf_ctx *ctx, int64_t n, const double *x, const double *y,
                          const double *z, const double *mx, const double *my,
                          const double *mz, double h, double *fx, double *fy, double *fz) {
  if (!ctx) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "ctx is null"));
  auto run = [&]() -> Status {
    MF_ARG(n >= 0, "n must be >= 0");
    MF_ARG(std::isfinite(h) && h > 0.0, "h must be finite and > 0");
    MF_HIP(hipSetDevice(ctx->device));
    return mf::force_allpairs_f64(ctx, n, x, y, z, mx, my, mz, h, fx, fy, fz);
  };
  return mf::finish(ctx, run());
}

int mf_force_subset_f64(mf_ctx *ctx, int64_t n, const double *x, const double *y,
                        const double *z, const double *mx, const double *my, const double *mz,
                        double h, int64_t n_t, const int64_t *idx, double *fx, double *fy,
                        double *fz) {
  if (!ctx) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "ctx is null"));
  auto run = [&]() -> Status {
    MF_ARG(n >= 0 && n_t >= 0, "sizes must be >= 0");
    MF_ARG(std::isfinite(h) && h > 0.0, "h must be finite and > 0");
    MF_ARG(n_t == 0 || (idx && fx && fy && fz), "pointers are null");
    MF_HIP(hipSetDevice(ctx->device));
    return mf::force_subset_f64(ctx, n, x, y, z, mx, my, mz, h, n_t, idx, fx, fy, fz);
  };
  return mf::finish(ctx, run());
}

int mf_plan_create(mf_ctx *ctx, int64_t n, const float *x, const float *y, const float *z,
                   float h, const mf_plan_opts *opts, mf_plan **out) {
  if (!ctx || !out) return mf::finish(ctx, Status::Err(MF_ERR_INVALID_ARG, "ctx/out is null"));
  *out = nullptr;
  auto run = [&]() -> Status {
    MF_ARG(std::isfinite(h) && h > 0.f, "h must be finite and > 0");
    MF_HIP(hipSetDevice(ctx->device));
    mf_plan *plan = new (std::nothrow) mf_plan();
    if (!plan) return Status::Err(MF_ERR_ALLOC, "out of host memory");
    plan->ctx = ctx;
    if (opts) plan->opts = *opts;
    if (plan->opts.mode != MF_MODE_ALLPAIRS && plan->opts.mode != MF_MODE_FMM) {
      delete plan;
      return Status::Err(MF_ERR_INVALID_ARG, "unknown plan mode");
    }
    if (plan->opts.m2l_precision < MF_M2L_FP32 || plan->opts.m2l_precision > MF_M2L_BF16X1_CUT) {
      delete plan;
      return Status::Err(MF_ERR_INVALID_ARG, "unknown M2L precision");
    }
    if (plan->opts.serial_near != 0 && plan->opts.serial_near != 1) {
      delete plan;
      return Status::Err(MF_ERR_INVALID_ARG, "serial_near must be 0 or 1");
    }
    plan->h = h;
    plan->fc = mf::make_field_consts(h);
    plan->hc = mf::make_hess_consts(h);
    Status s = mf::plan_set_positions(plan, n, x, y, z);
    if (!s.ok()) {
      delete plan;
      return s;
    }
    *out = plan;
    return Status::Ok();
  };
  return mf::finish(ctx, run());
}

int mf_plan_create_batch(mf_ctx *ctx, int nsys, const int64_t *offsets, const float *x,
                         const float *y, const float *z, float h, const mf_plan_opts *opts,
                         mf_plan **out) {
  if (!ctx || !out) return mf::finish(ctx, Status::Err(MF_ERR_INVALID_ARG, "ctx/out is null"));
  *out = nullptr;
  auto run = [&]() -> Status {
    MF_ARG(nsys >= 1 && offsets, "nsys must be >= 1 with nsys + 1 offsets");
    MF_ARG(offsets[0] == 0, "offsets[0] must be 0");
    for (int s = 0; s < nsys; ++s) MF_ARG(offsets[s + 1] >= offsets[s], "offsets must not decrease");
    MF_ARG(offsets[nsys] <= INT32_MAX, "batch too large");
    MF_ARG(opts && opts->mode == MF_MODE_FMM, "batch plans need mode MF_MODE_FMM");
    MF_ARG(std::isfinite(h) && h > 0.f, "h must be finite and > 0");
    MF_ARG(opts->m2l_precision >= MF_M2L_FP32 && opts->m2l_precision <= MF_M2L_BF16X1_CUT,
           "unknown M2L precision");
    MF_ARG(opts->serial_near == 0 || opts->serial_near == 1, "serial_near must be 0 or 1");
    MF_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<mf_plan> plan(new (std::nothrow) mf_plan());
    if (!plan) return Status::Err(MF_ERR_ALLOC, "out of host memory");
    plan->ctx = ctx;
    plan->opts = *opts;
    plan->h = h;
    plan->fc = mf::make_field_consts(h);
    plan->hc = mf::make_hess_consts(h);
    plan->nsys = nsys;
    plan->sys_offsets.assign(offsets, offsets + nsys + 1);
    const int64_t n = offsets[nsys];
    std::vector<int32_t> sys(static_cast<size_t>(n));
    for (int s = 0; s < nsys; ++s)
      for (int64_t i = offsets[s]; i < offsets[s + 1]; ++i) sys[static_cast<size_t>(i)] = s;
    MF_TRY(plan->sys_of.reserve(sizeof(int32_t) * static_cast<size_t>(n > 0 ? n : 1)));
    if (n > 0)
      MF_HIP(hipMemcpy(plan->sys_of.get(), sys.data(), sizeof(int32_t) * sys.size(),
                       hipMemcpyHostToDevice));
    MF_TRY(mf::plan_set_positions(plan.get(), n, x, y, z));
    *out = plan.release();
    return Status::Ok();
  };
  return mf::finish(ctx, run());
}

int mf_plan_solve_batch(mf_plan *plan, const double *kappa, const float *bx, const float *by,
                        const float *bz, float *xx, float *xy, float *xz,
                        const mf_solve_opts *opts, mf_solve_stats *stats,
                        double *sys_rel_residual) {
  if (!plan) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "plan is null"));
  auto run = [&]() -> Status {
    MF_ARG(kappa, "kappa (one per system) is null");
    MF_ARG(plan->ready, "the plan has no valid positions (its last update failed)");
    for (int s = 0; s < plan->nsys; ++s) MF_ARG(std::isfinite(kappa[s]), "kappa must be finite");
    MF_ARG(plan->n == 0 || (bx && by && bz && xx && xy && xz), "pointers are null");
    MF_ARG(plan->opts.mode == MF_MODE_FMM, "batch solves need an FMM plan");
    MF_TRY(mf::check_solve_opts(opts));
    MF_HIP(hipSetDevice(plan->ctx->device));
    mf_solve_opts o{};
    o.method = MF_SOLVE_CG;
    o.maxit = 1000;
    o.rtol = 1e-6;
    if (opts) o = *opts;
    mf_solve_stats st{};
    MF_TRY(mf::solve_batch(plan, kappa, bx, by, bz, xx, xy, xz, o, &st, sys_rel_residual));
    if (stats) *stats = st;
    return Status::Ok();
  };
  return mf::finish(plan->ctx, run());
}

int mf_plan_update_positions(mf_plan *plan, int64_t n, const float *x, const float *y,
                             const float *z) {
  if (!plan) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "plan is null"));
  auto run = [&]() -> Status {
    MF_HIP(hipSetDevice(plan->ctx->device));
    return mf::plan_set_positions(plan, n, x, y, z);
  };
  return mf::finish(plan->ctx, run());
}

int mf_plan_destroy(mf_plan *plan) {
  if (!plan) return MF_OK;
  (void)hipSetDevice(plan->ctx->device);
  (void)hipStreamSynchronize(plan->ctx->stream);
  delete plan;
  return MF_OK;
}

int64_t mf_plan_size(const mf_plan *plan) { return plan ? plan->n : -1; }

int mf_plan_get_info(const mf_plan *plan, mf_plan_info *info) {
  if (!plan || !info) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "plan/info is null"));
  if (!plan->ready)
    return mf::finish(plan->ctx, Status::Err(MF_ERR_INVALID_ARG,
                                             "the plan has no valid positions (its last update failed)"));
  *info = mf_plan_info{};
  info->n = plan->n;
  info->mode = plan->opts.mode;
  if (plan->opts.mode == MF_MODE_FMM && plan->fmm) {
    // the plan is logically const: fmm_info only fills its cached work counts
    mf_plan *p = const_cast<mf_plan *>(plan);
    auto run = [&]() -> Status {
      MF_HIP(hipSetDevice(p->ctx->device));
      return mf::fmm_info(p, info);
    };
    return mf::finish(p->ctx, run());
  }
  info->p2p_field_pairs = plan->n * plan->n;
  info->p2p_force_pairs = plan->n * plan->n;
  return MF_OK;
}

int mf_plan_field(mf_plan *plan, const float *mx, const float *my, const float *mz,
                  const float *hconst, float *ox, float *oy, float *oz) {
  if (!plan) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "plan is null"));
  auto run = [&]() -> Status {
    MF_ARG(plan->n == 0 || (mx && my && mz && ox && oy && oz), "pointers are null");
    MF_ARG(plan->ready, "the plan has no valid positions (its last update failed)");
    MF_HIP(hipSetDevice(plan->ctx->device));
    return mf::plan_apply_G(plan, 1.f, mx, my, mz, hconst, ox, oy, oz);
  };
  return mf::finish(plan->ctx, run());
}

int mf_plan_force(mf_plan *plan, const float *mx, const float *my, const float *mz,
                  const float *const *text9, float *fx, float *fy, float *fz) {
  if (!plan) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "plan is null"));
  auto run = [&]() -> Status {
    MF_ARG(plan->n == 0 || (mx && my && mz && fx && fy && fz), "pointers are null");
    MF_ARG(plan->ready, "the plan has no valid positions (its last update failed)");
    MF_HIP(hipSetDevice(plan->ctx->device));
    return mf::plan_force(plan, mx, my, mz, text9, fx, fy, fz);
  };
  return mf::finish(plan->ctx, run());
}

int mf_plan_solve(mf_plan *plan, double kappa, const float *bx, const float *by,
                  const float *bz, float *xx, float *xy, float *xz, const mf_solve_opts *opts,
                  mf_solve_stats *stats) {
  if (!plan) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "plan is null"));
  auto run = [&]() -> Status {
    MF_ARG(plan->n == 0 || (bx && by && bz && xx && xy && xz), "pointers are null");
    MF_ARG(std::isfinite(kappa), "kappa must be finite");
    MF_ARG(plan->ready, "the plan has no valid positions (its last update failed)");
    MF_TRY(mf::check_solve_opts(opts));
    MF_HIP(hipSetDevice(plan->ctx->device));
    mf_solve_opts o{};
    o.method = MF_SOLVE_CG;
    o.restart = 30;
    o.maxit = 1000;
    o.rtol = 1e-6;
    if (opts) o = *opts;
    mf_solve_stats st{};
    MF_TRY(mf::solve(plan, kappa, bx, by, bz, xx, xy, xz, o, &st));
    if (stats) *stats = st;
    return Status::Ok();
  };
  return mf::finish(plan->ctx, run());
}

int mf_plan_solve_f64(mf_plan *plan, double kappa, const double *b3n, double *x3n,
                      const mf_solve_opts *opts, mf_solve_stats *stats) {
  if (!plan) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "plan is null"));
  auto run = [&]() -> Status {
    const int64_t n = plan->n;
    MF_ARG(n == 0 || (b3n && x3n), "pointers are null");
    MF_ARG(plan->ready, "the plan has no valid positions (its last update failed)");
    MF_ARG(plan->nsys == 1, "mf_plan_solve_f64 takes single-system plans");
    MF_ARG(std::isfinite(kappa), "kappa must be finite");
    MF_TRY(mf::check_solve_opts(opts));
    MF_HIP(hipSetDevice(plan->ctx->device));
    mf_solve_opts o{};
    o.rtol = 1e-12;
    o.refine_fp64 = 10;
    if (opts) o = *opts;
    mf_solve_stats st{};
    MF_TRY(mf::solve_f64(plan, kappa, b3n, b3n + n, b3n + 2 * n, x3n, x3n + n, x3n + 2 * n, o, &st));
    if (stats) *stats = st;
    return Status::Ok();
  };
  return mf::finish(plan->ctx, run());
}

}  // extern "C"

namespace {
template <class T>
int vec_axpby_t(mf_ctx *ctx, int64_t n, T alpha, const T *ax, const T *ay, const T *az, T beta,
                const T *bx, const T *by, const T *bz, T *ox, T *oy, T *oz) {
  if (!ctx) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "ctx is null"));
  auto run = [&]() -> Status {
    MF_ARG(n >= 0, "n must be >= 0");
    MF_ARG(n == 0 || (ax && ay && az && ox && oy && oz), "pointers are null");
    MF_ARG(!bx == !by && !by == !bz, "b must be all null or all non-null");
    return mf::vec_axpby<T>(ctx, n, alpha, ax, ay, az, beta, bx, by, bz, ox, oy, oz);
  };
  return mf::finish(ctx, run());
}

template <class T>
int vec_mul_t(mf_ctx *ctx, int64_t n, const T *ax, const T *ay, const T *az, const T *bx,
              const T *by, const T *bz, T *ox, T *oy, T *oz) {
  if (!ctx) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "ctx is null"));
  auto run = [&]() -> Status {
    MF_ARG(n >= 0, "n must be >= 0");
    MF_ARG(n == 0 || (ax && ay && az && bx && by && bz && ox && oy && oz), "pointers are null");
    return mf::vec_mul<T>(ctx, n, ax, ay, az, bx, by, bz, ox, oy, oz);
  };
  return mf::finish(ctx, run());
}

template <class T>
int vec_dot_t(mf_ctx *ctx, int64_t n, const T *ax, const T *ay, const T *az, const T *bx,
              const T *by, const T *bz, double *result) {
  if (!ctx) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "ctx is null"));
  auto run = [&]() -> Status {
    MF_ARG(result, "result is null");
    MF_ARG(n >= 0, "n must be >= 0");
    if (n == 0) {
      *result = 0.0;
      return Status::Ok();
    }
    MF_ARG(ax && ay && az && bx && by && bz, "pointers are null");
    return mf::vec_dot<T>(ctx, n, ax, ay, az, bx, by, bz, result);
  };
  return mf::finish(ctx, run());
}

template <class T>
int vec_fill_t(mf_ctx *ctx, int64_t n, T v, T *x) {
  if (!ctx) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "ctx is null"));
  auto run = [&]() -> Status {
    MF_ARG(n >= 0, "n must be >= 0");
    MF_ARG(n == 0 || x, "x is null");
    return mf::vec_fill<T>(ctx, n, v, x);
  };
  return mf::finish(ctx, run());
}
}  // namespace

extern "C" {

int mf_vec_axpby(mf_ctx *ctx, int64_t n, float alpha, const float *ax, const float *ay,
                 const float *az, float beta, const float *bx, const float *by,
                 const float *bz, float *ox, float *oy, float *oz) {
  return vec_axpby_t<float>(ctx, n, alpha, ax, ay, az, beta, bx, by, bz, ox, oy, oz);
}

int mf_vec_mul(mf_ctx *ctx, int64_t n, const float *ax, const float *ay, const float *az,
               const float *bx, const float *by, const float *bz, float *ox, float *oy,
               float *oz) {
  return vec_mul_t<float>(ctx, n, ax, ay, az, bx, by, bz, ox, oy, oz);
}

int mf_vec_dot(mf_ctx *ctx, int64_t n, const float *ax, const float *ay, const float *az,
               const float *bx, const float *by, const float *bz, double *result) {
  return vec_dot_t<float>(ctx, n, ax, ay, az, bx, by, bz, result);
}

int mf_vec_fill(mf_ctx *ctx, int64_t n, float v, float *x) {
  return vec_fill_t<float>(ctx, n, v, x);
}

int mf_vec_axpby_f64(mf_ctx *ctx, int64_t n, double alpha, const double *ax, const double *ay,
                     const double *az, double beta, const double *bx, const double *by,
                     const double *bz, double *ox, double *oy, double *oz) {
  return vec_axpby_t<double>(ctx, n, alpha, ax, ay, az, beta, bx, by, bz, ox, oy, oz);
}

int mf_vec_mul_f64(mf_ctx *ctx, int64_t n, const double *ax, const double *ay, const double *az,
                   const double *bx, const double *by, const double *bz, double *ox, double *oy,
                   double *oz) {
  return vec_mul_t<double>(ctx, n, ax, ay, az, bx, by, bz, ox, oy, oz);
}

int mf_vec_dot_f64(mf_ctx *ctx, int64_t n, const double *ax, const double *ay, const double *az,
                   const double *bx, const double *by, const double *bz, double *result) {
  return vec_dot_t<double>(ctx, n, ax, ay, az, bx, by, bz, result);
}

int mf_vec_fill_f64(mf_ctx *ctx, int64_t n, double v, double *x) {
  return vec_fill_t<double>(ctx, n, v, x);
}

int mf_vec_convert(mf_ctx *ctx, int64_t n, int src_dtype, const void *src, int dst_dtype,
                   void *dst) {
  if (!ctx) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "ctx is null"));
  auto run = [&]() -> Status {
    MF_ARG(n >= 0, "n must be >= 0");
    MF_ARG(n == 0 || (src && dst), "pointers are null");
    return mf::vec_convert(ctx, n, src_dtype, src, dst_dtype, dst);
  };
  return mf::finish(ctx, run());
}

}  // extern "C"
