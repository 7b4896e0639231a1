// mf_noalloc.hip -- the all-pairs field product without any allocation.
//
// mf_field_allpairs_noalloc runs the same kernels as mf_field_allpairs (pack
// the sources and targets into records, tile boxes, the chunked field kernel,
// the fixed-order chunk reduction), but every scratch buffer lives in one
// file-scope device array, `scratch`, instead of the context's workspaces.
// It is for callers that cannot call hipMalloc: code under stream capture,
// tight serving loops, harnesses that forbid allocation.
//
// The array is part of the code object, so every GPU the library is loaded
// on holds kScratchBytes of it until the process ends.  Its device address is
// looked up once per device, with that device current, when a context is
// created on it (mf_ctx_create calls noalloc_setup); host code then uses only
// that address, never the array's name (which is a host-side stand-in).
//
// All calls on one GPU share the array: calls that overlap in time on one GPU
// (two streams, two contexts, two host threads) race and must be ordered by
// the caller.  Everything a call reads from the array it has written earlier
// in the same call, so no state carries over from the previous call.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mf_allpairs.hpp"
#include "mf_capi.h"
#include "mf_internal.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;  // 32 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch_addr[kMaxDevices] = {};  // device address of `scratch`, per device

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

namespace mf {

// Looks up the address of `scratch` on `device`, which must be current.
Status noalloc_setup(int device) {
  MF_ARG(device >= 0 && device < kMaxDevices, "device index out of range for the no-alloc scratch");
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch_addr[device]) return Status::Ok();
  void *p = nullptr;
  MF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return Status::Err(MF_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
  g_scratch_addr[device] = p;
  return Status::Ok();
}

namespace {

void *scratch_addr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_mu);
  return g_scratch_addr[device];
}

// Byte layout of one call inside the scratch array.
struct NoallocLayout {
  size_t src, boxes, tgt, partial, end;
};

NoallocLayout noalloc_layout(int64_t n_t, int64_t n_s, int64_t nchunk) {
  NoallocLayout l{};
  const size_t ns_pad = static_cast<size_t>(padded_sources(n_s));
  const size_t nt_pad = static_cast<size_t>(padded_sources(n_t));
  l.src = 0;
  l.boxes = align256(l.src + sizeof(SrcRec) * (ns_pad + 1));
  l.tgt = align256(l.boxes + sizeof(float4) * 2 * (ns_pad / kTile + 1));
  l.partial = align256(l.tgt + sizeof(SrcRec) * (nt_pad + 1));
  l.end = l.partial + sizeof(double) * static_cast<size_t>(field_partial_doubles(n_t, nchunk));
  return l;
}

}  // namespace
}  // namespace mf

extern "C" int mf_noalloc_scratch_bytes(int64_t *bytes) {
  if (!bytes) return MF_ERR_INVALID_ARG;
  *bytes = static_cast<int64_t>(kScratchBytes);
  return MF_OK;
}

extern "C" int mf_field_allpairs_noalloc(mf_ctx *ctx, int64_t n_t, const float *tx,
                                         const float *ty, const float *tz, int64_t n_s,
                                         const float *sx, const float *sy, const float *sz,
                                         const float *mx, const float *my, const float *mz,
                                         float h, const float *hconst, float *ox, float *oy,
                                         float *oz) {
  using mf::Status;
  if (!ctx) return mf::finish(nullptr, Status::Err(MF_ERR_INVALID_ARG, "ctx is null"));
  auto run = [&]() -> Status {
    MF_ARG(n_t >= 0 && n_s >= 0, "sizes must be >= 0");
    MF_ARG(h > 0.f, "h must be > 0");
    MF_ARG(n_t == 0 || (tx && ty && tz && ox && oy && oz), "target/output pointers are null");
    MF_ARG(n_s == 0 || (sx && sy && sz && mx && my && mz), "source pointers are null");
    if (n_t == 0) return Status::Ok();
    const mf::NoallocLayout need = mf::noalloc_layout(n_t, n_s, 1);
    if (need.end > kScratchBytes)
      return Status::Err(MF_ERR_INVALID_ARG,
                         "mf_field_allpairs_noalloc: n_t = " + std::to_string(n_t) +
                             ", n_s = " + std::to_string(n_s) + " need " +
                             std::to_string(need.end) + " bytes of scratch; the maximum is " +
                             std::to_string(kScratchBytes));
    MF_HIP(hipSetDevice(ctx->device));
    unsigned char *base = static_cast<unsigned char *>(mf::scratch_addr(ctx->device));
    if (!base) return Status::Err(MF_ERR_HIP, "no-alloc scratch is not set up on this device");
    // the rest of the array holds as many source-chunk partials as fit
    const int64_t cap = static_cast<int64_t>((kScratchBytes - need.partial) / sizeof(double));
    mf::SrcRec *src = reinterpret_cast<mf::SrcRec *>(base + need.src);
    float4 *boxes = reinterpret_cast<float4 *>(base + need.boxes);
    mf::SrcRec *tgt = reinterpret_cast<mf::SrcRec *>(base + need.tgt);
    double *partial = reinterpret_cast<double *>(base + need.partial);
    MF_TRY(mf::pack_sources(ctx, n_s, sx, sy, sz, mx, my, mz, 1.f, src));
    MF_TRY(mf::tile_boxes(ctx, n_s, src, boxes));
    MF_TRY(mf::pack_sources(ctx, n_t, tx, ty, tz, nullptr, nullptr, nullptr, 0.f, tgt));
    return mf::field_allpairs_launch_ws(ctx, reinterpret_cast<const float4 *>(tgt), 2, n_t, src,
                                        boxes, n_s, mf::make_field_consts(h), hconst, ox, oy, oz,
                                        partial, cap);
  };
  return mf::finish(ctx, run());
}
