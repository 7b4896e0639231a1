// mf_sort.hip -- device radix sort and scans for the octree build (rocPRIM).
// Kept in its own translation unit: the rocPRIM templates are slow to compile.
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "mf_sort.hpp"

namespace mf {

Status sort_pairs_u64_i32(mf_ctx *ctx, const uint64_t *keys_in, uint64_t *keys_out,
                          const int32_t *vals_in, int32_t *vals_out, int64_t n, int begin_bit,
                          int end_bit) {
  if (n == 0) return Status::Ok();
  size_t bytes = 0;
  MF_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out,
                                   static_cast<size_t>(n), begin_bit, end_bit, ctx->stream));
  void *tmp = nullptr;
  MF_TRY(workspace(ctx, "sort_tmp", bytes + 16, &tmp));
  MF_HIP(rocprim::radix_sort_pairs(tmp, bytes, keys_in, keys_out, vals_in, vals_out,
                                   static_cast<size_t>(n), begin_bit, end_bit, ctx->stream));
  return Status::Ok();
}

Status exclusive_scan_i32(mf_ctx *ctx, const int32_t *in, int32_t *out, int64_t n) {
  if (n == 0) return Status::Ok();
  size_t bytes = 0;
  MF_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, static_cast<size_t>(n),
                                 rocprim::plus<int32_t>(), ctx->stream));
  void *tmp = nullptr;
  MF_TRY(workspace(ctx, "scan_tmp", bytes + 16, &tmp));
  MF_HIP(rocprim::exclusive_scan(tmp, bytes, in, out, 0, static_cast<size_t>(n),
                                 rocprim::plus<int32_t>(), ctx->stream));
  return Status::Ok();
}

}  // namespace mf
