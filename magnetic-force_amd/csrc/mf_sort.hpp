// mf_sort.hpp -- device sort/scan entry points used by the octree build.
#pragma once

#include <cstdint>

#include "mf_internal.hpp"

namespace mf {

// Stable LSD radix sort of (key, value) pairs on bits [begin_bit, end_bit).
Status sort_pairs_u64_i32(mf_ctx *ctx, const uint64_t *keys_in, uint64_t *keys_out,
                          const int32_t *vals_in, int32_t *vals_out, int64_t n, int begin_bit,
                          int end_bit);
Status exclusive_scan_i32(mf_ctx *ctx, const int32_t *in, int32_t *out, int64_t n);

}  // namespace mf
