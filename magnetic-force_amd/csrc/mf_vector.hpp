// mf_vector.hpp -- fused SoA 3-vector operations (host entry points).
#pragma once

#include "mf_internal.hpp"

namespace mf {

constexpr int kRedBlock = 256;
constexpr int kRedMaxBlocks = 1024;

int red_blocks(int64_t n);

Status vec_axpby(mf_ctx *ctx, int64_t n, float alpha, const float *ax, const float *ay,
                 const float *az, float beta, const float *bx, const float *by, const float *bz,
                 float *ox, float *oy, float *oz);
Status vec_mul(mf_ctx *ctx, int64_t n, const float *ax, const float *ay, const float *az,
               const float *bx, const float *by, const float *bz, float *ox, float *oy,
               float *oz);
Status vec_fill(mf_ctx *ctx, int64_t n, float v, float *x);
// Device-side result (no synchronization).
Status vec_dot_device(mf_ctx *ctx, int64_t n, const float *ax, const float *ay, const float *az,
                      const float *bx, const float *by, const float *bz, double *d_out);
// Host result (synchronizes the context stream).
Status vec_dot(mf_ctx *ctx, int64_t n, const float *ax, const float *ay, const float *az,
               const float *bx, const float *by, const float *bz, double *result);
Status sum_partials(mf_ctx *ctx, const double *partials, int count, double *d_out);

}  // namespace mf
