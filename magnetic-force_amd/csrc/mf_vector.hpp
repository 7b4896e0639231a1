// mf_vector.hpp -- fused SoA 3-vector operations (host entry points).
#pragma once

#include "mf_internal.hpp"

namespace mf {

constexpr int kRedBlock = 256;
constexpr int kRedMaxBlocks = 1024;

int red_blocks(int64_t n);

// T = float or double.
template <class T>
Status vec_axpby(mf_ctx *ctx, int64_t n, T alpha, const T *ax, const T *ay, const T *az, T beta,
                 const T *bx, const T *by, const T *bz, T *ox, T *oy, T *oz);
template <class T>
Status vec_mul(mf_ctx *ctx, int64_t n, const T *ax, const T *ay, const T *az, const T *bx,
               const T *by, const T *bz, T *ox, T *oy, T *oz);
template <class T>
Status vec_fill(mf_ctx *ctx, int64_t n, T v, T *x);
// Device-side result (no synchronization).
template <class T>
Status vec_dot_device(mf_ctx *ctx, int64_t n, const T *ax, const T *ay, const T *az, const T *bx,
                      const T *by, const T *bz, double *d_out);
// Host result (synchronizes the context stream).
template <class T>
Status vec_dot(mf_ctx *ctx, int64_t n, const T *ax, const T *ay, const T *az, const T *bx,
               const T *by, const T *bz, double *result);
// Element-wise type conversion between mf_dtype arrays.
Status vec_convert(mf_ctx *ctx, int64_t n, int src_dtype, const void *src, int dst_dtype,
                   void *dst);
Status sum_partials(mf_ctx *ctx, const double *partials, int count, double *d_out);

}  // namespace mf
