// mf_vector.hip -- fused SoA 3-vector kernels and deterministic reductions.
//
// The reference's cuda_abstract_vector launches one kernel per component and
// allocates a new vector for every value-returning operation
// (cuda_abstract_vector.cuh:12-86); its reductions sum 64 FP32 block partials
// on the host and contain a barrier inside a divergent branch (:89-161).
// Here every operation covers x, y, z in one launch, reductions keep FP64
// partials in a fixed-size grid and finish on the device in a fixed order
// (bitwise reproducible), and nothing allocates after the first call.
#include <hip/hip_runtime.h>

#include "mf_capi.h"
#include "mf_internal.hpp"
#include "mf_vector.hpp"

namespace mf {
namespace {

template <class T>
__global__ void axpby_kernel(int64_t n, T alpha, const T *__restrict__ ax, const T *__restrict__ ay,
                             const T *__restrict__ az, T beta, const T *bx, const T *by, const T *bz,
                             T *ox, T *oy, T *oz) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T x = alpha * ax[i], y = alpha * ay[i], z = alpha * az[i];
  if (bx) {
    x = fma(beta, bx[i], x);
    y = fma(beta, by[i], y);
    z = fma(beta, bz[i], z);
  }
  ox[i] = x;
  oy[i] = y;
  oz[i] = z;
}

template <class T>
__global__ void mul_kernel(int64_t n, const T *__restrict__ ax, const T *__restrict__ ay,
                           const T *__restrict__ az, const T *bx, const T *by, const T *bz, T *ox,
                           T *oy, T *oz) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ox[i] = ax[i] * bx[i];
  oy[i] = ay[i] * by[i];
  oz[i] = az[i] * bz[i];
}

template <class T>
__global__ void fill_kernel(int64_t n, T v, T *x) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// dst[i] = (D) src[i]; float -> int truncates toward zero, as a C cast.
template <class S, class D>
__global__ void convert_kernel(int64_t n, const S *__restrict__ src, D *__restrict__ dst) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = static_cast<D>(src[i]);
}

__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < static_cast<int>(blockDim.x >> 6); ++k) s += sh[k];
  return s;  // valid in thread 0
}

// Per-element product over the three components: FP32 inputs multiply in
// FP32 (one rounding per element before the FP64 sum), FP64 inputs in FP64.
__device__ __forceinline__ double dot3(float ax, float ay, float az, float bx, float by, float bz) {
  float t = ax * bx;
  t = fmaf(ay, by, t);
  t = fmaf(az, bz, t);
  return static_cast<double>(t);
}
__device__ __forceinline__ double dot3(double ax, double ay, double az, double bx, double by,
                                       double bz) {
  return fma(az, bz, fma(ay, by, ax * bx));
}

// Grid-stride FP64 partial of sum a.b over 3 components; one value per block.
template <class T>
__global__ __launch_bounds__(kRedBlock) void dot_partial_kernel(
    int64_t n, const T *__restrict__ ax, const T *__restrict__ ay, const T *__restrict__ az,
    const T *__restrict__ bx, const T *__restrict__ by, const T *__restrict__ bz,
    double *__restrict__ partials) {
  __shared__ double sh[kRedBlock / 64];
  double acc = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kRedBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kRedBlock)
    acc += dot3(ax[i], ay[i], az[i], bx[i], by[i], bz[i]);
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Sums `count` partials in a fixed order into out[0].
__global__ __launch_bounds__(kRedBlock) void sum_partials_kernel(const double *__restrict__ partials,
                                                                 int count,
                                                                 double *__restrict__ out) {
  __shared__ double sh[kRedBlock / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < count; i += kRedBlock) acc += partials[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[0] = s;
}

}  // namespace

template <class T>
Status vec_axpby(mf_ctx *ctx, int64_t n, T alpha, const T *ax, const T *ay, const T *az, T beta,
                 const T *bx, const T *by, const T *bz, T *ox, T *oy, T *oz) {
  if (n == 0) return Status::Ok();
  hipLaunchKernelGGL(axpby_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, alpha,
                     ax, ay, az, beta, bx, by, bz, ox, oy, oz);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

template <class T>
Status vec_mul(mf_ctx *ctx, int64_t n, const T *ax, const T *ay, const T *az, const T *bx,
               const T *by, const T *bz, T *ox, T *oy, T *oz) {
  if (n == 0) return Status::Ok();
  hipLaunchKernelGGL(mul_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, ax, ay,
                     az, bx, by, bz, ox, oy, oz);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

template <class T>
Status vec_fill(mf_ctx *ctx, int64_t n, T v, T *x) {
  if (n == 0) return Status::Ok();
  hipLaunchKernelGGL(fill_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, v, x);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

template Status vec_axpby<float>(mf_ctx *, int64_t, float, const float *, const float *,
                                 const float *, float, const float *, const float *,
                                 const float *, float *, float *, float *);
template Status vec_axpby<double>(mf_ctx *, int64_t, double, const double *, const double *,
                                  const double *, double, const double *, const double *,
                                  const double *, double *, double *, double *);
template Status vec_mul<float>(mf_ctx *, int64_t, const float *, const float *, const float *,
                               const float *, const float *, const float *, float *, float *,
                               float *);
template Status vec_mul<double>(mf_ctx *, int64_t, const double *, const double *,
                                const double *, const double *, const double *, const double *,
                                double *, double *, double *);
template Status vec_fill<float>(mf_ctx *, int64_t, float, float *);
template Status vec_fill<double>(mf_ctx *, int64_t, double, double *);

Status vec_convert(mf_ctx *ctx, int64_t n, int src_dtype, const void *src, int dst_dtype,
                   void *dst) {
  if (n == 0) return Status::Ok();
  const dim3 g(grid_for(n, 256)), b(256);
#define MF_CONV(ST, SC, DT, DC)                                                             \
  if (src_dtype == SC && dst_dtype == DC) {                                                 \
    hipLaunchKernelGGL((convert_kernel<ST, DT>), g, b, 0, ctx->stream, n,                   \
                       static_cast<const ST *>(src), static_cast<DT *>(dst));               \
    MF_LAUNCH_CHECK();                                                                      \
    return Status::Ok();                                                                    \
  }
  MF_CONV(float, MF_DTYPE_F32, float, MF_DTYPE_F32)
  MF_CONV(float, MF_DTYPE_F32, double, MF_DTYPE_F64)
  MF_CONV(float, MF_DTYPE_F32, int32_t, MF_DTYPE_I32)
  MF_CONV(double, MF_DTYPE_F64, float, MF_DTYPE_F32)
  MF_CONV(double, MF_DTYPE_F64, double, MF_DTYPE_F64)
  MF_CONV(double, MF_DTYPE_F64, int32_t, MF_DTYPE_I32)
  MF_CONV(int32_t, MF_DTYPE_I32, float, MF_DTYPE_F32)
  MF_CONV(int32_t, MF_DTYPE_I32, double, MF_DTYPE_F64)
  MF_CONV(int32_t, MF_DTYPE_I32, int32_t, MF_DTYPE_I32)
#undef MF_CONV
  return Status::Err(MF_ERR_INVALID_ARG, "unknown mf_dtype");
}

int red_blocks(int64_t n) {
  int64_t b = (n + kRedBlock - 1) / kRedBlock;
  if (b < 1) b = 1;
  if (b > kRedMaxBlocks) b = kRedMaxBlocks;
  return static_cast<int>(b);
}

template <class T>
Status vec_dot_device(mf_ctx *ctx, int64_t n, const T *ax, const T *ay, const T *az, const T *bx,
                      const T *by, const T *bz, double *d_out) {
  void *pw = nullptr;
  MF_TRY(workspace(ctx, "dot_partials", sizeof(double) * kRedMaxBlocks, &pw));
  double *partials = static_cast<double *>(pw);
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(dot_partial_kernel<T>, dim3(nb), dim3(kRedBlock), 0, ctx->stream, n, ax, ay,
                     az, bx, by, bz, partials);
  MF_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kRedBlock), 0, ctx->stream, partials, nb,
                     d_out);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

template <class T>
Status vec_dot(mf_ctx *ctx, int64_t n, const T *ax, const T *ay, const T *az, const T *bx,
               const T *by, const T *bz, double *result) {
  void *pw = nullptr;
  MF_TRY(workspace(ctx, "dot_result", sizeof(double) * 8, &pw));
  double *d = static_cast<double *>(pw);
  MF_TRY(vec_dot_device<T>(ctx, n, ax, ay, az, bx, by, bz, d));
  MF_HIP(hipMemcpyAsync(ctx->host_scalars, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  MF_HIP(hipStreamSynchronize(ctx->stream));
  *result = ctx->host_scalars[0];
  return Status::Ok();
}

template Status vec_dot_device<float>(mf_ctx *, int64_t, const float *, const float *,
                                      const float *, const float *, const float *, const float *,
                                      double *);
template Status vec_dot_device<double>(mf_ctx *, int64_t, const double *, const double *,
                                       const double *, const double *, const double *,
                                       const double *, double *);
template Status vec_dot<float>(mf_ctx *, int64_t, const float *, const float *, const float *,
                               const float *, const float *, const float *, double *);
template Status vec_dot<double>(mf_ctx *, int64_t, const double *, const double *,
                                const double *, const double *, const double *, const double *,
                                double *);

Status sum_partials(mf_ctx *ctx, const double *partials, int count, double *d_out) {
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kRedBlock), 0, ctx->stream, partials,
                     count, d_out);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

}  // namespace mf
