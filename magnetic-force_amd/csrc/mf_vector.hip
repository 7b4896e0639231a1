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

#include "mf_internal.hpp"
#include "mf_vector.hpp"

namespace mf {
namespace {

__global__ void axpby_kernel(int64_t n, float alpha, const float *__restrict__ ax,
                             const float *__restrict__ ay, const float *__restrict__ az,
                             float beta, const float *bx, const float *by, const float *bz,
                             float *ox, float *oy, float *oz) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x = alpha * ax[i], y = alpha * ay[i], z = alpha * az[i];
  if (bx) {
    x = fmaf(beta, bx[i], x);
    y = fmaf(beta, by[i], y);
    z = fmaf(beta, bz[i], z);
  }
  ox[i] = x;
  oy[i] = y;
  oz[i] = z;
}

__global__ void mul_kernel(int64_t n, const float *__restrict__ ax, const float *__restrict__ ay,
                           const float *__restrict__ az, const float *bx, const float *by,
                           const float *bz, float *ox, float *oy, float *oz) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ox[i] = ax[i] * bx[i];
  oy[i] = ay[i] * by[i];
  oz[i] = az[i] * bz[i];
}

__global__ void fill_kernel(int64_t n, float v, float *x) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
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

// Grid-stride FP64 partial of sum a.b over 3 components; one value per block.
__global__ __launch_bounds__(kRedBlock) void dot_partial_kernel(
    int64_t n, const float *__restrict__ ax, const float *__restrict__ ay,
    const float *__restrict__ az, const float *__restrict__ bx, const float *__restrict__ by,
    const float *__restrict__ bz, double *__restrict__ partials) {
  __shared__ double sh[kRedBlock / 64];
  double acc = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kRedBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kRedBlock) {
    float t = ax[i] * bx[i];
    t = fmaf(ay[i], by[i], t);
    t = fmaf(az[i], bz[i], t);
    acc += static_cast<double>(t);
  }
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

Status vec_axpby(mf_ctx *ctx, int64_t n, float alpha, const float *ax, const float *ay,
                 const float *az, float beta, const float *bx, const float *by, const float *bz,
                 float *ox, float *oy, float *oz) {
  if (n == 0) return Status::Ok();
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, alpha,
                     ax, ay, az, beta, bx, by, bz, ox, oy, oz);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

Status vec_mul(mf_ctx *ctx, int64_t n, const float *ax, const float *ay, const float *az,
               const float *bx, const float *by, const float *bz, float *ox, float *oy,
               float *oz) {
  if (n == 0) return Status::Ok();
  hipLaunchKernelGGL(mul_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, ax, ay,
                     az, bx, by, bz, ox, oy, oz);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

Status vec_fill(mf_ctx *ctx, int64_t n, float v, float *x) {
  if (n == 0) return Status::Ok();
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, v, x);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

int red_blocks(int64_t n) {
  int64_t b = (n + kRedBlock - 1) / kRedBlock;
  if (b < 1) b = 1;
  if (b > kRedMaxBlocks) b = kRedMaxBlocks;
  return static_cast<int>(b);
}

Status vec_dot_device(mf_ctx *ctx, int64_t n, const float *ax, const float *ay, const float *az,
                      const float *bx, const float *by, const float *bz, double *d_out) {
  void *pw = nullptr;
  MF_TRY(workspace(ctx, "dot_partials", sizeof(double) * kRedMaxBlocks, &pw));
  double *partials = static_cast<double *>(pw);
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(kRedBlock), 0, ctx->stream, n, ax, ay, az,
                     bx, by, bz, partials);
  MF_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kRedBlock), 0, ctx->stream, partials, nb,
                     d_out);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

Status vec_dot(mf_ctx *ctx, int64_t n, const float *ax, const float *ay, const float *az,
               const float *bx, const float *by, const float *bz, double *result) {
  void *pw = nullptr;
  MF_TRY(workspace(ctx, "dot_result", sizeof(double) * 8, &pw));
  double *d = static_cast<double *>(pw);
  MF_TRY(vec_dot_device(ctx, n, ax, ay, az, bx, by, bz, d));
  MF_HIP(hipMemcpyAsync(ctx->host_scalars, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  MF_HIP(hipStreamSynchronize(ctx->stream));
  *result = ctx->host_scalars[0];
  return Status::Ok();
}

Status sum_partials(mf_ctx *ctx, const double *partials, int count, double *d_out) {
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kRedBlock), 0, ctx->stream, partials,
                     count, d_out);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

}  // namespace mf
