// mf_solver.hip -- device-resident magnetization solve.
//
// Solves (I - kappa G) x = b, x = B/mu0 per particle, m = kappa x
// (the linear system behind magnetic_force.cu:1016-1017).  A = I - kappa G is
// symmetric positive definite (G is symmetric: K(d) = K(-d) = K(d)^T, and
// kappa G has spectrum in [0, chi/(1+chi)) for a physical particle set), so
// conjugate gradients is the default; restarted GMRES and the reference's own
// nonlinear CG (magnetic_force.cu:1008-1095, 1275-1315) are also provided.
//
// All vectors and scalars stay on the device.  The host reads one scalar per
// iteration (the residual norm, for the stopping test) and nothing else.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "mf_internal.hpp"
#include "mf_plan.hpp"
#include "mf_vector.hpp"

namespace mf {
namespace {

enum { S_RR = 0, S_PAP = 1, S_ALPHA = 2, S_BETA = 3, S_BB = 4, S_TMP = 5, S_COUNT = 8 };

__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < static_cast<int>(blockDim.x >> 6); ++k) s += sh[k];
  return s;
}

struct V3 {
  float *x, *y, *z;
};
struct CV3 {
  const float *x, *y, *z;
};

// r = b - x + g (g = G(kappa x)); p = r
__global__ void cg_init_kernel(int64_t n, CV3 b, CV3 x, CV3 g, V3 r, V3 p) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float rx = b.x[i] - x.x[i] + g.x[i];
  const float ry = b.y[i] - x.y[i] + g.y[i];
  const float rz = b.z[i] - x.z[i] + g.z[i];
  r.x[i] = rx;
  r.y[i] = ry;
  r.z[i] = rz;
  if (p.x) {
    p.x[i] = rx;
    p.y[i] = ry;
    p.z[i] = rz;
  }
}

// Ap = p - g ; partial p.Ap
__global__ __launch_bounds__(kRedBlock) void cg_ap_kernel(int64_t n, CV3 p, CV3 g, V3 ap,
                                                          double *__restrict__ partials) {
  __shared__ double sh[kRedBlock / 64];
  double acc = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kRedBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kRedBlock) {
    const float px = p.x[i], py = p.y[i], pz = p.z[i];
    const float ax = px - g.x[i], ay = py - g.y[i], az = pz - g.z[i];
    ap.x[i] = ax;
    ap.y[i] = ay;
    ap.z[i] = az;
    acc += static_cast<double>(px) * ax + static_cast<double>(py) * ay +
           static_cast<double>(pz) * az;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedBlock) void cg_alpha_kernel(const double *__restrict__ partials,
                                                             int nb, double *__restrict__ sc) {
  __shared__ double sh[kRedBlock / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kRedBlock) acc += partials[i];
  const double pap = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    sc[S_PAP] = pap;
    sc[S_ALPHA] = (pap != 0.0) ? sc[S_RR] / pap : 0.0;
  }
}

// x += alpha p ; r -= alpha Ap ; partial r.r
__global__ __launch_bounds__(kRedBlock) void cg_xr_kernel(int64_t n,
                                                          const double *__restrict__ sc, CV3 p,
                                                          CV3 ap, V3 x, V3 r,
                                                          double *__restrict__ partials) {
  __shared__ double sh[kRedBlock / 64];
  const float alpha = static_cast<float>(sc[S_ALPHA]);
  double acc = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kRedBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kRedBlock) {
    x.x[i] = fmaf(alpha, p.x[i], x.x[i]);
    x.y[i] = fmaf(alpha, p.y[i], x.y[i]);
    x.z[i] = fmaf(alpha, p.z[i], x.z[i]);
    const float rx = fmaf(-alpha, ap.x[i], r.x[i]);
    const float ry = fmaf(-alpha, ap.y[i], r.y[i]);
    const float rz = fmaf(-alpha, ap.z[i], r.z[i]);
    r.x[i] = rx;
    r.y[i] = ry;
    r.z[i] = rz;
    acc += static_cast<double>(rx) * rx + static_cast<double>(ry) * ry +
           static_cast<double>(rz) * rz;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedBlock) void cg_beta_kernel(const double *__restrict__ partials,
                                                            int nb, double *__restrict__ sc) {
  __shared__ double sh[kRedBlock / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kRedBlock) acc += partials[i];
  const double rr = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    sc[S_BETA] = (sc[S_RR] != 0.0) ? rr / sc[S_RR] : 0.0;
    sc[S_RR] = rr;
  }
}

// p = r + beta p
__global__ void cg_p_kernel(int64_t n, const double *__restrict__ sc, CV3 r, V3 p) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float beta = static_cast<float>(sc[S_BETA]);
  p.x[i] = fmaf(beta, p.x[i], r.x[i]);
  p.y[i] = fmaf(beta, p.y[i], r.y[i]);
  p.z[i] = fmaf(beta, p.z[i], r.z[i]);
}

// ---- GMRES helpers --------------------------------------------------------
// partials[k * nb + block] = sum V_k . w   for k < nk  (V stored [k][3][n])
__global__ __launch_bounds__(kRedBlock) void multidot_kernel(int64_t n, int nk,
                                                             const float *__restrict__ V,
                                                             const float *__restrict__ w,
                                                             double *__restrict__ partials) {
  __shared__ double sh[kRedBlock / 64];
  const int64_t n3 = 3 * n;
  for (int k = 0; k < nk; ++k) {
    const float *v = V + static_cast<int64_t>(k) * n3;
    double acc = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kRedBlock + threadIdx.x; i < n3;
         i += static_cast<int64_t>(gridDim.x) * kRedBlock)
      acc += static_cast<double>(v[i]) * w[i];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[static_cast<int64_t>(k) * gridDim.x + blockIdx.x] = s;
    __syncthreads();
  }
}

// out[k] = sum_b partials[k * nb + b]
__global__ __launch_bounds__(kRedBlock) void multisum_kernel(const double *__restrict__ partials,
                                                             int nb, double *__restrict__ out) {
  __shared__ double sh[kRedBlock / 64];
  const int k = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kRedBlock) acc += partials[static_cast<int64_t>(k) * nb + i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[k] = s;
}

// w -= sum_k coef[k] V_k   (coef in device memory, nk <= 64)
__global__ void combine_kernel(int64_t n3, int nk, const float *__restrict__ V,
                               const double *__restrict__ coef, float sign, float *__restrict__ w) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  double acc = 0.0;
  for (int k = 0; k < nk; ++k) acc += coef[k] * static_cast<double>(V[static_cast<int64_t>(k) * n3 + i]);
  w[i] = static_cast<float>(static_cast<double>(w[i]) + sign * acc);
}

// out = s * a (out may alias a)
__global__ void scale_kernel(int64_t n3, float s, const float *a, float *out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n3) out[i] = s * a[i];
}

// x (SoA, three arrays) += sum_k y[k] V_k
__global__ void gmres_update_kernel(int64_t n, int nk, const float *__restrict__ V,
                                    const double *__restrict__ y, V3 x) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  const int64_t n3 = 3 * n;
  for (int k = 0; k < nk; ++k) {
    const float *v = V + static_cast<int64_t>(k) * n3;
    a0 += y[k] * v[i];
    a1 += y[k] * v[n + i];
    a2 += y[k] * v[2 * n + i];
  }
  x.x[i] = static_cast<float>(x.x[i] + a0);
  x.y[i] = static_cast<float>(x.y[i] + a1);
  x.z[i] = static_cast<float>(x.z[i] + a2);
}

// ---- FP64 refinement helpers ----------------------------------------------
__global__ void f64_moments_kernel(int64_t n, double kappa, const double *__restrict__ x,
                                   double *__restrict__ m) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < 3 * n) m[i] = kappa * x[i];
}
// r = b - x + g  (all double, [3][n]), partial r.r and b.b
__global__ __launch_bounds__(kRedBlock) void f64_residual_kernel(int64_t n3,
                                                                 const double *__restrict__ b,
                                                                 const double *__restrict__ x,
                                                                 const double *__restrict__ g,
                                                                 double *__restrict__ r,
                                                                 double *__restrict__ partials) {
  __shared__ double sh[kRedBlock / 64];
  double acc = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kRedBlock + threadIdx.x; i < n3;
       i += static_cast<int64_t>(gridDim.x) * kRedBlock) {
    const double v = b[i] - x[i] + g[i];
    r[i] = v;
    acc += v * v;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
__global__ __launch_bounds__(kRedBlock) void f64_norm2_kernel(int64_t n3,
                                                              const double *__restrict__ a,
                                                              double *__restrict__ partials) {
  __shared__ double sh[kRedBlock / 64];
  double acc = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kRedBlock + threadIdx.x; i < n3;
       i += static_cast<int64_t>(gridDim.x) * kRedBlock)
    acc += a[i] * a[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
__global__ void f64_to_f32_scaled_kernel(int64_t n3, double s, const double *__restrict__ a,
                                         float *__restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n3) out[i] = static_cast<float>(s * a[i]);
}
__global__ void f64_axpy_f32_kernel(int64_t n3, double s, const float *__restrict__ d,
                                    double *__restrict__ x) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n3) x[i] += s * static_cast<double>(d[i]);
}
// g64 = (double) g32 (the FMM product of the refinement residual)
__global__ void f32_to_f64_kernel(int64_t n3, const float *__restrict__ a, double *__restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n3) out[i] = static_cast<double>(a[i]);
}

}  // namespace

// ---------------------------------------------------------------------------

namespace {

struct Work {
  float *r, *p, *ap, *g;  // each [3][n]
  double *sc, *partials;
};

Status get_work(mf_ctx *ctx, int64_t n, Work &w) {
  void *p = nullptr;
  MF_TRY(workspace(ctx, "solve_vecs", sizeof(float) * 12 * static_cast<size_t>(n), &p));
  float *f = static_cast<float *>(p);
  w.r = f;
  w.p = f + 3 * n;
  w.ap = f + 6 * n;
  w.g = f + 9 * n;
  MF_TRY(workspace(ctx, "solve_scalars", sizeof(double) * S_COUNT, &p));
  w.sc = static_cast<double *>(p);
  MF_TRY(workspace(ctx, "solve_partials", sizeof(double) * kRedMaxBlocks * 64, &p));
  w.partials = static_cast<double *>(p);
  return Status::Ok();
}

inline V3 v3(float *base, int64_t n) { return {base, base + n, base + 2 * n}; }
inline CV3 cv3(const float *base, int64_t n) { return {base, base + n, base + 2 * n}; }

Status read_scalar(mf_ctx *ctx, const double *d, double *h) {
  MF_HIP(hipMemcpyAsync(ctx->host_scalars, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  MF_HIP(hipStreamSynchronize(ctx->stream));
  *h = ctx->host_scalars[0];
  return Status::Ok();
}

// r = b - A x with A x = x - G(kappa x); also writes p = r if p != null.
Status residual(mf_plan *plan, double kappa, CV3 b, V3 x, const Work &w, bool set_p,
                bool zero_x = false) {
  mf_ctx *ctx = plan->ctx;
  const int64_t n = plan->n;
  float *g = w.g;
  if (zero_x) {
    // x = 0: G(kappa x) = 0 and r = b, no product needed
    MF_TRY(vec_fill(ctx, n, 0.f, x.x));
    MF_TRY(vec_fill(ctx, n, 0.f, x.y));
    MF_TRY(vec_fill(ctx, n, 0.f, x.z));
    MF_TRY(vec_fill(ctx, 3 * n, 0.f, g));
  } else {
    MF_TRY(plan_apply_G(plan, static_cast<float>(kappa), x.x, x.y, x.z, nullptr, g, g + n, g + 2 * n));
  }
  V3 r = v3(w.r, n);
  V3 p = set_p ? v3(w.p, n) : V3{nullptr, nullptr, nullptr};
  hipLaunchKernelGGL(cg_init_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, b,
                     CV3{x.x, x.y, x.z}, cv3(g, n), r, p);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

// The plan's M2L precision override, restored when the solve exits.
struct M2LOverrideGuard {
  mf_plan *p;
  int saved;
  explicit M2LOverrideGuard(mf_plan *plan) : p(plan), saved(plan->m2l_override) {}
  ~M2LOverrideGuard() { p->m2l_override = saved; }
};

// Relaxed products of a CG solve (inexact Krylov): once the relative residual
// is at or below relax_rtol, the remaining products of an FMM plan run at
// relax_m2l_precision.  A product error e made at residual r perturbs the
// final residual by about e r, so a coarse M2L late in the solve costs no
// accuracy at the solve's tolerance; returns whether the switch happened.
bool relax_products(mf_plan *plan, const mf_solve_opts &o, double rel_residual) {
  if (o.relax_m2l_precision <= 0 || plan->opts.mode != MF_MODE_FMM) return false;
  const double tau = o.relax_rtol > 0.0 ? o.relax_rtol : 1e-3;
  if (rel_residual > tau) return false;
  plan->m2l_override = o.relax_m2l_precision - 1;
  return true;
}

Status solve_cg(mf_plan *plan, double kappa, CV3 b, V3 x, const mf_solve_opts &o,
                mf_solve_stats *st) {
  mf_ctx *ctx = plan->ctx;
  const int64_t n = plan->n;
  Work w{};
  MF_TRY(get_work(ctx, n, w));
  const int nb = red_blocks(n);
  MF_TRY(vec_dot_device(ctx, n, b.x, b.y, b.z, b.x, b.y, b.z, w.sc + S_BB));
  MF_TRY(residual(plan, kappa, b, x, w, true, o.zero_start != 0));
  if (!o.zero_start) st->matvecs += 1;
  MF_TRY(vec_dot_device(ctx, n, w.r, w.r + n, w.r + 2 * n, w.r, w.r + n, w.r + 2 * n, w.sc + S_RR));
  double bb = 0.0, rr = 0.0;
  MF_TRY(read_scalar(ctx, w.sc + S_BB, &bb));
  MF_TRY(read_scalar(ctx, w.sc + S_RR, &rr));
  const double bn = std::sqrt(bb);
  const double tol = o.rtol * (bn > 0.0 ? bn : 1.0);
  M2LOverrideGuard guard(plan);
  bool relaxed = false;
  int it = 0;
  while (std::sqrt(rr) > tol && st->matvecs < o.maxit) {
    if (!relaxed) relaxed = relax_products(plan, o, std::sqrt(rr) / (bn > 0.0 ? bn : 1.0));
    float *g = w.g;
    MF_TRY(plan_apply_G(plan, static_cast<float>(kappa), w.p, w.p + n, w.p + 2 * n, nullptr, g,
                        g + n, g + 2 * n));
    st->matvecs += 1;
    if (relaxed) st->relaxed_matvecs += 1;
    hipLaunchKernelGGL(cg_ap_kernel, dim3(nb), dim3(kRedBlock), 0, ctx->stream, n, cv3(w.p, n),
                       cv3(g, n), v3(w.ap, n), w.partials);
    MF_LAUNCH_CHECK();
    hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(kRedBlock), 0, ctx->stream, w.partials, nb,
                       w.sc);
    MF_LAUNCH_CHECK();
    hipLaunchKernelGGL(cg_xr_kernel, dim3(nb), dim3(kRedBlock), 0, ctx->stream, n, w.sc,
                       cv3(w.p, n), cv3(w.ap, n), x, v3(w.r, n), w.partials);
    MF_LAUNCH_CHECK();
    hipLaunchKernelGGL(cg_beta_kernel, dim3(1), dim3(kRedBlock), 0, ctx->stream, w.partials, nb,
                       w.sc);
    MF_LAUNCH_CHECK();
    hipLaunchKernelGGL(cg_p_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, w.sc,
                       cv3(w.r, n), v3(w.p, n));
    MF_LAUNCH_CHECK();
    ++it;
    MF_TRY(read_scalar(ctx, w.sc + S_RR, &rr));
    if (!std::isfinite(rr)) return Status::Err(MF_ERR_NOT_CONVERGED, "CG residual is not finite");
  }
  st->iterations = it;
  st->rel_residual = std::sqrt(rr) / (bn > 0.0 ? bn : 1.0);
  st->converged = std::sqrt(rr) <= tol;
  return Status::Ok();
}

// Givens-rotation GMRES(m) with classical Gram-Schmidt and one
// re-orthogonalisation pass (CGS2).  The basis is kept as FP32 [m+1][3][n].
Status solve_gmres(mf_plan *plan, double kappa, CV3 b, V3 x, const mf_solve_opts &o,
                   mf_solve_stats *st) {
  mf_ctx *ctx = plan->ctx;
  const int64_t n = plan->n, n3 = 3 * n;
  const int m = std::max(1, std::min(o.restart > 0 ? o.restart : 30, 60));
  Work w{};
  MF_TRY(get_work(ctx, n, w));
  void *pv = nullptr;
  MF_TRY(workspace(ctx, "gmres_basis", sizeof(float) * static_cast<size_t>(m + 1) * n3, &pv));
  float *V = static_cast<float *>(pv);
  void *pc = nullptr;
  MF_TRY(workspace(ctx, "gmres_coef", sizeof(double) * 2 * 64, &pc));
  double *dcoef = static_cast<double *>(pc);
  const int nb = red_blocks(n3);

  double bb = 0.0;
  MF_TRY(vec_dot(ctx, n, b.x, b.y, b.z, b.x, b.y, b.z, &bb));
  const double bn = std::sqrt(bb) > 0.0 ? std::sqrt(bb) : 1.0;
  const double tol = o.rtol * bn;
  std::vector<double> H((m + 1) * m), cs(m), sn(m), gvec(m + 1), y(m), hcol(m + 1);
  std::vector<double> hostbuf(64);
  double resid = 0.0;
  int total_it = 0;

  auto multidot = [&](const float *wv, int nk, double *out_host) -> Status {
    hipLaunchKernelGGL(multidot_kernel, dim3(nb), dim3(kRedBlock), 0, ctx->stream, n, nk, V, wv,
                       w.partials);
    MF_LAUNCH_CHECK();
    hipLaunchKernelGGL(multisum_kernel, dim3(nk), dim3(kRedBlock), 0, ctx->stream, w.partials, nb,
                       dcoef);
    MF_LAUNCH_CHECK();
    MF_HIP(hipMemcpyAsync(ctx->host_scalars, dcoef, sizeof(double) * nk, hipMemcpyDeviceToHost,
                          ctx->stream));
    MF_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < nk; ++k) out_host[k] = ctx->host_scalars[k];
    return Status::Ok();
  };

  bool first_cycle = true;
  while (true) {
    const bool zero_x = first_cycle && o.zero_start != 0;
    first_cycle = false;
    MF_TRY(residual(plan, kappa, b, x, w, false, zero_x));
    if (!zero_x) st->matvecs += 1;
    double rr = 0.0;
    MF_TRY(vec_dot(ctx, n, w.r, w.r + n, w.r + 2 * n, w.r, w.r + n, w.r + 2 * n, &rr));
    resid = std::sqrt(rr);
    if (resid <= tol || st->matvecs >= o.maxit) break;
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n3, 256)), dim3(256), 0, ctx->stream, n3,
                       static_cast<float>(1.0 / resid), w.r, V);
    MF_LAUNCH_CHECK();
    std::fill(gvec.begin(), gvec.end(), 0.0);
    gvec[0] = resid;
    int j = 0;
    for (; j < m && st->matvecs < o.maxit; ++j) {
      float *vj = V + static_cast<int64_t>(j) * n3;
      float *wv = V + static_cast<int64_t>(j + 1) * n3;
      // wv = A vj = vj - G(kappa vj)
      MF_TRY(plan_apply_G(plan, static_cast<float>(kappa), vj, vj + n, vj + 2 * n, nullptr, w.g,
                          w.g + n, w.g + 2 * n));
      st->matvecs += 1;
      MF_TRY(vec_axpby<float>(ctx, n, 1.f, vj, vj + n, vj + 2 * n, -1.f, w.g, w.g + n, w.g + 2 * n, wv,
                       wv + n, wv + 2 * n));
      std::fill(hcol.begin(), hcol.end(), 0.0);
      for (int pass = 0; pass < 2; ++pass) {
        MF_TRY(multidot(wv, j + 1, hostbuf.data()));
        for (int k = 0; k <= j; ++k) hcol[k] += hostbuf[k];
        MF_HIP(hipMemcpyAsync(dcoef, hostbuf.data(), sizeof(double) * (j + 1),
                              hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(combine_kernel, dim3(grid_for(n3, 256)), dim3(256), 0, ctx->stream, n3,
                           j + 1, V, dcoef, -1.f, wv);
        MF_LAUNCH_CHECK();
      }
      double ww = 0.0;
      MF_TRY(vec_dot(ctx, n, wv, wv + n, wv + 2 * n, wv, wv + n, wv + 2 * n, &ww));
      hcol[j + 1] = std::sqrt(ww);
      if (hcol[j + 1] > 0.0) {
        hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n3, 256)), dim3(256), 0, ctx->stream, n3,
                           static_cast<float>(1.0 / hcol[j + 1]), wv, wv);
        MF_LAUNCH_CHECK();
      }
      for (int k = 0; k < j; ++k) {
        const double t = cs[k] * hcol[k] + sn[k] * hcol[k + 1];
        hcol[k + 1] = -sn[k] * hcol[k] + cs[k] * hcol[k + 1];
        hcol[k] = t;
      }
      const double den = std::hypot(hcol[j], hcol[j + 1]);
      cs[j] = den > 0.0 ? hcol[j] / den : 1.0;
      sn[j] = den > 0.0 ? hcol[j + 1] / den : 0.0;
      hcol[j] = den;
      hcol[j + 1] = 0.0;
      gvec[j + 1] = -sn[j] * gvec[j];
      gvec[j] = cs[j] * gvec[j];
      for (int k = 0; k <= j; ++k) H[k * m + j] = hcol[k];
      ++total_it;
      if (std::fabs(gvec[j + 1]) <= tol) {
        ++j;
        break;
      }
    }
    // back substitution H y = g
    for (int k = j - 1; k >= 0; --k) {
      double s = gvec[k];
      for (int l = k + 1; l < j; ++l) s -= H[k * m + l] * y[l];
      y[k] = H[k * m + k] != 0.0 ? s / H[k * m + k] : 0.0;
    }
    MF_HIP(hipMemcpyAsync(dcoef, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(gmres_update_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n,
                       j, V, dcoef, x);
    MF_LAUNCH_CHECK();
    if (std::fabs(gvec[j]) <= tol && j < m) {
      // converged inside the cycle; the loop head recomputes the true residual
    }
  }
  st->iterations = total_it;
  st->rel_residual = resid / bn;
  st->converged = resid <= tol;
  return Status::Ok();
}

// The reference's nonlinear CG, magnetic_force.cu:1008-1095 with
// calc_alpha :1287-1302 and calc_beta :1305-1315.  rtol is the reference's
// absolute RMS tolerance (its default 0.1) and maxit its max_iteration.
Status solve_ref_nlcg(mf_plan *plan, double kappa, CV3 b, V3 x, const mf_solve_opts &o,
                      mf_solve_stats *st) {
  mf_ctx *ctx = plan->ctx;
  const int64_t n = plan->n;
  void *p = nullptr;
  MF_TRY(workspace(ctx, "nlcg_vecs", sizeof(float) * 18 * static_cast<size_t>(n), &p));
  float *f = static_cast<float *>(p);
  float *res = f, *grad = f + 3 * n, *ograd = f + 6 * n, *sd = f + 9 * n, *osd = f + 12 * n,
        *tmp = f + 15 * n;
  const float k = static_cast<float>(kappa);
  auto V = [&](float *a) { return v3(a, n); };
  // residual = G(kappa x) + b - x
  auto residual_fn = [&]() -> Status {
    MF_TRY(plan_apply_G(plan, k, x.x, x.y, x.z, nullptr, tmp, tmp + n, tmp + 2 * n));
    st->matvecs += 1;
    MF_TRY(vec_axpby<float>(ctx, n, 1.f, tmp, tmp + n, tmp + 2 * n, 1.f, b.x, b.y, b.z, res, res + n,
                     res + 2 * n));
    return vec_axpby<float>(ctx, n, 1.f, res, res + n, res + 2 * n, -1.f, x.x, x.y, x.z, res, res + n,
                     res + 2 * n);
  };
  // gradient = kappa G(res) - res
  auto gradient_fn = [&]() -> Status {
    MF_TRY(plan_apply_G(plan, 1.f, res, res + n, res + 2 * n, nullptr, tmp, tmp + n, tmp + 2 * n));
    st->matvecs += 1;
    return vec_axpby<float>(ctx, n, k, tmp, tmp + n, tmp + 2 * n, -1.f, res, res + n, res + 2 * n, grad,
                     grad + n, grad + 2 * n);
  };
  // alpha = -(D(sd).res)/(|D(sd)|^2 + 1e-10), D(v) = G(kappa v) - v
  auto alpha_fn = [&](double *alpha) -> Status {
    MF_TRY(plan_apply_G(plan, k, sd, sd + n, sd + 2 * n, nullptr, tmp, tmp + n, tmp + 2 * n));
    st->matvecs += 1;
    MF_TRY(vec_axpby<float>(ctx, n, 1.f, tmp, tmp + n, tmp + 2 * n, -1.f, sd, sd + n, sd + 2 * n, tmp,
                     tmp + n, tmp + 2 * n));
    double nt = 0.0, tr = 0.0;
    MF_TRY(vec_dot(ctx, n, tmp, tmp + n, tmp + 2 * n, tmp, tmp + n, tmp + 2 * n, &nt));
    if (nt < 1e-10) {
      *alpha = 0.0;
      return Status::Ok();
    }
    MF_TRY(vec_dot(ctx, n, tmp, tmp + n, tmp + 2 * n, res, res + n, res + 2 * n, &tr));
    *alpha = -(tr / (nt + 1e-10));
    return Status::Ok();
  };
  auto err_fn = [&](double *e) -> Status {
    double rr = 0.0;
    MF_TRY(vec_dot(ctx, n, res, res + n, res + 2 * n, res, res + n, res + 2 * n, &rr));
    *e = 0.5 * rr;
    return Status::Ok();
  };
  const double tol = o.rtol > 0.0 ? o.rtol : 0.1;
  const int max_iteration = o.maxit > 0 ? o.maxit : 80;
  double err = 0.0, last = 0.0, alpha = 0.0;
  MF_TRY(residual_fn());
  MF_TRY(err_fn(&err));
  last = err;
  int i_step = 0, it = 0;
  if (std::sqrt(err / static_cast<double>(n)) >= tol) {
    MF_TRY(gradient_fn());
    MF_TRY(vec_axpby<float>(ctx, n, -1.f, grad, grad + n, grad + 2 * n, 0.f, nullptr, nullptr, nullptr,
                     sd, sd + n, sd + 2 * n));
    MF_TRY(alpha_fn(&alpha));
    MF_TRY(vec_axpby<float>(ctx, n, 1.f, x.x, x.y, x.z, static_cast<float>(alpha), sd, sd + n,
                     sd + 2 * n, x.x, x.y, x.z));
    for (; it < max_iteration; ++it) {
      MF_HIP(hipMemcpyAsync(ograd, grad, sizeof(float) * 3 * n, hipMemcpyDeviceToDevice, ctx->stream));
      MF_HIP(hipMemcpyAsync(osd, sd, sizeof(float) * 3 * n, hipMemcpyDeviceToDevice, ctx->stream));
      MF_TRY(residual_fn());
      MF_TRY(err_fn(&err));
      if (std::sqrt(err / static_cast<double>(n)) < tol) break;
      if (err > last) break;
      last = err;
      ++i_step;
      MF_TRY(gradient_fn());
      double beta = 0.0;
      if (i_step % 10 != 0) {
        double gg = 0.0, den = 0.0;
        MF_TRY(vec_dot(ctx, n, grad, grad + n, grad + 2 * n, grad, grad + n, grad + 2 * n, &gg));
        MF_TRY(vec_axpby<float>(ctx, n, 1.f, ograd, ograd + n, ograd + 2 * n, -1.f, grad, grad + n,
                         grad + 2 * n, tmp, tmp + n, tmp + 2 * n));
        MF_TRY(vec_dot(ctx, n, osd, osd + n, osd + 2 * n, tmp, tmp + n, tmp + 2 * n, &den));
        beta = -gg / (den + 1e-10);
        if (beta < 0.0) beta = 0.0;
      }
      MF_TRY(vec_axpby<float>(ctx, n, -1.f, grad, grad + n, grad + 2 * n, static_cast<float>(beta), osd,
                       osd + n, osd + 2 * n, sd, sd + n, sd + 2 * n));
      MF_TRY(alpha_fn(&alpha));
      MF_TRY(vec_axpby<float>(ctx, n, 1.f, x.x, x.y, x.z, static_cast<float>(alpha), sd, sd + n,
                       sd + 2 * n, x.x, x.y, x.z));
    }
  }
  (void)V;
  double bb = 0.0;
  MF_TRY(vec_dot(ctx, n, b.x, b.y, b.z, b.x, b.y, b.z, &bb));
  st->iterations = it;
  st->rel_residual = std::sqrt(2.0 * err) / (bb > 0.0 ? std::sqrt(bb) : 1.0);
  st->converged = std::sqrt(err / static_cast<double>(n)) < tol;
  return Status::Ok();
}

}  // namespace

Status solve(mf_plan *plan, double kappa, const float *bx, const float *by, const float *bz,
             float *xx, float *xy, float *xz, const mf_solve_opts &opts, mf_solve_stats *st) {
  *st = mf_solve_stats{};
  if (plan->n == 0) {
    st->converged = 1;
    return Status::Ok();
  }
  mf_solve_opts o = opts;
  if (o.maxit <= 0) o.maxit = 1000;
  if (o.rtol <= 0.0 && o.method != MF_SOLVE_REF_NLCG) o.rtol = 1e-6;
  const CV3 b{bx, by, bz};
  V3 x{xx, xy, xz};
  switch (o.method) {
    case MF_SOLVE_CG:
      MF_TRY(solve_cg(plan, kappa, b, x, o, st));
      break;
    case MF_SOLVE_GMRES:
      MF_TRY(solve_gmres(plan, kappa, b, x, o, st));
      break;
    case MF_SOLVE_REF_NLCG:
      if (o.zero_start) {
        MF_TRY(vec_fill(plan->ctx, plan->n, 0.f, xx));
        MF_TRY(vec_fill(plan->ctx, plan->n, 0.f, xy));
        MF_TRY(vec_fill(plan->ctx, plan->n, 0.f, xz));
      }
      MF_TRY(solve_ref_nlcg(plan, kappa, b, x, o, st));
      break;
    default:
      return Status::Err(MF_ERR_INVALID_ARG, "unknown solve method");
  }
  if (opts.true_residual && o.method != MF_SOLVE_REF_NLCG) {
    Work w{};
    MF_TRY(get_work(plan->ctx, plan->n, w));
    MF_TRY(residual(plan, kappa, b, x, w, false));
    st->matvecs += 1;
    double rr = 0.0, bb = 0.0;
    const int64_t n = plan->n;
    MF_TRY(vec_dot(plan->ctx, n, w.r, w.r + n, w.r + 2 * n, w.r, w.r + n, w.r + 2 * n, &rr));
    MF_TRY(vec_dot(plan->ctx, n, bx, by, bz, bx, by, bz, &bb));
    st->rel_residual_true = std::sqrt(rr) / (bb > 0.0 ? std::sqrt(bb) : 1.0);
  }
  return Status::Ok();
}

// Iterative refinement with x and r = b - A x in FP64; corrections from the
// plan's own CG.  All-pairs plans form the residual with the FP64 all-pairs
// product; FMM plans with the plan's FMM product (its own M2L precision), the
// inner solves running at opts.inner_m2l_precision.
Status solve_f64(mf_plan *plan, double kappa, const double *bx, const double *by,
                 const double *bz, double *xx, double *xy, double *xz, const mf_solve_opts &opts,
                 mf_solve_stats *st) {
  *st = mf_solve_stats{};
  mf_ctx *ctx = plan->ctx;
  const int64_t n = plan->n, n3 = 3 * n;
  if (n == 0) return Status::Ok();
  MF_ARG(by == bx + n && bz == bx + 2 * n && xy == xx + n && xz == xx + 2 * n,
         "solve_f64 expects contiguous [3][n] arrays");
  const double *px = nullptr, *py = nullptr, *pz = nullptr;
  if (plan->opts.mode != MF_MODE_FMM) MF_TRY(plan_pos64(plan, &px, &py, &pz));
  void *p = nullptr;
  MF_TRY(workspace(ctx, "f64_vecs", sizeof(double) * 3 * static_cast<size_t>(n3), &p));
  double *m64 = static_cast<double *>(p), *g64 = m64 + n3, *r64 = g64 + n3;
  MF_TRY(workspace(ctx, "f64_f32", sizeof(float) * 6 * static_cast<size_t>(n), &p));
  float *rhs = static_cast<float *>(p), *d = rhs + n3;
  MF_TRY(workspace(ctx, "f64_partials", sizeof(double) * (kRedMaxBlocks + 8), &p));
  double *partials = static_cast<double *>(p), *dsum = partials + kRedMaxBlocks;
  const int nb = red_blocks(n3);
  const bool fmm = plan->opts.mode == MF_MODE_FMM;
  const double rtol = opts.rtol > 0.0 ? opts.rtol : 1e-12;
  const int max_ref = opts.refine_fp64 > 0 ? opts.refine_fp64 : 10;
  const double inner_rtol = opts.inner_rtol > 0.0 ? opts.inner_rtol : (fmm ? 1e-3 : 1e-5);
  // FMM plans: the residual product runs at the plan's own M2L precision, the
  // inner solves at inner_m2l_precision (the override is restored on every exit)
  M2LOverrideGuard guard(plan);
  const int outer_prec = plan->m2l_override;
  const int inner_prec = opts.inner_m2l_precision > 0 ? opts.inner_m2l_precision - 1 : outer_prec;
  double bn = 0.0;
  for (int ref = 0; ref <= max_ref; ++ref) {
    if (fmm) {
      // g = G (kappa x) from the FP32 rounding of kappa x, promoted to FP64
      hipLaunchKernelGGL(f64_to_f32_scaled_kernel, dim3(grid_for(n3, 256)), dim3(256), 0,
                         ctx->stream, n3, kappa, xx, rhs);
      MF_LAUNCH_CHECK();
      plan->m2l_override = outer_prec;
      MF_TRY(plan_apply_G(plan, 1.f, rhs, rhs + n, rhs + 2 * n, nullptr, d, d + n, d + 2 * n));
      hipLaunchKernelGGL(f32_to_f64_kernel, dim3(grid_for(n3, 256)), dim3(256), 0, ctx->stream,
                         n3, d, g64);
      MF_LAUNCH_CHECK();
    } else {
      hipLaunchKernelGGL(f64_moments_kernel, dim3(grid_for(n3, 256)), dim3(256), 0, ctx->stream, n,
                         kappa, xx, m64);
      MF_LAUNCH_CHECK();
      MF_TRY(field_allpairs_f64(ctx, n, px, py, pz, n, px, py, pz, m64, m64 + n, m64 + 2 * n,
                                plan->h, nullptr, g64, g64 + n, g64 + 2 * n));
    }
    st->matvecs += 1;
    if (ref == 0) {
      hipLaunchKernelGGL(f64_norm2_kernel, dim3(nb), dim3(kRedBlock), 0, ctx->stream, n3, bx,
                         partials);
      MF_LAUNCH_CHECK();
      MF_TRY(sum_partials(ctx, partials, nb, dsum));
      MF_TRY(read_scalar(ctx, dsum, &bn));
      bn = std::sqrt(bn);
      if (bn == 0.0) bn = 1.0;
    }
    hipLaunchKernelGGL(f64_residual_kernel, dim3(nb), dim3(kRedBlock), 0, ctx->stream, n3, bx, xx,
                       g64, r64, partials);
    MF_LAUNCH_CHECK();
    MF_TRY(sum_partials(ctx, partials, nb, dsum));
    double rr = 0.0;
    MF_TRY(read_scalar(ctx, dsum, &rr));
    const double rn = std::sqrt(rr);
    st->rel_residual_true = rn / bn;
    st->refinements = ref;
    if (rn <= rtol * bn) {
      st->converged = 1;
      break;
    }
    if (ref == max_ref) break;
    // d solves A d = r / |r| in FP32, x += |r| d
    hipLaunchKernelGGL(f64_to_f32_scaled_kernel, dim3(grid_for(n3, 256)), dim3(256), 0,
                       ctx->stream, n3, 1.0 / rn, r64, rhs);
    MF_LAUNCH_CHECK();
    mf_solve_opts inner{};
    inner.method = MF_SOLVE_CG;
    inner.zero_start = 1;
    inner.maxit = opts.maxit > 0 ? opts.maxit : 500;
    inner.rtol = inner_rtol;
    mf_solve_stats ist{};
    if (fmm) plan->m2l_override = inner_prec;
    MF_TRY(solve(plan, kappa, rhs, rhs + n, rhs + 2 * n, d, d + n, d + 2 * n, inner, &ist));
    st->matvecs += ist.matvecs;
    st->iterations += ist.iterations;
    hipLaunchKernelGGL(f64_axpy_f32_kernel, dim3(grid_for(n3, 256)), dim3(256), 0, ctx->stream,
                       n3, rn, d, xx);
    MF_LAUNCH_CHECK();
  }
  st->rel_residual = st->rel_residual_true;
  return Status::Ok();
}

// ---------------------------------------------------------------------------
// Batched CG over the independent systems of a batch plan: the CG recurrences
// of every system run in lockstep, with per-system scalars (alpha, beta, r.r)
// and one shared FMM product per iteration.  A system whose residual reaches
// its tolerance is frozen (alpha = beta = 0) while the others continue.

namespace {

struct SegState {
  double *bb, *rr, *rrn, *pap, *alpha, *beta, *partial, *maxratio;
  int32_t *active;
};

// Per-chunk partial dot products: partial[c] = sum over chunk c of a . b.
__global__ __launch_bounds__(kRedBlock) void seg_dot_kernel(const int64_t *__restrict__ seg, CV3 a,
                                                            CV3 b, double *__restrict__ partial) {
  __shared__ double sh[kRedBlock / 64];
  const int64_t i0 = seg[2 * blockIdx.x], i1 = seg[2 * blockIdx.x + 1];
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kRedBlock)
    acc += static_cast<double>(a.x[i]) * b.x[i] + static_cast<double>(a.y[i]) * b.y[i] +
           static_cast<double>(a.z[i]) * b.z[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[s] = sum of the partials of system s's chunks (one thread per system).
__global__ void seg_sum_kernel(int nsys, const int32_t *__restrict__ seg_first,
                               const double *__restrict__ partial, double *__restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  double acc = 0.0;
  for (int c = seg_first[s]; c < seg_first[s + 1]; ++c) acc += partial[c];
  out[s] = acc;
}

__global__ void fill_i32_kernel(int n, int32_t v, int32_t *__restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void bcg_ap_kernel(int64_t n, CV3 p, CV3 g, V3 ap) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ap.x[i] = p.x[i] - g.x[i];
  ap.y[i] = p.y[i] - g.y[i];
  ap.z[i] = p.z[i] - g.z[i];
}

__global__ void bcg_alpha_kernel(int nsys, SegState st) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  st.alpha[s] = (st.active[s] && st.pap[s] != 0.0) ? st.rr[s] / st.pap[s] : 0.0;
}

__global__ void bcg_xr_kernel(int64_t n, const int32_t *__restrict__ sys_of,
                              const double *__restrict__ alpha, CV3 p, CV3 ap, V3 x, V3 r) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float a = static_cast<float>(alpha[sys_of[i]]);
  if (a == 0.f) return;  // frozen system: its product entries may be stale
  x.x[i] = fmaf(a, p.x[i], x.x[i]);
  x.y[i] = fmaf(a, p.y[i], x.y[i]);
  x.z[i] = fmaf(a, p.z[i], x.z[i]);
  r.x[i] = fmaf(-a, ap.x[i], r.x[i]);
  r.y[i] = fmaf(-a, ap.y[i], r.y[i]);
  r.z[i] = fmaf(-a, ap.z[i], r.z[i]);
}

// beta, r.r update, activity and the largest relative residual over active
// systems (one block).
__global__ __launch_bounds__(kRedBlock) void bcg_beta_kernel(int nsys, double rtol, int first,
                                                             SegState st) {
  __shared__ double sh[kRedBlock];
  double mx = 0.0;
  for (int s = threadIdx.x; s < nsys; s += kRedBlock) {
    const double bn = st.bb[s] > 0.0 ? sqrt(st.bb[s]) : 1.0;
    if (first) {
      st.rr[s] = st.rrn[s];
      st.beta[s] = 0.0;
    } else if (st.active[s]) {
      st.beta[s] = st.rr[s] != 0.0 ? st.rrn[s] / st.rr[s] : 0.0;
      st.rr[s] = st.rrn[s];
    } else {
      st.beta[s] = 0.0;
    }
    const double ratio = sqrt(st.rr[s]) / bn;
    st.active[s] = (st.active[s] || first) && ratio > rtol ? 1 : 0;
    mx = fmax(mx, ratio);
  }
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int o = kRedBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) st.maxratio[0] = sh[0];
}

__global__ void bcg_p_kernel(int64_t n, const int32_t *__restrict__ sys_of,
                             const double *__restrict__ beta, CV3 r, V3 p) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float b = static_cast<float>(beta[sys_of[i]]);
  p.x[i] = fmaf(b, p.x[i], r.x[i]);
  p.y[i] = fmaf(b, p.y[i], r.y[i]);
  p.z[i] = fmaf(b, p.z[i], r.z[i]);
}

__global__ void kappa_of_system_kernel(int64_t n, const int32_t *__restrict__ sys_of,
                                       const float *__restrict__ kappa_sys,
                                       float *__restrict__ kappa_part) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) kappa_part[i] = kappa_sys[sys_of[i]];
}

Status build_segments(mf_plan *plan) {
  if (plan->nseg > 0) return Status::Ok();
  std::vector<int64_t> seg;
  std::vector<int32_t> first;
  constexpr int64_t kSeg = 16384;
  for (int s = 0; s < plan->nsys; ++s) {
    first.push_back(static_cast<int32_t>(seg.size() / 2));
    for (int64_t b0 = plan->sys_offsets[s]; b0 < plan->sys_offsets[s + 1]; b0 += kSeg) {
      seg.push_back(b0);
      seg.push_back(std::min(plan->sys_offsets[s + 1], b0 + kSeg));
    }
  }
  first.push_back(static_cast<int32_t>(seg.size() / 2));
  const int nseg = static_cast<int>(seg.size() / 2);
  MF_TRY(plan->seg.reserve(sizeof(int64_t) * std::max<size_t>(seg.size(), 2)));
  MF_TRY(plan->seg_first.reserve(sizeof(int32_t) * first.size()));
  mf_ctx *ctx = plan->ctx;
  if (!seg.empty())
    MF_HIP(hipMemcpyAsync(plan->seg.get(), seg.data(), sizeof(int64_t) * seg.size(),
                          hipMemcpyHostToDevice, ctx->stream));
  MF_HIP(hipMemcpyAsync(plan->seg_first.get(), first.data(), sizeof(int32_t) * first.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  MF_HIP(hipStreamSynchronize(ctx->stream));
  plan->nseg = nseg;
  return Status::Ok();
}

}  // namespace

Status solve_batch(mf_plan *plan, const double *kappa, const float *bx, const float *by,
                   const float *bz, float *xx, float *xy, float *xz, const mf_solve_opts &opts,
                   mf_solve_stats *st, double *sys_rel_residual) {
  *st = mf_solve_stats{};
  mf_ctx *ctx = plan->ctx;
  const int64_t n = plan->n;
  const int nsys = plan->nsys;
  if (n == 0) {
    st->converged = 1;
    for (int s = 0; s < nsys && sys_rel_residual; ++s) sys_rel_residual[s] = 0.0;
    return Status::Ok();
  }
  MF_ARG(opts.method == MF_SOLVE_CG, "batch solves use CG");
  const double rtol = opts.rtol > 0.0 ? opts.rtol : 1e-6;
  const int maxit = opts.maxit > 0 ? opts.maxit : 1000;
  MF_TRY(build_segments(plan));
  // per-particle kappa (caller order), applied to the moments of every
  // product: expanded on the device from the nsys values
  {
    std::vector<float> ks(static_cast<size_t>(nsys));
    for (int s = 0; s < nsys; ++s) ks[static_cast<size_t>(s)] = static_cast<float>(kappa[s]);
    void *pk = nullptr;
    MF_TRY(workspace(ctx, "batch_kappa", sizeof(float) * ks.size(), &pk));
    MF_HIP(hipMemcpyAsync(pk, ks.data(), sizeof(float) * ks.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    MF_TRY(plan->kappa_part.reserve(sizeof(float) * static_cast<size_t>(n)));
    hipLaunchKernelGGL(kappa_of_system_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n,
                       plan->sys_of.as<int32_t>(), static_cast<const float *>(pk),
                       plan->kappa_part.as<float>());
    MF_LAUNCH_CHECK();
    // the host copy's source must outlive it
    MF_HIP(hipStreamSynchronize(ctx->stream));
  }
  Work w{};
  MF_TRY(get_work(ctx, n, w));
  void *pv = nullptr;
  MF_TRY(workspace(ctx, "batch_scalars", sizeof(double) * (7 * static_cast<size_t>(nsys) + plan->nseg + 1) +
                                             sizeof(int32_t) * static_cast<size_t>(nsys), &pv));
  SegState S{};
  double *d = static_cast<double *>(pv);
  S.bb = d;
  S.rr = d + nsys;
  S.rrn = d + 2 * nsys;
  S.pap = d + 3 * nsys;
  S.alpha = d + 4 * nsys;
  S.beta = d + 5 * nsys;
  S.maxratio = d + 6 * nsys;
  S.partial = d + 7 * nsys;
  S.active = reinterpret_cast<int32_t *>(d + 7 * nsys + plan->nseg + 1);
  const int32_t *sys_of = plan->sys_of.as<int32_t>();
  const int64_t *seg = plan->seg.as<int64_t>();
  const int32_t *seg_first = plan->seg_first.as<int32_t>();
  const int nseg = plan->nseg;
  const CV3 b{bx, by, bz};
  V3 x{xx, xy, xz};
  auto seg_dot = [&](CV3 a, CV3 c, double *out) -> Status {
    hipLaunchKernelGGL(seg_dot_kernel, dim3(nseg), dim3(kRedBlock), 0, ctx->stream, seg, a, c,
                       S.partial);
    MF_LAUNCH_CHECK();
    hipLaunchKernelGGL(seg_sum_kernel, dim3((nsys + 255) / 256), dim3(256), 0, ctx->stream, nsys,
                       seg_first, static_cast<const double *>(S.partial), out);
    MF_LAUNCH_CHECK();
    return Status::Ok();
  };
  hipLaunchKernelGGL(fill_i32_kernel, dim3((nsys + 255) / 256), dim3(256), 0, ctx->stream, nsys, 0,
                     S.active);
  MF_LAUNCH_CHECK();
  MF_TRY(seg_dot(b, b, S.bb));
  // products with per-particle kappa for the span of the solve
  struct WeightGuard {
    mf_plan *p;
    ~WeightGuard() {
      p->weights = nullptr;
      p->active_sys = nullptr;
    }
  } guard{plan};
  plan->weights = plan->kappa_part.as<float>();
  MF_TRY(residual(plan, 1.0, b, x, w, true, opts.zero_start != 0));
  // products of the iterations skip converged systems, whose entries of g
  // then keep earlier values: start g finite (the fill is a kernel, so the
  // captured iteration holds no memset)
  MF_TRY(vec_fill<float>(ctx, 3 * n, 0.f, w.g));
  if (!opts.zero_start) st->matvecs += 1;
  const CV3 r = cv3(w.r, n), p = cv3(w.p, n), ap = cv3(w.ap, n), g = cv3(w.g, n);
  MF_TRY(seg_dot(r, r, S.rrn));
  hipLaunchKernelGGL(bcg_beta_kernel, dim3(1), dim3(kRedBlock), 0, ctx->stream, nsys, rtol, 1, S);
  MF_LAUNCH_CHECK();
  double mx = 0.0;
  MF_TRY(read_scalar(ctx, S.maxratio, &mx));
  plan->active_sys = S.active;
  // one CG iteration, enqueued on ctx->stream without host reads
  auto body = [&]() -> Status {
    MF_TRY(plan_apply_G(plan, 1.f, w.p, w.p + n, w.p + 2 * n, nullptr, w.g, w.g + n, w.g + 2 * n));
    hipLaunchKernelGGL(bcg_ap_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, p, g,
                       v3(w.ap, n));
    MF_LAUNCH_CHECK();
    MF_TRY(seg_dot(p, ap, S.pap));
    hipLaunchKernelGGL(bcg_alpha_kernel, dim3((nsys + 255) / 256), dim3(256), 0, ctx->stream, nsys, S);
    MF_LAUNCH_CHECK();
    hipLaunchKernelGGL(bcg_xr_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, sys_of,
                       static_cast<const double *>(S.alpha), p, ap, x, v3(w.r, n));
    MF_LAUNCH_CHECK();
    MF_TRY(seg_dot(r, r, S.rrn));
    hipLaunchKernelGGL(bcg_beta_kernel, dim3(1), dim3(kRedBlock), 0, ctx->stream, nsys, rtol, 0, S);
    MF_LAUNCH_CHECK();
    hipLaunchKernelGGL(bcg_p_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, sys_of,
                       static_cast<const double *>(S.beta), r, v3(w.p, n));
    MF_LAUNCH_CHECK();
    return Status::Ok();
  };
  // Graph mode: after one eager iteration (every workspace exists by then) one
  // iteration is captured and replayed graph_iters times per convergence
  // check.  The capture and the replays run on a non-null stream (the
  // context's own when it works on the null stream), ordered with the
  // context's stream by events.  Converged systems are frozen and their boxes
  // are skipped by the products, so replays past convergence cost little.
  struct GraphState {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~GraphState() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    }
  } gs;
  const int graph_iters = opts.graph_iters;
  hipStream_t work = ctx->stream ? ctx->stream : ctx->own_stream;
  M2LOverrideGuard m2l_guard(plan);
  bool relaxed = false;
  int it = 0;
  while (mx > rtol && st->matvecs < maxit) {
    if (!relaxed && relax_products(plan, opts, mx)) {
      relaxed = true;
      // a captured iteration holds the old precision's kernels: capture anew
      if (gs.exec) MF_HIP(hipGraphExecDestroy(gs.exec));
      if (gs.graph) MF_HIP(hipGraphDestroy(gs.graph));
      gs.exec = nullptr;
      gs.graph = nullptr;
    }
    const int m0 = st->matvecs;
    if (graph_iters > 0 && it >= 1) {
      if (!gs.exec) {
        if (!gs.ev[0]) MF_HIP(hipEventCreateWithFlags(&gs.ev[0], hipEventDisableTiming));
        if (!gs.ev[1]) MF_HIP(hipEventCreateWithFlags(&gs.ev[1], hipEventDisableTiming));
        hipStream_t saved = ctx->stream;
        const bool timing = ctx->timing;
        MF_HIP(hipEventRecord(gs.ev[0], saved));
        MF_HIP(hipStreamWaitEvent(work, gs.ev[0], 0));
        ctx->stream = work;
        ctx->timing = false;  // no timing events inside the graph
        hipError_t e0 = hipStreamBeginCapture(work, hipStreamCaptureModeThreadLocal);
        Status sb = e0 == hipSuccess ? body() : Status::Ok();
        hipError_t e1 = e0 == hipSuccess ? hipStreamEndCapture(work, &gs.graph) : e0;
        ctx->stream = saved;
        ctx->timing = timing;
        MF_TRY(sb);
        MF_HIP(e1);
        MF_HIP(hipGraphInstantiate(&gs.exec, gs.graph, nullptr, nullptr, 0));
      } else {
        MF_HIP(hipEventRecord(gs.ev[0], ctx->stream));
        MF_HIP(hipStreamWaitEvent(work, gs.ev[0], 0));
      }
      const int k = std::min(graph_iters, maxit - st->matvecs);
      for (int j = 0; j < k; ++j) MF_HIP(hipGraphLaunch(gs.exec, work));
      MF_HIP(hipEventRecord(gs.ev[1], work));
      MF_HIP(hipStreamWaitEvent(ctx->stream, gs.ev[1], 0));
      st->matvecs += k;
      it += k;
    } else {
      MF_TRY(body());
      st->matvecs += 1;
      ++it;
    }
    if (relaxed) st->relaxed_matvecs += st->matvecs - m0;
    MF_TRY(read_scalar(ctx, S.maxratio, &mx));
    if (!std::isfinite(mx)) return Status::Err(MF_ERR_NOT_CONVERGED, "CG residual is not finite");
  }
  st->iterations = it;
  st->rel_residual = mx;
  st->converged = mx <= rtol;
  if (sys_rel_residual) {
    std::vector<double> rr(nsys), bb(nsys);
    MF_HIP(hipMemcpyAsync(rr.data(), S.rr, sizeof(double) * nsys, hipMemcpyDeviceToHost, ctx->stream));
    MF_HIP(hipMemcpyAsync(bb.data(), S.bb, sizeof(double) * nsys, hipMemcpyDeviceToHost, ctx->stream));
    MF_HIP(hipStreamSynchronize(ctx->stream));
    for (int s = 0; s < nsys; ++s)
      sys_rel_residual[s] = std::sqrt(rr[s]) / (bb[s] > 0.0 ? std::sqrt(bb[s]) : 1.0);
  }
  return Status::Ok();
}

}  // namespace mf
