// mf_check64.hip -- FP64 all-pairs field and force on the GPU.
//
// This is the FP64 check path: it evaluates the reference's formulas
// literally (Dipole_B.kern:46-104 for the field, Dipole_hessian.kern:53-238
// with the original Horner curves for the force tensor) in double precision,
// so FP32 and mixed-precision results can be scored at sizes the CPU oracle
// cannot reach.  It is validated against the CPU oracle at small N.
// One target per lane; sources staged through LDS in tiles of 256.
#include <hip/hip_runtime.h>

#include "mf_allpairs.hpp"
#include "mf_internal.hpp"

namespace mf {
namespace {

constexpr int kB = 256;

__constant__ double c_C1[4][5] = {
    {9.97813616438174e-09, -2.97897856524718e-08, 2.38918644566813e-09, 4.53199938857366e-08,
     2.44617454752747e-11},
    {-2.76473728643294e-09, 2.86975546540539e-08, -9.94582836806651e-08, 1.25129924573675e-07,
     -2.37010166723652e-08},
    {-1.09679990621465e-09, 9.77055663264614e-09, -2.54781238661150e-08, 2.65020634884934e-09,
     5.00787562417835e-08},
    {3.79927162333632e-10, -6.26368404962679e-09, 3.94760528277489e-08, -1.13580541622200e-07,
     1.27491333574323e-07}};
__constant__ double c_C2[4][5] = {
    {6.69550479838731e-08, -1.61753307173877e-07, 1.68213714992711e-08, 1.34558143036838e-07,
     1.10976027980100e-10},
    {-3.08460139955194e-08, 2.29192245602275e-07, -5.88399621128587e-07, 5.61170054591844e-07,
     -1.14421132829680e-07},
    {3.50477408060213e-09, -5.25956271895141e-08, 2.78876509535747e-07, -6.24199554212217e-07,
     4.91807818904985e-07},
    {7.33485346367840e-10, -9.58855788627803e-09, 4.37085309763591e-08, -7.48004594092261e-08,
     2.34161209605651e-08}};

__device__ inline double horner4(const double *c, double x) {
  return x * (x * (x * (x * c[0] + c[1]) + c[2]) + c[3]) + c[4];
}

// Field pair, Dipole_B.kern:46-104 in FP64.
__device__ inline void field_pair64(double dx, double dy, double dz, double mx, double my,
                                    double mz, double h, double &ox, double &oy, double &oz) {
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  const double q = r / h;
  const double inv_pi = 1.0 / kPi;
  const double sigma = inv_pi / (h * h * h);
  double wa, wd;
  if (q >= 2.0) {
    wa = 0.75 * inv_pi / (r * r * r);
    wd = 0.0;
  } else if (q >= 1.0) {
    const double q2 = q * q, q3 = q2 * q;
    const double poly = q3 * q3 / 6.0 - 6.0 * q3 * q2 / 5.0 + 3.0 * q2 * q2 - 8.0 * q3 / 3.0 +
                        1.0 / 15.0;
    wa = -3.0 * poly * inv_pi / (4.0 * r * r * r);
    const double t = 2.0 - q;
    wd = sigma * 0.25 * t * t * t;
  } else {
    wa = (q * q * (0.375 * q - 0.9) + 1.0) * sigma;
    const double t = 2.0 - q, u = 1.0 - q;
    wd = sigma * (0.25 * t * t * t - u * u * u);
  }
  const double invr = (r == 0.0) ? 0.0 : 1.0 / r;
  const double ux = dx * invr, uy = dy * invr, uz = dz * invr;
  const double v1 = (ux * mx + uy * my + uz * mz) * (wa - wd);
  const double v2 = wa / 3.0 - wd;
  ox += ux * v1 - v2 * mx;
  oy += uy * v1 - v2 * my;
  oz += uz * v1 - v2 * mz;
}

__global__ __launch_bounds__(kB) void field64_kernel(int64_t nt, const double *__restrict__ tx,
                                                     const double *__restrict__ ty,
                                                     const double *__restrict__ tz, int64_t ns,
                                                     const double *__restrict__ sx,
                                                     const double *__restrict__ sy,
                                                     const double *__restrict__ sz,
                                                     const double *__restrict__ mx,
                                                     const double *__restrict__ my,
                                                     const double *__restrict__ mz, double h,
                                                     double h0, double h1, double h2,
                                                     double *__restrict__ ox,
                                                     double *__restrict__ oy,
                                                     double *__restrict__ oz) {
  __shared__ double s[6][kB];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  const bool valid = i < nt;
  const double px = valid ? tx[i] : 0.0, py = valid ? ty[i] : 0.0, pz = valid ? tz[i] : 0.0;
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int64_t base = 0; base < ns; base += kB) {
    const int64_t j = base + threadIdx.x;
    __syncthreads();
    if (j < ns) {
      s[0][threadIdx.x] = sx[j];
      s[1][threadIdx.x] = sy[j];
      s[2][threadIdx.x] = sz[j];
      s[3][threadIdx.x] = mx[j];
      s[4][threadIdx.x] = my[j];
      s[5][threadIdx.x] = mz[j];
    }
    __syncthreads();
    const int cnt = static_cast<int>(ns - base < kB ? ns - base : kB);
    for (int k = 0; k < cnt; ++k)
      field_pair64(px - s[0][k], py - s[1][k], pz - s[2][k], s[3][k], s[4][k], s[5][k], h, ax,
                   ay, az);
  }
  if (valid) {
    ox[i] = ax + h0;
    oy[i] = ay + h1;
    oz[i] = az + h2;
  }
}

// Force: F_i = sum_j B(x_i - x_j)[m_j] m_i with the isotropic closed form of
// Dipole_hessian.kern (far q > 4, near (0, 4] with the original curves).
__global__ __launch_bounds__(kB) void force64_kernel(int64_t n, int64_t nt,
                                                     const int64_t *__restrict__ idx,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ y,
                                                     const double *__restrict__ z,
                                                     const double *__restrict__ mx,
                                                     const double *__restrict__ my,
                                                     const double *__restrict__ mz, double h,
                                                     double *__restrict__ fx,
                                                     double *__restrict__ fy,
                                                     double *__restrict__ fz) {
  __shared__ double s[6][kB];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  const bool valid = i < nt;
  int64_t ii = valid ? (idx ? idx[i] : i) : 0;
  ii = (ii >= 0 && ii < n) ? ii : 0;
  const double px = x[ii], py = y[ii], pz = z[ii];
  const double nx = mx[ii], ny = my[ii], nz = mz[ii];
  const double h4 = h * h * h * h;
  const double cfar = kMu0 / (4.0 * kPi);
  double F0 = 0.0, F1 = 0.0, F2 = 0.0;
  for (int64_t base = 0; base < n; base += kB) {
    const int64_t j = base + threadIdx.x;
    __syncthreads();
    if (j < n) {
      s[0][threadIdx.x] = x[j];
      s[1][threadIdx.x] = y[j];
      s[2][threadIdx.x] = z[j];
      s[3][threadIdx.x] = mx[j];
      s[4][threadIdx.x] = my[j];
      s[5][threadIdx.x] = mz[j];
    }
    __syncthreads();
    const int cnt = static_cast<int>(n - base < kB ? n - base : kB);
    for (int k = 0; k < cnt; ++k) {
      const double dx = px - s[0][k], dy = py - s[1][k], dz = pz - s[2][k];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      const double q = r * (1.0 / h);
      if (q == 0.0) continue;
      double a, b;
      if (q > 4.0) {
        const double r4 = r * r * r * r;
        a = 3.0 * cfar / r4;
        b = -15.0 * cfar / r4;
      } else {
        const int kk = (q <= 1.0) ? 0 : (q <= 2.0) ? 1 : (q <= 3.0) ? 2 : 3;
        const double c1 = horner4(c_C1[kk], q), c2 = horner4(c_C2[kk], q);
        a = c1 / h4;
        b = (c2 - 3.0 * c1) / h4;
      }
      const double ir = 1.0 / r;
      const double ux = dx * ir, uy = dy * ir, uz = dz * ir;
      const double smx = s[3][k], smy = s[4][k], smz = s[5][k];
      const double um = ux * smx + uy * smy + uz * smz;
      const double un = ux * nx + uy * ny + uz * nz;
      const double mn = smx * nx + smy * ny + smz * nz;
      // F_a = sum_b [a (u_a m_b + m_a u_b + d_ab um) + b u_a u_b um] n_b
      const double e = a * mn + b * um * un;
      F0 += ux * e + a * (smx * un + nx * um);
      F1 += uy * e + a * (smy * un + ny * um);
      F2 += uz * e + a * (smz * un + nz * um);
    }
  }
  if (valid) {
    fx[i] = F0;
    fy[i] = F1;
    fz[i] = F2;
  }
}

}  // namespace

Status field_allpairs_f64(mf_ctx *ctx, int64_t nt, const double *tx, const double *ty,
                          const double *tz, int64_t ns, const double *sx, const double *sy,
                          const double *sz, const double *mx, const double *my,
                          const double *mz, double h, const double *hconst, double *ox,
                          double *oy, double *oz) {
  if (nt == 0) return Status::Ok();
  KernelTimer timer(ctx, "field64");
  hipLaunchKernelGGL(field64_kernel, dim3(grid_for(nt, kB)), dim3(kB), 0, ctx->stream, nt, tx, ty,
                     tz, ns, sx, sy, sz, mx, my, mz, h, hconst ? hconst[0] : 0.0,
                     hconst ? hconst[1] : 0.0, hconst ? hconst[2] : 0.0, ox, oy, oz);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

Status force_allpairs_f64(mf_ctx *ctx, int64_t n, const double *x, const double *y,
                          const double *z, const double *mx, const double *my, const double *mz,
                          double h, double *fx, double *fy, double *fz) {
  if (n == 0) return Status::Ok();
  KernelTimer timer(ctx, "force64");
  hipLaunchKernelGGL(force64_kernel, dim3(grid_for(n, kB)), dim3(kB), 0, ctx->stream, n, n,
                     static_cast<const int64_t *>(nullptr), x, y, z, mx, my, mz, h, fx, fy, fz);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

Status force_subset_f64(mf_ctx *ctx, int64_t n, const double *x, const double *y,
                        const double *z, const double *mx, const double *my, const double *mz,
                        double h, int64_t nt, const int64_t *idx, double *fx, double *fy,
                        double *fz) {
  if (nt == 0 || n == 0) return Status::Ok();
  KernelTimer timer(ctx, "force64");
  hipLaunchKernelGGL(force64_kernel, dim3(grid_for(nt, kB)), dim3(kB), 0, ctx->stream, n, nt, idx,
                     x, y, z, mx, my, mz, h, fx, fy, fz);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

}  // namespace mf
