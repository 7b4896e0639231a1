// mf_expansion.hpp -- solid-harmonic expansions of the dipole far field.
//
// The far field of the reference's kernels (q >= 2 for the field, q > 4 for
// the force tensor) is that of point dipoles, i.e. derivatives of the Laplace
// potential psi(x) = sum_j m_j . grad_y 1/|x - y_j|, with
//   H = -grad psi / (4 pi),   B(d)[m] = mu0 grad H = -mu0/(4 pi) grad grad psi.
// The reference expands it with fmmtl's spherical multipoles
// (fmmtl_minimum/kernel/DipoleBSpherical.hpp:95-230,
// DipoleHessianSpherical.hpp:83-451, Util/SphericalMultipole3D.hpp:46-538);
// this file is an independent formulation in scaled regular/irregular solid
// harmonics (recursions of Dehnen 2014, "A fast multipole method for stellar
// dynamics"):
//   R_0^0 = 1,  R_n^n = -(x + i y)/(2n) R_{n-1}^{n-1},
//   R_n^m = ((2n-1) z R_{n-1}^m - r^2 R_{n-2}^m) / ((n+m)(n-m)),
//   I_0^0 = 1/r, I_n^n = -(2n-1)(x + i y)/r^2 I_{n-1}^{n-1},
//   I_n^m = ((2n-1) z I_{n-1}^m - (n-1+m)(n-1-m) I_{n-2}^m) / r^2,
//   X_n^{-m} = (-1)^m conj(X_n^m),
// so that 1/|x - y| = sum_{n,m} conj(R_n^m(y)) I_n^m(x) for |y| < |x|, and
//   d/dz R_n^m = R_{n-1}^m, (d/dx + i d/dy) R_n^m = R_{n-1}^{m+1},
//   (d/dx - i d/dy) R_n^m = -R_{n-1}^{m-1}.
// These derivative identities give the dipole P2M and a gradient/Hessian L2P
// without the 1/sin(theta) factors of the reference's spherical L2T.
//
// Expansions are stored per box as (P+1)^2 real numbers ("dofs"): degree n
// occupies [n^2, n^2 + 2n]: Re X_n^0, then Re X_n^m, Im X_n^m for m = 1..n.
// Box-size scaling: with box side s, Mt_n = M_n / s^n and Lt_n = L_n s^(n+1),
// which makes M2M, M2L and L2L independent of the tree level.
#pragma once

#include <hip/hip_runtime.h>

namespace mf {

struct cplx {
  float re, im;
};

__host__ __device__ constexpr int dof_re(int n, int m) { return m == 0 ? n * n : n * n + 2 * m - 1; }
__host__ __device__ constexpr int dof_im(int n, int m) { return n * n + 2 * m; }

// Regular harmonics R_n^m(x), 0 <= m <= n <= P, packed triangle index
// t(n, m) = n (n + 1) / 2 + m.
template <int P>
__device__ __forceinline__ void regular_tri(float x, float y, float z, cplx (&R)[(P + 1) * (P + 2) / 2]) {
  const float r2 = x * x + y * y + z * z;
  R[0] = {1.f, 0.f};
#pragma unroll
  for (int n = 1; n <= P; ++n) {
    const cplx a = R[(n - 1) * n / 2 + (n - 1)];
    const float f = -1.f / (2.f * n);
    // -(x + i y)/(2n) * a
    R[n * (n + 1) / 2 + n] = {f * (x * a.re - y * a.im), f * (x * a.im + y * a.re)};
  }
#pragma unroll
  for (int m = 0; m <= P; ++m) {
#pragma unroll
    for (int n = m + 1; n <= P; ++n) {
      const float inv = 1.f / static_cast<float>((n + m) * (n - m));
      const cplx a = R[(n - 1) * n / 2 + m];
      cplx v = {(2 * n - 1) * z * a.re, (2 * n - 1) * z * a.im};
      if (n - 2 >= m) {
        const cplx b = R[(n - 2) * (n - 1) / 2 + m];
        v.re -= r2 * b.re;
        v.im -= r2 * b.im;
      }
      R[n * (n + 1) / 2 + m] = {v.re * inv, v.im * inv};
    }
  }
}

template <int P>
__device__ __forceinline__ cplx rget(const cplx (&R)[(P + 1) * (P + 2) / 2], int n, int m) {
  // R_n^m for |m| <= n (negative m through the conjugate symmetry), 0 outside.
  if (n < 0 || m > n || -m > n) return {0.f, 0.f};
  if (m >= 0) return R[n * (n + 1) / 2 + m];
  const cplx a = R[n * (n + 1) / 2 - m];
  return ((-m) & 1) ? cplx{-a.re, a.im} : cplx{a.re, -a.im};
}

}  // namespace mf
