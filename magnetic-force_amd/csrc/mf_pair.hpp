// mf_pair.hpp -- per-pair field and force-tensor coefficients for gfx950.
//
// Both pair kernels of the reference reduce to two radial scalars per pair:
//
// Field (fmmtl_minimum/kernel/Dipole_B.kern:46-104, == magnetic_force.cu:63-98)
//   K(d) m = u (u.m)(w_avr - w_dens) - (w_avr/3 - w_dens) m,   u = d/|d|
//          = c [ 3 alpha d (d.m) - beta m ],                    c = 1/(4 pi)
// with (w~ = w / c)
//   far  (q >= 2): alpha = |d|^-5,                 beta = |d|^-3
//   q < 1       : alpha = (0.8 - 0.5 q)/h^5,       beta = (-8/3 + 4.8 q^2 - 2.5 q^3)/h^3
//   1 <= q < 2  : alpha = (g(q)/r^3 - ((2-q)/h)^3)/(3 r^2),
//                 beta  =  g(q)/(3 r^3) - ((2-q)/h)^3,
//                 g(q) = -3 (q^6/6 - 6q^5/5 + 3q^4 - 8q^3/3 + 1/15)
// The q < 1 forms are the reference's w_avr/w_dens polynomials with the
// common factors cancelled exactly (w_avr - w_dens = sigma q^2 (0.6 - 0.375 q)),
// which also removes the 0/0 at r = 0 (self term (2/3) sigma m).
// g is evaluated re-centred at q = 1.5 (exact rational re-expansion), which
// lowers its FP32 conditioning from 118 to 2.1.
//
// Force tensor (Dipole_hessian.kern:53-238, == magnetic_force.cu:239-427)
//   B(d)[m]_ab = P (d_a m_b + m_a d_b + delta_ab d.m) + R d_a d_b (d.m)
// scaled by s = 3 mu0/(4 pi):
//   far  (q > 4) : P = |d|^-5,            R = -5 |d|^-7
//   0 < q <= 4   : P = C1(q)/(s h^4 r),   R = (C2(q) - 3 C1(q))/(s h^4 r^3)
//   q == 0       : 0
// The rotation-frame product M U M^T of the reference equals this isotropic
// form.  C1, C2 are the reference's fitted quartics, evaluated re-centred at
// the middle of their interval (k + 0.5) so that FP32 Horner does not cancel
// near q = 4 (absolute conditioning <= 2 on every interval).
#pragma once

#include <hip/hip_runtime.h>

namespace mf {

constexpr double kPi = 3.14159265358979323846;
constexpr double kMu0 = 1.25663706e-6;  // the reference's truncated value

// Field constants passed by value (kernel arguments live in SGPRs).
struct FieldConsts {
  float h;
  float inv_h;
  float near2;   // (2h)^2
  float inv_h3;
  float inv_h5;
  float g[7];    // g(1.5 + t) = sum g[k] t^k
  float tab_scale;  // kFieldTabN / (2h): r -> table coordinate q kFieldTabN / 2
};

// Near-field coefficient table of the FMM near-field kernels.  Interval k of
// kFieldTabN covers u = q kFieldTabN / 2 in [k, k + 1) (q = 1 is a boundary,
// so no interval straddles the two forms) and holds two cubics in t = u - k,
// for alpha and for beta, each interpolating the exact FP64 form at the four
// Chebyshev points of the interval: tab[8k + 0..3] = alpha coefficients,
// tab[8k + 4..7] = beta coefficients (ascending powers, h included).  Entry
// kFieldTabN continues the last interval's cubics one interval further, so a
// u that rounds up to kFieldTabN at q -> 2 reads a finite value with no
// clamp.  At FP32 inputs q in [0, 2) the error is at most 1.5e-7 of
// max(|alpha|, (2h)^-5) and of max(|beta|, (2h)^-3); the FP32 closed form of
// field_near_coeffs stays within 7e-7 and 1.6e-6.
//
// Device copies of both tables (split_near_table) hold the first curve's
// cubic of every entry, then the second's: at a 16-byte stride, lanes that
// read neighbouring entries meet in half as many LDS banks as at 32 bytes.
constexpr int kFieldTabN = 64;
constexpr int kFieldTabEntries = kFieldTabN + 1;

// Force-tensor constants; the curve table goes to LDS for per-lane lookup.
struct HessConsts {
  float h;
  float inv_h;
  float near2;       // (4h)^2
  float tab_scale;   // kHessTabN / (4h): r -> table coordinate q kHessTabN / 4
  // curve[k][0..4] = C1 re-centred at k+0.5 (ascending powers),
  // curve[k][5..9] = C2 - 3 C1 re-centred likewise, both times 1/(s h^4),
  // s = 3 mu0/(4 pi), in FP64 before rounding; [10..11] padding.
  float curve[4][12];
};

// Force-curve table of the FMM near field: kHessTabN intervals over q in
// (0, 4], 16 per unit, each inside one of the reference's quartic pieces
// (piece p covers (p, p + 1], so an integer q stays with the lower piece).
// Entry e = ceil(u), u = q kHessTabN / 4, covers u in (e - 1, e] and holds two
// cubics in t = u - e, in (-1, 0]: [0..3] C1 / (s h^4), [4..7]
// -(C2 - 3 C1) / (5 s h^4), the R / -5 the FMM force pair uses.  Entries
// 1 .. kHessTabN interpolate at their interval's four Chebyshev points;
// entry 0 (u = 0, a pair at d = 0) continues entry 1's cubics and entry
// kHessTabN + 1 (u rounded above kHessTabN at q -> 4) entry kHessTabN's, so
// the lookup needs no clamp.  Within 9e-8 of the largest curve value (the
// FP32 rounding of the curves is 6e-8).
constexpr int kHessTabN = 64;
constexpr int kHessTabEntries = kHessTabN + 2;

FieldConsts make_field_consts(double h);
HessConsts make_hess_consts(double h);
void make_field_table(double h, float *tab);  // kFieldTabEntries x 8 floats
void make_hess_table(double h, float *tab);   // kHessTabEntries x 8 floats
// [entries][8] -> [2][entries][4]: the device layout of both tables.
void split_near_table(const float *tab, int entries, float *out);

// ---------------------------------------------------------------------------
// Field pair coefficients (FP32).  Returns alpha, beta such that the pair adds
// c (3 alpha d (d.m) - beta m).  `ir` = 1/|d| (may be inf at r2 == 0).
__device__ __forceinline__ void field_near_coeffs(const FieldConsts &fc, float r2, float ir,
                                                  float &alpha, float &beta) {
  const float r = r2 * ir;                 // |d|; NaN when r2 == 0, handled below
  const float q = (r2 > 0.f) ? r * fc.inv_h : 0.f;
  // inner interval q < 1
  const float a_in = fmaf(-0.5f, q, 0.8f) * fc.inv_h5;
  const float q2 = q * q;
  const float b_in = fmaf(q2, fmaf(-2.5f, q, 4.8f), -2.6666666666666667f) * fc.inv_h3;
  // outer interval 1 <= q < 2
  const float t = q - 1.5f;
  float g = fc.g[6];
  g = fmaf(g, t, fc.g[5]);
  g = fmaf(g, t, fc.g[4]);
  g = fmaf(g, t, fc.g[3]);
  g = fmaf(g, t, fc.g[2]);
  g = fmaf(g, t, fc.g[1]);
  g = fmaf(g, t, fc.g[0]);
  const float tmq = (2.f - q) * fc.inv_h;
  const float wd = tmq * tmq * tmq;
  const float ir3 = ir * ir * ir;
  const float wa = g * ir3;
  const float a_out = (wa - wd) * (ir * ir) * 0.33333333333333333f;
  const float b_out = wa * 0.33333333333333333f - wd;
  const bool inner = q < 1.f;
  alpha = inner ? a_in : a_out;
  beta = inner ? b_in : b_out;
}

// The same coefficients from the table (split layout, in LDS), for
// r2 < (2h)^2: 12 VALU operations and two LDS reads instead of about 30.
__device__ __forceinline__ void field_near_coeffs_tab(const float4 *tab, float tab_scale, float r2,
                                                      float ir, float &alpha, float &beta) {
  // u = q kFieldTabN / 2.  The clamp keeps the pair at d = 0 (ir = inf) at
  // u = 0 instead of 0 x inf; u < kFieldTabN up to rounding at q -> 2, and
  // entry kFieldTabN takes a u rounded up to it.
  const float irc = __builtin_amdgcn_fmed3f(ir, 0.f, 1e30f);
  const float u = r2 * tab_scale * irc;
  const int k = static_cast<int>(u);
  const float t = __builtin_amdgcn_fractf(u);
  const float4 a = tab[k], b = tab[kFieldTabEntries + k];
  alpha = fmaf(fmaf(fmaf(a.w, t, a.z), t, a.y), t, a.x);
  beta = fmaf(fmaf(fmaf(b.w, t, b.z), t, b.y), t, b.x);
}

// Force pair coefficients for the near field (q <= 4).  `tab` points at the
// 4 x 12 curve table in LDS.  P, R as in the header comment (scaled by 1/s).
__device__ __forceinline__ void hess_near_coeffs(const float *tab, float inv_h, float r2,
                                                 float ir, float &P, float &R) {
  const float r = r2 * ir;
  const float q = r * inv_h;
  int k = static_cast<int>(ceilf(q)) - 1;
  k = k < 0 ? 0 : (k > 3 ? 3 : k);
  const float t = q - (static_cast<float>(k) + 0.5f);
  const float4 *row = reinterpret_cast<const float4 *>(tab + 12 * k);
  const float4 v0 = row[0], v1 = row[1], v2 = row[2];
  // C1: v0.x v0.y v0.z v0.w v1.x ; D = C2 - 3 C1: v1.y v1.z v1.w v2.x v2.y
  float c1 = fmaf(fmaf(fmaf(fmaf(v1.x, t, v0.w), t, v0.z), t, v0.y), t, v0.x);
  float dd = fmaf(fmaf(fmaf(fmaf(v2.y, t, v2.x), t, v1.w), t, v1.z), t, v1.y);
  const bool live = r2 > 0.f;
  P = live ? c1 * ir : 0.f;
  R = live ? dd * (ir * ir * ir) : 0.f;
}

// The two curves of a force pair within 4h (r2 <= near2) from the table
// (LDS): P = c1 irc and R / -5 = dq irc^3.  The caller clamps the inverse
// distance irc to a finite value, so a pair at d = 0 (a particle and itself,
// or two coincident ones) gets finite coefficients, whose terms all vanish
// with d.
__device__ __forceinline__ void hess_near_curves_tab(const float4 *tab, float tab_scale, float r2,
                                                     float irc, float &c1, float &dq) {
  const float u = r2 * irc * tab_scale;  // in [0, kHessTabN], up to rounding
  const float kc = ceilf(u);              // entry 0 .. kHessTabN + 1
  const float t = u - kc;
  const int e = static_cast<int>(kc);
  const float4 a = tab[e], b = tab[kHessTabEntries + e];
  c1 = fmaf(fmaf(fmaf(a.w, t, a.z), t, a.y), t, a.x);
  dq = fmaf(fmaf(fmaf(b.w, t, b.z), t, b.y), t, b.x);
}

}  // namespace mf
