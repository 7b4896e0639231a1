// mf_allpairs.hip -- tiled all-pairs field and force kernels for gfx950.
//
// Replaces the reference's naive kernels (magnetic_solver/magnetic_force.cu):
//   cuda_Apply_G            :8-110    -> field_allpairs_kernel
//   cuda_ext_field_kernel   :120-237  -> field_allpairs_kernel (separate sources)
//   fitted_force_ofM_kernel :516-605  -> force_allpairs_kernel
//   cuda_ext_fitted_force_tensor_kernel :432-512 -> hessian_allpairs_kernel
//
// Design (one wavefront = 64 lanes x TPT targets):
//  * Sources live in HBM as 32-byte records {x,y,z,0, mx,my,mz,0}, padded to a
//    whole tile with far-away zero-moment records.  The source index is
//    uniform across the wave, so record loads are scalar loads and the pair
//    math takes source values straight from SGPRs: no LDS traffic and no
//    barriers in the inner loop.
//  * Sources are cut into tiles of 64 with a precomputed bounding box.  Each
//    wave computes the bounding box of its own targets once; a tile whose
//    box-to-box distance exceeds the smoothing support runs a branch-free
//    far-field loop (pure dipole, one v_rsq_f32 per pair).  Other tiles run
//    the mixed loop that also evaluates the smoothed near-field polynomials.
//  * Per-tile sums are FP32; tile sums are accumulated in FP64, so the
//    summation error does not grow with N.
//  * The source range is split into chunks along grid.y to give the 256 CUs
//    enough waves; chunk partials (FP64) are summed in a fixed order by a
//    separate reduction kernel, so results are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "mf_allpairs.hpp"
#include "mf_internal.hpp"
#include "mf_pair.hpp"

namespace mf {

namespace {

constexpr int kBlock = 256;
constexpr float kPadCoord = 1.0e18f;  // far-away position of padding records

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// ---------------------------------------------------------------------------
// Packing and tile boxes.

__global__ void pack_sources_kernel(int64_t n, int64_t n_pad, const float *__restrict__ x,
                                    const float *__restrict__ y, const float *__restrict__ z,
                                    const float *__restrict__ mx, const float *__restrict__ my,
                                    const float *__restrict__ mz, float scale,
                                    SrcRec *__restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_pad) return;
  SrcRec r;
  if (i < n) {
    r.x = x[i];
    r.y = y[i];
    r.z = z[i];
    r.mx = mx ? scale * mx[i] : 0.f;
    r.my = my ? scale * my[i] : 0.f;
    r.mz = mz ? scale * mz[i] : 0.f;
  } else {
    r.x = kPadCoord;
    r.y = kPadCoord;
    r.z = kPadCoord;
    r.mx = r.my = r.mz = 0.f;
  }
  r.w = 0.f;
  r.mw = 0.f;
  out[i] = r;
}

__global__ void pack_moments_kernel(int64_t n, const float *__restrict__ mx,
                                    const float *__restrict__ my, const float *__restrict__ mz,
                                    float scale, SrcRec *__restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 m = make_float4(scale * mx[i], scale * my[i], scale * mz[i], 0.f);
  reinterpret_cast<float4 *>(out + i)[1] = m;
}

// One wave per tile of kTile sources.  Boxes ignore padding records.
__global__ void tile_box_kernel(int64_t n, int64_t ntiles, const SrcRec *__restrict__ src,
                                float4 *__restrict__ boxes) {
  const int lane = threadIdx.x & 63;
  const int64_t tile = static_cast<int64_t>(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (tile >= ntiles) return;
  const int64_t j = tile * kTile + lane;
  float lx = INFINITY, ly = INFINITY, lz = INFINITY;
  float hx = -INFINITY, hy = -INFINITY, hz = -INFINITY;
  if (j < n) {
    const SrcRec s = src[j];
    lx = hx = s.x;
    ly = hy = s.y;
    lz = hz = s.z;
  }
  lx = wave_min(lx);
  ly = wave_min(ly);
  lz = wave_min(lz);
  hx = wave_max(hx);
  hy = wave_max(hy);
  hz = wave_max(hz);
  if (lane == 0) {
    boxes[2 * tile] = make_float4(lx, ly, lz, 0.f);
    boxes[2 * tile + 1] = make_float4(hx, hy, hz, 0.f);
  }
}

// Distance^2 between two axis-aligned boxes (0 when they overlap).
__device__ __forceinline__ float box_dist2(float tlx, float tly, float tlz, float thx, float thy,
                                           float thz, float4 slo, float4 shi) {
  const float dx = fmaxf(0.f, fmaxf(tlx - shi.x, slo.x - thx));
  const float dy = fmaxf(0.f, fmaxf(tly - shi.y, slo.y - thy));
  const float dz = fmaxf(0.f, fmaxf(tlz - shi.z, slo.z - thz));
  return dx * dx + dy * dy + dz * dz;
}

// ---------------------------------------------------------------------------
// Field: partial[chunk][c][i] = c * sum_{j in chunk} (3 alpha d (d.m) - beta m)_c

template <int TPT>
__global__ __launch_bounds__(kBlock) void field_allpairs_kernel(
    const float4 *__restrict__ tpos, int tstride, int64_t nt, const SrcRec *__restrict__ src,
    const float4 *__restrict__ boxes, int64_t ntiles, int64_t tiles_per_chunk, FieldConsts fc,
    int64_t nt_ld, double *__restrict__ partial) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
  const int64_t tbase = wave * 64 * TPT;
  if (tbase >= nt) return;  // whole wave idle (uniform)

  float tx[TPT], ty[TPT], tz[TPT];
  float lx = INFINITY, ly = INFINITY, lz = INFINITY, hx = -INFINITY, hy = -INFINITY,
        hz = -INFINITY;
#pragma unroll
  for (int k = 0; k < TPT; ++k) {
    int64_t i = tbase + k * 64 + lane;
    const bool valid = i < nt;
    i = valid ? i : nt - 1;
    const float4 p = tpos[i * tstride];
    tx[k] = p.x;
    ty[k] = p.y;
    tz[k] = p.z;
    if (valid) {
      lx = fminf(lx, p.x);
      ly = fminf(ly, p.y);
      lz = fminf(lz, p.z);
      hx = fmaxf(hx, p.x);
      hy = fmaxf(hy, p.y);
      hz = fmaxf(hz, p.z);
    }
  }
  lx = uniform_f(wave_min(lx));
  ly = uniform_f(wave_min(ly));
  lz = uniform_f(wave_min(lz));
  hx = uniform_f(wave_max(hx));
  hy = uniform_f(wave_max(hy));
  hz = uniform_f(wave_max(hz));

  double acc[TPT][3];
#pragma unroll
  for (int k = 0; k < TPT; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;

  const int64_t chunk = blockIdx.y;
  const int64_t t0 = chunk * tiles_per_chunk;
  const int64_t t1 = t0 + tiles_per_chunk < ntiles ? t0 + tiles_per_chunk : ntiles;

  for (int64_t tile = t0; tile < t1; ++tile) {
    const float4 slo = boxes[2 * tile];
    const float4 shi = boxes[2 * tile + 1];
    const float dist2 = uniform_f(box_dist2(lx, ly, lz, hx, hy, hz, slo, shi));
    float a1[TPT][3], a2[TPT][3];
#pragma unroll
    for (int k = 0; k < TPT; ++k) a1[k][0] = a1[k][1] = a1[k][2] = a2[k][0] = a2[k][1] = a2[k][2] = 0.f;
    const SrcRec *__restrict__ s = src + tile * kTile;
    if (dist2 >= fc.near2) {
      // far tile: every pair has q >= 2 (pure dipole)
#pragma unroll 8
      for (int j = 0; j < kTile; ++j) {
        const float sx = s[j].x, sy = s[j].y, sz = s[j].z;
        const float mx = s[j].mx, my = s[j].my, mz = s[j].mz;
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
          const float dx = tx[k] - sx, dy = ty[k] - sy, dz = tz[k] - sz;
          const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          const float ir = __builtin_amdgcn_rsqf(r2);
          const float ir2 = ir * ir;
          const float ir3 = ir2 * ir;
          const float dm = fmaf(dx, mx, fmaf(dy, my, dz * mz));
          const float s1 = dm * (ir3 * ir2);
          a1[k][0] = fmaf(dx, s1, a1[k][0]);
          a1[k][1] = fmaf(dy, s1, a1[k][1]);
          a1[k][2] = fmaf(dz, s1, a1[k][2]);
          a2[k][0] = fmaf(mx, ir3, a2[k][0]);
          a2[k][1] = fmaf(my, ir3, a2[k][1]);
          a2[k][2] = fmaf(mz, ir3, a2[k][2]);
        }
      }
    } else {
      // mixed tile: smoothed near field where q < 2
#pragma unroll 2
      for (int j = 0; j < kTile; ++j) {
        const float sx = s[j].x, sy = s[j].y, sz = s[j].z;
        const float mx = s[j].mx, my = s[j].my, mz = s[j].mz;
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
          const float dx = tx[k] - sx, dy = ty[k] - sy, dz = tz[k] - sz;
          const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          const float ir = __builtin_amdgcn_rsqf(r2);
          const float ir2 = ir * ir;
          float beta = ir2 * ir;
          float alpha = beta * ir2;
          if (r2 < fc.near2) field_near_coeffs(fc, r2, ir, alpha, beta);
          const float dm = fmaf(dx, mx, fmaf(dy, my, dz * mz));
          const float s1 = dm * alpha;
          a1[k][0] = fmaf(dx, s1, a1[k][0]);
          a1[k][1] = fmaf(dy, s1, a1[k][1]);
          a1[k][2] = fmaf(dz, s1, a1[k][2]);
          a2[k][0] = fmaf(mx, beta, a2[k][0]);
          a2[k][1] = fmaf(my, beta, a2[k][1]);
          a2[k][2] = fmaf(mz, beta, a2[k][2]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < TPT; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        acc[k][c] += 3.0 * static_cast<double>(a1[k][c]) - static_cast<double>(a2[k][c]);
  }

  const double cfac = 1.0 / (4.0 * kPi);
  double *out = partial + chunk * 3 * nt_ld;
#pragma unroll
  for (int k = 0; k < TPT; ++k) {
    const int64_t i = tbase + k * 64 + lane;
    if (i < nt) {
      out[i] = acc[k][0] * cfac;
      out[nt_ld + i] = acc[k][1] * cfac;
      out[2 * nt_ld + i] = acc[k][2] * cfac;
    }
  }
}

// ---------------------------------------------------------------------------
// Force: targets are the sources themselves (same records), target moment n.
//   F = s sum_j [ d (P (m.n) + R (d.m)(d.n)) + m P (d.n) + n P (d.m) ]
// with s = 3 mu0/(4 pi), accumulated as A_d (3), A_m (3), A_n (1).

template <int TPT>
__global__ __launch_bounds__(kBlock) void force_allpairs_kernel(
    const SrcRec *__restrict__ tgt, int64_t nt, const SrcRec *__restrict__ src,
    const float4 *__restrict__ boxes, int64_t ntiles, int64_t tiles_per_chunk, HessConsts hc,
    int64_t nt_ld, double *__restrict__ partial) {
  __shared__ float4 s_tab[4 * 3];
  if (threadIdx.x < 12) {
    const float *row = &hc.curve[threadIdx.x / 3][(threadIdx.x % 3) * 4];
    s_tab[threadIdx.x] = make_float4(row[0], row[1], row[2], row[3]);
  }
  __syncthreads();
  const float *tab = reinterpret_cast<const float *>(s_tab);

  const int lane = threadIdx.x & 63;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
  const int64_t tbase = wave * 64 * TPT;
  if (tbase >= nt) return;

  float tx[TPT], ty[TPT], tz[TPT], nx[TPT], ny[TPT], nz[TPT];
  float lx = INFINITY, ly = INFINITY, lz = INFINITY, hx = -INFINITY, hy = -INFINITY,
        hz = -INFINITY;
#pragma unroll
  for (int k = 0; k < TPT; ++k) {
    int64_t i = tbase + k * 64 + lane;
    const bool valid = i < nt;
    i = valid ? i : nt - 1;
    const SrcRec t = tgt[i];
    tx[k] = t.x;
    ty[k] = t.y;
    tz[k] = t.z;
    nx[k] = t.mx;
    ny[k] = t.my;
    nz[k] = t.mz;
    if (valid) {
      lx = fminf(lx, t.x);
      ly = fminf(ly, t.y);
      lz = fminf(lz, t.z);
      hx = fmaxf(hx, t.x);
      hy = fmaxf(hy, t.y);
      hz = fmaxf(hz, t.z);
    }
  }
  lx = uniform_f(wave_min(lx));
  ly = uniform_f(wave_min(ly));
  lz = uniform_f(wave_min(lz));
  hx = uniform_f(wave_max(hx));
  hy = uniform_f(wave_max(hy));
  hz = uniform_f(wave_max(hz));

  double acc[TPT][4];
#pragma unroll
  for (int k = 0; k < TPT; ++k) acc[k][0] = acc[k][1] = acc[k][2] = acc[k][3] = 0.0;
  double accm[TPT][3];
#pragma unroll
  for (int k = 0; k < TPT; ++k) accm[k][0] = accm[k][1] = accm[k][2] = 0.0;

  const int64_t chunk = blockIdx.y;
  const int64_t t0 = chunk * tiles_per_chunk;
  const int64_t t1 = t0 + tiles_per_chunk < ntiles ? t0 + tiles_per_chunk : ntiles;

  for (int64_t tile = t0; tile < t1; ++tile) {
    const float4 slo = boxes[2 * tile];
    const float4 shi = boxes[2 * tile + 1];
    const float dist2 = uniform_f(box_dist2(lx, ly, lz, hx, hy, hz, slo, shi));
    float ad[TPT][3], am[TPT][3], an[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
      ad[k][0] = ad[k][1] = ad[k][2] = 0.f;
      am[k][0] = am[k][1] = am[k][2] = 0.f;
      an[k] = 0.f;
    }
    const SrcRec *__restrict__ s = src + tile * kTile;
    if (dist2 > hc.near2) {
#pragma unroll 4
      for (int j = 0; j < kTile; ++j) {
        const float sx = s[j].x, sy = s[j].y, sz = s[j].z;
        const float mx = s[j].mx, my = s[j].my, mz = s[j].mz;
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
          const float dx = tx[k] - sx, dy = ty[k] - sy, dz = tz[k] - sz;
          const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          const float ir = __builtin_amdgcn_rsqf(r2);
          const float ir2 = ir * ir;
          const float P = ir2 * ir2 * ir;            // r^-5
          const float dm = fmaf(dx, mx, fmaf(dy, my, dz * mz));
          const float dn = fmaf(dx, nx[k], fmaf(dy, ny[k], dz * nz[k]));
          const float mn = fmaf(mx, nx[k], fmaf(my, ny[k], mz * nz[k]));
          // e = P mn + R dm dn, R = -5 P ir2
          const float e = P * fmaf(-5.f * ir2, dm * dn, mn);
          const float f = P * dn;
          ad[k][0] = fmaf(dx, e, ad[k][0]);
          ad[k][1] = fmaf(dy, e, ad[k][1]);
          ad[k][2] = fmaf(dz, e, ad[k][2]);
          am[k][0] = fmaf(mx, f, am[k][0]);
          am[k][1] = fmaf(my, f, am[k][1]);
          am[k][2] = fmaf(mz, f, am[k][2]);
          an[k] = fmaf(P, dm, an[k]);
        }
      }
    } else {
#pragma unroll 2
      for (int j = 0; j < kTile; ++j) {
        const float sx = s[j].x, sy = s[j].y, sz = s[j].z;
        const float mx = s[j].mx, my = s[j].my, mz = s[j].mz;
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
          const float dx = tx[k] - sx, dy = ty[k] - sy, dz = tz[k] - sz;
          const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          const float ir = __builtin_amdgcn_rsqf(r2);
          const float ir2 = ir * ir;
          float P = ir2 * ir2 * ir;
          float R = -5.f * P * ir2;
          if (!(r2 > hc.near2)) hess_near_coeffs(tab, hc.inv_h, r2, ir, P, R);
          const float dm = fmaf(dx, mx, fmaf(dy, my, dz * mz));
          const float dn = fmaf(dx, nx[k], fmaf(dy, ny[k], dz * nz[k]));
          const float mn = fmaf(mx, nx[k], fmaf(my, ny[k], mz * nz[k]));
          const float e = fmaf(R, dm * dn, P * mn);
          const float f = P * dn;
          ad[k][0] = fmaf(dx, e, ad[k][0]);
          ad[k][1] = fmaf(dy, e, ad[k][1]);
          ad[k][2] = fmaf(dz, e, ad[k][2]);
          am[k][0] = fmaf(mx, f, am[k][0]);
          am[k][1] = fmaf(my, f, am[k][1]);
          am[k][2] = fmaf(mz, f, am[k][2]);
          an[k] = fmaf(P, dm, an[k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc[k][c] += static_cast<double>(ad[k][c]);
        accm[k][c] += static_cast<double>(am[k][c]);
      }
      acc[k][3] += static_cast<double>(an[k]);
    }
  }

  const double sfac = 3.0 * kMu0 / (4.0 * kPi);
  double *out = partial + chunk * 3 * nt_ld;
#pragma unroll
  for (int k = 0; k < TPT; ++k) {
    const int64_t i = tbase + k * 64 + lane;
    if (i < nt) {
      out[i] = sfac * (acc[k][0] + accm[k][0] + static_cast<double>(nx[k]) * acc[k][3]);
      out[nt_ld + i] = sfac * (acc[k][1] + accm[k][1] + static_cast<double>(ny[k]) * acc[k][3]);
      out[2 * nt_ld + i] = sfac * (acc[k][2] + accm[k][2] + static_cast<double>(nz[k]) * acc[k][3]);
    }
  }
}

// ---------------------------------------------------------------------------
// Force tensor (6 components).  One target per lane; mixed loop only (this is
// the external-source product, not on the solve path).

// Force-tensor field at separate targets (the external tensor T_ext of
// eval_hext), tiled like the field kernel: one target per lane, sources in
// 64-record tiles through scalar loads; a tile whose box lies beyond 4h of the
// wave's targets runs the branch-free dipole form.  Chunks of source tiles
// along grid.y write FP64 partials [chunk][6][nt_ld] (xx yy zz xy xz yz).
__global__ __launch_bounds__(kBlock) void hessian_allpairs_kernel(
    const float4 *__restrict__ tpos, int64_t nt, const SrcRec *__restrict__ src,
    const float4 *__restrict__ boxes, int64_t ntiles, int64_t tiles_per_chunk, HessConsts hc,
    int64_t nt_ld, double *__restrict__ partial) {
  __shared__ float4 s_tab[12];
  if (threadIdx.x < 12) {
    const float *row = &hc.curve[threadIdx.x / 3][(threadIdx.x % 3) * 4];
    s_tab[threadIdx.x] = make_float4(row[0], row[1], row[2], row[3]);
  }
  __syncthreads();
  const float *tab = reinterpret_cast<const float *>(s_tab);
  const int lane = threadIdx.x & 63;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
  int64_t i = wave * 64 + lane;
  if (wave * 64 >= nt) return;  // whole wave idle (uniform)
  const bool valid = i < nt;
  i = valid ? i : nt - 1;
  const float4 t = tpos[i];
  float lx = valid ? t.x : INFINITY, ly = valid ? t.y : INFINITY, lz = valid ? t.z : INFINITY;
  float hx = valid ? t.x : -INFINITY, hy = valid ? t.y : -INFINITY, hz = valid ? t.z : -INFINITY;
  lx = uniform_f(wave_min(lx));
  ly = uniform_f(wave_min(ly));
  lz = uniform_f(wave_min(lz));
  hx = uniform_f(wave_max(hx));
  hy = uniform_f(wave_max(hy));
  hz = uniform_f(wave_max(hz));
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int64_t chunk = blockIdx.y;
  const int64_t t0 = chunk * tiles_per_chunk;
  const int64_t t1 = t0 + tiles_per_chunk < ntiles ? t0 + tiles_per_chunk : ntiles;
  for (int64_t tile = t0; tile < t1; ++tile) {
    const float4 slo = boxes[2 * tile];
    const float4 shi = boxes[2 * tile + 1];
    const float dist2 = uniform_f(box_dist2(lx, ly, lz, hx, hy, hz, slo, shi));
    float a[6] = {0, 0, 0, 0, 0, 0};
    const SrcRec *__restrict__ s = src + tile * kTile;
    auto pair = [&](int j, bool near) {
      const float dx = t.x - s[j].x, dy = t.y - s[j].y, dz = t.z - s[j].z;
      const float mx = s[j].mx, my = s[j].my, mz = s[j].mz;
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      const float ir = __builtin_amdgcn_rsqf(r2);
      const float ir2 = ir * ir;
      float P = ir2 * ir2 * ir;
      float R = -5.f * P * ir2;
      if (near && !(r2 > hc.near2)) hess_near_coeffs(tab, hc.inv_h, r2, ir, P, R);
      const float dm = fmaf(dx, mx, fmaf(dy, my, dz * mz));
      const float Rdm = R * dm;
      const float Pdm = P * dm;
      a[0] = fmaf(2.f * P * dx, mx, fmaf(Rdm * dx, dx, a[0] + Pdm));
      a[1] = fmaf(2.f * P * dy, my, fmaf(Rdm * dy, dy, a[1] + Pdm));
      a[2] = fmaf(2.f * P * dz, mz, fmaf(Rdm * dz, dz, a[2] + Pdm));
      a[3] = fmaf(P, fmaf(dx, my, mx * dy), fmaf(Rdm * dx, dy, a[3]));
      a[4] = fmaf(P, fmaf(dx, mz, mx * dz), fmaf(Rdm * dx, dz, a[4]));
      a[5] = fmaf(P, fmaf(dy, mz, my * dz), fmaf(Rdm * dy, dz, a[5]));
    };
    if (dist2 > hc.near2) {
#pragma unroll 4
      for (int j = 0; j < kTile; ++j) pair(j, false);
    } else {
#pragma unroll 2
      for (int j = 0; j < kTile; ++j) pair(j, true);
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[c] += static_cast<double>(a[c]);
  }
  if (valid) {
    double *out = partial + chunk * 6 * nt_ld;
#pragma unroll
    for (int c = 0; c < 6; ++c) out[c * nt_ld + i] = acc[c];
  }
}

// Destinations of the six tensor components: out[k] (may be null) receives
// component sym[k] (0..5 = xx yy zz xy xz yz) times mu0 3/(4 pi).
struct TensorOut {
  float *p[9];
  int sym[9];
};

__global__ void reduce_hessian_kernel(int64_t n, int64_t nt_ld, int nchunk,
                                      const double *__restrict__ partial, TensorOut out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int c = 0; c < nchunk; ++c) {
    const double *p = partial + static_cast<int64_t>(c) * 6 * nt_ld;
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] += p[k * nt_ld + i];
  }
  const double sfac = 3.0 * kMu0 / (4.0 * kPi);
#pragma unroll
  for (int k = 0; k < 9; ++k)
    if (out.p[k]) out.p[k][i] = static_cast<float>(sfac * s[out.sym[k]]);
}

// ---------------------------------------------------------------------------
// Chunk reductions (fixed order => deterministic).

__global__ void reduce_field_kernel(int64_t n, int64_t nt_ld, int nchunk,
                                    const double *__restrict__ partial, float hc0, float hc1,
                                    float hc2, float *__restrict__ ox, float *__restrict__ oy,
                                    float *__restrict__ oz) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int c = 0; c < nchunk; ++c) {
    const double *p = partial + static_cast<int64_t>(c) * 3 * nt_ld;
    s0 += p[i];
    s1 += p[nt_ld + i];
    s2 += p[2 * nt_ld + i];
  }
  ox[i] = static_cast<float>(s0 + hc0);
  oy[i] = static_cast<float>(s1 + hc1);
  oz[i] = static_cast<float>(s2 + hc2);
}

// F_i = sum_chunks partial + Text_i n_i  (Text given as 9 SoA arrays, may be null).
struct Text9 {
  const float *p[9];
};
__global__ void reduce_force_kernel(int64_t n, int64_t nt_ld, int nchunk,
                                    const double *__restrict__ partial,
                                    const SrcRec *__restrict__ tgt, Text9 text, int has_text,
                                    float *__restrict__ fx, float *__restrict__ fy,
                                    float *__restrict__ fz) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < nchunk; ++c) {
    const double *p = partial + static_cast<int64_t>(c) * 3 * nt_ld;
    s[0] += p[i];
    s[1] += p[nt_ld + i];
    s[2] += p[2 * nt_ld + i];
  }
  if (has_text) {
    // Text rows: dHextdx = (p0,p1,p2), dHextdy = (p3,p4,p5), dHextdz = (p6,p7,p8),
    // glm column a holds (d/dx_a H_x, H_y, H_z); the tensor is symmetric for a
    // curl-free external field, F_r = sum_a T[a][r] n_a (magnetic_force.cu:968-976).
    const double n3[3] = {tgt[i].mx, tgt[i].my, tgt[i].mz};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) v += static_cast<double>(text.p[3 * a + r][i]) * n3[a];
      s[r] += v;
    }
  }
  fx[i] = static_cast<float>(s[0]);
  fy[i] = static_cast<float>(s[1]);
  fz[i] = static_cast<float>(s[2]);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side.

FieldConsts make_field_consts(double h) {
  FieldConsts fc{};
  fc.h = static_cast<float>(h);
  fc.inv_h = static_cast<float>(1.0 / h);
  fc.near2 = static_cast<float>(4.0 * h * h);
  fc.inv_h3 = static_cast<float>(1.0 / (h * h * h));
  fc.inv_h5 = static_cast<float>(1.0 / (h * h * h * h * h));
  // g(q) = -3 (q^6/6 - 6 q^5/5 + 3 q^4 - 8 q^3/3 + 1/15) re-expanded exactly
  // about q = 1.5 (ascending powers of t = q - 1.5).
  const double g[7] = {1843.0 / 640.0, 27.0 / 32.0, -63.0 / 32.0, 5.0 / 4.0,
                       9.0 / 8.0,      -9.0 / 10.0, -1.0 / 2.0};
  for (int k = 0; k < 7; ++k) fc.g[k] = static_cast<float>(g[k]);
  fc.tab_scale = static_cast<float>(kFieldTabN / (2.0 * h));
  return fc;
}

namespace {
// alpha h^5 and beta h^3 of a near pair at q < 2, in FP64 (the forms of the
// mf_pair.hpp header; g unexpanded, which is accurate in FP64).
void near_coeffs_f64(double q, double &a, double &b) {
  if (q < 1.0) {
    a = 0.8 - 0.5 * q;
    b = -8.0 / 3.0 + q * q * (4.8 - 2.5 * q);
    return;
  }
  const double q3 = q * q * q;
  const double g = -3.0 * (q3 * (((q / 6.0 - 1.2) * q + 3.0) * q - 8.0 / 3.0) + 1.0 / 15.0);
  const double w = (2.0 - q) * (2.0 - q) * (2.0 - q);
  a = (g / q3 - w) / (3.0 * q * q);
  b = g / (3.0 * q3) - w;
}

// Power-basis coefficients c[0..3] of the cubic through (t[i], y[i]), i < 4
// (Gaussian elimination with partial pivoting on the Vandermonde system).
void cubic_through(const double (&t)[4], const double (&y)[4], double (&c)[4]) {
  double m[4][5];
  for (int i = 0; i < 4; ++i) {
    double p = 1.0;
    for (int j = 0; j < 4; ++j, p *= t[i]) m[i][j] = p;
    m[i][4] = y[i];
  }
  for (int col = 0; col < 4; ++col) {
    int piv = col;
    for (int i = col + 1; i < 4; ++i)
      if (std::fabs(m[i][col]) > std::fabs(m[piv][col])) piv = i;
    for (int j = 0; j < 5; ++j) std::swap(m[col][j], m[piv][j]);
    for (int i = col + 1; i < 4; ++i) {
      const double f = m[i][col] / m[col][col];
      for (int j = col; j < 5; ++j) m[i][j] -= f * m[col][j];
    }
  }
  for (int i = 3; i >= 0; --i) {
    double s = m[i][4];
    for (int j = i + 1; j < 4; ++j) s -= m[i][j] * c[j];
    c[i] = s / m[i][i];
  }
}
}  // namespace

// Coefficients (ascending powers) of c(t + s) from those of the cubic c(t).
static void shift_cubic(const double *c, double s, double *out) {
  out[0] = ((c[3] * s + c[2]) * s + c[1]) * s + c[0];
  out[1] = (3.0 * c[3] * s + 2.0 * c[2]) * s + c[1];
  out[2] = 3.0 * c[3] * s + c[2];
  out[3] = c[3];
}

static void store_entry(float *tab, int e, const double *ca, const double *cb) {
  for (int j = 0; j < 4; ++j) {
    tab[8 * e + j] = static_cast<float>(ca[j]);
    tab[8 * e + 4 + j] = static_cast<float>(cb[j]);
  }
}

void make_field_table(double h, float *tab) {
  const double ih3 = 1.0 / (h * h * h), ih5 = ih3 / (h * h);
  double t[4];
  for (int i = 0; i < 4; ++i) t[i] = 0.5 + 0.5 * std::cos((2 * i + 1) * kPi / 8.0);
  double ca[4], cb[4];
  for (int k = 0; k < kFieldTabN; ++k) {
    double ya[4], yb[4];
    for (int i = 0; i < 4; ++i) {
      near_coeffs_f64((k + t[i]) * 2.0 / kFieldTabN, ya[i], yb[i]);
      ya[i] *= ih5;
      yb[i] *= ih3;
    }
    cubic_through(t, ya, ca);
    cubic_through(t, yb, cb);
    store_entry(tab, k, ca, cb);
  }
  double sa[4], sb[4];  // the last interval's cubics, one interval on
  shift_cubic(ca, 1.0, sa);
  shift_cubic(cb, 1.0, sb);
  store_entry(tab, kFieldTabN, sa, sb);
}

// Fitted curve table of the reference (Dipole_hessian.kern:53-61), highest
// power first.  Kept as data: the force must use exactly these curves.
static const double kC1[4][5] = {
    {9.97813616438174e-09, -2.97897856524718e-08, 2.38918644566813e-09, 4.53199938857366e-08,
     2.44617454752747e-11},
    {-2.76473728643294e-09, 2.86975546540539e-08, -9.94582836806651e-08, 1.25129924573675e-07,
     -2.37010166723652e-08},
    {-1.09679990621465e-09, 9.77055663264614e-09, -2.54781238661150e-08, 2.65020634884934e-09,
     5.00787562417835e-08},
    {3.79927162333632e-10, -6.26368404962679e-09, 3.94760528277489e-08, -1.13580541622200e-07,
     1.27491333574323e-07}};
static const double kC2[4][5] = {
    {6.69550479838731e-08, -1.61753307173877e-07, 1.68213714992711e-08, 1.34558143036838e-07,
     1.10976027980100e-10},
    {-3.08460139955194e-08, 2.29192245602275e-07, -5.88399621128587e-07, 5.61170054591844e-07,
     -1.14421132829680e-07},
    {3.50477408060213e-09, -5.25956271895141e-08, 2.78876509535747e-07, -6.24199554212217e-07,
     4.91807818904985e-07},
    {7.33485346367840e-10, -9.58855788627803e-09, 4.37085309763591e-08, -7.48004594092261e-08,
     2.34161209605651e-08}};

// Taylor shift of a quartic given highest-power-first to ascending powers
// about q0 (exact up to FP64 rounding: binomial re-expansion).
static void shift_quartic(const double *desc, double q0, double *asc_out) {
  double a[5];
  for (int j = 0; j < 5; ++j) a[j] = desc[4 - j];
  static const double binom[5][5] = {
      {1, 0, 0, 0, 0}, {1, 1, 0, 0, 0}, {1, 2, 1, 0, 0}, {1, 3, 3, 1, 0}, {1, 4, 6, 4, 1}};
  for (int k = 0; k < 5; ++k) {
    double s = 0.0;
    for (int j = k; j < 5; ++j) s += a[j] * binom[j][k] * std::pow(q0, j - k);
    asc_out[k] = s;
  }
}

void make_hess_table(double h, float *tab) {
  const double scale = 1.0 / (3.0 * kMu0 / (4.0 * kPi) * h * h * h * h);
  constexpr int per_unit = kHessTabN / 4;
  double t[4];
  for (int i = 0; i < 4; ++i) t[i] = -0.5 + 0.5 * std::cos((2 * i + 1) * kPi / 8.0);
  double first_c[4], first_d[4];
  for (int e = 1; e <= kHessTabN; ++e) {  // entry e: u in (e - 1, e]
    const int piece = (e - 1) / per_unit;
    double yc[4], yd[4], cc[4], cd[4];
    for (int i = 0; i < 4; ++i) {
      const double q = (e + t[i]) / per_unit;
      double c1 = 0.0, c2 = 0.0;
      for (int j = 0; j < 5; ++j) {
        c1 = c1 * q + kC1[piece][j];
        c2 = c2 * q + kC2[piece][j];
      }
      yc[i] = scale * c1;
      yd[i] = -scale * (c2 - 3.0 * c1) / 5.0;
    }
    cubic_through(t, yc, cc);
    cubic_through(t, yd, cd);
    store_entry(tab, e, cc, cd);
    if (e == 1) {
      std::copy(cc, cc + 4, first_c);
      std::copy(cd, cd + 4, first_d);
    }
    if (e == kHessTabN) {  // entry kHessTabN + 1 continues this one
      double sc[4], sd[4];
      shift_cubic(cc, 1.0, sc);
      shift_cubic(cd, 1.0, sd);
      store_entry(tab, kHessTabN + 1, sc, sd);
    }
  }
  double sc[4], sd[4];  // entry 0 (u = 0) continues entry 1
  shift_cubic(first_c, -1.0, sc);
  shift_cubic(first_d, -1.0, sd);
  store_entry(tab, 0, sc, sd);
}

void split_near_table(const float *tab, int entries, float *out) {
  for (int e = 0; e < entries; ++e)
    for (int j = 0; j < 4; ++j) {
      out[4 * e + j] = tab[8 * e + j];
      out[4 * (entries + e) + j] = tab[8 * e + 4 + j];
    }
}

const double (*reference_c1_table())[5] { return kC1; }
const double (*reference_c2_table())[5] { return kC2; }

HessConsts make_hess_consts(double h) {
  HessConsts hc{};
  hc.h = static_cast<float>(h);
  hc.inv_h = static_cast<float>(1.0 / h);
  hc.near2 = static_cast<float>(16.0 * h * h);
  hc.tab_scale = static_cast<float>(kHessTabN / (4.0 * h));
  const double s = 3.0 * kMu0 / (4.0 * kPi);
  const double scale = 1.0 / (s * h * h * h * h);
  for (int k = 0; k < 4; ++k) {
    double d[5];
    for (int j = 0; j < 5; ++j) d[j] = kC2[k][j] - 3.0 * kC1[k][j];
    double c1s[5], ds[5];
    shift_quartic(kC1[k], k + 0.5, c1s);
    shift_quartic(d, k + 0.5, ds);
    for (int j = 0; j < 5; ++j) {
      hc.curve[k][j] = static_cast<float>(scale * c1s[j]);
      hc.curve[k][5 + j] = static_cast<float>(scale * ds[j]);
    }
    hc.curve[k][10] = hc.curve[k][11] = 0.f;
  }
  return hc;
}

int64_t padded_sources(int64_t n) { return ((n + kTile - 1) / kTile) * kTile; }

Status pack_sources(mf_ctx *ctx, int64_t n, const float *x, const float *y, const float *z,
                    const float *mx, const float *my, const float *mz, float scale,
                    SrcRec *out) {
  const int64_t np = padded_sources(n);
  if (np == 0) return Status::Ok();
  hipLaunchKernelGGL(pack_sources_kernel, dim3(grid_for(np, 256)), dim3(256), 0, ctx->stream, n,
                     np, x, y, z, mx, my, mz, scale, out);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

Status pack_moments(mf_ctx *ctx, int64_t n, const float *mx, const float *my, const float *mz,
                    float scale, SrcRec *out) {
  if (n == 0) return Status::Ok();
  hipLaunchKernelGGL(pack_moments_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n,
                     mx, my, mz, scale, out);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

Status tile_boxes(mf_ctx *ctx, int64_t n, const SrcRec *src, float4 *boxes) {
  const int64_t ntiles = padded_sources(n) / kTile;
  if (ntiles == 0) return Status::Ok();
  hipLaunchKernelGGL(tile_box_kernel, dim3(grid_for(ntiles, 4)), dim3(256), 0, ctx->stream, n,
                     ntiles, src, boxes);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

namespace {
constexpr int kFieldTPT = 2;
constexpr int kForceTPT = 2;

// Splits the source tiles so that the launch has >= ~8 waves per SIMD.
int64_t choose_chunks(int64_t target_waves, int64_t ntiles, int num_cu) {
  const int64_t want = static_cast<int64_t>(num_cu) * 4 * 8;
  int64_t nchunk = (want + target_waves - 1) / std::max<int64_t>(target_waves, 1);
  nchunk = std::max<int64_t>(1, std::min<int64_t>(nchunk, 64));
  nchunk = std::min<int64_t>(nchunk, std::max<int64_t>(ntiles, 1));
  return nchunk;
}
}  // namespace

namespace {
struct FieldSplit {
  int64_t ntiles, nchunk, per;
  unsigned gx;
};

// Launch shape of the field kernel: source chunks along grid.y, at most
// max_chunks of them.
FieldSplit field_split(mf_ctx *ctx, int64_t nt, int64_t ns, int64_t max_chunks) {
  FieldSplit f{};
  f.ntiles = padded_sources(ns) / kTile;
  const int64_t waves = (nt + 64 * kFieldTPT - 1) / (64 * kFieldTPT);
  f.gx = static_cast<unsigned>((waves + 3) / 4);
  int64_t nchunk = std::min(choose_chunks(waves, f.ntiles, ctx->num_cu), std::max<int64_t>(max_chunks, 1));
  f.per = f.ntiles > 0 ? (f.ntiles + nchunk - 1) / nchunk : 0;
  f.nchunk = f.per > 0 ? (f.ntiles + f.per - 1) / f.per : 1;
  return f;
}

__global__ void zero_f64_kernel(int64_t n, double *__restrict__ a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}
}  // namespace

int64_t field_partial_doubles(int64_t nt, int64_t nchunk) { return 3 * nt * nchunk; }

Status field_allpairs_launch_ws(mf_ctx *ctx, const float4 *tpos, int tstride, int64_t nt,
                                const SrcRec *src, const float4 *boxes, int64_t ns,
                                const FieldConsts &fc, const float *hconst, float *ox, float *oy,
                                float *oz, double *partial, int64_t partial_cap) {
  if (nt == 0) return Status::Ok();
  MF_ARG(partial_cap >= field_partial_doubles(nt, 1), "field partial buffer too small");
  const FieldSplit f = field_split(ctx, nt, ns, partial_cap / field_partial_doubles(nt, 1));
  if (f.ntiles == 0) {
    // a fill kernel, not hipMemsetAsync: callers may capture this in a graph
    hipLaunchKernelGGL(zero_f64_kernel, dim3(grid_for(3 * nt, 256)), dim3(256), 0, ctx->stream,
                       3 * nt, partial);
    MF_LAUNCH_CHECK();
  } else {
    KernelTimer timer(ctx, "field");
    hipLaunchKernelGGL(field_allpairs_kernel<kFieldTPT>, dim3(f.gx, static_cast<unsigned>(f.nchunk)),
                       dim3(kBlock), 0, ctx->stream, tpos, tstride, nt, src, boxes, f.ntiles,
                       f.per, fc, nt, partial);
    MF_LAUNCH_CHECK();
  }
  const float h0 = hconst ? hconst[0] : 0.f, h1 = hconst ? hconst[1] : 0.f,
              h2 = hconst ? hconst[2] : 0.f;
  hipLaunchKernelGGL(reduce_field_kernel, dim3(grid_for(nt, 256)), dim3(256), 0, ctx->stream, nt,
                     nt, static_cast<int>(f.nchunk), partial, h0, h1, h2, ox, oy, oz);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

Status field_allpairs_launch(mf_ctx *ctx, const float4 *tpos, int tstride, int64_t nt,
                             const SrcRec *src, const float4 *boxes, int64_t ns,
                             const FieldConsts &fc, const float *hconst, float *ox, float *oy,
                             float *oz) {
  if (nt == 0) return Status::Ok();
  const FieldSplit f = field_split(ctx, nt, ns, 64);
  const int64_t cap = field_partial_doubles(nt, f.nchunk);
  void *pw = nullptr;
  MF_TRY(workspace(ctx, "field_partial", sizeof(double) * cap, &pw));
  return field_allpairs_launch_ws(ctx, tpos, tstride, nt, src, boxes, ns, fc, hconst, ox, oy, oz,
                                  static_cast<double *>(pw), cap);
}

Status force_allpairs_launch(mf_ctx *ctx, const SrcRec *rec, const float4 *boxes, int64_t n,
                             const HessConsts &hc, const float *const *text9, float *fx,
                             float *fy, float *fz) {
  if (n == 0) return Status::Ok();
  const int64_t ntiles = padded_sources(n) / kTile;
  const int64_t waves = (n + 64 * kForceTPT - 1) / (64 * kForceTPT);
  const unsigned gx = static_cast<unsigned>((waves + 3) / 4);
  int64_t nchunk = choose_chunks(waves, ntiles, ctx->num_cu);
  const int64_t per = (ntiles + nchunk - 1) / nchunk;
  nchunk = (ntiles + per - 1) / per;
  void *pw = nullptr;
  MF_TRY(workspace(ctx, "force_partial", sizeof(double) * 3 * n * nchunk, &pw));
  double *partial = static_cast<double *>(pw);
  {
    KernelTimer timer(ctx, "force");
    hipLaunchKernelGGL(force_allpairs_kernel<kForceTPT>, dim3(gx, static_cast<unsigned>(nchunk)),
                       dim3(kBlock), 0, ctx->stream, rec, n, rec, boxes, ntiles, per, hc, n,
                       partial);
    MF_LAUNCH_CHECK();
  }
  Text9 t{};
  int has_text = 0;
  if (text9) {
    has_text = 1;
    for (int k = 0; k < 9; ++k) {
      MF_ARG(text9[k] != nullptr, "text9 entries must be non-null device pointers");
      t.p[k] = text9[k];
    }
  }
  hipLaunchKernelGGL(reduce_force_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, n,
                     static_cast<int>(nchunk), partial, rec, t, has_text, fx, fy, fz);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

Status hessian_allpairs_launch(mf_ctx *ctx, const float4 *tpos, int64_t nt, const SrcRec *src,
                               const float4 *boxes, int64_t ns, const HessConsts &hc,
                               float *const dst[9], const int sym[9]) {
  if (nt == 0) return Status::Ok();
  const int64_t ntiles = padded_sources(ns) / kTile;
  const int64_t waves = (nt + 63) / 64;
  const unsigned gx = static_cast<unsigned>((waves + 3) / 4);
  int64_t nchunk = choose_chunks(waves, ntiles, ctx->num_cu);
  const int64_t per = ntiles > 0 ? (ntiles + nchunk - 1) / nchunk : 0;
  nchunk = per > 0 ? (ntiles + per - 1) / per : 1;
  void *pw = nullptr;
  MF_TRY(workspace(ctx, "hessian_partial", sizeof(double) * 6 * nt * nchunk, &pw));
  double *partial = static_cast<double *>(pw);
  if (ntiles == 0) {
    hipLaunchKernelGGL(zero_f64_kernel, dim3(grid_for(6 * nt, 256)), dim3(256), 0, ctx->stream,
                       6 * nt, partial);
    MF_LAUNCH_CHECK();
  } else {
    KernelTimer timer(ctx, "hessian");
    hipLaunchKernelGGL(hessian_allpairs_kernel, dim3(gx, static_cast<unsigned>(nchunk)),
                       dim3(kBlock), 0, ctx->stream, tpos, nt, src, boxes, ntiles, per, hc, nt,
                       partial);
    MF_LAUNCH_CHECK();
  }
  TensorOut out{};
  for (int k = 0; k < 9; ++k) {
    out.p[k] = dst[k];
    out.sym[k] = sym[k];
  }
  hipLaunchKernelGGL(reduce_hessian_kernel, dim3(grid_for(nt, 256)), dim3(256), 0, ctx->stream,
                     nt, nt, static_cast<int>(nchunk), partial, out);
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

}  // namespace mf

// Internal: the near-pair coefficient tables, for the CPU tests (not part of
// mf_capi.h).  kind 0 = field (kFieldTabEntries x 8 floats), 1 = force
// (kHessTabEntries x 8).
extern "C" int mf_internal_near_table(int kind, double h, float *out) {
  if (!out || !(h > 0.0)) return 1;
  if (kind == 0) mf::make_field_table(h, out);
  else if (kind == 1) mf::make_hess_table(h, out);
  else return 1;
  return 0;
}
