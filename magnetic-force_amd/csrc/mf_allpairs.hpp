// mf_allpairs.hpp -- host entry points of the all-pairs kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "mf_internal.hpp"
#include "mf_pair.hpp"

namespace mf {

// 32-byte source record: position and moment, both float4-aligned.
struct __attribute__((aligned(32))) SrcRec {
  float x, y, z, w;
  float mx, my, mz, mw;
};

constexpr int kTile = 64;  // sources per tile (bounding-box granularity)

int64_t padded_sources(int64_t n);

// Packs SoA positions and scale * moments (moments may be null => 0) into
// n_pad = padded_sources(n) records; padding records are far away, zero moment.
Status pack_sources(mf_ctx *ctx, int64_t n, const float *x, const float *y, const float *z,
                    const float *mx, const float *my, const float *mz, float scale, SrcRec *out);
// Rewrites only the moments of the first n records.
Status pack_moments(mf_ctx *ctx, int64_t n, const float *mx, const float *my, const float *mz,
                    float scale, SrcRec *out);
// Bounding box (lo, hi float4 pairs) of every tile of kTile records.
Status tile_boxes(mf_ctx *ctx, int64_t n, const SrcRec *src, float4 *boxes);

// out = hconst + G m over targets tpos[i * tstride] (float4 views).
Status field_allpairs_launch(mf_ctx *ctx, const float4 *tpos, int tstride, int64_t nt,
                             const SrcRec *src, const float4 *boxes, int64_t ns,
                             const FieldConsts &fc, const float *hconst, float *ox, float *oy,
                             float *oz);
// The same with the chunk partials in caller memory (partial_cap doubles, at
// least field_partial_doubles(nt, 1)); the source chunk count is cut to fit.
// Allocates nothing.
Status field_allpairs_launch_ws(mf_ctx *ctx, const float4 *tpos, int tstride, int64_t nt,
                                const SrcRec *src, const float4 *boxes, int64_t ns,
                                const FieldConsts &fc, const float *hconst, float *ox, float *oy,
                                float *oz, double *partial, int64_t partial_cap);
int64_t field_partial_doubles(int64_t nt, int64_t nchunk);
// F_i = (sum_j B(x_i - x_j)[m_j] + Text_i) m_i over one record set.
Status force_allpairs_launch(mf_ctx *ctx, const SrcRec *rec, const float4 *boxes, int64_t n,
                             const HessConsts &hc, const float *const *text9, float *fx,
                             float *fy, float *fz);
// Force tensor at separate targets: dst[k] (null = skipped) receives the
// symmetric component sym[k] (0..5 = xx yy zz xy xz yz).  boxes: tile boxes
// of the source records.
Status hessian_allpairs_launch(mf_ctx *ctx, const float4 *tpos, int64_t nt, const SrcRec *src,
                               const float4 *boxes, int64_t ns, const HessConsts &hc,
                               float *const dst[9], const int sym[9]);

// FP64 check path (mf_check64.hip).
Status field_allpairs_f64(mf_ctx *ctx, int64_t nt, const double *tx, const double *ty,
                          const double *tz, int64_t ns, const double *sx, const double *sy,
                          const double *sz, const double *mx, const double *my,
                          const double *mz, double h, const double *hconst, double *ox,
                          double *oy, double *oz);
Status force_allpairs_f64(mf_ctx *ctx, int64_t n, const double *x, const double *y,
                          const double *z, const double *mx, const double *my, const double *mz,
                          double h, double *fx, double *fy, double *fz);

Status force_subset_f64(mf_ctx *ctx, int64_t n, const double *x, const double *y,
                        const double *z, const double *mx, const double *my, const double *mz,
                        double h, int64_t nt, const int64_t *idx, double *fx, double *fy,
                        double *fz);

const double (*reference_c1_table())[5];
const double (*reference_c2_table())[5];

}  // namespace mf
