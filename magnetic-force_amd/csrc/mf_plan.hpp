// mf_plan.hpp -- the particle-set plan shared by products and solves.
#pragma once

#include <memory>
#include <vector>

#include "mf_allpairs.hpp"
#include "mf_internal.hpp"
#include "mf_pair.hpp"

namespace mf {
struct FmmState;  // defined in mf_fmm.hip
void destroy_fmm(FmmState *s);
struct FmmDeleter {
  void operator()(FmmState *s) const { destroy_fmm(s); }
};
}  // namespace mf

struct mf_plan {
  mf_ctx *ctx = nullptr;
  mf_plan_opts opts{};
  int64_t n = 0;
  double h = 0.0;
  // positions set and (FMM) the tree built; false from the start of an update
  // until it completes, so that a failed update leaves no half-built state
  // for a product to read
  bool ready = false;
  mf::FieldConsts fc{};
  mf::HessConsts hc{};
  // All-pairs state: packed records (positions + current moments) and tile boxes.
  mf::DeviceBuffer rec;
  mf::DeviceBuffer boxes;
  // Double copies of the positions for the FP64 check/refinement path (lazy).
  mf::DeviceBuffer pos64;
  bool pos64_valid = false;
  std::unique_ptr<mf::FmmState, mf::FmmDeleter> fmm;
  // M2L precision of the FMM products; >= 0 overrides opts.m2l_precision
  // (the mixed-precision solve switches its inner products this way).
  int m2l_override = -1;
  // Independent systems (batch plans): system s owns particles
  // [sys_offsets[s], sys_offsets[s + 1]) in caller order; sys_of holds each
  // particle's system (device int32, empty for a single system).
  int nsys = 1;
  std::vector<int64_t> sys_offsets;
  mf::DeviceBuffer sys_of;
  // Optional per-particle factor (caller order) applied to the moments of FMM
  // field products; set by the batch solve for the span of its products.
  const float *weights = nullptr;
  // Optional per-system flags (device int32, nsys): field products of a batch
  // plan skip the boxes of systems whose flag is 0.  Set by the batch solve
  // for the span of its iterations (converged systems are frozen there).
  const int32_t *active_sys = nullptr;
  // Segments of the per-system reductions (built on first batch solve):
  // chunk c covers particles [seg[2c], seg[2c + 1]); system s owns chunks
  // [seg_first[s], seg_first[s + 1]).
  mf::DeviceBuffer seg, seg_first, kappa_part;
  int nseg = 0;
  int m2l_precision() const { return m2l_override >= 0 ? m2l_override : opts.m2l_precision; }
};

namespace mf {

// out = G(scale * p) (+ hconst); p, out are SoA device arrays of plan->n.
Status plan_apply_G(mf_plan *plan, float scale, const float *px, const float *py,
                    const float *pz, const float *hconst, float *ox, float *oy, float *oz);
// F = (B(m) + Text) m.
Status plan_force(mf_plan *plan, const float *mx, const float *my, const float *mz,
                  const float *const *text9, float *fx, float *fy, float *fz);
// (Re)builds all position-dependent state.
Status plan_set_positions(mf_plan *plan, int64_t n, const float *x, const float *y,
                          const float *z);
// Double positions [3][n] for the FP64 path.
Status plan_pos64(mf_plan *plan, const double **px, const double **py, const double **pz);

Status solve(mf_plan *plan, double kappa, const float *bx, const float *by, const float *bz,
             float *xx, float *xy, float *xz, const mf_solve_opts &opts, mf_solve_stats *stats);
// Every system s of a batch plan: (I - kappa[s] G_s) x = b (kappa: host, nsys).
Status solve_batch(mf_plan *plan, const double *kappa, const float *bx, const float *by,
                   const float *bz, float *xx, float *xy, float *xz, const mf_solve_opts &opts,
                   mf_solve_stats *stats, double *sys_rel_residual);
Status solve_f64(mf_plan *plan, double kappa, const double *bx, const double *by,
                 const double *bz, double *xx, double *xy, double *xz,
                 const mf_solve_opts &opts, mf_solve_stats *stats);

// FMM (mf_fmm.hip).
Status fmm_build(mf_plan *plan);
Status fmm_apply_G(mf_plan *plan, float scale, const float *px, const float *py, const float *pz,
                   const float *hconst, float *ox, float *oy, float *oz);
Status fmm_force(mf_plan *plan, const float *mx, const float *my, const float *mz,
                 const float *const *text9, float *fx, float *fy, float *fz);
// Geometry and work counts (the counts are computed on the first call after a build).
Status fmm_info(mf_plan *plan, mf_plan_info *info);

}  // namespace mf
