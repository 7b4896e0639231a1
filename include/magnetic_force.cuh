// magnetic_force.cuh -- drop-in host API of the magnetic hot path (gfx950).
//
// Source-compatible with the reference's magnetic_solver/magnetic_force.cuh:36-174:
// the same namespace, class, nested parameter structs, methods (including the
// misspelt evaluate_fitted_near_foece) and public data members.  The bodies
// (magnetic-force_amd/csrc/magnetic_force.cpp) call the C ABI in mf_capi.h.
//
// Differences a caller can observe:
//  * the fmmtl-typed members keep their names; the position/charge vectors are
//    std::vector<float3> and the evaluators are opaque handles (mf_evaluator);
//  * total_B_over_mu0 starts at zero (the reference left it uninitialised);
//  * the external field uses ext_mag_h in both modes and its force tensor
//    carries mu0 once in both modes (the reference's naive mode used h and
//    multiplied the external tensor by mu0 twice, magnetic_force.cu:586-600);
//  * errors throw std::runtime_error instead of being stored and ignored;
//  * solve_linear() is an addition: the device-resident CG/GMRES solve to a
//    relative residual, next to the reference's solve() loop.
#pragma once
#ifndef LIBO_MAGNETIC_FORCE_SOLVER
#define LIBO_MAGNETIC_FORCE_SOLVER

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <memory>
#include <vector>

#include "cuda_abstract_vector.cuh"
#include "mf_capi.h"

namespace Libo {

// Owns one mf_plan (a particle set with its all-pairs records or octree).
class mf_evaluator {
 public:
  // leaf_size: FMM particle budget per leaf (mf_plan_opts.leaf_size; 0 = leaf side 2h)
  mf_evaluator(int64_t n, const float *dx, const float *dy, const float *dz, float h, int mode,
               int leaf_size = 0);
  ~mf_evaluator();
  mf_evaluator(const mf_evaluator &) = delete;
  mf_evaluator &operator=(const mf_evaluator &) = delete;
  mf_plan *plan() const { return plan_; }
  int mode() const { return mode_; }
  float h() const { return h_; }
  int leaf_size() const { return leaf_size_; }

 private:
  mf_plan *plan_ = nullptr;
  int mode_ = MF_MODE_ALLPAIRS;
  float h_ = 0.f;
  int leaf_size_ = 0;
};

// cuda_magnet_cgOptimizer: magnetization solve and magnetic force for one
// particle set.  Minimises 0.5 ||G(kappa x) + H_ext - x||^2 like the reference
// (solve / start / step_iteration), or solves (I - kappa G) x = H_ext directly
// (solve_linear).  x = total_B_over_mu0, M = kappa x, kappa = V chi/(1+chi).
class cuda_magnet_cgOptimizer {
 public:
  struct CtorParam {
    CtorParam(size_t _num_particles, float _h, float _ext_mag_h, float _volume, float _chi,
              float *_const_ext_field, const std::vector<float3> &_ext_source_pos,
              const std::vector<float3> &_ext_source_M, const std::vector<float3> &_target_pos)
        : num_particles(_num_particles), h(_h), ext_mag_h(_ext_mag_h), volume(_volume), chi(_chi),
          ext_source_pos(_ext_source_pos), ext_source_M(_ext_source_M), target_pos(_target_pos) {
      const_ext_field[0] = _const_ext_field[0];
      const_ext_field[1] = _const_ext_field[1];
      const_ext_field[2] = _const_ext_field[2];
    }
    CtorParam(size_t _num_particles, float _h, float _ext_mag_h, float _volume, float _chi,
              float3 _const_ext_field, const std::vector<float3> &&_ext_source_pos,
              const std::vector<float3> &&_ext_source_M, const std::vector<float3> &_target_pos)
        : num_particles(_num_particles), h(_h), ext_mag_h(_ext_mag_h), volume(_volume), chi(_chi),
          ext_source_pos(_ext_source_pos), ext_source_M(_ext_source_M), target_pos(_target_pos) {
      const_ext_field[0] = _const_ext_field.x;
      const_ext_field[1] = _const_ext_field.y;
      const_ext_field[2] = _const_ext_field.z;
    }
    size_t num_particles;
    float h;
    float ext_mag_h;
    float volume;
    float chi;
    float const_ext_field[3];
    std::vector<float3> ext_source_pos;
    std::vector<float3> ext_source_M;
    std::vector<float3> target_pos;
  };

  struct setHextParams {
    float const_ext_field[3];
    std::vector<float3> ext_source_pos;
    std::vector<float3> ext_source_M;
    float ext_mag_h;
  };

  cuda_magnet_cgOptimizer(const typename cuda_magnet_cgOptimizer::CtorParam &_in_arg);
  ~cuda_magnet_cgOptimizer() {}

  void Resize(size_t _np);
  void setHext(const typename cuda_magnet_cgOptimizer::setHextParams &);
  void setTargetpos(const std::vector<float3> &);
  void solve(size_t);
  int step_iteration();
  int start();
  void evaluate_fitted_near_foece();

  void clear();

  void eval_hext();

  float calc_alpha();
  float calc_beta();

  // G: H_ferro + M (density)
  cuda_abstract_vector<float> G(const cuda_abstract_vector<float> &);
  cuda_abstract_vector<float> D_transpose(const cuda_abstract_vector<float> &);
  cuda_abstract_vector<float> D(const cuda_abstract_vector<float> &);
  void export_data(const char *);
  void export_force(const char *);

  // Device-resident solve of (I - kappa G) x = H_ext to ||r|| <= rtol ||H_ext||
  // (CG by default; method = MF_SOLVE_GMRES for restarted GMRES).  Evaluates
  // H_ext first.  Returns the solver statistics.
  mf_solve_stats solve_linear(double rtol = 1e-6, int method = MF_SOLVE_CG, int maxit = 1000);

  // optimization variables; index 0 is the initial state
  float tolerance;
  size_t max_iteration;
  size_t i_step;
  std::vector<float> total_error;
  std::vector<float> line_search_record;

  // Magnetic field variables
  size_t np;
  float chi;
  float h;

  // Use this density if M calculated as density
  float volume;

  // optimization variables
  cuda_abstract_vector<float> step_dir;
  cuda_abstract_vector<float> old_step_dir;
  cuda_abstract_vector<float> gradient;
  cuda_abstract_vector<float> old_gradient;
  cuda_abstract_vector<float> residual;

  // magnetic field part
  cuda_abstract_vector<float> xyz;
  cuda_abstract_vector<float> Hferro;
  cuda_abstract_vector<float> Hext;

  cuda_abstract_vector<float> M;
  cuda_abstract_vector<float> total_B_over_mu0;
  // external magnetic field as particles
  cuda_abstract_vector<float> Hext_source_pos;
  cuda_abstract_vector<float> Hext_source_M;
  float3 const_hext;
  // force part: gradient of the external field used for force calculation
  cuda_abstract_vector<float> dHextdx;
  cuda_abstract_vector<float> dHextdy;
  cuda_abstract_vector<float> dHextdz;

  cuda_abstract_vector<float> magnetic_bodyforce;

  // evaluation mode: naive = all-pairs kernels, otherwise the GPU FMM
  bool naive;
  int crit_size;
  std::vector<float3> fmm_hext_pos;
  std::vector<float3> fmm_object_pos;
  std::vector<float3> fmm_hext_charge;
  std::vector<float3> fmm_object_charge;

  std::shared_ptr<mf_evaluator> fmm_field_evaluator;
  std::shared_ptr<mf_evaluator> fmm_hferro_evaluator;
  std::shared_ptr<mf_evaluator> fmm_hext_evaluator;

  std::shared_ptr<mf_evaluator> fmm_hessian_evaluator;
  std::shared_ptr<mf_evaluator> fmm_gradhext_evaluator;

  // smoothing length of the external sources
  float ext_mag_h;

 private:
  mf_evaluator &self_evaluator();
};

}  // namespace Libo
#endif  // LIBO_MAGNETIC_FORCE_SOLVER
