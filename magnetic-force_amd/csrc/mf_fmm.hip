// mf_fmm.hip -- GPU octree fast multipole method (placeholder until built).
#include <hip/hip_runtime.h>

#include "mf_internal.hpp"
#include "mf_plan.hpp"

namespace mf {

struct FmmState {
  int dummy = 0;
};

void destroy_fmm(FmmState *s) { delete s; }

Status fmm_build(mf_plan *) {
  return Status::Err(MF_ERR_UNSUPPORTED, "FMM mode is not available in this build");
}
Status fmm_apply_G(mf_plan *, float, const float *, const float *, const float *, const float *,
                   float *, float *, float *) {
  return Status::Err(MF_ERR_UNSUPPORTED, "FMM mode is not available in this build");
}
Status fmm_force(mf_plan *, const float *, const float *, const float *, const float *const *,
                 float *, float *, float *) {
  return Status::Err(MF_ERR_UNSUPPORTED, "FMM mode is not available in this build");
}

}  // namespace mf
