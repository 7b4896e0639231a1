// magnetic_force.cpp -- bodies of the drop-in Libo::cuda_magnet_cgOptimizer.
//
// Method-by-method counterpart of magnetic_solver/magnetic_force.cu:608-1410.
// Every product goes through the C ABI (mf_capi.h) on the default context.
#include "magnetic_force.cuh"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace Libo {

namespace {
void zero(cuda_abstract_vector<float> &v) {
  if (v.np == 0) return;
  mf_ctx *ctx = detail::default_ctx();
  detail::check(mf_vec_fill(ctx, static_cast<int64_t>(3 * v.np), 0.f, v.dev_x), ctx);
}
void fill3(cuda_abstract_vector<float> &v, float a, float b, float c) {
  if (v.np == 0) return;
  mf_ctx *ctx = detail::default_ctx();
  detail::check(mf_vec_fill(ctx, static_cast<int64_t>(v.np), a, v.dev_x), ctx);
  detail::check(mf_vec_fill(ctx, static_cast<int64_t>(v.np), b, v.dev_y), ctx);
  detail::check(mf_vec_fill(ctx, static_cast<int64_t>(v.np), c, v.dev_z), ctx);
}
}  // namespace

mf_evaluator::mf_evaluator(int64_t n, const float *dx, const float *dy, const float *dz, float h,
                           int mode, int leaf_size)
    : mode_(mode), h_(h), leaf_size_(leaf_size) {
  mf_ctx *ctx = detail::default_ctx();
  mf_plan_opts o{};
  o.mode = mode;
  o.leaf_size = mode == MF_MODE_FMM ? leaf_size : 0;
  detail::check(mf_plan_create(ctx, n, dx, dy, dz, h, &o, &plan_), ctx);
}

mf_evaluator::~mf_evaluator() { mf_plan_destroy(plan_); }

// magnetic_force.cu:608-703
cuda_magnet_cgOptimizer::cuda_magnet_cgOptimizer(
    const typename cuda_magnet_cgOptimizer::CtorParam &a)
    : step_dir(a.num_particles),
      old_step_dir(a.num_particles),
      gradient(a.num_particles),
      old_gradient(a.num_particles),
      residual(a.num_particles),
      xyz(a.num_particles),
      Hferro(a.num_particles),
      Hext(a.num_particles),
      M(a.num_particles),
      total_B_over_mu0(a.num_particles),
      dHextdx(a.num_particles),
      dHextdy(a.num_particles),
      dHextdz(a.num_particles),
      magnetic_bodyforce(a.num_particles) {
  if (a.target_pos.size() != a.num_particles)
    throw std::invalid_argument("CtorParam: target_pos.size() != num_particles");
  if (a.ext_source_pos.size() != a.ext_source_M.size())
    throw std::invalid_argument("CtorParam: ext_source_pos and ext_source_M sizes differ");
  chi = a.chi;
  np = a.num_particles;
  tolerance = 0.1f;
  max_iteration = 80;
  i_step = 0;
  total_error.reserve(max_iteration);
  volume = a.volume;
  h = a.h;
  ext_mag_h = a.ext_mag_h;
  xyz.import_data(a.target_pos);
  Hext_source_pos.import_data(a.ext_source_pos);
  Hext_source_M.import_data(a.ext_source_M);
  zero(total_B_over_mu0);  // the reference left x uninitialised
  zero(magnetic_bodyforce);
  // FMM particle budget per leaf (the reference's ncrit, 384 there): 0 = leaves
  // of side 2h, the library's default; a caller may set it before solve()
  crit_size = 0;
  fmm_object_pos = a.target_pos;
  fmm_hext_pos = a.ext_source_pos;
  fmm_hext_charge = a.ext_source_M;
  const_hext.x = a.const_ext_field[0];
  const_hext.y = a.const_ext_field[1];
  const_hext.z = a.const_ext_field[2];
  naive = false;
}

mf_evaluator &cuda_magnet_cgOptimizer::self_evaluator() {
  const int mode = naive ? MF_MODE_ALLPAIRS : MF_MODE_FMM;
  if (!fmm_field_evaluator || fmm_field_evaluator->mode() != mode ||
      fmm_field_evaluator->h() != h || fmm_field_evaluator->leaf_size() != crit_size) {
    fmm_field_evaluator = std::make_shared<mf_evaluator>(static_cast<int64_t>(np), xyz.dev_x,
                                                         xyz.dev_y, xyz.dev_z, h, mode, crit_size);
    fmm_hessian_evaluator = fmm_field_evaluator;  // one plan serves field and force
    fmm_hferro_evaluator = fmm_field_evaluator;
  }
  return *fmm_field_evaluator;
}

// magnetic_force.cu:706-743: keep the first min(old, new) entries of x.
void cuda_magnet_cgOptimizer::Resize(size_t _np) {
  if (_np == np) return;
  const size_t old_np = np;
  np = _np;
  cuda_abstract_vector<float> old_init_data;
  old_init_data = total_B_over_mu0;
  xyz.reserve(np);
  Hferro.reserve(np);
  Hext.reserve(np);
  M.reserve(np);
  total_B_over_mu0.reserve(np);
  dHextdx.reserve(np);
  dHextdy.reserve(np);
  dHextdz.reserve(np);
  magnetic_bodyforce.reserve(np);
  step_dir.reserve(np);
  old_step_dir.reserve(np);
  gradient.reserve(np);
  old_gradient.reserve(np);
  residual.reserve(np);
  zero(total_B_over_mu0);
  const size_t copy_count = np > old_np ? old_np : np;
  if (copy_count > 0) {
    detail::check_hip(hipMemcpy(total_B_over_mu0.dev_x, old_init_data.dev_x,
                                copy_count * sizeof(float), hipMemcpyDeviceToDevice),
                      "Resize");
    detail::check_hip(hipMemcpy(total_B_over_mu0.dev_y, old_init_data.dev_y,
                                copy_count * sizeof(float), hipMemcpyDeviceToDevice),
                      "Resize");
    detail::check_hip(hipMemcpy(total_B_over_mu0.dev_z, old_init_data.dev_z,
                                copy_count * sizeof(float), hipMemcpyDeviceToDevice),
                      "Resize");
  }
}

// magnetic_force.cu:746-803
void cuda_magnet_cgOptimizer::setHext(const typename cuda_magnet_cgOptimizer::setHextParams &a) {
  if (a.ext_source_pos.size() != a.ext_source_M.size())
    throw std::invalid_argument("setHext: ext_source_pos and ext_source_M sizes differ");
  Hext_source_pos.import_data(a.ext_source_pos);
  Hext_source_M.import_data(a.ext_source_M);
  fmm_hext_pos = a.ext_source_pos;
  fmm_hext_charge = a.ext_source_M;
  ext_mag_h = a.ext_mag_h;
  const_hext.x = a.const_ext_field[0];
  const_hext.y = a.const_ext_field[1];
  const_hext.z = a.const_ext_field[2];
}

// magnetic_force.cu:806-871
void cuda_magnet_cgOptimizer::setTargetpos(const std::vector<float3> &in) {
  Resize(in.size());
  xyz.import_data(in);
  fmm_object_pos = in;
  if (fmm_field_evaluator) {
    mf_ctx *ctx = detail::default_ctx();
    detail::check(mf_plan_update_positions(fmm_field_evaluator->plan(), static_cast<int64_t>(np),
                                           xyz.dev_x, xyz.dev_y, xyz.dev_z),
                  ctx);
  }
}

// magnetic_force.cu:874-888
void cuda_magnet_cgOptimizer::solve(size_t niter) {
  clear();
  if (start()) return;
  for (size_t i = 0; i < niter && i < max_iteration; ++i) {
    if (step_iteration()) break;
  }
}

// magnetic_force.cu:890-997 (FMM-path semantics for the external tensor)
void cuda_magnet_cgOptimizer::evaluate_fitted_near_foece() {
  mf_ctx *ctx = detail::default_ctx();
  M = volume * chi / (chi + 1) * total_B_over_mu0;
  const float *text9[9] = {dHextdx.dev_x, dHextdx.dev_y, dHextdx.dev_z,
                           dHextdy.dev_x, dHextdy.dev_y, dHextdy.dev_z,
                           dHextdz.dev_x, dHextdz.dev_y, dHextdz.dev_z};
  detail::check(mf_plan_force(self_evaluator().plan(), M.dev_x, M.dev_y, M.dev_z,
                              np > 0 ? text9 : nullptr, magnetic_bodyforce.dev_x,
                              magnetic_bodyforce.dev_y, magnetic_bodyforce.dev_z),
                ctx);
}

void cuda_magnet_cgOptimizer::clear() {
  total_error.clear();
  i_step = 0;
}

// magnetic_force.cu:1008-1040
int cuda_magnet_cgOptimizer::start() {
  i_step = 0;
  eval_hext();
  M = volume * chi / (chi + 1) * total_B_over_mu0;
  residual = G(M) + Hext - total_B_over_mu0;
  const float curr_error = 0.5f * residual.norm2();
  total_error.push_back(curr_error);
  if (std::sqrt(curr_error / static_cast<float>(np)) < tolerance) return 1;
  gradient = D_transpose(residual);
  step_dir = -gradient;
  const float step = calc_alpha();
  total_B_over_mu0 = total_B_over_mu0 + (step * step_dir);
  return 0;
}

// magnetic_force.cu:1043-1095
int cuda_magnet_cgOptimizer::step_iteration() {
  old_gradient = gradient;
  old_step_dir = step_dir;
  M = volume * chi / (chi + 1) * total_B_over_mu0;
  residual = G(M) + Hext - total_B_over_mu0;
  const float curr_error = 0.5f * residual.norm2();
  total_error.push_back(curr_error);
  if (std::sqrt(curr_error / static_cast<float>(np)) < tolerance) return 1;
  if (total_error.size() >= 2 && curr_error > *(total_error.rbegin() + 1)) return 1;
  i_step++;
  gradient = D_transpose(residual);
  float beta = 0.f;
  if (i_step % 10 != 0) {
    beta = calc_beta();
    if (beta < 0.f) beta = 0.f;
  }
  step_dir = -gradient + beta * old_step_dir;
  const float step = calc_alpha();
  total_B_over_mu0 = total_B_over_mu0 + (step * step_dir);
  return 0;
}

// magnetic_force.cu:1102-1192
void cuda_magnet_cgOptimizer::eval_hext() {
  mf_ctx *ctx = detail::default_ctx();
  const size_t next = Hext_source_pos.np;
  if (np == 0) return;
  if (next == 0) {
    fill3(Hext, const_hext.x, const_hext.y, const_hext.z);
    zero(dHextdx);
    zero(dHextdy);
    zero(dHextdz);
    return;
  }
  const float hc[3] = {const_hext.x, const_hext.y, const_hext.z};
  detail::check(mf_field_allpairs(ctx, static_cast<int64_t>(np), xyz.dev_x, xyz.dev_y, xyz.dev_z,
                                  static_cast<int64_t>(next), Hext_source_pos.dev_x,
                                  Hext_source_pos.dev_y, Hext_source_pos.dev_z,
                                  Hext_source_M.dev_x, Hext_source_M.dev_y, Hext_source_M.dev_z,
                                  ext_mag_h, hc, Hext.dev_x, Hext.dev_y, Hext.dev_z),
                ctx);
  // the nine dHextd{x,y,z} arrays straight from the tensor kernel
  float *const out9[9] = {dHextdx.dev_x, dHextdx.dev_y, dHextdx.dev_z,
                          dHextdy.dev_x, dHextdy.dev_y, dHextdy.dev_z,
                          dHextdz.dev_x, dHextdz.dev_y, dHextdz.dev_z};
  detail::check(mf_hessian_allpairs9(ctx, static_cast<int64_t>(np), xyz.dev_x, xyz.dev_y,
                                     xyz.dev_z, static_cast<int64_t>(next), Hext_source_pos.dev_x,
                                     Hext_source_pos.dev_y, Hext_source_pos.dev_z,
                                     Hext_source_M.dev_x, Hext_source_M.dev_y,
                                     Hext_source_M.dev_z, ext_mag_h, out9),
                ctx);
}

// magnetic_force.cu:1196-1271
cuda_abstract_vector<float> cuda_magnet_cgOptimizer::G(const cuda_abstract_vector<float> &in) {
  cuda_abstract_vector<float> result(np);
  if (np == 0) return result;
  mf_ctx *ctx = detail::default_ctx();
  detail::check(mf_plan_field(self_evaluator().plan(), in.dev_x, in.dev_y, in.dev_z, nullptr,
                              result.dev_x, result.dev_y, result.dev_z),
                ctx);
  return result;
}

// magnetic_force.cu:1275-1278
cuda_abstract_vector<float> cuda_magnet_cgOptimizer::D(const cuda_abstract_vector<float> &v) {
  return G(volume * (chi) / (1 + chi) * v) - v;
}

// magnetic_force.cu:1281-1284
cuda_abstract_vector<float> cuda_magnet_cgOptimizer::D_transpose(const cuda_abstract_vector<float> &v) {
  return volume * (chi) / (1 + chi) * G(v) - v;
}

// magnetic_force.cu:1287-1302
float cuda_magnet_cgOptimizer::calc_alpha() {
  const float epsl = 1e-10f;
  cuda_abstract_vector<float> temp = D(step_dir);
  const float normtemp = temp.dot(temp);
  if (normtemp < epsl) return 0.f;
  return -(temp.dot(residual) / (normtemp + epsl));
}

// magnetic_force.cu:1305-1315
float cuda_magnet_cgOptimizer::calc_beta() {
  if (i_step == 0) return 0.f;
  const float numer = gradient.norm2();
  const float denom = old_step_dir.dot(old_gradient - gradient);
  return -numer / (denom + 1e-10f);
}

mf_solve_stats cuda_magnet_cgOptimizer::solve_linear(double rtol, int method, int maxit) {
  eval_hext();
  mf_solve_opts o{};
  o.method = method;
  o.restart = 30;
  o.maxit = maxit;
  o.rtol = rtol;
  mf_solve_stats st{};
  mf_ctx *ctx = detail::default_ctx();
  detail::check(mf_plan_solve(self_evaluator().plan(), static_cast<double>(volume) * chi / (1.0 + chi),
                              Hext.dev_x, Hext.dev_y, Hext.dev_z, total_B_over_mu0.dev_x,
                              total_B_over_mu0.dev_y, total_B_over_mu0.dev_z, &o, &st),
                ctx);
  return st;
}

namespace {
void write_rows(const char *file_name, size_t np, const cuda_abstract_vector<float> &pos,
                const cuda_abstract_vector<float> &vec, const cuda_abstract_vector<float> &hext) {
  std::vector<float> px, py, pz, vx, vy, vz, hx, hy, hz;
  pos.export_data(px, py, pz);
  vec.export_data(vx, vy, vz);
  hext.export_data(hx, hy, hz);
  std::ofstream out(file_name);
  if (!out) throw std::runtime_error(std::string("cannot open ") + file_name);
  for (size_t i = 0; i < np; ++i) {
    out << i << " ";
    out << px[i] << " " << py[i] << " " << pz[i] << " ";
    out << vx[i] << " " << vy[i] << " " << vz[i] << " ";
    out << hx[i] << " " << hy[i] << " " << hz[i] << std::endl;
  }
}
}  // namespace

// magnetic_force.cu:1318-1363: i x y z (B/mu0)xyz Hext_xyz
void cuda_magnet_cgOptimizer::export_data(const char *file_name) {
  write_rows(file_name, np, xyz, total_B_over_mu0, Hext);
}

// magnetic_force.cu:1365-1410: i x y z F_xyz Hext_xyz
void cuda_magnet_cgOptimizer::export_force(const char *file_name) {
  write_rows(file_name, np, xyz, magnetic_bodyforce, Hext);
}

}  // namespace Libo
