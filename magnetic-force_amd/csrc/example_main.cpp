// example_main.cpp -- the reference's example driver (magnetic_solver/main.cu:3-62)
// on the drop-in API: a 2x2x2 lattice with one external dipole on a particle,
// solve -> force -> move the particles -> setHext -> solve -> force -> export.
//
// usage: example [naive|fmm] [out_dir]
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "magnetic_force.cuh"

int main(int argc, char **argv) {
  const bool naive = !(argc > 1 && std::strcmp(argv[1], "fmm") == 0);
  const std::string out_dir = argc > 2 ? argv[2] : ".";
  try {
    const int nx = 2, ny = 2, nz = 2;
    const float dx = 1;
    // one external dipole sitting on the particle at the origin
    auto ext_source_pos = std::vector<float3>{float3{0.f, 0.f, 0.f}};
    auto ext_source_M = std::vector<float3>{float3{0.f, 1.f, 0.f}};
    std::vector<float3> target_pos, updated_target_pos;
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          target_pos.push_back(float3{i * dx, j * dx, k * dx});
          if (k < nz - 1) updated_target_pos.push_back(float3{i * dx, j * dx, k * dx + 1});
        }
    const size_t num_particles = target_pos.size();
    const float h = dx, ext_mag_h = dx, volume = dx * dx * dx, chi = 1;
    float const_ext_field[3] = {0, 1, 0};

    Libo::cuda_magnet_cgOptimizer::CtorParam ctor{num_particles, h, ext_mag_h, volume, chi,
                                                 const_ext_field, ext_source_pos, ext_source_M,
                                                 target_pos};
    Libo::cuda_magnet_cgOptimizer a{ctor};
    a.naive = naive;
    a.solve(100);
    a.evaluate_fitted_near_foece();
    std::printf("first solve: %zu error evaluations, final 0.5|r|^2 = %.6e\n",
                a.total_error.size(), a.total_error.empty() ? 0.0 : a.total_error.back());

    a.setTargetpos(updated_target_pos);
    Libo::cuda_magnet_cgOptimizer::setHextParams seth;
    seth.const_ext_field[0] = const_ext_field[0];
    seth.const_ext_field[1] = const_ext_field[1];
    seth.const_ext_field[2] = const_ext_field[2];
    seth.ext_source_M = ext_source_M;
    seth.ext_source_pos = ext_source_pos;
    seth.ext_mag_h = h;
    a.setHext(seth);
    std::printf("particle position updated\n");
    a.solve(100);
    a.evaluate_fitted_near_foece();
    a.export_data((out_dir + "/temp.txt").c_str());
    a.export_force((out_dir + "/force.txt").c_str());
    std::printf("second solve: %zu error evaluations, final 0.5|r|^2 = %.6e\n",
                a.total_error.size(), a.total_error.empty() ? 0.0 : a.total_error.back());
  } catch (const std::exception &e) {
    std::fprintf(stderr, "example failed: %s\n", e.what());
    return 1;
  }
  return 0;
}
