// mf_fmm_tables.hpp -- dense real translation matrices (host, FP64).
#pragma once

#include <vector>

namespace mf {

// ld x ld column-major matrices on the (P+1)^2 real dofs (rows/cols beyond
// (P+1)^2 are zero).  Offsets/displacements in box units (see the .cpp).
std::vector<double> m2l_matrix(int P, int ld, double ox, double oy, double oz);
std::vector<double> m2m_matrix(int P, int ld, double dx, double dy, double dz);
std::vector<double> l2l_matrix(int P, int ld, double dx, double dy, double dz);

}  // namespace mf
