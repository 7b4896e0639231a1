// mf_fmm_tables.cpp -- translation operators of the FMM as dense real matrices.
//
// All operators act on the scaled real dofs of mf_expansion.hpp and are built
// once in FP64 on the host, then rounded to FP32:
//   M2L(o): Lt_k^l = (-1)^(k+l) sum_{n,m} Mt_n^m I_{n+k}^{m-l}(o), o in box units
//   M2M(c): Mt_n^m(parent) = sum_{k,l} Mt_k^l(child) 2^-k conj(R_{n-k}^{m-l}(d))
//   L2L(c): Lt_k^l(child) = sum_{n>=k,m} Lt_n^m(parent) 2^-(k+1) R_{n-k}^{m-l}(d)
// with d = child centre - parent centre in parent-box units ((+-1/4)^3).
// Layout: column-major, t[col * ld + row] (row = output dof, col = input dof).
#include "mf_fmm_tables.hpp"

#include <cmath>
#include <complex>
#include <vector>

namespace mf {

namespace {

using cd = std::complex<double>;

// X(n, m) for all |m| <= n <= P from the m >= 0 triangle.
struct Harm {
  int P = 0;
  std::vector<cd> tri;  // n (n+1)/2 + m
  cd get(int n, int m) const {
    if (n < 0 || n > P || m > n || -m > n) return cd(0.0, 0.0);
    if (m >= 0) return tri[n * (n + 1) / 2 + m];
    const cd a = tri[n * (n + 1) / 2 - m];
    return ((-m) & 1) ? -std::conj(a) : std::conj(a);
  }
};

Harm regular(double x, double y, double z, int P) {
  Harm h;
  h.P = P;
  h.tri.assign((P + 1) * (P + 2) / 2, cd(0.0, 0.0));
  const double r2 = x * x + y * y + z * z;
  h.tri[0] = 1.0;
  for (int n = 1; n <= P; ++n)
    h.tri[n * (n + 1) / 2 + n] = -cd(x, y) / (2.0 * n) * h.tri[(n - 1) * n / 2 + n - 1];
  for (int m = 0; m <= P; ++m)
    for (int n = m + 1; n <= P; ++n) {
      cd v = (2.0 * n - 1.0) * z * h.tri[(n - 1) * n / 2 + m];
      if (n - 2 >= m) v -= r2 * h.tri[(n - 2) * (n - 1) / 2 + m];
      h.tri[n * (n + 1) / 2 + m] = v / static_cast<double>((n + m) * (n - m));
    }
  return h;
}

Harm irregular(double x, double y, double z, int P) {
  Harm h;
  h.P = P;
  h.tri.assign((P + 1) * (P + 2) / 2, cd(0.0, 0.0));
  const double r2 = x * x + y * y + z * z;
  h.tri[0] = 1.0 / std::sqrt(r2);
  for (int n = 1; n <= P; ++n)
    h.tri[n * (n + 1) / 2 + n] =
        -(2.0 * n - 1.0) * cd(x, y) / r2 * h.tri[(n - 1) * n / 2 + n - 1];
  for (int m = 0; m <= P; ++m)
    for (int n = m + 1; n <= P; ++n) {
      cd v = (2.0 * n - 1.0) * z * h.tri[(n - 1) * n / 2 + m];
      if (n - 2 >= m)
        v -= static_cast<double>((n - 1 + m) * (n - 1 - m)) * h.tri[(n - 2) * (n - 1) / 2 + m];
      h.tri[n * (n + 1) / 2 + m] = v / r2;
    }
  return h;
}

inline int dre(int n, int m) { return m == 0 ? n * n : n * n + 2 * m - 1; }
inline int dim_(int n, int m) { return n * n + 2 * m; }
inline double sgn(int k) { return (k & 1) ? -1.0 : 1.0; }

// Writes the complex coefficient c of input dof into output row(s) of (k, l):
// Re row gets Re c, Im row (l > 0) gets Im c.
inline void put(std::vector<double> &t, int ld, int k, int l, int col, cd c) {
  t[static_cast<size_t>(col) * ld + dre(k, l)] += c.real();
  if (l > 0) t[static_cast<size_t>(col) * ld + dim_(k, l)] += c.imag();
}

}  // namespace

std::vector<double> m2l_matrix(int P, int ld, double ox, double oy, double oz) {
  std::vector<double> t(static_cast<size_t>(ld) * ld, 0.0);
  const Harm I = irregular(ox, oy, oz, 2 * P);
  for (int k = 0; k <= P; ++k)
    for (int l = 0; l <= k; ++l) {
      const double s = sgn(k + l);
      for (int n = 0; n <= P; ++n) {
        put(t, ld, k, l, dre(n, 0), s * I.get(n + k, -l));
        for (int m = 1; m <= n; ++m) {
          const cd a = I.get(n + k, m - l), b = sgn(m) * I.get(n + k, -m - l);
          put(t, ld, k, l, dre(n, m), s * (a + b));
          put(t, ld, k, l, dim_(n, m), s * cd(0.0, 1.0) * (a - b));
        }
      }
    }
  return t;
}

std::vector<double> m2m_matrix(int P, int ld, double dx, double dy, double dz) {
  std::vector<double> t(static_cast<size_t>(ld) * ld, 0.0);
  const Harm R = regular(dx, dy, dz, P);
  for (int n = 0; n <= P; ++n)
    for (int m = 0; m <= n; ++m)
      for (int k = 0; k <= n; ++k) {
        const double w = std::ldexp(1.0, -k);
        // input (k, 0)
        {
          const cd c = w * std::conj(R.get(n - k, m));
          t[static_cast<size_t>(dre(k, 0)) * ld + dre(n, m)] += c.real();
          if (m > 0) t[static_cast<size_t>(dre(k, 0)) * ld + dim_(n, m)] += c.imag();
        }
        for (int l = 1; l <= k; ++l) {
          const cd a = std::conj(R.get(n - k, m - l)), b = sgn(l) * std::conj(R.get(n - k, m + l));
          const cd ca = w * (a + b), cb = w * cd(0.0, 1.0) * (a - b);
          t[static_cast<size_t>(dre(k, l)) * ld + dre(n, m)] += ca.real();
          t[static_cast<size_t>(dim_(k, l)) * ld + dre(n, m)] += cb.real();
          if (m > 0) {
            t[static_cast<size_t>(dre(k, l)) * ld + dim_(n, m)] += ca.imag();
            t[static_cast<size_t>(dim_(k, l)) * ld + dim_(n, m)] += cb.imag();
          }
        }
      }
  return t;
}

std::vector<double> l2l_matrix(int P, int ld, double dx, double dy, double dz) {
  std::vector<double> t(static_cast<size_t>(ld) * ld, 0.0);
  const Harm R = regular(dx, dy, dz, P);
  for (int k = 0; k <= P; ++k)
    for (int l = 0; l <= k; ++l) {
      const double w = std::ldexp(1.0, -(k + 1));
      for (int n = k; n <= P; ++n) {
        put(t, ld, k, l, dre(n, 0), w * R.get(n - k, -l));
        for (int m = 1; m <= n; ++m) {
          const cd a = R.get(n - k, m - l), b = sgn(m) * R.get(n - k, -m - l);
          put(t, ld, k, l, dre(n, m), w * (a + b));
          put(t, ld, k, l, dim_(n, m), w * cd(0.0, 1.0) * (a - b));
        }
      }
    }
  return t;
}

}  // namespace mf

// Internal: the tables above, for the CPU tests (not part of mf_capi.h).
// kind 0 = M2L, 1 = M2M, 2 = L2L; out has ld * ld doubles (column-major).
extern "C" int mf_internal_fmm_matrix(int kind, int P, int ld, double dx, double dy, double dz,
                                      double *out) {
  if (!out || P < 0 || ld < (P + 1) * (P + 1)) return 1;
  std::vector<double> t;
  switch (kind) {
    case 0: t = mf::m2l_matrix(P, ld, dx, dy, dz); break;
    case 1: t = mf::m2m_matrix(P, ld, dx, dy, dz); break;
    case 2: t = mf::l2l_matrix(P, ld, dx, dy, dz); break;
    default: return 1;
  }
  for (size_t i = 0; i < t.size(); ++i) out[i] = t[i];
  return 0;
}
