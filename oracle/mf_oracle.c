/*
 * mf_oracle.c -- FP64 (and reference-order FP32) CPU restatement of the
 * reference's magnetic hot path.  TEST INFRASTRUCTURE ONLY (see mf_oracle.h).
 *
 * Citations are to files under the reference tree:
 *   magnetic_solver/magnetic_force.cu, fmmtl_minimum/kernel/Dipole_B.kern,
 *   fmmtl_minimum/kernel/Dipole_hessian.kern.
 */
#include "mf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MFO_PI 3.14159265358979323846

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* ---- A1: weights, Dipole_B.kern:46-76 ------------------------------------ */
void mfo_weights(double r, double h, double *w_avr, double *w_dens) {
  const double inv_pi = 1.0 / MFO_PI;
  const double q = r / h;
  const double q2 = q * q, q3 = q2 * q;
  const double sigma = inv_pi / (h * h * h);
  if (q >= 2.0) {
    *w_avr = 0.75 * inv_pi / (r * r * r);
    *w_dens = 0.0;
  } else if (q >= 1.0) {
    const double poly = q3 * q3 / 6.0 - 6.0 * q3 * q2 / 5.0 + 3.0 * q2 * q2 - 8.0 * q3 / 3.0 +
                        1.0 / 15.0;
    *w_avr = -3.0 * poly * inv_pi / (4.0 * r * r * r);
    const double t = 2.0 - q;
    *w_dens = sigma * 0.25 * t * t * t;
  } else {
    *w_avr = (q * q * (15.0 / 40.0 * q - 36.0 / 40.0) + 1.0) * sigma;
    const double t = 2.0 - q, u = 1.0 - q;
    *w_dens = sigma * (0.25 * t * t * t - u * u * u);
  }
}

/* ---- A2: pair field, Dipole_B.kern:79-104 --------------------------------- */
void mfo_field_pair(const double d[3], const double m[3], double h, double out[3]) {
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double invr = (r == 0.0) ? 1.0 : 1.0 / r;
  const double u[3] = {d[0] * invr, d[1] * invr, d[2] * invr}; /* 0 when r == 0 */
  double wa, wd;
  mfo_weights(r, h, &wa, &wd);
  const double v1 = (u[0] * m[0] + u[1] * m[1] + u[2] * m[2]) * (wa - wd);
  const double v2 = wa / 3.0 - wd;
  for (int a = 0; a < 3; ++a) out[a] = u[a] * v1 - v2 * m[a];
}

/* ---- A4: the fitted curves, Dipole_hessian.kern:53-85 ---------------------
 * Coefficient table (reference data, highest power first):
 * Dipole_hessian.kern:53-61 == magnetic_force.cu:301-309. */
static const double C1TAB[4][5] = {
    {9.97813616438174e-09, -2.97897856524718e-08, 2.38918644566813e-09, 4.53199938857366e-08,
     2.44617454752747e-11},
    {-2.76473728643294e-09, 2.86975546540539e-08, -9.94582836806651e-08, 1.25129924573675e-07,
     -2.37010166723652e-08},
    {-1.09679990621465e-09, 9.77055663264614e-09, -2.54781238661150e-08, 2.65020634884934e-09,
     5.00787562417835e-08},
    {3.79927162333632e-10, -6.26368404962679e-09, 3.94760528277489e-08, -1.13580541622200e-07,
     1.27491333574323e-07}};
static const double C2TAB[4][5] = {
    {6.69550479838731e-08, -1.61753307173877e-07, 1.68213714992711e-08, 1.34558143036838e-07,
     1.10976027980100e-10},
    {-3.08460139955194e-08, 2.29192245602275e-07, -5.88399621128587e-07, 5.61170054591844e-07,
     -1.14421132829680e-07},
    {3.50477408060213e-09, -5.25956271895141e-08, 2.78876509535747e-07, -6.24199554212217e-07,
     4.91807818904985e-07},
    {7.33485346367840e-10, -9.58855788627803e-09, 4.37085309763591e-08, -7.48004594092261e-08,
     2.34161209605651e-08}};

static double horner4(const double *c, double x) {
  return x * (x * (x * (x * c[0] + c[1]) + c[2]) + c[3]) + c[4];
}

void mfo_curves(double q, double *c1, double *c2) {
  int k = -1;
  if (0.0 < q && q <= 1.0) k = 0;
  else if (1.0 < q && q <= 2.0) k = 1;
  else if (2.0 < q && q <= 3.0) k = 2;
  else if (3.0 < q && q <= 4.0) k = 3;
  if (k < 0) {
    *c1 = 0.0;
    *c2 = 0.0;
    return;
  }
  *c1 = horner4(C1TAB[k], q);
  *c2 = horner4(C2TAB[k], q);
}

/* ---- A4: far branch, Dipole_hessian.kern:154-207 (raw d, not normalized) -- */
static void hess_far(const double d[3], const double m[3], double r, double out9[9]) {
  const double w_avr = 3.0 / (4.0 * MFO_PI * r * r * r);
  const double invr = (r != 0.0) ? 1.0 / r : 1.0;
  const double A = MFO_MU0 * invr * invr * w_avr;
  const double rdotm = d[0] * m[0] + d[1] * m[1] + d[2] * m[2];
  const double dAdr = MFO_MU0 * invr * invr * (5.0 * invr * (-w_avr)) * rdotm * invr;
  const double dBdr = MFO_MU0 * invr * (-w_avr) * invr;
  /* out9[a*3+b]: the reference writes Bij[b][a] (glm column b, row a) for
   * "G_ab"; the tensor is symmetric so the storage order does not matter. */
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double v = d[a] * m[b] * A + d[a] * dAdr * d[b] - m[a] * dBdr * d[b];
      if (a == b) v += rdotm * A;
      out9[a * 3 + b] += v;
    }
}

/* ---- A4: near branch in the reference's rotation frame,
 * Dipole_hessian.kern:88-152 (glm column-major semantics restated). -------- */
static void normalize3(double v[3]) {
  const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= n;
  v[1] /= n;
  v[2] /= n;
}
static void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

static void hess_near_frame(const double d[3], const double m[3], double q, double h,
                            double out9[9]) {
  if (q == 0.0) return; /* coincident particles contribute nothing, :101-104 */
  double c1, c2;
  mfo_curves(q, &c1, &c2);
  const double h4 = 1.0 / (h * h * h * h);
  c1 *= h4;
  c2 *= h4;
  double zeta[3] = {d[0], d[1], d[2]};
  normalize3(zeta);
  const double ez[3] = {0.0, 0.0, 1.0};
  double eta[3];
  cross3(zeta, ez, eta);
  eta[1] += 1e-10; /* the reference's pole patch, :116 */
  normalize3(eta);
  double xi[3];
  cross3(eta, zeta, xi);
  /* Columns of R are xi, eta, zeta.  s' = R^T m. */
  const double *col[3] = {xi, eta, zeta};
  double sp[3];
  for (int c = 0; c < 3; ++c) sp[c] = col[c][0] * m[0] + col[c][1] * m[1] + col[c][2] * m[2];
  /* U in the frame (symmetric):
   *   [ s'z C1   0        s'x C1 ]
   *   [ 0        s'z C1   s'y C1 ]
   *   [ s'x C1   s'y C1   s'z C2 ]                                        */
  double U[3][3] = {{sp[2] * c1, 0.0, sp[0] * c1},
                    {0.0, sp[2] * c1, sp[1] * c1},
                    {sp[0] * c1, sp[1] * c1, sp[2] * c2}};
  /* B = R U R^T, R[a][c] = col[c][a]. */
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double v = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) v += col[i][a] * U[i][j] * col[j][b];
      out9[a * 3 + b] += v;
    }
}

/* Dispatch, Dipole_hessian.kern:212-238: q > 4 far, else near. */
void mfo_hessian_pair_frame(const double d[3], const double m[3], double h, double out9[9]) {
  for (int i = 0; i < 9; ++i) out9[i] = 0.0;
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double q = r * (1.0 / h);
  if (q > 4.0) hess_far(d, m, r, out9);
  else hess_near_frame(d, m, q, h, out9);
}

/* Closed isotropic form of the same tensor:
 *   B_ab = a (u_a m_b + m_a u_b + delta_ab u.m) + b u_a u_b (u.m),  u = d/r
 * far (q > 4): a = 3 mu0/(4 pi r^4), b = -15 mu0/(4 pi r^4)
 * near (0 < q <= 4): a = C1(q)/h^4, b = (C2(q) - 3 C1(q))/h^4.            */
void mfo_hessian_pair_iso(const double d[3], const double m[3], double h, double out9[9]) {
  for (int i = 0; i < 9; ++i) out9[i] = 0.0;
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double q = r / h;
  if (q == 0.0) return;
  double a, b;
  if (q > 4.0) {
    const double r4 = r * r * r * r;
    a = 3.0 * MFO_MU0 / (4.0 * MFO_PI * r4);
    b = -15.0 * MFO_MU0 / (4.0 * MFO_PI * r4);
  } else {
    double c1, c2;
    mfo_curves(q, &c1, &c2);
    const double h4 = h * h * h * h;
    a = c1 / h4;
    b = (c2 - 3.0 * c1) / h4;
  }
  const double u[3] = {d[0] / r, d[1] / r, d[2] / r};
  const double um = u[0] * m[0] + u[1] * m[1] + u[2] * m[2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      out9[i * 3 + j] = a * (u[i] * m[j] + m[i] * u[j] + (i == j ? um : 0.0)) + b * u[i] * u[j] * um;
}

/* ---- A3 / A8: all-pairs field ---------------------------------------------- */
void mfo_field_allpairs(int64_t nt, const double *tpos, int64_t ns, const double *spos,
                        const double *smom, double h, const double *hconst, double *out,
                        int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < nt; ++i) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 0; j < ns; ++j) {
      const double d[3] = {tpos[3 * i] - spos[3 * j], tpos[3 * i + 1] - spos[3 * j + 1],
                           tpos[3 * i + 2] - spos[3 * j + 2]};
      double o[3];
      mfo_field_pair(d, smom + 3 * j, h, o);
      acc[0] += o[0];
      acc[1] += o[1];
      acc[2] += o[2];
    }
    for (int a = 0; a < 3; ++a) out[3 * i + a] = acc[a] + (hconst ? hconst[a] : 0.0);
  }
}

/* ---- A4: all-pairs tensor ---------------------------------------------------- */
void mfo_hessian_allpairs(int64_t nt, const double *tpos, int64_t ns, const double *spos,
                          const double *smom, double h, double *out9, int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < nt; ++i) {
    double acc[9] = {0};
    for (int64_t j = 0; j < ns; ++j) {
      const double d[3] = {tpos[3 * i] - spos[3 * j], tpos[3 * i + 1] - spos[3 * j + 1],
                           tpos[3 * i + 2] - spos[3 * j + 2]};
      double o[9];
      mfo_hessian_pair_frame(d, smom + 3 * j, h, o);
      for (int k = 0; k < 9; ++k) acc[k] += o[k];
    }
    for (int k = 0; k < 9; ++k) out9[9 * i + k] = acc[k];
  }
}

/* ---- A5 / A9: force ------------------------------------------------------------ */
void mfo_force_allpairs(int64_t n, const double *pos, const double *mom, double h,
                        const double *text9, double *out, int nthreads) {
  double *T = (double *)malloc(sizeof(double) * 9 * (size_t)n);
  mfo_hessian_allpairs(n, pos, n, pos, mom, h, T, nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const double *m = mom + 3 * i;
    for (int a = 0; a < 3; ++a) {
      double f = 0.0;
      for (int b = 0; b < 3; ++b) {
        double t = T[9 * i + 3 * a + b];
        if (text9) t += text9[9 * i + 3 * a + b];
        f += t * m[b];
      }
      out[3 * i + a] = f;
    }
  }
  free(T);
}

/* ---- A7: the linear operator and solvers -------------------------------------- */
void mfo_apply_A(int64_t n, const double *pos, double h, double kappa, const double *x,
                 double *y, int nthreads) {
  double *m = (double *)calloc(3 * (size_t)n, sizeof(double));
  for (int64_t i = 0; i < 3 * n; ++i) m[i] = kappa * x[i];
  mfo_field_allpairs(n, pos, n, pos, m, h, NULL, y, nthreads);
  for (int64_t i = 0; i < 3 * n; ++i) y[i] = x[i] - y[i];
  free(m);
}

void mfo_dense_G(int64_t n, const double *pos, double h, double *G) {
  const int64_t N3 = 3 * n;
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < n; ++j) {
      const double d[3] = {pos[3 * i] - pos[3 * j], pos[3 * i + 1] - pos[3 * j + 1],
                           pos[3 * i + 2] - pos[3 * j + 2]};
      for (int b = 0; b < 3; ++b) {
        double e[3] = {0.0, 0.0, 0.0}, o[3];
        e[b] = 1.0;
        mfo_field_pair(d, e, h, o);
        for (int a = 0; a < 3; ++a) G[(3 * i + a) * N3 + 3 * j + b] = o[a];
      }
    }
}

static double dot(int64_t n, const double *a, const double *b) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

int mfo_cg_solve(int64_t n, const double *pos, double h, double kappa, const double *b,
                 double *x, double rtol, int maxit, double *relres, int nthreads) {
  const int64_t N3 = 3 * n;
  double *r = (double *)malloc(sizeof(double) * N3);
  double *p = (double *)malloc(sizeof(double) * N3);
  double *Ap = (double *)malloc(sizeof(double) * N3);
  mfo_apply_A(n, pos, h, kappa, x, Ap, nthreads);
  for (int64_t i = 0; i < N3; ++i) {
    r[i] = b[i] - Ap[i];
    p[i] = r[i];
  }
  const double bn = sqrt(dot(N3, b, b));
  double rr = dot(N3, r, r);
  int it = 0;
  while (it < maxit && sqrt(rr) > rtol * bn) {
    mfo_apply_A(n, pos, h, kappa, p, Ap, nthreads);
    const double alpha = rr / dot(N3, p, Ap);
    for (int64_t i = 0; i < N3; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    const double rr2 = dot(N3, r, r);
    const double beta = rr2 / rr;
    rr = rr2;
    for (int64_t i = 0; i < N3; ++i) p[i] = r[i] + beta * p[i];
    ++it;
  }
  if (relres) *relres = (bn > 0.0) ? sqrt(rr) / bn : sqrt(rr);
  free(r);
  free(p);
  free(Ap);
  return it;
}

/* The reference's nonlinear CG (magnetic_force.cu:874-888, 1008-1095,
 * 1275-1315), restated in FP64.  D(v) = G(kappa v) - v, D^T(v) = kappa G(v) - v. */
int mfo_ref_nlcg_solve(int64_t n, const double *pos, double h, double kappa,
                       const double *hext, double *x, int niter, double tol,
                       int max_iteration, double *err_hist, int max_hist, int nthreads) {
  const int64_t N3 = 3 * n;
  double *res = (double *)malloc(sizeof(double) * N3);
  double *grad = (double *)malloc(sizeof(double) * N3);
  double *ograd = (double *)malloc(sizeof(double) * N3);
  double *sd = (double *)malloc(sizeof(double) * N3);
  double *osd = (double *)malloc(sizeof(double) * N3);
  double *tmp = (double *)calloc((size_t)N3, sizeof(double));
  double *Gv = (double *)malloc(sizeof(double) * N3);
  int nh = 0;
  int i_step = 0;
  double last_err = 0.0;

/* residual = G(kappa x) + Hext - x   (:1016-1017) */
#define MFO_RESIDUAL()                                                   \
  do {                                                                   \
    for (int64_t k = 0; k < N3; ++k) tmp[k] = kappa * x[k];              \
    mfo_field_allpairs(n, pos, n, pos, tmp, h, NULL, Gv, nthreads);      \
    for (int64_t k = 0; k < N3; ++k) res[k] = Gv[k] + hext[k] - x[k];    \
  } while (0)
/* gradient = D^T(res) = kappa G(res) - res   (:1281-1284) */
#define MFO_GRADIENT()                                                   \
  do {                                                                   \
    mfo_field_allpairs(n, pos, n, pos, res, h, NULL, Gv, nthreads);      \
    for (int64_t k = 0; k < N3; ++k) grad[k] = kappa * Gv[k] - res[k];   \
  } while (0)
/* alpha = -(D(sd).res)/(|D(sd)|^2 + 1e-10), 0 if |D(sd)|^2 < 1e-10 (:1287-1302) */
#define MFO_ALPHA(out_alpha)                                             \
  do {                                                                   \
    for (int64_t k = 0; k < N3; ++k) tmp[k] = kappa * sd[k];             \
    mfo_field_allpairs(n, pos, n, pos, tmp, h, NULL, Gv, nthreads);      \
    for (int64_t k = 0; k < N3; ++k) Gv[k] -= sd[k];                     \
    const double nt_ = dot(N3, Gv, Gv);                                  \
    out_alpha = (nt_ < 1e-10) ? 0.0 : -(dot(N3, Gv, res) / (nt_ + 1e-10)); \
  } while (0)

  /* start(), :1008-1040 */
  MFO_RESIDUAL();
  double err = 0.5 * dot(N3, res, res);
  if (nh < max_hist) err_hist[nh++] = err;
  last_err = err;
  if (sqrt(err / (double)n) < tol) goto done;
  MFO_GRADIENT();
  for (int64_t k = 0; k < N3; ++k) sd[k] = -grad[k];
  {
    double alpha;
    MFO_ALPHA(alpha);
    for (int64_t k = 0; k < N3; ++k) x[k] += alpha * sd[k];
  }
  /* step_iteration(), :1043-1095 */
  for (int it = 0; it < niter && it < max_iteration; ++it) {
    memcpy(ograd, grad, sizeof(double) * N3);
    memcpy(osd, sd, sizeof(double) * N3);
    MFO_RESIDUAL();
    err = 0.5 * dot(N3, res, res);
    if (nh < max_hist) err_hist[nh++] = err;
    if (sqrt(err / (double)n) < tol) break;
    if (err > last_err) break;
    last_err = err;
    ++i_step;
    MFO_GRADIENT();
    double beta = 0.0;
    if (i_step % 10 != 0) {
      /* calc_beta, :1305-1315: -|g|^2 / (s_old.(g_old - g) + 1e-10) */
      double den = 0.0;
      for (int64_t k = 0; k < N3; ++k) den += osd[k] * (ograd[k] - grad[k]);
      beta = -dot(N3, grad, grad) / (den + 1e-10);
      if (beta < 0.0) beta = 0.0;
    }
    for (int64_t k = 0; k < N3; ++k) sd[k] = -grad[k] + beta * osd[k];
    double alpha;
    MFO_ALPHA(alpha);
    for (int64_t k = 0; k < N3; ++k) x[k] += alpha * sd[k];
  }
done:
#undef MFO_RESIDUAL
#undef MFO_GRADIENT
#undef MFO_ALPHA
  free(res);
  free(grad);
  free(ograd);
  free(sd);
  free(osd);
  free(tmp);
  free(Gv);
  return nh;
}

/* ---- FP32 restatements in the reference's operation order ------------------ */
void mfo_field_allpairs_f32ref(int64_t n, const float *x, const float *y, const float *z,
                               const float *mx, const float *my, const float *mz, float h,
                               float *ox, float *oy, float *oz, int nthreads) {
  set_threads(nthreads);
  const float PI_F = 3.14159265358979f;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    float hx = 0.f, hy = 0.f, hz = 0.f;
    float w_avr = 0.f, w_dens = 0.f;
    for (int64_t off = 0; off < n; off += 512) {
      int64_t ls = (n - off > 512) ? 512 : n - off;
      for (int64_t l = ls - 1; l >= 0; --l) { /* :62, walked downwards */
        const int64_t j = off + l;
        float rx = x[i] - x[j], ry = y[i] - y[j], rz = z[i] - z[j];
        float r = sqrtf(rx * rx + ry * ry + rz * rz);
        float q = r / h;
        float coe = 1 / PI_F / (h * h * h);
        if (q >= 2) {
          w_avr = 3.f / (4 * PI_F * r * r * r);
          w_dens = 0;
        }
        if (q >= 1 && q < 2) {
          w_avr = -(3.f * (q * q * q * q * q * q / 6.f - (6.f * q * q * q * q * q) / 5.f +
                           3.f * q * q * q * q - (8.f * q * q * q) / 3.f + 1.f / 15)) /
                  (4 * PI_F * r * r * r);
          w_dens = (float)(coe * 0.25 * (2 - q) * (2 - q) * (2 - q));
        }
        if (q < 1) {
          w_avr = (15.f * q * q * q - 36.f * q * q + 40.f) / (40 * PI_F * h * h * h);
          w_dens = (float)(coe * (0.25 * (2 - q) * (2 - q) * (2 - q) - (1 - q) * (1 - q) * (1 - q)));
        }
        if (r != 0) {
          float invr = 1 / r;
          rx *= invr;
          ry *= invr;
          rz *= invr;
        }
        float v1 = (rx * mx[j] + ry * my[j] + rz * mz[j]) * (w_avr - w_dens);
        float v2 = w_avr * 0.33333333333333333333333f - w_dens;
        hx += rx * v1 - v2 * mx[j];
        hy += ry * v1 - v2 * my[j];
        hz += rz * v1 - v2 * mz[j];
      }
    }
    ox[i] = hx;
    oy[i] = hy;
    oz[i] = hz;
  }
}

static float horner4f(const double *c, float x) {
  return x * (x * (x * (x * (float)c[0] + (float)c[1]) + (float)c[2]) + (float)c[3]) + (float)c[4];
}

void mfo_force_allpairs_f32ref(int64_t n, const float *x, const float *y, const float *z,
                               const float *mx, const float *my, const float *mz, float h,
                               float *fx, float *fy, float *fz, int nthreads) {
  set_threads(nthreads);
  const float PI_F = 3.14159265358979f;
  const float mu0 = 1.25663706e-6f;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    float B[3][3] = {{0}}; /* B[a][b] */
    for (int64_t off = 0; off < n; off += 512) {
      int64_t ls = (n - off > 512) ? 512 : n - off;
      for (int64_t l = ls - 1; l >= 0; --l) {
        const int64_t j = off + l;
        const float d[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
        const float s[3] = {mx[j], my[j], mz[j]};
        const float r = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const float q = r * (1 / h);
        if (q > 4) {
          const float rr = q * h;
          const float w_avr = 3.f / (4 * PI_F * rr * rr * rr);
          const float invr = (rr != 0) ? 1.f / rr : 1.f;
          const float A = mu0 * invr * invr * w_avr;
          const float rdm = d[0] * s[0] + d[1] * s[1] + d[2] * s[2];
          const float dAdr = mu0 * invr * invr * (5 * invr * (-w_avr)) * rdm * invr;
          const float dBdr = mu0 * invr * (-w_avr) * invr;
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
              B[a][b] += d[a] * s[b] * A + d[a] * dAdr * d[b] - s[a] * dBdr * d[b];
              if (a == b) B[a][b] += rdm * A;
            }
        } else if (q != 0) {
          int k = (q <= 1) ? 0 : (q <= 2) ? 1 : (q <= 3) ? 2 : 3;
          float c1 = horner4f(C1TAB[k], q) * (1 / (h * h * h * h));
          float c2 = horner4f(C2TAB[k], q) * (1 / (h * h * h * h));
          float zeta[3] = {d[0] / r, d[1] / r, d[2] / r};
          float eta[3] = {zeta[1], -zeta[0], 0.f}; /* cross(zeta, z) */
          eta[1] += 1e-10f;
          float en = sqrtf(eta[0] * eta[0] + eta[1] * eta[1] + eta[2] * eta[2]);
          eta[0] /= en;
          eta[1] /= en;
          eta[2] /= en;
          float xi[3] = {eta[1] * zeta[2] - eta[2] * zeta[1], eta[2] * zeta[0] - eta[0] * zeta[2],
                         eta[0] * zeta[1] - eta[1] * zeta[0]};
          const float *col[3] = {xi, eta, zeta};
          float sp[3];
          for (int c = 0; c < 3; ++c) sp[c] = col[c][0] * s[0] + col[c][1] * s[1] + col[c][2] * s[2];
          float U[3][3] = {{sp[2] * c1, 0.f, sp[0] * c1},
                           {0.f, sp[2] * c1, sp[1] * c1},
                           {sp[0] * c1, sp[1] * c1, sp[2] * c2}};
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
              float v = 0.f;
              for (int p = 0; p < 3; ++p)
                for (int qq = 0; qq < 3; ++qq) v += col[p][a] * U[p][qq] * col[qq][b];
              B[a][b] += v;
            }
        }
      }
    }
    const float m[3] = {mx[i], my[i], mz[i]};
    fx[i] = B[0][0] * m[0] + B[0][1] * m[1] + B[0][2] * m[2];
    fy[i] = B[1][0] * m[0] + B[1][1] * m[1] + B[1][2] * m[2];
    fz[i] = B[2][0] * m[0] + B[2][1] * m[1] + B[2][2] * m[2];
  }
}
