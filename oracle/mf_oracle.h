/*
 * mf_oracle.h -- CPU restatement of the reference's magnetic hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for the CPU baseline, as the thing timed beside the GPU).  The product
 * path (magnetic-force_amd) never links, loads or calls it.
 *
 * Layout: positions, moments and fields are AoS double[n][3] unless the name
 * says f32.  Tensors are double[n][9], row-major (a,b) = d_a H_b.
 *
 * Every function names the reference file:line it restates.  Parity pins
 * (closed forms, invariants, brute force) live in tests/test_oracle_*.py;
 * functions with no such pin are marked "parity unpinned" below (none today).
 */
#ifndef MF_ORACLE_H
#define MF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* mu0 exactly as the reference truncates it:
 * magnetic_solver/magnetic_force.cu:254, fmmtl_minimum/kernel/Dipole_hessian.kern:170 */
#define MFO_MU0 1.25663706e-6

/* SPH weights w_avr (ball average of the M4 spline) and w_dens (the spline).
 * fmmtl_minimum/kernel/Dipole_B.kern:46-76 (== magnetic_force.cu:69-83). */
void mfo_weights(double r, double h, double *w_avr, double *w_dens);

/* Pair field K(t-s) m (B/mu0 form, density term included).
 * Dipole_B.kern:79-104, 114-128.  r = 0 gives the self term (2/3) sigma m. */
void mfo_field_pair(const double d[3], const double m[3], double h, double out[3]);

/* Pair force tensor B(t-s)[m] = mu0 dH/dx as 9 comps, row-major.
 * _frame follows the reference literally (far: Dipole_hessian.kern:154-207;
 * near: rotation frame M U M^T, :88-152; dispatch :212-238).
 * _iso is the closed isotropic form of the same tensor. */
void mfo_hessian_pair_frame(const double d[3], const double m[3], double h, double out9[9]);
void mfo_hessian_pair_iso(const double d[3], const double m[3], double h, double out9[9]);

/* The two fitted radial curves C1, C2 (already divided by nothing: raw table
 * values at q), Dipole_hessian.kern:53-85.  Returns 0 outside (0,4]. */
void mfo_curves(double q, double *c1, double *c2);

/* All-pairs field: out_i = Hc + sum_j K(t_i - s_j) m_j  (cuda_Apply_G
 * magnetic_force.cu:8-110 when t == s; cuda_ext_field_kernel :120-237 with a
 * separate source set).  hconst may be NULL.  nthreads <= 0: all cores. */
void mfo_field_allpairs(int64_t nt, const double *tpos, int64_t ns, const double *spos,
                        const double *smom, double h, const double *hconst, double *out,
                        int nthreads);

/* All-pairs force tensor T_i = sum_j B(t_i - s_j)[m_j] (9 comps).
 * cuda_ext_fitted_force_tensor_kernel magnetic_force.cu:432-512 / the
 * DipoleHessianKernel direct sum. */
void mfo_hessian_allpairs(int64_t nt, const double *tpos, int64_t ns, const double *spos,
                          const double *smom, double h, double *out9, int nthreads);

/* Force F_i = (sum_j B(x_i-x_j)[m_j] + Text_i) m_i over one particle set.
 * FMM-path semantics (magnetic_force.cu:968-980): Text carries mu0 once.
 * text9 may be NULL. */
void mfo_force_allpairs(int64_t n, const double *pos, const double *mom, double h,
                        const double *text9, double *out, int nthreads);

/* y = x - kappa * G x over one particle set (the linear operator of the
 * magnetization solve, magnetic_force.cu:1016-1017, 1275-1284). */
void mfo_apply_A(int64_t n, const double *pos, double h, double kappa, const double *x,
                 double *y, int nthreads);

/* Dense G (3n x 3n, row-major) for small n: block (i,j) = K(x_i - x_j). */
void mfo_dense_G(int64_t n, const double *pos, double h, double *G);

/* FP64 conjugate gradients on (I - kappa G) x = b (A is symmetric positive
 * definite).  x holds the start on entry.  Stops at ||b-Ax|| <= rtol ||b||.
 * Returns iterations; *relres gets the final relative residual. */
int mfo_cg_solve(int64_t n, const double *pos, double h, double kappa, const double *b,
                 double *x, double rtol, int maxit, double *relres, int nthreads);

/* The reference's own solver: nonlinear CG on 0.5||G(kappa x) + Hext - x||^2
 * with Dai-Yuan beta clamped >= 0, reset every 10 steps, exact line step
 * alpha = -(Ds.r)/(|Ds|^2 + 1e-10), stop on RMS residual < tol or error
 * increase or max_iteration.  magnetic_force.cu:874-888, 1008-1095,
 * 1275-1315.  Runs in FP64.  err_hist receives 0.5||r||^2 per evaluation
 * (at most max_hist entries).  Returns the number of entries written. */
int mfo_ref_nlcg_solve(int64_t n, const double *pos, double h, double kappa,
                       const double *hext, double *x, int niter, double tol,
                       int max_iteration, double *err_hist, int max_hist, int nthreads);

/* FP32 restatement of cuda_Apply_G in its own operation order (tiles of 512
 * sources, each tile walked from its last source to its first, plain FP32
 * accumulation, magnetic_force.cu:44-99).  SoA float arrays. */
void mfo_field_allpairs_f32ref(int64_t n, const float *x, const float *y, const float *z,
                               const float *mx, const float *my, const float *mz, float h,
                               float *ox, float *oy, float *oz, int nthreads);

/* FP32 restatement of fitted_force_ofM_kernel (magnetic_force.cu:516-605)
 * with the rotation-frame near field, no external tensor. */
void mfo_force_allpairs_f32ref(int64_t n, const float *x, const float *y, const float *z,
                               const float *mx, const float *my, const float *mz, float h,
                               float *fx, float *fy, float *fz, int nthreads);

/* FP64 CPU restatement of the reference's FMM field product (mf_cpufmm.c):
 * out_i = hconst + (G m)_i with an ncrit-split octree (NDTree.hpp:413-522),
 * the breadth-first dual traversal (DualTraversal.hpp:79-125) under the MAC
 * theta^2 r0^2 > (r1 + r2)^2 and r0 - r1 - r2 > gap (FMMOptions.hpp:56-78),
 * P = 6 expansion terms (DipoleBSpherical.hpp:95-230,
 * SphericalMultipole3D.hpp:46-336) and the smoothed pair kernel over near
 * leaf pairs.  The reference configures ncrit 384, theta 0.5, gap 3h
 * (magnetic_force.cu:646-648).  Returns 0.  Its accuracy is pinned against
 * mfo_field_allpairs in the CPU tests. */
typedef struct mfo_fmm_stats {
  int64_t boxes, leaves, levels;
  int64_t near_box_pairs, far_box_pairs, near_particle_pairs;
} mfo_fmm_stats;
int mfo_fmm_field(int64_t n, const double *pos, const double *mom, double h, int ncrit,
                  double theta, double gap, const double *hconst, double *out, int nthreads,
                  mfo_fmm_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
