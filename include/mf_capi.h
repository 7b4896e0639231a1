/*
 * mf_capi.h -- C ABI of the MI355X (gfx950) magnetic hot path.
 *
 * The reference has no C ABI: its boundary is the C++ class in
 * magnetic_solver/magnetic_force.cuh:36-174.  This ABI is what that class
 * (re-declared in include/magnetic_force.cuh) calls, and what any FFI
 * (ctypes, cgo, JNI) binds.  Each entry point names the reference function
 * it replaces.
 *
 * Conventions
 *  - Every function returns an int status (mf_status); 0 = MF_OK.
 *    mf_ctx_last_error()/mf_last_error() return the message of the last
 *    failure (per context / per host thread).
 *  - Particle arrays are device pointers in SoA form (x, y, z arrays of
 *    float), the layout of the reference's cuda_abstract_vector
 *    (cuda_abstract_vector.cuh:166-243: dev_x, dev_y, dev_z).
 *  - Small parameter arrays (hconst[3], options, stats) are host memory.
 *  - All work is enqueued on the context's stream (mf_ctx_set_stream);
 *    functions that return host scalars synchronize that stream.
 *  - The library never frees caller memory.  Handles are released by their
 *    destroy function.  A context is not thread-safe: use one per thread.
 *  - Units: x = B/mu0 per particle, m = kappa x, kappa = V chi/(1+chi)
 *    (magnetic_force.cu:1016); fields in the B/mu0 form of the reference's
 *    G (G m = H_ferro + M, magnetic_force.cuh:107); forces carry mu0 once.
 */
#ifndef MF_CAPI_H
#define MF_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MF_ABI_VERSION 1

typedef enum mf_status {
  MF_OK = 0,
  MF_ERR_INVALID_ARG = 1,
  MF_ERR_HIP = 2,
  MF_ERR_ALLOC = 3,
  MF_ERR_NOT_CONVERGED = 4,
  MF_ERR_UNSUPPORTED = 5,
  MF_ERR_NO_DEVICE = 6
} mf_status;

typedef enum mf_precision {
  MF_PREC_FP32 = 0,      /* FP32 pair math, FP32 tile partials, FP64 cross-tile sums */
  MF_PREC_FP64 = 1       /* FP64 throughout (the GPU FP64 check path) */
} mf_precision;

/* Arithmetic of the FMM translations (M2L, and the M2M and L2L passes with it). */
typedef enum mf_m2l_precision {
  MF_M2L_FP32 = 0,   /* FP32 MFMA (v_mfma_f32_16x16x4_f32) */
  MF_M2L_BF16X6 = 1, /* BF16 MFMA, operands split in 3 BF16 parts, 6 products,
                        FP32 accumulation: FP32-level accuracy */
  MF_M2L_BF16X3 = 2, /* BF16 MFMA, 2 parts, 3 products: ~1e-5 relative per
                        translation (inner products of the mixed solve) */
  MF_M2L_BF16X1 = 3, /* BF16 MFMA, 1 part, 1 product: ~4e-3 relative per
                        translation (inner corrections of the refined solve) */
  MF_M2L_BF16X1_CUT = 4 /* MF_M2L_BF16X1 with the translation cut to the
                        16 x 32 blocks whose lowest local and multipole
                        degrees add up to at most min(order, 7) (5 of 32
                        blocks at order 10): ~2e-4 relative per product, for the
                        late products of a relaxed solve (relax_m2l_precision).
                        M2M and L2L run as MF_M2L_BF16X1. */
} mf_m2l_precision;

typedef enum mf_mode {
  MF_MODE_ALLPAIRS = 0,  /* reference "naive" mode, magnetic_force.cu:1200-1222 */
  MF_MODE_FMM = 1        /* reference default FMM mode, magnetic_force.cu:1223-1270 */
} mf_mode;

typedef enum mf_method {
  MF_SOLVE_CG = 0,       /* conjugate gradients on the SPD system (I - kappa G) x = b */
  MF_SOLVE_GMRES = 1,    /* restarted GMRES(restart) */
  MF_SOLVE_REF_NLCG = 2  /* the reference's nonlinear CG, magnetic_force.cu:1008-1095 */
} mf_method;

typedef struct mf_ctx mf_ctx;
typedef struct mf_plan mf_plan;

/* ---- context ------------------------------------------------------------- */
int mf_abi_version(void);
const char *mf_status_string(int status);
/* Thread-local message of the last failure on this host thread. */
const char *mf_last_error(void);

/* Binds HIP device `device`, creates a default-priority stream. */
int mf_ctx_create(int device, mf_ctx **out);
int mf_ctx_destroy(mf_ctx *ctx);
const char *mf_ctx_last_error(const mf_ctx *ctx);
/* Process-wide context of the current HIP device, created on first use and
 * kept until exit.  It runs on the HIP null stream, so its work is ordered
 * with synchronous hipMemcpy calls, as the reference's kernels were with its
 * cudaMemcpy calls (all on the legacy default stream).  Used by the C++
 * drop-in classes in magnetic_force.cuh / cuda_abstract_vector.cuh. */
int mf_ctx_get_default(mf_ctx **out);
/* Enqueue all further work on hip_stream (a hipStream_t passed as void*).
 * NULL selects the HIP null stream.  A new context starts on a stream of its
 * own (non-blocking, default priority); mf_ctx_get_stream returns the current
 * one, so a caller can switch back later. */
int mf_ctx_set_stream(mf_ctx *ctx, void *hip_stream);
void *mf_ctx_get_stream(mf_ctx *ctx);
int mf_ctx_synchronize(mf_ctx *ctx);
/* Per-kernel timing with HIP events on the context stream (off by default).
 * mf_ctx_kernel_time returns the summed milliseconds and launch count of the
 * named kernel family since the last reset ("field", "force", "hessian",
 * "p2p_field", "p2p_force", "m2l", ...). */
int mf_ctx_enable_timing(mf_ctx *ctx, int enable);
int mf_ctx_kernel_time(mf_ctx *ctx, const char *name, double *ms_total, int64_t *launches);
int mf_ctx_reset_timing(mf_ctx *ctx);

/* ---- stateless all-pairs products (separate target and source sets) ------ */
/* out_i = hconst + sum_j K(t_i - s_j) m_j.
 * Replaces cuda_Apply_G (magnetic_force.cu:8-110, targets == sources) and
 * cuda_ext_field_kernel (:120-237, external sources + constant field).
 * K is DipoleBKernel (fmmtl_minimum/kernel/Dipole_B.kern:46-128).
 * hconst (host, 3 floats) may be NULL.  precision: MF_PREC_FP32 only here. */
int mf_field_allpairs(mf_ctx *ctx, int64_t n_t, const float *tx, const float *ty,
                      const float *tz, int64_t n_s, const float *sx, const float *sy,
                      const float *sz, const float *mx, const float *my, const float *mz,
                      float h, const float *hconst, float *ox, float *oy, float *oz);

/* FP64 all-pairs field (device double SoA in and out): the GPU FP64 check
 * path used to score FP32 results at sizes the CPU oracle cannot reach. */
int mf_field_allpairs_f64(mf_ctx *ctx, int64_t n_t, const double *tx, const double *ty,
                          const double *tz, int64_t n_s, const double *sx, const double *sy,
                          const double *sz, const double *mx, const double *my,
                          const double *mz, double h, const double *hconst, double *ox,
                          double *oy, double *oz);

/* Force-tensor field T_i = sum_j B(t_i - s_j)[m_j], 6 symmetric components
 * written component-major to out6[6 * n_t] in the order xx, yy, zz, xy, xz, yz.
 * Replaces cuda_ext_fitted_force_tensor_kernel (magnetic_force.cu:432-512);
 * B is DipoleHessianKernel (fmmtl_minimum/kernel/Dipole_hessian.kern:19-278). */
int mf_hessian_allpairs(mf_ctx *ctx, int64_t n_t, const float *tx, const float *ty,
                        const float *tz, int64_t n_s, const float *sx, const float *sy,
                        const float *sz, const float *mx, const float *my, const float *mz,
                        float h, float *out6);

/* The same tensor written as the nine arrays the force takes: out9 is a host
 * array of 9 device pointers in the order of text9 below (the reference's
 * dHextdx.{x,y,z}, dHextdy.{x,y,z}, dHextdz.{x,y,z}, which
 * cuda_ext_fitted_force_tensor_kernel fills, magnetic_force.cu:432-512 and
 * :1164-1172), each symmetric component written to both of its places. */
int mf_hessian_allpairs9(mf_ctx *ctx, int64_t n_t, const float *tx, const float *ty,
                         const float *tz, int64_t n_s, const float *sx, const float *sy,
                         const float *sz, const float *mx, const float *my, const float *mz,
                         float h, float *const *out9);

/* Force on one particle set: F_i = (sum_j B(x_i - x_j)[m_j] + Text_i) m_i.
 * Replaces fitted_force_ofM_kernel (magnetic_force.cu:516-605).  text9 is a
 * host array of 9 device pointers (the reference's dHextdx.{x,y,z},
 * dHextdy.{x,y,z}, dHextdz.{x,y,z}) or NULL; Text carries mu0 once (the FMM
 * path's convention, magnetic_force.cu:968-980). */
int mf_force_allpairs(mf_ctx *ctx, int64_t n, const float *x, const float *y, const float *z,
                      const float *mx, const float *my, const float *mz, float h,
                      const float *const *text9, float *fx, float *fy, float *fz);

/* FP64 force (device double SoA), the FP64 check path for the force. */
int mf_force_allpairs_f64(mf_ctx *ctx, int64_t n, const double *x, const double *y,
                          const double *z, const double *mx, const double *my,
                          const double *mz, double h, double *fx, double *fy, double *fz);

/* FP64 force at a subset of the particles: F_t for t = idx[k], k < n_t
 * (idx: device int64 array), all n particles as sources.  Scores large runs
 * on a seeded subsample. */
int mf_force_subset_f64(mf_ctx *ctx, int64_t n, const double *x, const double *y,
                        const double *z, const double *mx, const double *my, const double *mz,
                        double h, int64_t n_t, const int64_t *idx, double *fx, double *fy,
                        double *fz);

/* ---- plans: one particle set, reused across products and solves ---------- */
typedef struct mf_plan_opts {
  int mode;          /* mf_mode */
  int leaf_size;     /* FMM: particle budget per leaf, the role of the reference's
                        crit_size / ncrit (magnetic_force.cu:647-648, NDTree.hpp:413-496):
                        with leaf_side 0, the leaf side is the largest 2h * 2^j whose
                        non-empty leaves hold at most leaf_size particles on average.
                        0 = leaf side 2h.  The octree has uniform depth. */
  int order;         /* FMM: expansion order p (degrees 0..p): 6, 8 or 10; 0 = 10 */
  int m2l_precision; /* FMM: mf_m2l_precision of mf_plan_field / mf_plan_force */
  int serial_near;   /* FMM: 0 = the near field runs on a second stream beside the M2L
                        (default); 1 = on the context stream, after the upward pass (for
                        timing each kernel alone) */
  float leaf_side;   /* FMM: leaf side in units of h, >= 2 (0 = from leaf_size).  Larger
                        leaves move work from the M2L to the near field. */
  int reserved[2];
} mf_plan_opts;

/* Replaces the construction of fmm_field_evaluator / fmm_hessian_evaluator
 * (magnetic_force.cu:650-672, 838-860) and, for the all-pairs mode, the
 * position upload of the constructor (:633).  Positions are copied. */
int mf_plan_create(mf_ctx *ctx, int64_t n, const float *x, const float *y, const float *z,
                   float h, const mf_plan_opts *opts, mf_plan **out);
/* New positions (same or different n): setTargetpos, magnetic_force.cu:806-871. */
int mf_plan_update_positions(mf_plan *plan, int64_t n, const float *x, const float *y,
                             const float *z);
int mf_plan_destroy(mf_plan *plan);
int64_t mf_plan_size(const mf_plan *plan);

/* Work counts of a plan, for throughput and roofline accounting.  All-pairs
 * plans report n^2 near pairs and no boxes. */
typedef struct mf_plan_info {
  int64_t n;
  int mode;
  int order;                /* FMM expansion order (0 for all-pairs) */
  int depth;                /* octree leaf level */
  int reserved0;
  int64_t leaves;
  int64_t boxes;
  int64_t p2p_field_pairs;  /* particle pairs evaluated directly per field product */
  int64_t p2p_force_pairs;  /* ... per force product */
  int64_t m2l_field_pairs;  /* box pairs translated per field product */
  int64_t m2l_force_pairs;
  int64_t m2l_field_tile_slots; /* (target tile, offset) steps the M2L kernel runs per
                                   field product: each applies one offset's operator
                                   to a tile of up to 64 target boxes */
  int64_t m2l_force_tile_slots;
  int64_t m2l_field_operators;  /* distinct offset operators of the field M2L */
} mf_plan_info;
int mf_plan_get_info(const mf_plan *plan, mf_plan_info *info);

/* out = G m (+ hconst): cuda_magnet_cgOptimizer::G, magnetic_force.cu:1196-1271. */
int mf_plan_field(mf_plan *plan, const float *mx, const float *my, const float *mz,
                  const float *hconst, float *ox, float *oy, float *oz);
/* F = (B(m) + Text) m: evaluate_fitted_near_foece, magnetic_force.cu:890-997. */
int mf_plan_force(mf_plan *plan, const float *mx, const float *my, const float *mz,
                  const float *const *text9, float *fx, float *fy, float *fz);

typedef struct mf_solve_opts {
  int method;        /* mf_method */
  int restart;       /* GMRES restart length (default 30) */
  int maxit;         /* max matvecs (CG/GMRES) or max iterations (REF_NLCG) */
  double rtol;       /* ||b - A x|| <= rtol ||b|| (CG/GMRES); RMS tol for REF_NLCG */
  int refine_fp64;   /* mf_plan_solve_f64: max FP64 refinement sweeps (0 = 10) */
  int true_residual; /* 1: recompute ||b - A x|| with a fresh product at exit */
  int zero_start;    /* 1: start from x = 0 (x is overwritten); the first residual
                        is b itself, which saves one product */
  int graph_iters;   /* mf_plan_solve_batch: > 0 captures one CG iteration (after one
                        eager iteration) in a hipGraph and replays it graph_iters
                        times between convergence checks; 0 = plain launches */
  int inner_m2l_precision; /* mf_plan_solve_f64 on an FMM plan: 1 + the mf_m2l_precision
                              of the inner correction solves; 0 = the plan's own */
  double inner_rtol; /* mf_plan_solve_f64: relative tolerance of each inner solve;
                        0 = 1e-5 (all-pairs plans), 1e-3 (FMM plans) */
  int relax_m2l_precision; /* CG on an FMM plan (mf_plan_solve, mf_plan_solve_batch): 1 + the
                              mf_m2l_precision of the products made once the relative residual
                              (the largest over the systems of a batch) is at or below
                              relax_rtol; 0 = off.  An error of the product made at residual
                              r perturbs the final residual by about that error times r, so the
                              late products may be less accurate (inexact Krylov). */
  int reserved_solve;
  double relax_rtol; /* see relax_m2l_precision; 0 = 1e-3 */
} mf_solve_opts;

typedef struct mf_solve_stats {
  int iterations;    /* inner iterations */
  int matvecs;       /* G products */
  int converged;
  int refinements;
  double rel_residual;      /* ||b - A x|| / ||b|| from the solver recurrences */
  double rel_residual_true; /* recomputed with a fresh product at exit */
  int relaxed_matvecs;      /* products made at relax_m2l_precision */
  int reserved_stats;
} mf_solve_stats;

/* Solve (I - kappa G) x = b in place (x holds the start).  Replaces
 * cuda_magnet_cgOptimizer::solve / start / step_iteration
 * (magnetic_force.cu:874-888, 1008-1095) with a device-resident solver. */
int mf_plan_solve(mf_plan *plan, double kappa, const float *bx, const float *by,
                  const float *bz, float *xx, float *xy, float *xz, const mf_solve_opts *opts,
                  mf_solve_stats *stats);

/* Mixed-precision iterative refinement with x and the residual
 * r = b - A x accumulated in FP64; the corrections d (A d = r) come from the
 * plan's own solver (CG to opts->inner_rtol).  b and x are device double
 * arrays laid out as one contiguous [3][n] block each (x component first);
 * x holds the start.
 *  - All-pairs plans: the residual product is the FP64 all-pairs product, so
 *    x reaches the FP64 solution (rtol 1e-12): the reference solution that
 *    FP32 and FMM results are scored against.
 *  - FMM plans: the residual product is the plan's FMM product at its own M2L
 *    precision, applied to the FP32 rounding of kappa x; the inner solves run
 *    at opts->inner_m2l_precision (e.g. BF16x1 under a BF16x3 or FP32 plan).
 *    x converges to the solution of the plan's FMM operator, whose distance
 *    from the exact solution is set by the expansion order.
 * stats->refinements counts the correction sweeps, stats->matvecs every
 * product (residual and inner).  No counterpart in the reference, whose solve
 * stops at an RMS residual of 0.1 (magnetic_force.cu:1043-1095). */
int mf_plan_solve_f64(mf_plan *plan, double kappa, const double *b3n, double *x3n,
                      const mf_solve_opts *opts, mf_solve_stats *stats);

/* ---- batch plans: many independent particle systems in one plan ---------- */
/* nsys systems stored back to back: system s owns particles
 * [offsets[s], offsets[s + 1]) of x, y, z (offsets: host, nsys + 1 entries,
 * offsets[0] = 0).  Systems do not interact: every product of the plan is
 * block-diagonal.  FMM mode only (opts->mode = MF_MODE_FMM); one smoothing
 * length h for all systems.  The systems share one octree forest (the system
 * index sits above the Morton bits of every key), so each kernel of a product
 * is one launch for the whole batch.  mf_plan_update_positions keeps nsys and
 * the offsets (same total n).  No counterpart in the reference, which solves
 * one particle set per cuda_magnet_cgOptimizer. */
int mf_plan_create_batch(mf_ctx *ctx, int nsys, const int64_t *offsets, const float *x,
                         const float *y, const float *z, float h, const mf_plan_opts *opts,
                         mf_plan **out);
/* (I - kappa[s] G_s) x = b for every system s at once: conjugate gradients
 * with per-system scalars in lockstep and one shared product per iteration;
 * a system stops updating once ||r_s|| <= rtol ||b_s||.  kappa: host, nsys.
 * opts: method must be MF_SOLVE_CG (rtol, maxit, zero_start honoured).
 * stats->rel_residual is the largest per-system relative residual;
 * sys_rel_residual (host, nsys, may be NULL) receives each system's. */
int mf_plan_solve_batch(mf_plan *plan, const double *kappa, const float *bx, const float *by,
                        const float *bz, float *xx, float *xy, float *xz,
                        const mf_solve_opts *opts, mf_solve_stats *stats,
                        double *sys_rel_residual);

/* ---- fused SoA 3-vector utilities (cuda_abstract_vector.cuh:12-161, 502-563) */
/* out = alpha a + beta b (any of a, b may alias out; b may be NULL => beta ignored) */
int mf_vec_axpby(mf_ctx *ctx, int64_t n, float alpha, const float *ax, const float *ay,
                 const float *az, float beta, const float *bx, const float *by,
                 const float *bz, float *ox, float *oy, float *oz);
/* out = a * b elementwise per component */
int mf_vec_mul(mf_ctx *ctx, int64_t n, const float *ax, const float *ay, const float *az,
               const float *bx, const float *by, const float *bz, float *ox, float *oy,
               float *oz);
/* sum over i and components of a_i b_i, FP64 partials, deterministic order. */
int mf_vec_dot(mf_ctx *ctx, int64_t n, const float *ax, const float *ay, const float *az,
               const float *bx, const float *by, const float *bz, double *result);
/* x[i] = v for every i (one component array). */
int mf_vec_fill(mf_ctx *ctx, int64_t n, float v, float *x);

/* The same four operations on double arrays, for cuda_abstract_vector<double>
 * (the reference's vector is a template over T, cuda_abstract_vector.cuh:166).
 * mf_vec_dot_f64 multiplies and sums in FP64, in the same fixed order. */
int mf_vec_axpby_f64(mf_ctx *ctx, int64_t n, double alpha, const double *ax, const double *ay,
                     const double *az, double beta, const double *bx, const double *by,
                     const double *bz, double *ox, double *oy, double *oz);
int mf_vec_mul_f64(mf_ctx *ctx, int64_t n, const double *ax, const double *ay, const double *az,
                   const double *bx, const double *by, const double *bz, double *ox, double *oy,
                   double *oz);
int mf_vec_dot_f64(mf_ctx *ctx, int64_t n, const double *ax, const double *ay, const double *az,
                   const double *bx, const double *by, const double *bz, double *result);
int mf_vec_fill_f64(mf_ctx *ctx, int64_t n, double v, double *x);

/* Element types of mf_vec_convert. */
typedef enum mf_dtype {
  MF_DTYPE_F32 = 0,
  MF_DTYPE_F64 = 1,
  MF_DTYPE_I32 = 2
} mf_dtype;
/* dst[i] = (dst type) src[i] for one component array of n elements: the
 * element conversion of cuda_abstract_vector's templated conversion operator
 * (cuda_conversion_kernel, cuda_abstract_vector.cuh:367-379).  Floating
 * values convert to int32 by truncation toward zero, as a C cast. */
int mf_vec_convert(mf_ctx *ctx, int64_t n, int src_dtype, const void *src, int dst_dtype,
                   void *dst);

/* ---- allocation fault injection ------------------------------------------ */
/* No counterpart in the reference, which never checks its cudaMalloc results
 * (cuda_abstract_vector.cuh:171-174).  The hook is called with the size of
 * every device allocation the library is about to make; a nonzero return
 * makes that allocation fail with MF_ERR_ALLOC, as a failed hipMalloc does.
 * NULL removes it.  Process-wide.  For tests of the allocation-failure paths
 * of the library and of its callers. */
typedef int (*mf_alloc_hook)(uint64_t bytes, void *user);
int mf_set_alloc_hook(mf_alloc_hook hook, void *user);

/* ---- allocation-free variant ------------------------------------------------ */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph stream capture, tight serving loops, harnesses
 * that forbid hipMalloc): the same product as mf_field_allpairs, with every
 * scratch buffer (packed source and target records, tile boxes, source-chunk
 * partials) in one static device array of mf_noalloc_scratch_bytes() bytes
 * (32 MiB) per GPU, held from library load until the process ends.  The call
 * itself allocates nothing; the context must exist already (mf_ctx_create
 * looks up the array's address on its device).  Limit: n_t and n_s such that
 * 32 (pad64(n_s) + 1) + 32 (pad64(n_s) / 64 + 1) + 32 (pad64(n_t) + 1)
 * + 24 n_t bytes (each term rounded up to 256) fit in the array, e.g.
 * n_t = n_s = 262144; larger sizes return MF_ERR_INVALID_ARG.  Every call on
 * one GPU shares the array: calls that overlap on one GPU race. */
int mf_field_allpairs_noalloc(mf_ctx *ctx, int64_t n_t, const float *tx, const float *ty,
                              const float *tz, int64_t n_s, const float *sx, const float *sy,
                              const float *sz, const float *mx, const float *my,
                              const float *mz, float h, const float *hconst, float *ox,
                              float *oy, float *oz);
/* Size in bytes of that array. */
int mf_noalloc_scratch_bytes(int64_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
