/*
 * mf_cpufmm.c -- FP64 CPU restatement of the reference's FMM field product.
 *
 * TEST INFRASTRUCTURE ONLY (see mf_oracle.h): the CPU baseline that bench.py
 * times beside the GPU, and a checker of its own accuracy in the CPU tests.
 *
 * It follows the reference's fmmtl path for G m (KernelMatrix.hpp:127-131 ->
 * Context.hpp:159-174 -> EvalLists.hpp:85-129):
 *  - tree: fmmtl_minimum/fmmtl/tree/NDTree.hpp:413-522 -- a cube around the
 *    points (side max extent (1 + 1e-6), at least 1e-10), 10-level Morton
 *    codes, boxes split level by level while they hold more than ncrit
 *    points, box centre = Morton cell centre, radius = half the cell diagonal
 *    (NDTree.hpp:186-193);
 *  - lists: the breadth-first dual traversal of
 *    traversal/DualTraversal.hpp:79-125 (split the larger box, leaf x leaf =
 *    near) with the MAC of FMMOptions.hpp:56-78 as the reference configures
 *    it (magnetic_force.cu:646-648: theta = 0.5, the extra gap r0 - r1 - r2 >
 *    3h);
 *  - far field in FP64: S2M of dipoles (DipoleBSpherical.hpp:95-137), M2M,
 *    M2L, L2L (SphericalMultipole3D.hpp:46-336) and L2T of -grad phi
 *    (DipoleBSpherical.hpp:191-230), with P = 6 terms (degrees 0..5, as
 *    DipoleBSpherical's default P);
 *  - near field: the smoothed pair kernel (Dipole_B.kern:79-128) over the
 *    near leaf pairs, in FP64 (the reference runs it on its GPU in FP32).
 * OpenMP over target boxes for the M2L and the near field, over the boxes of
 * a level for the translations, as the reference's CPU evaluator does.
 *
 * The expansions are written with the Cartesian recurrences of the normalised
 * regular and irregular solid harmonics R_n^m, I_n^m (Dehnen 2014, "A fast
 * multipole method for stellar dynamics", Comput. Astrophys. Cosmol. 1:1):
 *   1/|x - y| = sum_{n,m} conj(R_n^m(y)) I_n^m(x),          |y| < |x|
 *   R_n^m(a + b) = sum_{k,l} R_k^l(a) R_{n-k}^{m-l}(b)
 *   I_n^m(X - Y) = sum_{k,l} conj(R_k^l(Y)) I_{n+k}^{m+l}(X), |Y| < |X|
 *   d/dz R_n^m = R_{n-1}^m, d/dx R_n^m = (R_{n-1}^{m-1} - R_{n-1}^{m+1}) / 2,
 *   d/dy R_n^m = i (R_{n-1}^{m-1} + R_{n-1}^{m+1}) / 2,
 * the same algebra as the reference's complex spherical harmonics evalZ and
 * evalW, free of its trigonometry.  Coefficients with m < 0 follow from
 * A_n^{-m} = (-1)^m conj(A_n^m).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "mf_oracle.h"

typedef double complex cplx;

#define CF_P 6                          /* expansion terms: degrees 0 .. CF_P - 1 */
#define CF_NC (CF_P * (CF_P + 1) / 2)   /* coefficients with m >= 0 */
#define CF_P2 (2 * CF_P - 1)            /* degrees of the M2L harmonics: 0 .. 2P - 2 */
#define CF_NC2 (CF_P2 * (CF_P2 + 1) / 2)
#define CF_LEVELS 10                    /* Morton levels (32-bit codes) */

static inline int nm_idx(int n, int m) { return n * (n + 1) / 2 + m; }

/* A_n^m for any m, |m| <= n, from the m >= 0 half; 0 outside. */
static inline cplx get_nm(const cplx *A, int n, int m) {
  if (n < 0 || m > n || -m > n) return 0.0;
  if (m >= 0) return A[nm_idx(n, m)];
  return (((-m) & 1) ? -1.0 : 1.0) * conj(A[nm_idx(n, -m)]);
}

static void regular_sh(int np, double x, double y, double z, cplx *R) {
  const double r2 = x * x + y * y + z * z;
  const cplx w = x + I * y;
  R[0] = 1.0;
  for (int n = 1; n < np; ++n) {
    R[nm_idx(n, n)] = w / (2.0 * n) * R[nm_idx(n - 1, n - 1)];
    for (int m = 0; m < n; ++m) {
      cplx v = (2.0 * n - 1.0) * z * R[nm_idx(n - 1, m)];
      if (m <= n - 2) v -= r2 * R[nm_idx(n - 2, m)];
      R[nm_idx(n, m)] = v / ((double)(n + m) * (double)(n - m));
    }
  }
}

static void irregular_sh(int np, double x, double y, double z, cplx *Ir) {
  const double r2 = x * x + y * y + z * z;
  const double ir2 = 1.0 / r2;
  const cplx w = x + I * y;
  Ir[0] = sqrt(ir2);
  for (int n = 1; n < np; ++n) {
    Ir[nm_idx(n, n)] = (2.0 * n - 1.0) * w * ir2 * Ir[nm_idx(n - 1, n - 1)];
    for (int m = 0; m < n; ++m) {
      cplx v = (2.0 * n - 1.0) * z * Ir[nm_idx(n - 1, m)];
      if (m <= n - 2) v -= ((double)(n - 1) * (n - 1) - (double)m * m) * Ir[nm_idx(n - 2, m)];
      Ir[nm_idx(n, m)] = v * ir2;
    }
  }
}

/* ---- tree (NDTree.hpp:413-522) ------------------------------------------- */

typedef struct {
  double c[3];     /* Morton cell centre */
  double r2;       /* squared radius: |extents|^2 / 4 */
  int64_t first, count;  /* points [first, first + count) in Morton order */
  int child0, nchild;    /* children are contiguous */
  int level;
} CBox;

typedef struct {
  uint32_t code;
  int64_t idx;
} CodeIdx;

static int cmp_code(const void *a, const void *b) {
  const CodeIdx *x = (const CodeIdx *)a, *y = (const CodeIdx *)b;
  if (x->code != y->code) return x->code < y->code ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

static uint32_t spread3(uint32_t v) { /* 10 bits -> every third bit */
  uint32_t x = v & 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

typedef struct {
  CBox *box;
  int nbox, cap;
  int level_start[CF_LEVELS + 2];
  int nlevels;
  double lo[3], side;
} CTree;

static int push_box(CTree *t, const CBox *b) {
  if (t->nbox == t->cap) {
    t->cap = t->cap ? 2 * t->cap : 1024;
    t->box = (CBox *)realloc(t->box, sizeof(CBox) * (size_t)t->cap);
  }
  t->box[t->nbox] = *b;
  return t->nbox++;
}

static void build_tree(CTree *t, int64_t n, const double *pos, int ncrit, int64_t *perm) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fmin(lo[a], pos[3 * i + a]);
      hi[a] = fmax(hi[a], pos[3 * i + a]);
    }
  double max_side = 0.0, ctr[3];
  for (int a = 0; a < 3; ++a) {
    max_side = fmax(max_side, hi[a] - lo[a]);
    ctr[a] = 0.5 * (hi[a] + lo[a]);
  }
  if (max_side < 1e-10) max_side = 1e-10;
  const double radius = (1.0 + 1e-6) * max_side / 2.0;
  for (int a = 0; a < 3; ++a) t->lo[a] = ctr[a] - radius;
  t->side = 2.0 * radius;
  const double cells = (double)(1u << CF_LEVELS);
  CodeIdx *ci = (CodeIdx *)malloc(sizeof(CodeIdx) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; ++i) {
    uint32_t cc[3];
    for (int a = 0; a < 3; ++a) {
      double u = (pos[3 * i + a] - t->lo[a]) / t->side * cells;
      int64_t k = (int64_t)floor(u);
      if (k < 0) k = 0;
      if (k >= (int64_t)cells) k = (int64_t)cells - 1;
      cc[a] = (uint32_t)k;
    }
    ci[i].code = spread3(cc[0]) | (spread3(cc[1]) << 1) | (spread3(cc[2]) << 2);
    ci[i].idx = i;
  }
  /* the reference bucket-sorts level by level; one full sort by code gives
     the same boxes (points of a leaf may be in another order) */
  qsort(ci, (size_t)n, sizeof(CodeIdx), cmp_code);
  for (int64_t i = 0; i < n; ++i) perm[i] = ci[i].idx;
  t->box = NULL;
  t->nbox = t->cap = 0;
  CBox root;
  for (int a = 0; a < 3; ++a) root.c[a] = ctr[a];
  root.r2 = 3.0 * t->side * t->side / 4.0;
  root.first = 0;
  root.count = n;
  root.child0 = -1;
  root.nchild = 0;
  root.level = 0;
  push_box(t, &root);
  t->level_start[0] = 0;
  t->level_start[1] = 1;
  t->nlevels = 1;
  for (int L = 1; L <= CF_LEVELS; ++L) {
    const int b0 = t->level_start[L - 1], b1 = t->level_start[L];
    const int shift = 3 * (CF_LEVELS - L);
    const double cs = t->side / (double)(1u << L);
    for (int k = b0; k < b1; ++k) {
      if (t->box[k].count <= ncrit) continue;
      const int64_t f = t->box[k].first, e = f + t->box[k].count;
      t->box[k].child0 = t->nbox;
      int64_t p = f;
      for (uint32_t oct = 0; oct < 8; ++oct) {
        const int64_t q0 = p;
        while (p < e && ((ci[p].code >> shift) & 7u) == oct) ++p;
        if (p == q0) continue;
        CBox c;
        const uint32_t code = ci[q0].code >> shift;  /* cell index at level L */
        uint32_t g[3] = {0, 0, 0};
        for (int bit = 0; bit < L; ++bit)
          for (int a = 0; a < 3; ++a) g[a] |= ((code >> (3 * bit + a)) & 1u) << bit;
        for (int a = 0; a < 3; ++a) c.c[a] = t->lo[a] + ((double)g[a] + 0.5) * cs;
        c.r2 = 3.0 * cs * cs / 4.0;
        c.first = q0;
        c.count = p - q0;
        c.child0 = -1;
        c.nchild = 0;
        c.level = L;
        push_box(t, &c);
        t->box[k].nchild++;
      }
    }
    if (t->nbox == b1) break;
    t->level_start[L + 1] = t->nbox;
    t->nlevels = L + 1;
  }
  free(ci);
}

/* ---- dual traversal (DualTraversal.hpp:79-125, FMMOptions.hpp:56-78) ------ */

typedef struct {
  int *t, *s;
  int64_t n, cap;
} PairList;

static void pl_push(PairList *l, int t, int s) {
  if (l->n == l->cap) {
    l->cap = l->cap ? 2 * l->cap : 4096;
    l->t = (int *)realloc(l->t, sizeof(int) * (size_t)l->cap);
    l->s = (int *)realloc(l->s, sizeof(int) * (size_t)l->cap);
  }
  l->t[l->n] = t;
  l->s[l->n] = s;
  l->n++;
}

static int mac_far(const CBox *a, const CBox *b, double theta2, double gap) {
  double d2 = 0.0;
  for (int k = 0; k < 3; ++k) d2 += (a->c[k] - b->c[k]) * (a->c[k] - b->c[k]);
  const double r0 = sqrt(d2), r1 = sqrt(a->r2), r2 = sqrt(b->r2);
  const double rs = a->r2 + b->r2 + 2.0 * sqrt(a->r2 * b->r2);
  return (theta2 * d2 > rs) && (r0 - r1 - r2 > gap);
}

static void traverse(const CTree *t, double theta, double gap, PairList *nearl, PairList *farl) {
  const double th2 = theta * theta;
  int64_t qcap = 1 << 16, qh = 0, qt = 0;
  int *qs = (int *)malloc(sizeof(int) * (size_t)qcap), *qtg = (int *)malloc(sizeof(int) * (size_t)qcap);
#define QPUSH(S, T)                                                              \
  do {                                                                           \
    if (qt == qcap) {                                                            \
      qcap *= 2;                                                                 \
      qs = (int *)realloc(qs, sizeof(int) * (size_t)qcap);                       \
      qtg = (int *)realloc(qtg, sizeof(int) * (size_t)qcap);                     \
    }                                                                            \
    qs[qt] = (S);                                                                \
    qtg[qt] = (T);                                                               \
    ++qt;                                                                        \
  } while (0)
  if (mac_far(&t->box[0], &t->box[0], th2, gap)) pl_push(farl, 0, 0);
  else QPUSH(0, 0);
  while (qh < qt) {
    const int s = qs[qh], tg = qtg[qh];
    ++qh;
    if (qh > (1 << 20) && qh * 2 > qt) { /* compact the consumed head */
      memmove(qs, qs + qh, sizeof(int) * (size_t)(qt - qh));
      memmove(qtg, qtg + qh, sizeof(int) * (size_t)(qt - qh));
      qt -= qh;
      qh = 0;
    }
    const CBox *sb = &t->box[s], *tb = &t->box[tg];
    const int sleaf = sb->nchild == 0, tleaf = tb->nchild == 0;
    if (sleaf && tleaf) {
      pl_push(nearl, tg, s);
      continue;
    }
    const int split_source = !sleaf && (tleaf || sb->r2 > tb->r2);
    if (split_source) {
      for (int c = sb->child0; c < sb->child0 + sb->nchild; ++c) {
        if (mac_far(&t->box[c], tb, th2, gap)) pl_push(farl, tg, c);
        else QPUSH(c, tg);
      }
    } else {
      for (int c = tb->child0; c < tb->child0 + tb->nchild; ++c) {
        if (mac_far(sb, &t->box[c], th2, gap)) pl_push(farl, c, s);
        else QPUSH(s, c);
      }
    }
  }
#undef QPUSH
  free(qs);
  free(qtg);
}

/* CSR of a pair list by target box. */
static void by_target(const PairList *l, int nbox, int64_t **off_out, int **src_out) {
  int64_t *off = (int64_t *)calloc((size_t)nbox + 1, sizeof(int64_t));
  int *src = (int *)malloc(sizeof(int) * (size_t)(l->n > 0 ? l->n : 1));
  for (int64_t k = 0; k < l->n; ++k) off[l->t[k] + 1]++;
  for (int b = 0; b < nbox; ++b) off[b + 1] += off[b];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nbox + 1));
  memcpy(fill, off, sizeof(int64_t) * (size_t)(nbox + 1));
  for (int64_t k = 0; k < l->n; ++k) src[fill[l->t[k]]++] = l->s[k];
  free(fill);
  *off_out = off;
  *src_out = src;
}

/* ---- expansions ----------------------------------------------------------- */

/* M_n^m += (1/4pi) p . grad conj(R_n^m)(y - c) for one dipole p at y. */
static void s2m_dipole(const double d[3], const double p[3], cplx *M) {
  cplx R[CF_NC];
  regular_sh(CF_P, d[0], d[1], d[2], R);
  const double s = 1.0 / (4.0 * 3.14159265358979323846);
  for (int n = 1; n < CF_P; ++n)
    for (int m = 0; m <= n; ++m) {
      const cplx gx = 0.5 * (get_nm(R, n - 1, m - 1) - get_nm(R, n - 1, m + 1));
      const cplx gy = 0.5 * I * (get_nm(R, n - 1, m - 1) + get_nm(R, n - 1, m + 1));
      const cplx gz = get_nm(R, n - 1, m);
      M[nm_idx(n, m)] += s * conj(p[0] * gx + p[1] * gy + p[2] * gz);
    }
}

/* M' = M shifted from centre c to centre c' (d = c - c'). */
static void m2m(const cplx *M, const double d[3], cplx *Mp) {
  cplx R[CF_NC];
  regular_sh(CF_P, d[0], d[1], d[2], R);
  for (int n = 0; n < CF_P; ++n)
    for (int m = 0; m <= n; ++m) {
      cplx acc = 0.0;
      for (int k = 0; k <= n; ++k)
        for (int l = -k; l <= k; ++l)
          acc += get_nm(M, k, l) * conj(get_nm(R, n - k, m - l));
      Mp[nm_idx(n, m)] += acc;
    }
}

/* L_n^m += (-1)^n sum_{k,l} M_k^l I_{n+k}^{m+l}(d), d = c_target - c_source. */
static void m2l(const cplx *M, const double d[3], cplx *L) {
  cplx Ir[CF_NC2];
  irregular_sh(CF_P2, d[0], d[1], d[2], Ir);
  for (int n = 0; n < CF_P; ++n)
    for (int m = 0; m <= n; ++m) {
      cplx acc = 0.0;
      for (int k = 0; k < CF_P; ++k)
        for (int l = -k; l <= k; ++l) acc += get_nm(M, k, l) * get_nm(Ir, n + k, m + l);
      L[nm_idx(n, m)] += (n & 1) ? -acc : acc;
    }
}

/* L' (about c') from L (about c): d = c' - c. */
static void l2l(const cplx *L, const double d[3], cplx *Lp) {
  cplx R[CF_NC];
  regular_sh(CF_P, d[0], d[1], d[2], R);
  for (int k = 0; k < CF_P; ++k)
    for (int l = 0; l <= k; ++l) {
      cplx acc = 0.0;
      for (int n = k; n < CF_P; ++n)
        for (int m = -n; m <= n; ++m) acc += get_nm(L, n, m) * conj(get_nm(R, n - k, m - l));
      Lp[nm_idx(k, l)] += acc;
    }
}

/* H += -grad phi at offset d from the local centre, phi = sum L conj(R). */
static void l2t(const cplx *L, const double d[3], double H[3]) {
  cplx R[CF_NC];
  regular_sh(CF_P, d[0], d[1], d[2], R);
  double hx = 0.0, hy = 0.0, hz = 0.0;
  for (int n = 1; n < CF_P; ++n)
    for (int m = -n; m <= n; ++m) {
      const cplx Ln = get_nm(L, n, m);
      const cplx gx = 0.5 * (get_nm(R, n - 1, m - 1) - get_nm(R, n - 1, m + 1));
      const cplx gy = 0.5 * I * (get_nm(R, n - 1, m - 1) + get_nm(R, n - 1, m + 1));
      const cplx gz = get_nm(R, n - 1, m);
      hx -= creal(Ln * conj(gx));
      hy -= creal(Ln * conj(gy));
      hz -= creal(Ln * conj(gz));
    }
  H[0] += hx;
  H[1] += hy;
  H[2] += hz;
}

/* ---- the product ------------------------------------------------------------ */

int mfo_fmm_field(int64_t n, const double *pos, const double *mom, double h, int ncrit,
                  double theta, double gap, const double *hconst, double *out, int nthreads,
                  mfo_fmm_stats *stats) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  if (n <= 0) return 0;
  if (ncrit < 1) ncrit = 384;
  int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  CTree t;
  build_tree(&t, n, pos, ncrit, perm);
  PairList nearl = {0}, farl = {0};
  traverse(&t, theta, gap, &nearl, &farl);
  int64_t *noff, *foff;
  int *nsrc, *fsrc;
  by_target(&nearl, t.nbox, &noff, &nsrc);
  by_target(&farl, t.nbox, &foff, &fsrc);
  const int nbox = t.nbox;
  double *sp = (double *)malloc(sizeof(double) * 3 * (size_t)n);  /* sorted positions */
  double *sm = (double *)malloc(sizeof(double) * 3 * (size_t)n);  /* sorted moments */
  double *so = (double *)calloc(3 * (size_t)n, sizeof(double));   /* sorted outputs */
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      sp[3 * i + a] = pos[3 * perm[i] + a];
      sm[3 * i + a] = mom[3 * perm[i] + a];
    }
  cplx *M = (cplx *)calloc((size_t)nbox * CF_NC, sizeof(cplx));
  cplx *L = (cplx *)calloc((size_t)nbox * CF_NC, sizeof(cplx));
  /* S2M at the leaves */
#pragma omp parallel for schedule(dynamic, 16)
  for (int b = 0; b < nbox; ++b) {
    const CBox *B = &t.box[b];
    if (B->nchild) continue;
    for (int64_t i = B->first; i < B->first + B->count; ++i) {
      double d[3] = {sp[3 * i] - B->c[0], sp[3 * i + 1] - B->c[1], sp[3 * i + 2] - B->c[2]};
      s2m_dipole(d, &sm[3 * i], &M[(size_t)b * CF_NC]);
    }
  }
  /* M2M, deepest level first */
  for (int lv = t.nlevels - 2; lv >= 0; --lv) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int b = t.level_start[lv]; b < t.level_start[lv + 1]; ++b) {
      const CBox *B = &t.box[b];
      for (int c = B->child0; c < B->child0 + B->nchild; ++c) {
        const double d[3] = {t.box[c].c[0] - B->c[0], t.box[c].c[1] - B->c[1],
                             t.box[c].c[2] - B->c[2]};
        m2m(&M[(size_t)c * CF_NC], d, &M[(size_t)b * CF_NC]);
      }
    }
  }
  /* M2L over target boxes */
#pragma omp parallel for schedule(dynamic, 8)
  for (int b = 0; b < nbox; ++b)
    for (int64_t k = foff[b]; k < foff[b + 1]; ++k) {
      const int s = fsrc[k];
      const double d[3] = {t.box[b].c[0] - t.box[s].c[0], t.box[b].c[1] - t.box[s].c[1],
                           t.box[b].c[2] - t.box[s].c[2]};
      m2l(&M[(size_t)s * CF_NC], d, &L[(size_t)b * CF_NC]);
    }
  /* L2L, root level first */
  for (int lv = 0; lv + 1 < t.nlevels; ++lv) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int b = t.level_start[lv]; b < t.level_start[lv + 1]; ++b) {
      const CBox *B = &t.box[b];
      for (int c = B->child0; c < B->child0 + B->nchild; ++c) {
        const double d[3] = {t.box[c].c[0] - B->c[0], t.box[c].c[1] - B->c[1],
                             t.box[c].c[2] - B->c[2]};
        l2l(&L[(size_t)b * CF_NC], d, &L[(size_t)c * CF_NC]);
      }
    }
  }
  /* L2T and the near field at the leaves */
#pragma omp parallel for schedule(dynamic, 4)
  for (int b = 0; b < nbox; ++b) {
    const CBox *B = &t.box[b];
    if (B->nchild) continue;
    for (int64_t i = B->first; i < B->first + B->count; ++i) {
      double H[3] = {0.0, 0.0, 0.0};
      const double d[3] = {sp[3 * i] - B->c[0], sp[3 * i + 1] - B->c[1], sp[3 * i + 2] - B->c[2]};
      l2t(&L[(size_t)b * CF_NC], d, H);
      for (int64_t k = noff[b]; k < noff[b + 1]; ++k) {
        const CBox *S = &t.box[nsrc[k]];
        for (int64_t j = S->first; j < S->first + S->count; ++j) {
          const double dd[3] = {sp[3 * i] - sp[3 * j], sp[3 * i + 1] - sp[3 * j + 1],
                                sp[3 * i + 2] - sp[3 * j + 2]};
          double f[3];
          mfo_field_pair(dd, &sm[3 * j], h, f);
          H[0] += f[0];
          H[1] += f[1];
          H[2] += f[2];
        }
      }
      for (int a = 0; a < 3; ++a) so[3 * i + a] = H[a];
    }
  }
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) out[3 * perm[i] + a] = so[3 * i + a] + (hconst ? hconst[a] : 0.0);
  if (stats) {
    int nleaf = 0;
    for (int b = 0; b < nbox; ++b) nleaf += t.box[b].nchild == 0;
    int64_t near_pairs = 0;
    for (int b = 0; b < nbox; ++b)
      for (int64_t k = noff[b]; k < noff[b + 1]; ++k) near_pairs += t.box[b].count * t.box[nsrc[k]].count;
    stats->boxes = nbox;
    stats->leaves = nleaf;
    stats->levels = t.nlevels;
    stats->near_box_pairs = nearl.n;
    stats->far_box_pairs = farl.n;
    stats->near_particle_pairs = near_pairs;
  }
  free(perm);
  free(t.box);
  free(nearl.t);
  free(nearl.s);
  free(farl.t);
  free(farl.s);
  free(noff);
  free(nsrc);
  free(foff);
  free(fsrc);
  free(sp);
  free(sm);
  free(so);
  free(M);
  free(L);
  return 0;
}
