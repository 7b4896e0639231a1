// mf_fmm.hip -- GPU octree fast multipole method for the dipole field and the
// force tensor (gfx950).
//
// Replaces the reference's FMM product (fmmtl kernel_matrix, KernelMatrix.hpp:
// 127-131 -> Context.hpp:159-174 -> EvalLists.hpp:85-129): the adaptive NDTree
// (NDTree.hpp:413-496), the dual-tree near/far split (DualTraversal.hpp:79-125,
// MAC FMMOptions.hpp:56-78), the GPU near field (blocked_p2p,
// S2T_Blocked_CSR_cu_Multi_GPU.inl:51-134) and the CPU far field
// (S2M/M2M/M2L/L2L/L2T of DipoleBSpherical / DipoleHessianSpherical).
//
// Design (everything on the device, one particle set per plan):
//  * Uniform-depth sparse octree.  The root cube has side 2h 2^D, so leaves
//    have side exactly 2h; particles are sorted by 63-bit Morton key of their
//    leaf (rocPRIM radix sort) and boxes of every level are the distinct key
//    prefixes.
//  * Near sets (in box units at one level): N = {o : |o|^2 <= 6} (81 boxes),
//    hereditary, worst M2L separation ratio 0.61.  The field near field is N
//    at the leaf level (covers every pair with q < 2); the force near field is
//    NH = {o : sum max(|o_i| - 1, 0)^2 <= 4} (179 boxes: every pair within 4h,
//    where the reference's fitted curves live, and the boxes exactly two leaf
//    sides away, whose pairs take the dipole form: translating those instead
//    costs the force twice its error for 13 % of its time).  Interaction
//    lists: children of the parent's N-neighbours minus the near set.
//  * Expansions in scaled real dofs (mf_expansion.hpp): M2M/L2L are 8 fixed
//    matrices, M2L one fixed matrix per offset (column-major tables in HBM).
//  * M2L runs on MFMA (FP32, or BF16 with split operands): a workgroup owns
//    64 target boxes of one (level, parity) class and walks that class's
//    offset list; per offset it gathers the 64 source multipoles into LDS and
//    multiplies by the offset's matrix, accumulating the 64 local expansions
//    in registers.
//  * Leaf pass: L2P (one wave per leaf, local expansion through scalar loads)
//    writes H_const + far field in the caller's particle order; the near-field
//    kernel then adds the sums over the leaf's near boxes, whose particle
//    ranges are merged into contiguous runs at plan build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <type_traits>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "mf_allpairs.hpp"
#include "mf_expansion.hpp"
#include "mf_fmm_tables.hpp"
#include "mf_internal.hpp"
#include "mf_pair.hpp"
#include "mf_plan.hpp"
#include "mf_sort.hpp"

namespace mf {

namespace {

constexpr int kNearF = 81;
constexpr int kNearFMixed = 27;  // slots with |o|_inf <= 1 may hold q < 2 pairs
// Force near set: the boxes at most two leaf sides apart (every pair within
// 4h for leaf sides >= 2h).  Its slots are ordered by that separation; the
// prefix that may hold pairs within 4h (at leaf side 2h the 125 boxes of
// {-2..2}^3) runs the near test, the rest (at 2h the 54 boxes exactly two
// leaf sides away) take the far form with no test.
constexpr int kNearH = 179;
constexpr int kMaxLevels = 21;

struct I3 {
  int x, y, z;
};

// ---------------------------------------------------------------------------
// Morton keys (21 bits per axis).

__host__ __device__ inline uint64_t spread3(uint64_t v) {
  v &= 0x1fffffULL;
  v = (v | (v << 32)) & 0x1f00000000ffffULL;
  v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
  v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}
__host__ __device__ inline uint64_t compact3(uint64_t v) {
  v &= 0x1249249249249249ULL;
  v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ULL;
  v = (v ^ (v >> 4)) & 0x100f00f00f00f00fULL;
  v = (v ^ (v >> 8)) & 0x1f0000ff0000ffULL;
  v = (v ^ (v >> 16)) & 0x1f00000000ffffULL;
  v = (v ^ (v >> 32)) & 0x1fffffULL;
  return v;
}
__host__ __device__ inline uint64_t morton(int x, int y, int z) {
  return spread3(static_cast<uint64_t>(x)) | (spread3(static_cast<uint64_t>(y)) << 1) |
         (spread3(static_cast<uint64_t>(z)) << 2);
}
__host__ __device__ inline I3 demorton(uint64_t k) {
  return {static_cast<int>(compact3(k)), static_cast<int>(compact3(k >> 1)),
          static_cast<int>(compact3(k >> 2))};
}

__device__ inline int find_key(const uint64_t *__restrict__ keys, int count, uint64_t key) {
  int lo = 0, hi = count - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const uint64_t v = keys[mid];
    if (v == key) return mid;
    if (v < key) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

// Box geometry of one level.  Keys carry the system index above the 3l
// Morton bits (a plan may hold several independent particle systems, each
// with its own root cube of the common side L); lo[3 sys + a] is that
// system's root corner.
// Products of a batch plan inside its batched solve skip the boxes of
// systems that have converged: box_sys[box] is the box's system, active[s]
// (device, written by the solver) whether system s still iterates.  active ==
// nullptr: every box runs.  The skipped boxes keep their last outputs, which
// the solver no longer reads.
struct SysMask {
  const int32_t *box_sys = nullptr;
  const int32_t *active = nullptr;
};
__device__ __forceinline__ bool sys_skip(const SysMask &m, int box) {
  return m.active != nullptr && m.active[m.box_sys[box]] == 0;
}

struct Frame {
  const double *__restrict__ lo;
  int shift;  // 3 l
};

__device__ inline void box_centre(const Frame f, uint64_t key, double s, double &cx, double &cy,
                                  double &cz) {
  const uint64_t sys = key >> f.shift;
  const I3 c = demorton(key & ((1ULL << f.shift) - 1ULL));
  cx = f.lo[3 * sys] + (c.x + 0.5) * s;
  cy = f.lo[3 * sys + 1] + (c.y + 0.5) * s;
  cz = f.lo[3 * sys + 2] + (c.z + 0.5) * s;
}

// ---------------------------------------------------------------------------
// Tree build kernels.

// Bounding box of each chunk [chunk[2c], chunk[2c + 1]) of the records.
__global__ void bbox_partial_kernel(const int64_t *__restrict__ chunk, const SrcRec *__restrict__ rec,
                                    float *__restrict__ part) {
  __shared__ float sh[6][256];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  const int64_t c0 = chunk[2 * blockIdx.x], c1 = chunk[2 * blockIdx.x + 1];
  int bad = 0;  // a NaN coordinate (fminf/fmaxf would drop it)
  for (int64_t i = c0 + threadIdx.x; i < c1; i += 256) {
    const SrcRec r = rec[i];
    bad |= (r.x != r.x) | (r.y != r.y) | (r.z != r.z);
    lo[0] = fminf(lo[0], r.x);
    lo[1] = fminf(lo[1], r.y);
    lo[2] = fminf(lo[2], r.z);
    hi[0] = fmaxf(hi[0], r.x);
    hi[1] = fmaxf(hi[1], r.y);
    hi[2] = fmaxf(hi[2], r.z);
  }
  for (int c = 0; c < 3; ++c) {
    sh[c][threadIdx.x] = lo[c];
    sh[3 + c][threadIdx.x] = hi[c];
  }
  bad = __syncthreads_or(bad);  // also orders the writes above
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int c = 0; c < 3; ++c) {
        sh[c][threadIdx.x] = fminf(sh[c][threadIdx.x], sh[c][threadIdx.x + s]);
        sh[3 + c][threadIdx.x] = fmaxf(sh[3 + c][threadIdx.x], sh[3 + c][threadIdx.x + s]);
      }
    __syncthreads();
  }
  // a NaN anywhere in the chunk: the partial box is NaN, which the host's
  // finiteness check of the bounding box rejects
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = bad ? NAN : sh[threadIdx.x][0];
}

// key = system << 3D | Morton(leaf cell); sys_of == nullptr: one system.
__global__ void keys_kernel(int64_t n, const SrcRec *__restrict__ rec,
                            const int32_t *__restrict__ sys_of, const double *__restrict__ sys_lo,
                            double inv_s, int cells, int shift, uint64_t *__restrict__ keys,
                            int32_t *__restrict__ idx) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SrcRec r = rec[i];
  const int sys = sys_of ? sys_of[i] : 0;
  int c[3];
  const double p[3] = {(r.x - sys_lo[3 * sys]) * inv_s, (r.y - sys_lo[3 * sys + 1]) * inv_s,
                       (r.z - sys_lo[3 * sys + 2]) * inv_s};
  for (int a = 0; a < 3; ++a) {
    int v = static_cast<int>(floor(p[a]));
    c[a] = v < 0 ? 0 : (v >= cells ? cells - 1 : v);
  }
  keys[i] = (static_cast<uint64_t>(sys) << shift) | morton(c[0], c[1], c[2]);
  idx[i] = static_cast<int32_t>(i);
}

// Sorted particle records (positions only; moments are packed per product).
__global__ void gather_sorted_kernel(int64_t n, const SrcRec *__restrict__ rec,
                                     const int32_t *__restrict__ perm, SrcRec *__restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  SrcRec r = rec[perm[i]];
  r.mx = r.my = r.mz = 0.f;
  out[i] = r;
}

// flags[i] = 1 where a new key starts (key shifted right by `shift` bits).
__global__ void run_flags_kernel(int64_t n, const uint64_t *__restrict__ keys, int shift,
                                 int32_t *__restrict__ flags) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = (i == 0 || (keys[i] >> shift) != (keys[i - 1] >> shift)) ? 1 : 0;
}

// Box counts of the levels above the leaves, from the sorted leaf keys: level
// l < D holds the distinct prefixes key >> 3 (D - l), so counts[l] (zero on
// entry) receives the number of prefix changes; the first box is the + 1.
__global__ void level_counts_kernel(int nleaf, const uint64_t *__restrict__ keys, int D,
                                    int32_t *__restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i > 0 && i < nleaf;
  const uint64_t k = live ? keys[i] : 0ULL, kp = live ? keys[i - 1] : 0ULL;
  const int lane = threadIdx.x & 63;
  for (int l = 0; l < D; ++l) {
    const int sh = 3 * (D - l);
    const uint64_t m = __ballot(live && (k >> sh) != (kp >> sh));
    if (m != 0 && lane == __ffsll(static_cast<long long>(m)) - 1) atomicAdd(&counts[l], __popcll(m));
  }
}

// Writes the distinct keys and the index of each run's first element.
__global__ void run_starts_kernel(int64_t n, const uint64_t *__restrict__ keys, int shift,
                                  const int32_t *__restrict__ flags, const int32_t *__restrict__ scan,
                                  uint64_t *__restrict__ out_keys, int32_t *__restrict__ out_first,
                                  int32_t *__restrict__ elem_run) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t run = scan[i] + flags[i] - 1;
  if (elem_run) elem_run[i] = run;
  if (flags[i]) {
    out_keys[run] = keys[i] >> shift;
    out_first[run] = static_cast<int32_t>(i);
  }
}

// count[j] = first[j+1] - first[j] (last: total - first).
__global__ void run_counts_kernel(int nrun, int total, const int32_t *__restrict__ first,
                                  int32_t *__restrict__ count) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nrun) return;
  count[j] = (j + 1 < nrun ? first[j + 1] : total) - first[j];
}

// Boxes of every level above the leaves at once.  Level l < D holds the
// distinct prefixes key >> 3 (D - l) of the sorted leaf keys and sits at
// [base[l], base[l] + cnt[l]) of the box arrays, the levels stored after the
// leaves from D - 1 down to 0.  Row r = D - 1 - l of the (level, leaf) flag
// array marks the leaves that start a box of level l; rows are laid out in
// that order, so an exclusive scan over all rows gives every flagged leaf its
// box's global index minus cnt[D].
struct LevelTable {
  int D;
  int base[kMaxLevels];
  int cnt[kMaxLevels];
};

__global__ void level_flags_kernel(int nleaf, const uint64_t *__restrict__ keys, int D,
                                   int32_t *__restrict__ flags) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<int64_t>(D) * nleaf) return;
  const int r = static_cast<int>(e / nleaf), i = static_cast<int>(e % nleaf);
  const int sh = 3 * (r + 1);  // level l = D - 1 - r: shift 3 (D - l)
  flags[e] = (i == 0 || (keys[i] >> sh) != (keys[i - 1] >> sh)) ? 1 : 0;
}

// Per (level l, leaf i): a box starting at i gets its key and first child
// (level-local index of leaf i's level l + 1 box); a level l + 1 box starting
// at i gets its parent (leaf i's level l box).
__global__ void level_boxes_kernel(int nleaf, const uint64_t *__restrict__ keys, LevelTable lt,
                                   const int32_t *__restrict__ flags,
                                   const int32_t *__restrict__ scan, uint64_t *__restrict__ box_keys,
                                   int32_t *__restrict__ box_first, int32_t *__restrict__ box_parent) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<int64_t>(lt.D) * nleaf) return;
  const int r = static_cast<int>(e / nleaf), i = static_cast<int>(e % nleaf);
  const int l = lt.D - 1 - r;
  const int g = nleaf + scan[e] + flags[e] - 1;  // leaf i's level l box
  // leaf i's level l + 1 box (at l + 1 = D, the leaf itself), and whether i starts it
  const int64_t ec = e - nleaf;
  const int gc = r == 0 ? i : nleaf + scan[ec] + flags[ec] - 1;
  const bool child_start = r == 0 || flags[ec] != 0;
  if (flags[e]) {
    box_keys[g] = keys[i] >> (3 * (r + 1));
    box_first[g] = gc - lt.base[l + 1];
  }
  if (child_start) box_parent[gc] = g;
}

// Child counts of the boxes above the leaves: first child of the next box of
// the level (or the child level's box count) minus this box's first child.
__global__ void level_child_counts_kernel(int nupper, LevelTable lt,
                                          const int32_t *__restrict__ box_first,
                                          int32_t *__restrict__ box_count) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x + lt.cnt[lt.D];
  if (j >= lt.cnt[lt.D] + nupper) return;
  int l = lt.D - 1;
  while (l > 0 && j >= lt.base[l] + lt.cnt[l]) --l;
  const int end = j + 1 < lt.base[l] + lt.cnt[l] ? box_first[j + 1] : lt.cnt[l + 1];
  box_count[j] = end - box_first[j];
}

// Near-field lists of the leaves: list[b * nslot + k] = leaf index or -1.
__global__ void near_list_kernel(int nleaf, const uint64_t *__restrict__ leaf_keys, int cells,
                                 int shift, const int3 *__restrict__ offs, int nslot,
                                 const int32_t *__restrict__ grid, int32_t *__restrict__ list) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(nleaf) * nslot) return;
  const int b = static_cast<int>(t / nslot), k = static_cast<int>(t % nslot);
  const uint64_t key = leaf_keys[b];
  const uint64_t sys_bits = (key >> shift) << shift;
  const I3 c = demorton(key & ((1ULL << shift) - 1ULL));
  const int3 o = offs[k];
  const int x = c.x + o.x, y = c.y + o.y, z = c.z + o.z;
  int r = -1;
  if (x >= 0 && y >= 0 && z >= 0 && x < cells && y < cells && z < cells)
    r = grid ? grid[sys_bits | morton(x, y, z)]
             : find_key(leaf_keys, nleaf, sys_bits | morton(x, y, z));
  list[t] = r;
}

// Dense cell -> box grid of one level: the level's keys are the system
// index above a Morton code of 3l bits, so grid[key] = box (level-local),
// -1 for an empty cell.  The list kernels then look a neighbour up with one
// load instead of a binary search over the level's keys.
__global__ void grid_clear_kernel(int64_t n, int32_t *__restrict__ grid) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) grid[i] = -1;
}

__global__ void grid_scatter_kernel(int nbox, const uint64_t *__restrict__ keys,
                                    int32_t *__restrict__ grid) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nbox) grid[keys[b]] = b;
}

// M2L source slots of the boxes of one level: src[b * nslot + k] = global
// index of the source box at offset offs[parity][k] (target - source), or -1.
__global__ void m2l_list_kernel(int nbox, int box_base, const uint64_t *__restrict__ keys,
                                int cells, int shift, const int3 *__restrict__ offs, int nslot,
                                const int32_t *__restrict__ grid, int32_t *__restrict__ src) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(nbox) * nslot) return;
  const int b = static_cast<int>(t / nslot), k = static_cast<int>(t % nslot);
  const uint64_t key = keys[b];
  const uint64_t sys_bits = (key >> shift) << shift;
  const I3 c = demorton(key & ((1ULL << shift) - 1ULL));
  const int parity = (c.x & 1) | ((c.y & 1) << 1) | ((c.z & 1) << 2);
  const int3 o = offs[parity * nslot + k];
  int r = -1;
  if (o.x != 0x7fffffff) {
    const int x = c.x - o.x, y = c.y - o.y, z = c.z - o.z;
    if (x >= 0 && y >= 0 && z >= 0 && x < cells && y < cells && z < cells) {
      const int f = grid ? grid[sys_bits | morton(x, y, z)]
                         : find_key(keys, nbox, sys_bits | morton(x, y, z));
      r = f >= 0 ? box_base + f : -1;
    }
  }
  src[static_cast<int64_t>(box_base + b) * nslot + k] = r;
}

// Class key (level << 3 | parity) of every box at levels >= 2.
__global__ void class_key_kernel(int nbox, int box_base, int level,
                                 const uint64_t *__restrict__ keys, uint64_t *__restrict__ ckey,
                                 int32_t *__restrict__ ids, int32_t *__restrict__ counts) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int parity = b < nbox ? static_cast<int>(keys[b] & 7ULL) : -1;
  if (b < nbox) {
    ckey[b] = (static_cast<uint64_t>(level) << 3) | static_cast<uint64_t>(parity);
    ids[b] = box_base + b;
  }
  // one counter update per parity present in the wave
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    const uint64_t m = __ballot(parity == p);
    if (m != 0 && lane == __ffsll(static_cast<long long>(m)) - 1)
      atomicAdd(&counts[level * 8 + p], __popcll(m));
  }
}

// ---------------------------------------------------------------------------
// Upward pass.

// P2M: one wave per leaf; each lane accumulates its particles' dipole
// multipoles, the wave reduces through LDS.
template <int P, int KP>
__global__ __launch_bounds__(64) void p2m_kernel(int nleaf, const int32_t *__restrict__ start,
                                                 const int32_t *__restrict__ count,
                                                 const uint64_t *__restrict__ keys,
                                                 const SrcRec *__restrict__ rec, Frame fr,
                                                 double s, float *__restrict__ Mt, SysMask mask) {
  constexpr int ND = (P + 1) * (P + 1);
  // The wave's sums go through LDS in two halves of the dofs: a 64 x ND
  // buffer (31 KB at p = 10) would hold P2M to 5 waves per CU.
  constexpr int NH = (ND + 1) / 2;
  static_assert(NH <= 64, "one dof per lane in each half");
  __shared__ float red[64][NH + 1];
  const int b = blockIdx.x;
  if (sys_skip(mask, b)) return;  // uniform over the wave, before any barrier
  const int lane = threadIdx.x;
  double cx, cy, cz;
  box_centre(fr, keys[b], s, cx, cy, cz);
  const float inv_s = static_cast<float>(1.0 / s);
  float acc[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) acc[d] = 0.f;
  const int s0 = start[b], cnt = count[b];
  for (int t = lane; t < cnt; t += 64) {
    const SrcRec r = rec[s0 + t];
    const float x = static_cast<float>(r.x - cx) * inv_s, y = static_cast<float>(r.y - cy) * inv_s,
                z = static_cast<float>(r.z - cz) * inv_s;
    cplx R[(P + 1) * (P + 2) / 2];
    regular_tri<P>(x, y, z, R);
    const float mx = r.mx, my = r.my, mz = r.mz;
    // v = mz R_{n-1}^m + (mx - i my)/2 R_{n-1}^{m+1} - (mx + i my)/2 R_{n-1}^{m-1};  M += conj(v)
#pragma unroll
    for (int n = 1; n <= P; ++n) {
#pragma unroll
      for (int m = 0; m <= n; ++m) {
        const cplx a = rget<P>(R, n - 1, m), bp = rget<P>(R, n - 1, m + 1), bm = rget<P>(R, n - 1, m - 1);
        const float vre = mz * a.re + 0.5f * (mx * bp.re + my * bp.im) - 0.5f * (mx * bm.re - my * bm.im);
        const float vim = mz * a.im + 0.5f * (mx * bp.im - my * bp.re) - 0.5f * (mx * bm.im + my * bm.re);
        acc[dof_re(n, m)] += vre;
        if (m > 0) acc[dof_im(n, m)] -= vim;
      }
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (h > 0) __syncthreads();  // the first half is read
#pragma unroll
    for (int d = 0; d < NH; ++d)
      if (h * NH + d < ND) red[lane][d] = acc[h * NH + d];
    __syncthreads();
    const int d = h * NH + lane;
    if (lane < NH && d < ND) {
      float sum = 0.f;
      for (int l = 0; l < 64; ++l) sum += red[l][lane];
      Mt[static_cast<int64_t>(b) * KP + d] = sum * inv_s;
    }
  }
  for (int d = ND + lane; d < KP; d += 64) Mt[static_cast<int64_t>(b) * KP + d] = 0.f;
}

// ---------------------------------------------------------------------------
// M2L on MFMA.  A workgroup owns a tile of up to kTileT = 64 target boxes of
// one (level, parity) class and walks that class's offset list.  Per offset
// it gathers the tile's source multipoles into LDS (skipping the offset when
// no target of the tile has a source there) and multiplies them by the
// offset's KP x KP matrix, accumulating the local expansions in registers.
// KP/32 waves: wave w owns rows [32 w, 32 w + 32) of all 64 columns, as
// 2 x 4 blocks of 16x16.  Workgroups are remapped so that each XCD
// (round-robin dispatch) walks a contiguous range of tiles: neighbouring tiles
// share classes, offset matrices and source multipoles, which then stay in
// that XCD's L2.
//
// Software pipeline: the gather of the next offset runs under the current
// offset's MFMAs.  The iteration for offset c stages the multipoles of c
// (held in registers) into LDS, issues the multipole loads of the next offset
// n1 (whose source indices were loaded one iteration earlier) and the
// source-index loads of n2, then multiplies offset c from LDS.  Each thread
// holds 8 float4 of the 64 x KP block: element e = tid + it * NT is target
// e / (KP / 4), dofs 4 (e % (KP / 4)) .. + 3.
//
// Precisions (NPART):
//  * 0: FP32 MFMA v_mfma_f32_16x16x4_f32, K stepped by 4.  Operand layout
//    (wave64): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16];
//    D[4 (l / 16) + r][l % 16] = acc[r].  Within a 16-wide k block, step j
//    takes k = 4 (l / 16) + j, so a lane's four steps read four consecutive
//    k: one 16-byte table load per row block (fragment order) and one
//    16-byte LDS read per column block.
//  * 2 or 3: BF16 MFMA v_mfma_f32_16x16x32_bf16, K stepped by 32; lane l
//    supplies A[l % 16][8 (l / 16) .. + 7] and B[8 (l / 16) .. + 7][l % 16].
//    Every FP32 operand x is split into BF16 parts x = x0 + x1 + x2 (each the
//    round-to-nearest-even BF16 of the remainder; three parts hold all 24
//    significand bits).  NPART = 3 sums the six products a_i b_j with
//    i + j <= 2 (~2^-24 relative per product, FP32 level); NPART = 2 the
//    three with i + j <= 1 (~2^-16).  Accumulation is FP32 in the MFMA.  A
//    BF16 MFMA does 16x the multiply-adds of the FP32 one per cycle, so six
//    products cost 3/8 of the FP32 MFMA time.  The offset matrices are stored
//    pre-split (from their FP64 values) in fragment order: a lane's eight
//    consecutive k of one row are one 16-byte load, and a wave's fragment is
//    1 KB of consecutive memory (whole cache lines).  The multipoles are
//    pre-split once per product (split_multipoles_kernel).
constexpr int kTileT = 64;
// Word 3 of a raw buffer resource on gfx950 (32-bit data format, no
// swizzle): loads past the range return zeros.
constexpr int kBufferRsrcWord3 = 0x00020000;
constexpr int kRowM2M = 16, kRowL2L = 17, kSlotRows = 25;  // slot-table rows

struct M2LTile {
  int first;   // position in the class-sorted box list
  int count;   // <= kTileT
  int slots;   // row of the slot -> table map (kind * 8 + parity)
  int nslot;
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint32_t bf16_bits_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

template <int NPART>
__device__ __forceinline__ void split_bf16(float x, uint32_t (&p)[NPART]) {
  float r = x;
#pragma unroll
  for (int i = 0; i < NPART; ++i) {
    p[i] = bf16_bits_rne(r);
    if (i + 1 < NPART) r -= __uint_as_float(p[i] << 16);
  }
}

// Source multipoles of one offset for the tile: the 64 x KP block as 16-byte
// chunks, FP32 float4 (NPART = 0) or BF16x8 of the pre-split parts (Mb,
// [box][3][KP]), of which the BF16 gather may take the first KG columns of
// each part only.  Chunk e (e = tid + it * 2 KP) is target e / PER_T, chunk
// e % PER_T of that target.
template <int KP, int NPART, int KG = KP, int KG1 = KG>
struct M2LGather {
  // chunks per target: KG columns of part 0, KG1 of each further part
  static constexpr int PER_T = NPART == 0 ? KP / 4 : (KG + (NPART - 1) * KG1) / 8;
  static constexpr int IT = 64 * PER_T / (2 * KP);                    // chunks per thread
  static_assert(IT * 2 * KP == 64 * PER_T, "the tile's chunks split evenly over the threads");
};

// The gather reads through a buffer resource whose range is the multipole
// array: a missing source (-1) becomes an offset past the range, which the
// hardware returns as zeros.  Every load is issued without a branch; a load
// under a branch makes the compiler wait for all outstanding loads (these
// and the operator prefetch) at the first use of any of them, which would
// serialize the gather with the MFMAs.
template <int KP, int NPART, int KG = KP, int KG1 = KG>
__device__ __forceinline__ void m2l_mult_load(__amdgpu_buffer_rsrc_t msrc,
                                              const int (&src)[M2LGather<KP, NPART, KG, KG1>::IT],
                                              int tid,
                                              float4 (&v)[M2LGather<KP, NPART, KG, KG1>::IT]) {
  using G = M2LGather<KP, NPART, KG, KG1>;
  constexpr uint32_t kBoxBytes = NPART == 0 ? KP * 4 : 3 * KP * 2;
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int it = 0; it < G::IT; ++it) {
    const uint32_t r = static_cast<uint32_t>((tid + it * 2 * KP) % G::PER_T);
    // BF16: chunk r of the box's [3][KP] parts block; src = -1 wraps to
    // 2^32 - kBoxBytes, beyond any range the host accepts
    // chunk r: part 0 for r < KG / 8, then KG1 / 8 chunks of each further part
    uint32_t cpart = 0, cq = r;
    if (NPART != 0 && r >= KG / 8) {
      cpart = 1 + (r - KG / 8) / (KG1 / 8);
      cq = (r - KG / 8) % (KG1 / 8);
    }
    const uint32_t off = static_cast<uint32_t>(src[it]) * kBoxBytes + cpart * (2u * KP) + 16u * cq;
    const u4 x = __builtin_amdgcn_raw_buffer_load_b128(msrc, static_cast<int>(off), 0, 0);
    v[it] = make_float4(__uint_as_float(x.x), __uint_as_float(x.y), __uint_as_float(x.z),
                        __uint_as_float(x.w));
  }
}

// Bytes of the multipole range the M2L gather may address: the wrapped offset
// of a missing source (2^32 - box bytes) must lie beyond it.
inline bool m2l_gather_range_ok(int64_t nbox, int KP, int npart) {
  const int64_t box_bytes = npart == 0 ? KP * 4 : 3 * KP * 2;
  return nbox * box_bytes <= (int64_t{1} << 32) - 2 * box_bytes;
}

// Pre-split multipoles: Mb[(box * 3 + p) * KP + d] = BF16 part p of Mt[box * KP + d].
__global__ void split_multipoles_kernel(int64_t nelem, int KP, const float *__restrict__ Mt,
                                        uint16_t *__restrict__ Mb) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nelem) return;
  const int64_t box = i / KP;
  const int d = static_cast<int>(i % KP);
  uint32_t p[3];
  split_bf16<3>(Mt[i], p);
#pragma unroll
  for (int k = 0; k < 3; ++k) Mb[(box * 3 + k) * KP + d] = static_cast<uint16_t>(p[k]);
}

// OCTANT marks the instances that run the M2M and L2L passes (octant
// matrices).  The code is the same; the separate instances let kernel traces
// list those passes apart from the M2L.
// Waves per SIMD the M2L instances are built for: three (at most 168
// registers) for the FP32 and one- and two-part kernels, whose operator
// prefetch would otherwise take the compiler to 196 registers and two waves
// (measured slower than without the prefetch); the three-part kernel and the
// p = 8 two-part kernel do not fit three without spilling and keep the
// compiler's choice.
template <int KP, int NPART>
constexpr int m2l_min_waves() {
  return (NPART == 3 || (NPART == 2 && KP != 128)) ? 1 : 3;
}

// Per-launch tile options of the M2L kernels.  src, ld, empty: the source
// slots of the tiles from first_tile on, when one launch covers tiles whose
// slot arrays differ (the force M2L: the leaf level reads the force lists,
// the levels above it the field's); first_tile = INT_MAX: none.  order: the
// launch order of the tiles (tile_order_kernel), or null for index order.
struct M2LSrc2 {
  const int32_t *src;
  int ld, empty, first_tile;
  const int32_t *order;  // tile order within each XCD's range, or null
};

// Tile order of an M2L launch: each XCD takes a contiguous range of tiles
// (per = ceil(ntiles / 8)); within it, the tiles with the most non-empty
// offsets go first, so the launch does not end on a long tile started last.
// order[r0 + rank] = t, ranked by descending count, ties by index.
__global__ void tile_order_kernel(int ntiles, const int32_t *__restrict__ tn,
                                  int32_t *__restrict__ order) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const int per = (ntiles + 7) / 8;
  const int r0 = (t / per) * per, r1 = min(r0 + per, ntiles);
  const int nt = tn[t];
  int rank = 0;
  for (int u = r0; u < r1; ++u) {
    const int nu = tn[u];
    rank += (nu > nt) || (nu == nt && u < t);
  }
  order[r0 + rank] = t;
}

template <int KP, int NPART, bool OCTANT = false, bool FENCE = false>
__global__ __launch_bounds__(2 * KP) __attribute__((amdgpu_waves_per_eu(m2l_min_waves<KP, NPART>())))
void m2l_kernel(const void *__restrict__ Msrc, uint32_t msrc_bytes,
                                                          float *__restrict__ Lt,
                                                          const void *__restrict__ Tv,
                                                          const M2LTile *__restrict__ tiles,
                                                          int ntiles,
                                                          const int32_t *__restrict__ sorted_boxes,
                                                          const int32_t *__restrict__ src_slots,
                                                          int src_ld,
                                                          int empty_idx,
                                                          const int32_t *__restrict__ slot_table,
                                                          int slot_ld,
                                                          const int32_t *__restrict__ tile_list,
                                                          const int32_t *__restrict__ tile_nlist,
                                                          int list_ld, int accumulate,
                                                          SysMask mask, M2LSrc2 src2) {
  constexpr int NT = 2 * KP;
  constexpr int LDF = KP + 4;  // FP32 row pitch (floats)
  // BF16 row pitch (elements).  Pitches of KP + 8, + 16 and + 24 are all
  // bank-conflict-free or nearly so for these 16-byte reads; + 24 measured
  // 2.5 % faster for the two-part kernel (the default), whose LDS stays
  // within four workgroups per CU.  The three-part kernel keeps + 8.
  constexpr int LDB = NPART == 2 ? KP + 24 : KP + 8;
  constexpr int kLds = NPART == 0 ? kTileT * LDF * 4 : NPART * kTileT * LDB * 2;
  __shared__ __attribute__((aligned(16))) unsigned char s_raw[kLds];
  __shared__ int s_tgt[kTileT];
  const int per_xcd = (ntiles + 7) / 8;
  int tile_id = static_cast<int>(blockIdx.x % 8) * per_xcd + static_cast<int>(blockIdx.x / 8);
  if (tile_id >= ntiles) return;  // uniform over the workgroup, before any barrier
  if (src2.order) tile_id = src2.order[tile_id];
  if (tile_id >= src2.first_tile) {
    src_slots = src2.src;
    src_ld = src2.ld;
    empty_idx = src2.empty;
  }
  const M2LTile tile = tiles[tile_id];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  for (int j = tid; j < kTileT; j += NT) s_tgt[j] = j < tile.count ? sorted_boxes[tile.first + j] : -1;
  __syncthreads();
  if (mask.active) {
    // a tile whose target boxes all belong to converged systems is skipped
    int live = 0;
    for (int j = tid; j < tile.count; j += NT) live |= mask.active[mask.box_sys[s_tgt[j]]];
    if (!__syncthreads_or(live)) return;  // uniform over the workgroup
  }
  typedef float f4 __attribute__((ext_vector_type(4)));
  f4 acc[2][4];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = f4{0.f, 0.f, 0.f, 0.f};
  const int32_t *__restrict__ smap = slot_table + static_cast<int64_t>(tile.slots) * slot_ld;
  const int r16 = lane & 15;
  const int kq = lane >> 4;
  // the tile's non-empty offsets (plan build), walked in slot order
  const int32_t *__restrict__ tl = tile_list + static_cast<int64_t>(tile_id) * list_ld;
  const int na = tile_nlist[tile_id];
  using GA = M2LGather<KP, NPART>;
  // the target box of each of this thread's gather chunks: fixed for the tile,
  // so the source-slot loads of a later offset wait on no LDS read
  int tgt_of[GA::IT];
#pragma unroll
  for (int it = 0; it < GA::IT; ++it) tgt_of[it] = s_tgt[(tid + it * NT) / GA::PER_T];
  const __amdgpu_buffer_rsrc_t msrc_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<void *>(Msrc), 0, static_cast<int>(msrc_bytes), kBufferRsrcWord3);
  int src[GA::IT];
  float4 v[GA::IT];
  // Unconditional loads, for the reason given at m2l_mult_load: an empty
  // column of a partial tile, and every column of an offset past the tile's
  // last one, reads row empty_idx of the slot array, all -1 (no source: the
  // gather returns zeros without touching memory).
  auto src_load = [&](int slot, bool live) {
#pragma unroll
    for (int it = 0; it < GA::IT; ++it) {
      const int row = live && tgt_of[it] >= 0 ? tgt_of[it] : empty_idx;
      src[it] = src_slots[static_cast<int64_t>(row) * src_ld + slot];
    }
  };
  // Operator fragments are loaded one k step ahead of the MFMAs that use them,
  // and the first step of the next offset during the last step of this one,
  // so the MFMAs of a step never wait on its own fragment loads.
  constexpr int NK = NPART == 0 ? KP / 16 : KP / 32;  // k steps per offset
  constexpr int NA = NPART == 0 ? 1 : NPART;         // fragments per row block
  typedef typename std::conditional<NPART == 0, float4, bf16x8>::type AFrag;
  auto a_load = [&](const void *__restrict__ T, int ks, AFrag (&dst)[2][NA]) {
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      if constexpr (NPART == 0) {
        // FP32: one 16-byte load per row block covers the step's four k
        const float4 *__restrict__ A = static_cast<const float4 *>(T);
        dst[rb][0] = A[(ks * (KP / 16) + 2 * wave + rb) * 64 + lane];
      } else {
        const uint16_t *__restrict__ A = static_cast<const uint16_t *>(T);
#pragma unroll
        for (int p = 0; p < NPART; ++p)
          dst[rb][p] = *reinterpret_cast<const bf16x8 *>(
              A + ((static_cast<int64_t>(p * (KP / 32) + ks) * (KP / 16) + 2 * wave + rb) * 64 +
                   lane) * 8);
      }
    }
  };
  auto table_of = [&](int pc) -> const void * {
    const int tix = smap[tl[pc]];
    return NPART == 0
               ? static_cast<const void *>(static_cast<const float *>(Tv) + static_cast<int64_t>(tix) * KP * KP)
               : static_cast<const void *>(static_cast<const uint16_t *>(Tv) +
                                           static_cast<int64_t>(tix) * 3 * KP * KP);
  };
  // The loads for offsets past the tile's last one read that offset's
  // operator and no sources (results unused): no load of the loop sits under
  // a branch.
  AFrag ac[2][NA];
  if (na > 0) {
    src_load(tl[0], true);
    m2l_mult_load<KP, NPART>(msrc_rsrc, src, tid, v);
    src_load(tl[min(1, na - 1)], na > 1);
    a_load(table_of(0), 0, ac);
  }
  for (int pc = 0; pc < na; ++pc) {
    __syncthreads();  // the previous offset's operands are consumed
#pragma unroll
    for (int it = 0; it < GA::IT; ++it) {
      const int e = tid + it * NT;
      const int j = e / GA::PER_T, r = e % GA::PER_T;
      if constexpr (NPART == 0) {
        float *sF = reinterpret_cast<float *>(s_raw);
        *reinterpret_cast<float4 *>(&sF[j * LDF + 4 * r]) = v[it];
      } else {
        uint16_t *sB = reinterpret_cast<uint16_t *>(s_raw);
        const int p = r / (KP / 8), q8 = r % (KP / 8);
        *reinterpret_cast<float4 *>(&sB[(p * kTileT + j) * LDB + 8 * q8]) = v[it];
      }
    }
    // a plain barrier: the tile lists keep only offsets with a source, so no
    // block-wide "any source" vote (an LDS atomic per thread) is needed
    __syncthreads();
    m2l_mult_load<KP, NPART>(msrc_rsrc, src, tid, v);
    src_load(tl[min(pc + 2, na - 1)], pc + 2 < na);
    // FENCE: scheduling fences keep the prefetches where they are written
    // (the scheduler would otherwise sink them next to their uses).  Measured
    // faster for the force M2L's tiles and slower for the field's.
    if constexpr (FENCE) __builtin_amdgcn_sched_barrier(0);
    const void *Tcur = table_of(pc);
    const void *Tnext = table_of(min(pc + 1, na - 1));
#pragma unroll
    for (int ks = 0; ks < NK; ++ks) {
      AFrag an[2][NA];
      if (ks + 1 < NK) a_load(Tcur, ks + 1, an);
      else a_load(Tnext, 0, an);
      if constexpr (FENCE) __builtin_amdgcn_sched_barrier(0);
      if constexpr (NPART == 0) {
        const float *sF = reinterpret_cast<const float *>(s_raw);
        // steps j = 0..3 take k = 16 ks + 4 kq + j (the tables' fragment order)
        float4 bv[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
          bv[cb] = *reinterpret_cast<const float4 *>(&sF[(r16 + 16 * cb) * LDF + 16 * ks + 4 * kq]);
        const float4 a0 = ac[0][0], a1 = ac[1][0];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
          acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.x, bv[cb].x, acc[0][cb], 0, 0, 0);
          acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.x, bv[cb].x, acc[1][cb], 0, 0, 0);
          acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.y, bv[cb].y, acc[0][cb], 0, 0, 0);
          acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.y, bv[cb].y, acc[1][cb], 0, 0, 0);
          acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.z, bv[cb].z, acc[0][cb], 0, 0, 0);
          acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.z, bv[cb].z, acc[1][cb], 0, 0, 0);
          acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.w, bv[cb].w, acc[0][cb], 0, 0, 0);
          acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.w, bv[cb].w, acc[1][cb], 0, 0, 0);
        }
      } else {
        const uint16_t *sB = reinterpret_cast<const uint16_t *>(s_raw);
        const int kk = 32 * ks + 8 * kq;
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
          bf16x8 b[NPART];
#pragma unroll
          for (int p = 0; p < NPART; ++p)
            b[p] = *reinterpret_cast<const bf16x8 *>(&sB[(p * kTileT + 16 * cb + r16) * LDB + kk]);
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) {
            if constexpr (NPART == 3) {
              acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ac[rb][0], b[2], acc[rb][cb], 0, 0, 0);
              acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ac[rb][1], b[1], acc[rb][cb], 0, 0, 0);
              acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ac[rb][2], b[0], acc[rb][cb], 0, 0, 0);
            }
            if constexpr (NPART >= 2) {
              acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ac[rb][0], b[1], acc[rb][cb], 0, 0, 0);
              acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ac[rb][1], b[0], acc[rb][cb], 0, 0, 0);
            }
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ac[rb][0], b[0], acc[rb][cb], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int p = 0; p < NA; ++p) ac[rb][p] = an[rb][p];
    }
  }
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const int tgt = s_tgt[r16 + 16 * cb];
    if (tgt < 0) continue;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 32 * wave + 16 * rb + 4 * kq + r;
        float *o = Lt + static_cast<int64_t>(tgt) * KP + row;
        *o = accumulate ? *o + acc[rb][cb][r] : acc[rb][cb][r];
      }
  }
}

// ---------------------------------------------------------------------------
// Degree-ordered M2L.  The translation L_j = sum_n T_jn M_n couples local
// degree j with multipole degree n; its terms fall off like rho^(j + n)
// (rho = box radius over separation, below 1/2 for the interaction lists
// here).  These kernels sort the 16 x 32 MFMA blocks (row block rb: local
// dofs 16 rb .. 16 rb + 15; k step ks: multipole dofs 32 ks .. 32 ks + 31)
// by s, the sum of the lowest degrees in them.  At P = 10, 13 of the 32
// blocks have s <= P; every term of the other 19 has j + n > P, and together
// they are ~4e-5 of a product on the 1.15 M pool (measured against FP64).
//  * MF_M2L_BF16X3 (SPLIT): the blocks with s <= P - 3 (5 of 32 at P = 10)
//    take the three-product split of m2l_kernel, the others one product of
//    the leading BF16 parts.  The one-part rounding (~4e-3) of terms that
//    small stays below the order-P truncation: on the pool the field error
//    went from 3.466e-6 to 3.494e-6 and the force error from 3.195e-6 to
//    3.219e-6 against three products on every s <= P block; at s <= P - 5
//    the force error grew 1.8 %, at s <= P - 6 21 %.  42 instead of 96
//    MFMAs per offset and column block, 37 instead of 64 KB of operator
//    fragments.
//  * MF_M2L_BF16X1_CUT (!SPLIT): the blocks with s <= P in one product, the
//    others dropped: ~1e-4 relative per product, for the late products of a
//    relaxed solve, whose errors may grow as the residual falls.
// Dofs of degree n are [n^2, (n + 1)^2); padding dofs (>= (P + 1)^2) are zero
// in every table and expansion.
constexpr int dof_degree(int i) {
  int d = 0;
  while ((d + 1) * (d + 1) <= i) ++d;
  return d;
}

template <int KP, int P>
struct DegreeBlocks {
  static constexpr int NRB = KP / 16, NK = KP / 32, NW = KP / 32;
  static constexpr int ND = (P + 1) * (P + 1);
  // blocks that hold real dofs (the others are zero padding, skipped)
  static constexpr bool real(int rb, int ks) { return 16 * rb < ND && 32 * ks < ND; }
  static constexpr int min_sum(int rb, int ks) { return dof_degree(16 * rb) + dof_degree(32 * ks); }
  // kept by the cut: lowest degrees add up to at most min(P, 7).  The cut
  // serves the relaxed products, made once the residual is small, whose error
  // budget admits a translation of order 7 (5 of 32 blocks at P = 10)
  static constexpr int kCutSum = P < 7 ? P : 7;
  static constexpr bool low(int rb, int ks) { return real(rb, ks) && min_sum(rb, ks) <= kCutSum; }
  // three products in the split: lowest degrees add up to at most P - 3
  static constexpr bool low3(int rb, int ks) { return real(rb, ks) && min_sum(rb, ks) <= P - 3; }
  // MFMA products of row block rb per column block
  static constexpr int cost(int rb, bool split) {
    int c = 0;
    for (int k = 0; k < NK; ++k)
      c += split ? (low3(rb, k) ? 3 : real(rb, k) ? 1 : 0) : (low(rb, k) ? 1 : 0);
    return c;
  }
};

// One wave's share: its row blocks, the blocks it multiplies (k-step order)
// and the operator fragments it loads.  Row blocks go to waves longest
// first, each to the least loaded wave (P = 10, KP = 128: 12, 10, 10, 10
// products per column block with SPLIT, against 24 per wave for m2l_kernel;
// 2, 1, 1, 1 for the cut, against 8).
struct DegWave {
  int nr;  // row blocks owned (all written; zeros where nothing is multiplied)
  int rb[8];
  int nb;  // blocks
  int b_row[32], b_ks[32], b_f0[32], b_f1[32];  // b_f1 = -1: one product
  int nf;  // operator fragments (16 x 32 BF16, one part)
  int f_rb[64], f_ks[64], f_part[64];
  bool ks_used[4], ks_part1[4];  // k steps read from LDS, and with part 1
};

template <int KP, int P, bool SPLIT>
constexpr DegWave deg_wave(int w) {
  using C = DegreeBlocks<KP, P>;
  int owner[8] = {};
  int load[4] = {};
  bool placed[8] = {};
  for (int step = 0; step < C::NRB; ++step) {
    int best = -1;
    for (int rb = 0; rb < C::NRB; ++rb)
      if (!placed[rb] && (best < 0 || C::cost(rb, SPLIT) > C::cost(best, SPLIT))) best = rb;
    int least = 0;
    for (int v = 1; v < C::NW; ++v)
      if (load[v] < load[least]) least = v;
    placed[best] = true;
    owner[best] = least;
    load[least] += C::cost(best, SPLIT);
  }
  DegWave s{};
  for (int rb = 0; rb < C::NRB; ++rb)
    if (owner[rb] == w) s.rb[s.nr++] = rb;
  for (int ks = 0; ks < C::NK; ++ks)
    for (int i = 0; i < s.nr; ++i) {
      const bool lo = SPLIT ? C::low3(s.rb[i], ks) : C::low(s.rb[i], ks);
      if (!lo && !(SPLIT && C::real(s.rb[i], ks))) continue;
      const int b = s.nb++;
      s.b_row[b] = i;
      s.b_ks[b] = ks;
      s.f_rb[s.nf] = s.rb[i];
      s.f_ks[s.nf] = ks;
      s.f_part[s.nf] = 0;
      s.b_f0[b] = s.nf++;
      s.b_f1[b] = -1;
      if (lo && SPLIT) {
        s.f_rb[s.nf] = s.rb[i];
        s.f_ks[s.nf] = ks;
        s.f_part[s.nf] = 1;
        s.b_f1[b] = s.nf++;
        s.ks_part1[ks] = true;
      }
      s.ks_used[ks] = true;
    }
  return s;
}

// The offset loop of wave W (the waves differ only in which fragments they
// load and multiply).  Same tiles, slot lists, gather and pipeline as
// m2l_kernel; the wave's operator fragments for the next offset are loaded
// whole while the current offset multiplies.
// Multipole columns the degree kernel reads: through the last k step that any
// wave multiplies (the cut at P = 10 reads dofs 0 .. 63 only).
// Part-1 columns of the split (the three-product blocks' k steps).
template <int KP, int P, bool SPLIT>
constexpr int deg_gather_width1() {
  using C = DegreeBlocks<KP, P>;
  if (!SPLIT) return KP;
  int last = 0;
  for (int ks = 0; ks < C::NK; ++ks)
    for (int rb = 0; rb < C::NRB; ++rb)
      if (C::low3(rb, ks)) last = ks;
  const int kg1 = 32 * (last + 1);
  // the whole width where the chunks would not split evenly over the threads
  return (64 * ((KP + kg1) / 8)) % (2 * KP) == 0 ? kg1 : KP;
}

template <int KP, int P, bool SPLIT>
constexpr int deg_gather_width() {
  using C = DegreeBlocks<KP, P>;
  int last = 0;
  for (int ks = 0; ks < C::NK; ++ks)
    for (int rb = 0; rb < C::NRB; ++rb)
      if (SPLIT ? C::real(rb, ks) : C::low(rb, ks)) last = ks;
  const int kg = 32 * (last + 1), npg = SPLIT ? 2 : 1;
  // the whole width where a prefix would not split evenly over the threads
  return (64 * npg * (kg / 8)) % (2 * KP) == 0 ? kg : KP;
}

template <int KP, int P, bool SPLIT, int W>
__device__ __forceinline__ void m2l_deg_wave(
    __amdgpu_buffer_rsrc_t msrc_rsrc, float *__restrict__ Lt, const uint16_t *__restrict__ Tb,
    const int32_t *__restrict__ src_slots, int src_ld, int empty_idx,
    const int32_t *__restrict__ smap, const int32_t *__restrict__ tl, int na, uint16_t *sB,
    const int *s_tgt, int tid) {
  constexpr DegWave S = deg_wave<KP, P, SPLIT>(W);
  constexpr int NPG = SPLIT ? 2 : 1;  // multipole parts gathered
  constexpr int KG = deg_gather_width<KP, P, SPLIT>();  // multipole columns gathered
  constexpr int KG1 = deg_gather_width1<KP, P, SPLIT>();  // of part 1 (split)
  constexpr int NT = 2 * KP, LDB = SPLIT ? KP + 24 : KP + 8;
  constexpr int NR = S.nr > 0 ? S.nr : 1, NF = S.nf > 0 ? S.nf : 1;
  using GA = M2LGather<KP, NPG, KG, KG1>;
  typedef float f4 __attribute__((ext_vector_type(4)));
  const int lane = tid & 63, r16 = lane & 15, kq = lane >> 4;
  f4 acc[NR][4];
#pragma unroll
  for (int i = 0; i < NR; ++i)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[i][cb] = f4{0.f, 0.f, 0.f, 0.f};
  int tgt_of[GA::IT];
#pragma unroll
  for (int it = 0; it < GA::IT; ++it) tgt_of[it] = s_tgt[(tid + it * NT) / GA::PER_T];
  int src[GA::IT];
  float4 v[GA::IT];
  auto src_load = [&](int slot, bool live) {
#pragma unroll
    for (int it = 0; it < GA::IT; ++it) {
      const int row = live && tgt_of[it] >= 0 ? tgt_of[it] : empty_idx;
      src[it] = src_slots[static_cast<int64_t>(row) * src_ld + slot];
    }
  };
  auto table_of = [&](int pc) -> const uint16_t * {
    return Tb + static_cast<int64_t>(smap[tl[pc]]) * 3 * KP * KP;
  };
  auto a_load = [&](const uint16_t *__restrict__ A, bf16x8 (&dst)[NF]) {
#pragma unroll
    for (int f = 0; f < S.nf; ++f)
      dst[f] = *reinterpret_cast<const bf16x8 *>(
          A + (((static_cast<int64_t>(S.f_part[f]) * (KP / 32) + S.f_ks[f]) * (KP / 16) + S.f_rb[f]) *
                   64 + lane) * 8);
  };
  // Operator fragments in two sets that swap roles every offset (the loop
  // takes offsets in pairs): a copy from the prefetch set into the working
  // set would make the compiler wait for the prefetch, and with it for every
  // load issued before it, inside the iteration that issued it.  An odd count
  // runs one more offset past the end, whose gathers read no sources (zeros).
  bf16x8 fa[NF], fb[NF];
  if (na > 0) {
    src_load(tl[0], true);
    m2l_mult_load<KP, NPG, KG, KG1>(msrc_rsrc, src, tid, v);
    src_load(tl[min(1, na - 1)], na > 1);
    a_load(table_of(0), fa);
  }
  auto offset_step = [&](int pc, const bf16x8 (&cur)[NF], bf16x8 (&nxt)[NF]) {
    __syncthreads();  // the previous offset's multipoles are consumed
#pragma unroll
    for (int it = 0; it < GA::IT; ++it) {
      const int e = tid + it * NT;
      const int j = e / GA::PER_T, q = e % GA::PER_T;
      const int p = q < KG / 8 ? 0 : 1 + (q - KG / 8) / (KG1 / 8);
      const int q8 = q < KG / 8 ? q : (q - KG / 8) % (KG1 / 8);
      *reinterpret_cast<float4 *>(&sB[(p * kTileT + j) * LDB + 8 * q8]) = v[it];
    }
    __syncthreads();
    m2l_mult_load<KP, NPG, KG, KG1>(msrc_rsrc, src, tid, v);
    src_load(tl[min(pc + 2, na - 1)], pc + 2 < na);
    a_load(table_of(min(pc + 1, na - 1)), nxt);
    // the scheduler would otherwise move these prefetches next to their
    // uses in the next step
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ks = 0; ks < DegreeBlocks<KP, P>::NK; ++ks) {
      if (!S.ks_used[ks]) continue;
      bf16x8 b0[4], b1[4];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        b0[cb] = *reinterpret_cast<const bf16x8 *>(&sB[(16 * cb + r16) * LDB + 32 * ks + 8 * kq]);
        if (S.ks_part1[ks])
          b1[cb] = *reinterpret_cast<const bf16x8 *>(
              &sB[(kTileT + 16 * cb + r16) * LDB + 32 * ks + 8 * kq]);
      }
#pragma unroll
      for (int b = 0; b < S.nb; ++b) {
        if (S.b_ks[b] != ks) continue;
        const int i = S.b_row[b];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
          if (S.b_f1[b] >= 0) {
            acc[i][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[S.b_f0[b]], b1[cb], acc[i][cb], 0, 0, 0);
            acc[i][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[S.b_f1[b]], b0[cb], acc[i][cb], 0, 0, 0);
          }
          acc[i][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[S.b_f0[b]], b0[cb], acc[i][cb], 0, 0, 0);
        }
      }
    }
  };
  for (int pc = 0; pc < na; pc += 2) {
    offset_step(pc, fa, fb);
    offset_step(pc + 1, fb, fa);
  }
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const int tgt = s_tgt[r16 + 16 * cb];
    if (tgt < 0) continue;
#pragma unroll
    for (int i = 0; i < S.nr; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        Lt[static_cast<int64_t>(tgt) * KP + 16 * S.rb[i] + 4 * kq + r] = acc[i][cb][r];
  }
}

// Two waves per SIMD for the split (two operator sets of up to 10
// fragments at P = 10), three for the cut.
template <int KP, int P, bool SPLIT>
__global__ __launch_bounds__(2 * KP) __attribute__((amdgpu_waves_per_eu(SPLIT ? 2 : 3)))
void m2l_deg_kernel(const void *__restrict__ Msrc, uint32_t msrc_bytes, float *__restrict__ Lt,
                    const uint16_t *__restrict__ Tb, const M2LTile *__restrict__ tiles, int ntiles,
                    const int32_t *__restrict__ sorted_boxes, const int32_t *__restrict__ src_slots,
                    int src_ld, int empty_idx, const int32_t *__restrict__ slot_table, int slot_ld,
                    const int32_t *__restrict__ tile_list, const int32_t *__restrict__ tile_nlist,
                    int list_ld, SysMask mask, M2LSrc2 src2) {
  constexpr int NT = 2 * KP, LDB = SPLIT ? KP + 24 : KP + 8;
  __shared__ __attribute__((aligned(16))) uint16_t sB[(SPLIT ? 2 : 1) * kTileT * LDB];
  __shared__ int s_tgt[kTileT];
  const int per_xcd = (ntiles + 7) / 8;
  int tile_id = static_cast<int>(blockIdx.x % 8) * per_xcd + static_cast<int>(blockIdx.x / 8);
  if (tile_id >= ntiles) return;  // uniform over the workgroup, before any barrier
  if (src2.order) tile_id = src2.order[tile_id];
  if (tile_id >= src2.first_tile) {
    src_slots = src2.src;
    src_ld = src2.ld;
    empty_idx = src2.empty;
  }
  const M2LTile tile = tiles[tile_id];
  const int tid = threadIdx.x;
  for (int j = tid; j < kTileT; j += NT) s_tgt[j] = j < tile.count ? sorted_boxes[tile.first + j] : -1;
  __syncthreads();
  if (mask.active) {
    int live = 0;
    for (int j = tid; j < tile.count; j += NT) live |= mask.active[mask.box_sys[s_tgt[j]]];
    if (!__syncthreads_or(live)) return;  // uniform over the workgroup
  }
  const int32_t *__restrict__ smap = slot_table + static_cast<int64_t>(tile.slots) * slot_ld;
  const int32_t *__restrict__ tl = tile_list + static_cast<int64_t>(tile_id) * list_ld;
  const int na = tile_nlist[tile_id];
  const __amdgpu_buffer_rsrc_t msrc_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<void *>(Msrc), 0, static_cast<int>(msrc_bytes), kBufferRsrcWord3);
  // the wave index as a scalar: each wave runs its own loop (no branch joins
  // inside a loop, so the compiler's load waits stay per load)
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
#define MF_DEG_ARGS msrc_rsrc, Lt, Tb, src_slots, src_ld, empty_idx, smap, tl, na, sB, s_tgt, tid
  if (wave == 0) m2l_deg_wave<KP, P, SPLIT, 0>(MF_DEG_ARGS);
  else if constexpr (KP / 32 > 1) {
    if (wave == 1) m2l_deg_wave<KP, P, SPLIT, 1>(MF_DEG_ARGS);
    else if constexpr (KP / 32 > 2) {
      if (wave == 2) m2l_deg_wave<KP, P, SPLIT, 2>(MF_DEG_ARGS);
      else if constexpr (KP / 32 > 3) m2l_deg_wave<KP, P, SPLIT, 3>(MF_DEG_ARGS);
    }
  }
#undef MF_DEG_ARGS
}

// Non-empty offsets of every M2L tile: list[tile][0 .. nlist[tile]) = the
// slots (in order) where at least one target of the tile has a source box.
// One workgroup per tile; nslot <= 1024.
__global__ __launch_bounds__(256) void tile_slots_kernel(const M2LTile *__restrict__ tiles,
                                                         const int32_t *__restrict__ sorted_boxes,
                                                         const int32_t *__restrict__ src_slots,
                                                         int src_ld,
                                                         const int32_t *__restrict__ slot_table,
                                                         int slot_ld, int list_ld,
                                                         int32_t *__restrict__ list,
                                                         int32_t *__restrict__ nlist) {
  __shared__ int s_flag[1024];
  const M2LTile tile = tiles[blockIdx.x];
  const int32_t *__restrict__ smap = slot_table + static_cast<int64_t>(tile.slots) * slot_ld;
  for (int k = threadIdx.x; k < tile.nslot; k += 256) s_flag[k] = 0;
  __syncthreads();
  for (int j = 0; j < tile.count; ++j) {
    const int tgt = sorted_boxes[tile.first + j];
    for (int k = threadIdx.x; k < tile.nslot; k += 256)
      if (src_slots[static_cast<int64_t>(tgt) * src_ld + k] >= 0) s_flag[k] = 1;
  }
  __syncthreads();
  // Compaction in slot order, 256 slots per pass: a slot's position is the
  // count of flagged slots before it (earlier passes, earlier waves, lower
  // lanes of its own wave).
  __shared__ int s_wave[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t *__restrict__ out = list + static_cast<int64_t>(blockIdx.x) * list_ld;
  int base = 0;
  for (int k0 = 0; k0 < tile.nslot; k0 += 256) {
    const int k = k0 + static_cast<int>(threadIdx.x);
    const bool f = k < tile.nslot && s_flag[k] != 0 && smap[k] >= 0;
    const uint64_t m = __ballot(f);
    if (lane == 0) s_wave[w] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int i = 0; i < w; ++i) off += s_wave[i];
    if (f) out[off + __popcll(m & ((1ULL << lane) - 1ULL))] = k;
    base += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();  // s_wave is rewritten by the next pass
  }
  if (threadIdx.x == 0) nlist[blockIdx.x] = base;
}

// box_sys[g] = system of box g of one level: the bits above the level's 3l
// Morton bits of its key.
__global__ void box_sys_kernel(int nbox, const uint64_t *__restrict__ keys, int bits,
                               int32_t *__restrict__ box_sys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbox) box_sys[i] = static_cast<int32_t>(keys[i] >> bits);
}

// child_tab[parent * 8 + octant] = global index of the child (one level).
__global__ void child_table_kernel(int nchild, int child_base, const int32_t *__restrict__ parent,
                                   const uint64_t *__restrict__ keys, int32_t *__restrict__ child_tab) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchild) return;
  child_tab[static_cast<int64_t>(parent[c]) * 8 + static_cast<int>(keys[c] & 7ULL)] = child_base + c;
}

__global__ void iota_kernel(int n, int32_t *__restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

// ---------------------------------------------------------------------------
// L2P helpers: far field and force tensor from a scaled local expansion Lt of
// a box with centre c and side s, at x_hat = (x - c)/s.  Lt is read through
// uniform (scalar) loads.

template <int P>
__device__ __forceinline__ void l2p_grad(const float *__restrict__ Lt, float x, float y, float z,
                                         float &gx, float &gy, float &gz) {
  cplx R[P * (P + 1) / 2];
  regular_tri<P - 1>(x, y, z, R);
  float sz = 0.f, spr = 0.f, spi = 0.f;
#pragma unroll
  for (int n = 1; n <= P; ++n) {
    // m = 0
    {
      const float L0 = Lt[dof_re(n, 0)];
      const cplx r0 = rget<P - 1>(R, n - 1, 0);
      const cplx r1 = rget<P - 1>(R, n - 1, 1);
      sz += L0 * r0.re;
      spr += L0 * r1.re;
      spi += L0 * r1.im;
    }
#pragma unroll
    for (int m = 1; m <= n; ++m) {
      const float Lr = Lt[dof_re(n, m)], Li = Lt[dof_im(n, m)];
      const cplx a = rget<P - 1>(R, n - 1, m);       // R_{n-1}^m
      const cplx bp = rget<P - 1>(R, n - 1, m + 1);  // R_{n-1}^{m+1}
      const cplx bm = rget<P - 1>(R, n - 1, m - 1);  // R_{n-1}^{m-1}
      // dz: 2 Re(L a)
      sz += 2.f * (Lr * a.re - Li * a.im);
      // (dx + i dy): L bp - conj(L bm)
      const float pr = Lr * bp.re - Li * bp.im, pi = Lr * bp.im + Li * bp.re;
      const float qr = Lr * bm.re - Li * bm.im, qi = Lr * bm.im + Li * bm.re;
      spr += pr - qr;
      spi += pi + qi;
    }
  }
  gx = spr;
  gy = spi;
  gz = sz;
}

template <int P>
__device__ __forceinline__ void l2p_hess(const float *__restrict__ Lt, float x, float y, float z,
                                         float H[6]) {
  cplx R[(P - 1) * P / 2];
  regular_tri<P - 2>(x, y, z, R);
  float A = 0.f, Br = 0.f, Bi = 0.f, Cr = 0.f, Ci = 0.f;
#pragma unroll
  for (int n = 2; n <= P; ++n) {
    {
      const float L0 = Lt[dof_re(n, 0)];
      const cplx a = rget<P - 2>(R, n - 2, 0), b = rget<P - 2>(R, n - 2, 1), c = rget<P - 2>(R, n - 2, 2);
      A += L0 * a.re;
      Br += L0 * b.re;
      Bi += L0 * b.im;
      Cr += L0 * c.re;
      Ci += L0 * c.im;
    }
#pragma unroll
    for (int m = 1; m <= n; ++m) {
      const float Lr = Lt[dof_re(n, m)], Li = Lt[dof_im(n, m)];
      const cplx a = rget<P - 2>(R, n - 2, m);
      A += 2.f * (Lr * a.re - Li * a.im);
      // B = sum_{m>=0} L R^{m+1} - sum_{m>=1} conj(L R^{m-1})
      const cplx bp = rget<P - 2>(R, n - 2, m + 1), bm = rget<P - 2>(R, n - 2, m - 1);
      Br += (Lr * bp.re - Li * bp.im) - (Lr * bm.re - Li * bm.im);
      Bi += (Lr * bp.im + Li * bp.re) + (Lr * bm.im + Li * bm.re);
      // C = sum_{m>=0} L R^{m+2} - conj(L_1) R^1 + sum_{m>=2} conj(L R^{m-2})
      const cplx cp = rget<P - 2>(R, n - 2, m + 2);
      Cr += Lr * cp.re - Li * cp.im;
      Ci += Lr * cp.im + Li * cp.re;
      if (m == 1) {
        const cplx r1 = rget<P - 2>(R, n - 2, 1);
        // - conj(L) R1 = -(Lr - i Li)(r1.re + i r1.im)
        Cr -= Lr * r1.re + Li * r1.im;
        Ci -= Lr * r1.im - Li * r1.re;
      } else {
        const cplx cm = rget<P - 2>(R, n - 2, m - 2);
        Cr += Lr * cm.re - Li * cm.im;
        Ci -= Lr * cm.im + Li * cm.re;
      }
    }
  }
  // Hessian of psi_hat: xx yy zz xy xz yz
  H[0] = 0.5f * (-A + Cr);
  H[1] = 0.5f * (-A - Cr);
  H[2] = A;
  H[3] = 0.5f * Ci;
  H[4] = Br;
  H[5] = Bi;
}

// ---------------------------------------------------------------------------
// Leaf kernels.  The far field (L2P) and the near field (P2P) are separate
// launches: L2P writes H_const + far field (or the far force + external
// tensor term) in the caller's particle order, then the near-field kernel adds
// its sums.  Apart, the near-field kernel needs a third of the registers the
// solid harmonics of L2P take, and runs at twice the occupancy.

// Near ranges of a leaf: the particle ranges of its near boxes, with boxes of
// consecutive index merged (leaf b + 1 holds the particles right after leaf
// b's, so a run of consecutive leaf indices is one contiguous range).  One
// wave per leaf: the slot entries, keyed (class << 28 | box) with slots
// [0, nmix) class 0 and [nmix, nslot) class 1, are bitonic-sorted in LDS;
// runs of consecutive keys are the ranges, emitted in key order (class 0
// first), so the summation order is fixed.  rng[leaf * nslot + r] =
// {first, end}; nr[leaf] = {class-0 ranges, class-1 ranges}.  kSortN, the
// sort size, is a power of two >= nslot (128 for the field's 81 slots, 256
// for the force's 179).
template <int kSortN>
__global__ __launch_bounds__(64) void near_ranges_kernel(int nleaf, const int32_t *__restrict__ list,
                                                         int nslot, int nmix,
                                                         const int32_t *__restrict__ first,
                                                         const int32_t *__restrict__ count,
                                                         int2 *__restrict__ rng,
                                                         int2 *__restrict__ nr) {
  __shared__ int s_key[kSortN];
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  const int32_t *__restrict__ l = list + static_cast<int64_t>(b) * nslot;
  for (int k = lane; k < kSortN; k += 64) {
    int key = 0x7fffffff;
    if (k < nslot) {
      const int sb = l[k];
      if (sb >= 0) key = ((k < nmix ? 0 : 1) << 28) | sb;
    }
    s_key[k] = key;
  }
  __syncthreads();
  // bitonic sort, ascending: kSortN / 2 compare-exchanges per step, 2 per lane
  for (int size = 2; size <= kSortN; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int c = lane; c < kSortN / 2; c += 64) {
        const int i = 2 * c - (c & (stride - 1));
        const int j = i + stride;
        const bool up = (i & size) == 0;
        const int a = s_key[i], d = s_key[j];
        if ((a > d) == up) {
          s_key[i] = d;
          s_key[j] = a;
        }
      }
      __syncthreads();
    }
  int2 *__restrict__ out = rng + static_cast<int64_t>(b) * nslot;
  int pos = 0, n0 = 0;
  for (int k0 = 0; k0 < kSortN; k0 += 64) {
    const int k = k0 + lane;
    const int key = s_key[k];
    const bool valid = key != 0x7fffffff;
    const bool is_start = valid && (k == 0 || s_key[k - 1] != key - 1);
    int j0 = 0, j1 = 0;
    if (is_start) {
      int e = k;
      while (e + 1 < kSortN && s_key[e + 1] == s_key[e] + 1) ++e;
      const int sb = key & 0x0fffffff, se = s_key[e] & 0x0fffffff;
      j0 = first[sb];
      j1 = first[se] + count[se];
    }
    const uint64_t mask = __ballot(is_start);
    const uint64_t cls0 = __ballot(is_start && (key >> 28) == 0);
    const int before = __popcll(mask & ((1ULL << lane) - 1ULL));
    if (is_start) out[pos + before] = make_int2(j0, j1);
    pos += __popcll(mask);
    n0 += __popcll(cls0);
  }
  if (lane == 0) nr[b] = make_int2(n0, pos - n0);
}

// out += far, per particle in caller order: the far part (L2P, and for the
// force the external-tensor term) was written in FP64 on the context stream
// while the near field ran beside it; the sum rounds once, as a single
// kernel adding FP64 terms would.  Systems that have converged (batch plans)
// are skipped, as in the other product kernels.
__global__ void add_far_kernel(int64_t n, const double *__restrict__ far3, float *__restrict__ ox,
                               float *__restrict__ oy, float *__restrict__ oz,
                               const int32_t *__restrict__ sys_of, const int32_t *__restrict__ active) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (active != nullptr && active[sys_of[i]] == 0) return;
  ox[i] = static_cast<float>(static_cast<double>(ox[i]) + far3[i]);
  oy[i] = static_cast<float>(static_cast<double>(oy[i]) + far3[n + i]);
  oz[i] = static_cast<float>(static_cast<double>(oz[i]) + far3[2 * n + i]);
}

// L2P, field: far3[.][perm] = far field (FP64; add_far_kernel adds it to
// H_const + near).  One wave per leaf.
template <int P, int KP>
__global__ __launch_bounds__(64) void l2p_field_kernel(
    int nleaf, const int32_t *__restrict__ start, const int32_t *__restrict__ count,
    const uint64_t *__restrict__ keys, const SrcRec *__restrict__ rec,
    const int32_t *__restrict__ perm, const float *__restrict__ Lt, Frame fr, double s,
    int64_t n, double *__restrict__ far3, SysMask mask) {
  const int b = blockIdx.x;
  if (sys_skip(mask, b)) return;
  const int lane = threadIdx.x;
  double cx, cy, cz;
  box_centre(fr, keys[b], s, cx, cy, cz);
  const float inv_s = static_cast<float>(1.0 / s);
  const double far_scale = -1.0 / (4.0 * kPi) / (s * s);
  const int s0 = start[b], cnt = count[b];
  const float *__restrict__ L = Lt + static_cast<int64_t>(b) * KP;
  for (int t0 = 0; t0 < cnt; t0 += 64) {
    const int t = t0 + lane;
    const SrcRec me = rec[s0 + (t < cnt ? t : cnt - 1)];
    float gx, gy, gz;
    l2p_grad<P>(L, static_cast<float>(me.x - cx) * inv_s, static_cast<float>(me.y - cy) * inv_s,
                static_cast<float>(me.z - cz) * inv_s, gx, gy, gz);
    if (t < cnt) {
      const int o = perm[s0 + t];
      far3[o] = far_scale * gx;
      far3[n + o] = far_scale * gy;
      far3[2 * n + o] = far_scale * gz;
    }
  }
}

// Near-field source loops over one contiguous source range, taking every
// STRIDE-th source from j0.  With STRIDE = 1 the source index is uniform
// over the wave and the records come through scalar loads.
template <int STRIDE>
__device__ __forceinline__ void field_range_far(const SrcRec *__restrict__ rec, int j0, int j1,
                                                float tx, float ty, float tz, float &a1x,
                                                float &a1y, float &a1z, float &a2x, float &a2y,
                                                float &a2z) {
#pragma unroll 2
  for (int j = j0; j < j1; j += STRIDE) {
    const float sx = rec[j].x, sy = rec[j].y, sz = rec[j].z;
    const float mx = rec[j].mx, my = rec[j].my, mz = rec[j].mz;
    const float dx = tx - sx, dy = ty - sy, dz = tz - sz;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float ir2 = ir * ir;
    const float ir3 = ir2 * ir;
    const float dm = fmaf(dx, mx, fmaf(dy, my, dz * mz));
    const float s1 = dm * (ir3 * ir2);
    a1x = fmaf(dx, s1, a1x);
    a1y = fmaf(dy, s1, a1y);
    a1z = fmaf(dz, s1, a1z);
    a2x = fmaf(mx, ir3, a2x);
    a2y = fmaf(my, ir3, a2y);
    a2z = fmaf(mz, ir3, a2z);
  }
}

template <int STRIDE>
__device__ __forceinline__ void field_range_mixed(const SrcRec *__restrict__ rec, int j0, int j1,
                                                  float tx, float ty, float tz,
                                                  const FieldConsts &fc, const float4 *tab,
                                                  float &a1x, float &a1y, float &a1z, float &a2x,
                                                  float &a2y, float &a2z) {
#pragma unroll 2
  for (int j = j0; j < j1; j += STRIDE) {
    const float sx = rec[j].x, sy = rec[j].y, sz = rec[j].z;
    const float mx = rec[j].mx, my = rec[j].my, mz = rec[j].mz;
    const float dx = tx - sx, dy = ty - sy, dz = tz - sz;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float ir2 = ir * ir;
    float beta = ir2 * ir;
    float alpha = beta * ir2;
    if (r2 < fc.near2) field_near_coeffs_tab(tab, fc.tab_scale, r2, ir, alpha, beta);
    const float dm = fmaf(dx, mx, fmaf(dy, my, dz * mz));
    const float s1 = dm * alpha;
    a1x = fmaf(dx, s1, a1x);
    a1y = fmaf(dy, s1, a1y);
    a1z = fmaf(dz, s1, a1z);
    a2x = fmaf(mx, beta, a2x);
    a2y = fmaf(my, beta, a2y);
    a2z = fmaf(mz, beta, a2z);
  }
}

// Targets [tb, te) of a leaf of cnt particles taken by the full-chunk path.
template <int T>
__device__ __forceinline__ void p2p_target_span(int cnt, int &tb, int &te) {
  const int full = cnt & ~63, r = cnt & 63;
  if (T == 64) {
    tb = 0;
    te = r > 32 ? cnt : full;
  } else {
    tb = full;
    te = r <= 32 ? cnt : full;
  }
}

constexpr int kLeafLists = 7;  // leaf shapes: full chunks (64) and remainders (32 .. 1)

// Stream offsets of a leaf's near ranges, concatenated: pre[r] = total
// length of ranges 0 .. r-1 (pre[nrange] = the stream's length), beg[r] =
// first record of range r.  One wave.
__device__ __forceinline__ void range_prefix(int *pre, int *beg, const int2 *__restrict__ rg,
                                             int nrange, int lane) {
  int carry = 0;
  for (int c = 0; c < nrange; c += 64) {
    const int r = c + lane;
    int len = 0;
    if (r < nrange) {
      const int2 q = rg[r];
      len = q.y - q.x;
      beg[r] = q.x;
    }
    int incl = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    if (r < nrange) pre[r] = carry + incl - len;
    carry += __shfl(incl, 63, 64);
  }
  if (lane == 0) pre[nrange] = carry;
}

// The range holding stream position i: the last r with pre[r] <= i (an
// empty range shares its successor's offset, so "last" skips it).
__device__ __forceinline__ int range_of(const int *pre, int nrange, int i) {
  int lo = 0, hi = nrange - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Sum over the 64 lanes of a wave.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One lane's share of a leaf's near-source stream, in blocks: block i0
// holds positions i0, i0 + 64, ..., i0 + 64 (kLaneBlock - 1) (consecutive
// lanes read consecutive records), and the lane loads all of them before
// evaluating any, so that kLaneBlock loads are in flight at once.  Two per
// block and four targets per pass keep the fused kernels at 64 registers
// or fewer (eight waves per SIMD for the full chunks), measured faster than
// larger blocks or passes.  The loads
// are issued without a branch (a load under a branch makes the compiler wait
// for every outstanding load at its join): a position past the stream's end
// re-reads the last record and becomes a zero moment 1e18 away, whose pair
// terms are all exactly zero.
constexpr int kLaneBlock = 2;
struct LaneStream {
  const float4 *__restrict__ rec4;
  const int *pre;
  const int *beg;
  int total, nmix;
  int lo, nxt, base;
  float4 pa[kLaneBlock], pb[kLaneBlock];  // records of the block
  bool mixed[kLaneBlock];  // their ranges may hold pairs inside the smoothing support
  __device__ __forceinline__ void start(int lane, int nrange) {
    lo = range_of(pre, nrange, min(lane, total - 1));
    nxt = pre[lo + 1];
    base = beg[lo] - pre[lo];
  }
  __device__ __forceinline__ void fetch(int i0) {
#pragma unroll
    for (int k = 0; k < kLaneBlock; ++k) {
      const int i = i0 + 64 * k;
      const int ic = min(i, total - 1);
      while (nxt <= ic) {
        ++lo;
        nxt = pre[lo + 1];
        base = beg[lo] - pre[lo];
      }
      const int64_t j = static_cast<int64_t>(base) + ic;
      const float4 a = rec4[2 * j], m = rec4[2 * j + 1];
      const bool live = i < total;
      const float away = live ? 0.f : 1e18f, w = live ? 1.f : 0.f;
      pa[k] = make_float4(a.x + away, a.y + away, a.z + away, 0.f);
      pb[k] = make_float4(m.x * w, m.y * w, m.z * w, 0.f);
      mixed[k] = live && lo < nmix;
    }
  }
};

// Leaf shapes for the near-field kernels.  List 0 holds every leaf with a
// full chunk of 64 or a remainder of more than 32 (the one-target-per-lane
// kernels); lists 1..6 hold the leaves whose remainder (cnt mod 64) is
// 17..32, 9..16, 5..8, 3..4, 2 or 1 (the lanes-over-sources kernels, each
// list with the pass size that fits it).  A leaf can be in list 0 and in one
// of the others: they cover disjoint targets.  Two (list, leaf) entries per leaf (list kLeafLists =
// none); a stable sort by list then gives every list in leaf (Morton) order.
__global__ void classify_leaves_kernel(int nleaf, const int32_t *__restrict__ count,
                                       uint64_t *__restrict__ key, int32_t *__restrict__ leaf,
                                       int32_t *__restrict__ sizes) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nleaf) return;
  const int c = count[b];
  const int r = c & 63;
  const int k0 = (c >= 64 || r > 32) ? 0 : kLeafLists;
  const int k1 = (r == 0 || r > 32) ? kLeafLists
                 : r > 16 ? 1 : r > 8 ? 2 : r > 4 ? 3 : r > 2 ? 4 : r > 1 ? 5 : 6;
  key[2 * b] = static_cast<uint64_t>(k0);
  key[2 * b + 1] = static_cast<uint64_t>(k1);
  leaf[2 * b] = b;
  leaf[2 * b + 1] = b;
  // one atomic per wave and list instead of one per leaf
#pragma unroll
  for (int k = 0; k < kLeafLists; ++k) {
    const uint64_t m = __ballot((k == 0 && k0 == 0) || k1 == k);
    if (m != 0 && (threadIdx.x & 63) == __ffsll(static_cast<long long>(m)) - 1)
      atomicAdd(&sizes[k], __popcll(m));
  }
}

struct Text9 {
  const float *p[9];
};

// L2P, force: far3[.][perm] = (far tensor + T_ext) . m (FP64;
// add_far_kernel adds it to the near force).  One wave per leaf.
template <int P, int KP>
__global__ __launch_bounds__(64) void l2p_force_kernel(
    int nleaf, const int32_t *__restrict__ start, const int32_t *__restrict__ count,
    const uint64_t *__restrict__ keys, const SrcRec *__restrict__ rec,
    const int32_t *__restrict__ perm, const float *__restrict__ Lt, int use_far, Frame fr,
    double s, Text9 text, int has_text, int64_t n, double *__restrict__ far3) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  double cx, cy, cz;
  box_centre(fr, keys[b], s, cx, cy, cz);
  const float inv_s = static_cast<float>(1.0 / s);
  const double far_scale = -kMu0 / (4.0 * kPi) / (s * s * s);
  const int s0 = start[b], cnt = count[b];
  const float *__restrict__ L = Lt + static_cast<int64_t>(b) * KP;
  for (int t0 = 0; t0 < cnt; t0 += 64) {
    const int t = t0 + lane;
    const SrcRec me = rec[s0 + (t < cnt ? t : cnt - 1)];
    const float nx = me.mx, ny = me.my, nz = me.mz;
    double F0 = 0.0, F1 = 0.0, F2 = 0.0;
    if (use_far) {
      float Hs[6];
      l2p_hess<P>(L, static_cast<float>(me.x - cx) * inv_s, static_cast<float>(me.y - cy) * inv_s,
                  static_cast<float>(me.z - cz) * inv_s, Hs);
      F0 = far_scale * (static_cast<double>(Hs[0]) * nx + static_cast<double>(Hs[3]) * ny +
                        static_cast<double>(Hs[4]) * nz);
      F1 = far_scale * (static_cast<double>(Hs[3]) * nx + static_cast<double>(Hs[1]) * ny +
                        static_cast<double>(Hs[5]) * nz);
      F2 = far_scale * (static_cast<double>(Hs[4]) * nx + static_cast<double>(Hs[5]) * ny +
                        static_cast<double>(Hs[2]) * nz);
    }
    if (t < cnt) {
      const int o = perm[s0 + t];
      if (has_text) {
        const double n3[3] = {nx, ny, nz};
        double v[3] = {0.0, 0.0, 0.0};
        for (int r = 0; r < 3; ++r)
          for (int a = 0; a < 3; ++a) v[r] += static_cast<double>(text.p[3 * a + r][o]) * n3[a];
        F0 += v[0];
        F1 += v[1];
        F2 += v[2];
      }
      far3[o] = F0;
      far3[n + o] = F1;
      far3[2 * n + o] = F2;
    }
  }
}

// Force near-field sums over one contiguous source range, every STRIDE-th
// source from j0.  NEAR: the range may hold pairs within 4h (fitted curves
// below q = 4); otherwise every pair takes the far form, with no test.
// The pair in the form with the far field's -5 folded away: with n5 = -5 n
// and Rq = R / -5 (far: P ir^2; near: the table's second curve),
//   e = Rq (d.m)(d.n5) + P (m.n) = R (d.m)(d.n) + P (m.n),
// and the m-sums collect P (d.n5) = -5 P (d.n): the caller scales them by
// -1/5.  One multiply per pair fewer than R = -5 P ir^2.
template <int STRIDE, bool NEAR>
__device__ __forceinline__ void force_range(const SrcRec *__restrict__ rec, int j0, int j1,
                                            float tx, float ty, float tz, float nx, float ny,
                                            float nz, float nx5, float ny5, float nz5,
                                            const HessConsts &hc, const float4 *tab, float &dx0,
                                            float &dy0, float &dz0, float &mx0, float &my0,
                                            float &mz0, float &n0) {
#pragma unroll 2
  for (int j = j0; j < j1; j += STRIDE) {
    const float sx = rec[j].x, sy = rec[j].y, sz = rec[j].z;
    const float mx = rec[j].mx, my = rec[j].my, mz = rec[j].mz;
    const float dx = tx - sx, dy = ty - sy, dz = tz - sz;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float dm = fmaf(dx, mx, fmaf(dy, my, dz * mz));
    const float dn5 = fmaf(dx, nx5, fmaf(dy, ny5, dz * nz5));
    const float mn = fmaf(mx, nx, fmaf(my, ny, mz * nz));
    float Pc, e;
    if constexpr (NEAR) {
      // Both curves default to their far form r^-4 (P = r^-5, R / -5 = r^-7)
      // and the table replaces them within 4h; one scaling serves both.
      // The clamp keeps a pair at d = 0 finite (its terms vanish with d).
      const float irc = __builtin_amdgcn_fmed3f(ir, 0.f, 1e12f);
      const float ir2 = irc * irc;
      float c1 = ir2 * ir2, dq = c1;
      if (!(r2 > hc.near2)) hess_near_curves_tab(tab, hc.tab_scale, r2, irc, c1, dq);
      Pc = c1 * irc;
      e = fmaf(dq * (ir2 * irc), dm * dn5, Pc * mn);
    } else {
      const float ir2 = ir * ir;
      Pc = ir2 * ir2 * ir;
      e = Pc * fmaf(dm * ir2, dn5, mn);  // R / -5 = P ir^2 factored out
    }
    const float f5 = Pc * dn5;
    dx0 = fmaf(dx, e, dx0);
    dy0 = fmaf(dy, e, dy0);
    dz0 = fmaf(dz, e, dz0);
    mx0 = fmaf(mx, f5, mx0);
    my0 = fmaf(my, f5, my0);
    mz0 = fmaf(mz, f5, mz0);
    n0 = fmaf(Pc, dm, n0);
  }
}

// ---------------------------------------------------------------------------
// P2P: the near field of every leaf, in one launch per product over two leaf
// lists.
//  * Full chunks (near_*_full): one wave per leaf, one target per lane,
//    every lane walking all sources (scalar loads), over the leaf's chunks
//    of 64 targets -- and over its remainder too when that holds more than
//    32.
//  * Remainders (near_*_remainder): the at most 32 targets left after a
//    leaf's full chunks (a sparse leaf, or the tail of a leaf of 65..96
//    particles).  One wave per leaf, lanes over sources: the leaf's near
//    ranges, concatenated, form one stream; lane l takes positions l, l +
//    64, ... (LaneStream) and evaluates each of its sources against R
//    remainder targets per pass.  The targets are wave-uniform (scalar
//    loads; R = 4 keeps them in scalar registers), every lane has work
//    whatever the remainder's size, and the lanes' partial sums are added
//    across the wave in FP64.
// Each XCD takes its share of the remainders before its share of the full
// chunks (xcd_two_lists): a remainder wave is long (its whole stream, one
// wave) but there are few, and started first they finish under the full
// chunks' work instead of trailing a launch of their own.  Ranges of the
// leaf's mixed slots (nr.x, first) may hold pairs inside the smoothing
// support.

// Workgroup -> position in one of two leaf lists (remainders first, then
// full chunks), each XCD (round-robin dispatch over 8) walking a contiguous
// range of each list: neighbouring leaves share most of their near sources,
// which then hit in that XCD's L2.  Returns -1 for the padding.
__device__ __forceinline__ int xcd_two_lists(int nrem, int nfull, bool &rem) {
  const int prem = (nrem + 7) / 8, pfull = (nfull + 7) / 8;
  const int x = static_cast<int>(blockIdx.x % 8), k = static_cast<int>(blockIdx.x / 8);
  rem = k < prem;
  const int pos = rem ? x * prem + k : x * pfull + (k - prem);
  return pos < (rem ? nrem : nfull) ? pos : -1;
}

// Remainder targets [tb, te) of a leaf of cnt particles (empty when the
// full-chunk path takes the remainder).
__device__ __forceinline__ void remainder_span(int cnt, int &tb, int &te) {
  tb = cnt & ~63;
  te = (cnt & 63) <= 32 ? cnt : tb;
}

struct FieldOut {
  int accumulate;
  float hc0, hc1, hc2;
  float *__restrict__ ox;
  float *__restrict__ oy;
  float *__restrict__ oz;
  __device__ __forceinline__ void put(int o, double g0, double g1, double g2) const {
    const double cf = 1.0 / (4.0 * kPi);
    const double b0 = accumulate ? static_cast<double>(ox[o]) : hc0;
    const double b1 = accumulate ? static_cast<double>(oy[o]) : hc1;
    const double b2 = accumulate ? static_cast<double>(oz[o]) : hc2;
    ox[o] = static_cast<float>(b0 + cf * g0);
    oy[o] = static_cast<float>(b1 + cf * g1);
    oz[o] = static_cast<float>(b2 + cf * g2);
  }
};

struct ForceOut {
  int accumulate;
  float *__restrict__ fx;
  float *__restrict__ fy;
  float *__restrict__ fz;
  __device__ __forceinline__ void put(int o, double f0, double f1, double f2) const {
    const double sfac = 3.0 * kMu0 / (4.0 * kPi);
    fx[o] = static_cast<float>((accumulate ? static_cast<double>(fx[o]) : 0.0) + sfac * f0);
    fy[o] = static_cast<float>((accumulate ? static_cast<double>(fy[o]) : 0.0) + sfac * f1);
    fz[o] = static_cast<float>((accumulate ? static_cast<double>(fz[o]) : 0.0) + sfac * f2);
  }
};

__device__ __forceinline__ void near_field_full(const SrcRec *__restrict__ rec,
                                                const int32_t *__restrict__ perm, int s0, int cnt,
                                                const int2 *__restrict__ rg, int2 nrb,
                                                const FieldConsts &fc, const float4 *s_ftab,
                                                int lane, const FieldOut &out) {
  int tb, te;
  p2p_target_span<64>(cnt, tb, te);
  for (int t0 = tb; t0 < te; t0 += 64) {
    const int t = t0 + lane;
    const SrcRec me = rec[s0 + (t < te ? t : te - 1)];
    const float tx = me.x, ty = me.y, tz = me.z;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    for (int r = 0; r < nrb.x + nrb.y; ++r) {
      const int2 q = rg[r];
      float a1x = 0.f, a1y = 0.f, a1z = 0.f, a2x = 0.f, a2y = 0.f, a2z = 0.f;
      if (r < nrb.x)
        field_range_mixed<1>(rec, q.x, q.y, tx, ty, tz, fc, s_ftab, a1x, a1y, a1z, a2x, a2y, a2z);
      else
        field_range_far<1>(rec, q.x, q.y, tx, ty, tz, a1x, a1y, a1z, a2x, a2y, a2z);
      acc0 += 3.0 * static_cast<double>(a1x) - static_cast<double>(a2x);
      acc1 += 3.0 * static_cast<double>(a1y) - static_cast<double>(a2y);
      acc2 += 3.0 * static_cast<double>(a1z) - static_cast<double>(a2z);
    }
    if (t < te) out.put(perm[s0 + t], acc0, acc1, acc2);
  }
}

template <int R>
__device__ __forceinline__ void near_field_remainder(const SrcRec *__restrict__ rec,
                                                     const int32_t *__restrict__ perm, int s0,
                                                     int cnt, const int *s_pre, const int *s_beg,
                                                     int nrange, int nmix, const FieldConsts &fc,
                                                     const float4 *s_ftab, int lane,
                                                     const FieldOut &out) {
  int tb, te;
  remainder_span(cnt, tb, te);
  for (int t0 = tb; t0 < te; t0 += R) {
    const int nt = min(R, te - t0);  // uniform
    float tx[R], ty[R], tz[R], gx[R], gy[R], gz[R];
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const SrcRec me = rec[s0 + t0 + min(t, nt - 1)];
      tx[t] = me.x;
      ty[t] = me.y;
      tz[t] = me.z;
      gx[t] = gy[t] = gz[t] = 0.f;
    }
    LaneStream st{reinterpret_cast<const float4 *>(rec), s_pre, s_beg, s_pre[nrange], nmix};
    st.start(lane, nrange);
    for (int i0 = lane; i0 < st.total; i0 += 64 * kLaneBlock) {
      st.fetch(i0);
#pragma unroll
      for (int k = 0; k < kLaneBlock; ++k) {
        const float4 pa = st.pa[k], pb = st.pb[k];
        const bool mixed = st.mixed[k];
        const float nmx = -pb.x, nmy = -pb.y, nmz = -pb.z;
#pragma unroll
        for (int t = 0; t < R; ++t) {
          if (t < nt) {
            const float dx = tx[t] - pa.x, dy = ty[t] - pa.y, dz = tz[t] - pa.z;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            const float ir = __builtin_amdgcn_rsqf(r2);
            const float ir2 = ir * ir;
            float beta = ir2 * ir;
            float alpha = beta * ir2;
            if (mixed && r2 < fc.near2)
              field_near_coeffs_tab(s_ftab, fc.tab_scale, r2, ir, alpha, beta);
            // g += 3 (d . m) alpha d - beta m
            const float s1 = 3.f * fmaf(dx, pb.x, fmaf(dy, pb.y, dz * pb.z)) * alpha;
            gx[t] = fmaf(dx, s1, fmaf(nmx, beta, gx[t]));
            gy[t] = fmaf(dy, s1, fmaf(nmy, beta, gy[t]));
            gz[t] = fmaf(dz, s1, fmaf(nmz, beta, gz[t]));
          }
        }
      }
    }
#pragma unroll
    for (int t = 0; t < R; ++t) {
      if (t < nt) {
        const double sx = wave_sum(static_cast<double>(gx[t]));
        const double sy = wave_sum(static_cast<double>(gy[t]));
        const double sz = wave_sum(static_cast<double>(gz[t]));
        if (lane == 0) out.put(perm[s0 + t0 + t], sx, sy, sz);
      }
    }
  }
}

// P2P, field: out[perm] (+)= sum over each leaf's near ranges; accumulate = 0
// writes H_const + near (trees without a far field).  full / remain: the two
// leaf lists (see above).
template <int T>
__global__ __launch_bounds__(64) void p2p_field_kernel(
    const int32_t *__restrict__ full, int nfull, const int32_t *__restrict__ remain, int nrem,
    const int32_t *__restrict__ start, const int32_t *__restrict__ count,
    const SrcRec *__restrict__ rec, const int32_t *__restrict__ perm,
    const int2 *__restrict__ rng, const int2 *__restrict__ nr, int nslot, FieldConsts fc,
    const float4 *__restrict__ ftab, FieldOut out, SysMask mask) {
  static_assert(T == 64, "one wave per leaf");
  bool rem;
  const int li = xcd_two_lists(nrem, nfull, rem);
  if (li < 0) return;
  const int b = rem ? remain[li] : full[li];
  if (sys_skip(mask, b)) return;  // uniform over the workgroup, before the barrier
  const int lane = threadIdx.x;
  // near-pair coefficients from the table (2 KB of LDS); the remainder's
  // stream offsets
  __shared__ float4 s_ftab[2 * kFieldTabEntries];
  __shared__ int s_pre[kNearF + 1], s_beg[kNearF];
  for (int i = lane; i < 2 * kFieldTabEntries; i += 64) s_ftab[i] = ftab[i];
  const int2 nrb = nr[b];
  const int2 *__restrict__ rg = rng + static_cast<int64_t>(b) * nslot;
  if (rem) range_prefix(s_pre, s_beg, rg, nrb.x + nrb.y, lane);
  __syncthreads();
  if (rem)
    near_field_remainder<4>(rec, perm, start[b], count[b], s_pre, s_beg, nrb.x + nrb.y, nrb.x, fc,
                            s_ftab, lane, out);
  else
    near_field_full(rec, perm, start[b], count[b], rg, nrb, fc, s_ftab, lane, out);
}

__device__ __forceinline__ void near_force_full(const SrcRec *__restrict__ rec,
                                                const int32_t *__restrict__ perm, int s0, int cnt,
                                                const int2 *__restrict__ rg, int2 nrb,
                                                const HessConsts &hc, const float4 *tab, int lane,
                                                const ForceOut &out) {
  int tb, te;
  p2p_target_span<64>(cnt, tb, te);
  const int nrange = nrb.x + nrb.y;
  for (int t0 = tb; t0 < te; t0 += 64) {
    const int t = t0 + lane;
    const SrcRec me = rec[s0 + (t < te ? t : te - 1)];
    const float tx = me.x, ty = me.y, tz = me.z, nx = me.mx, ny = me.my, nz = me.mz;
    const float nx5 = -5.f * nx, ny5 = -5.f * ny, nz5 = -5.f * nz;
    double ad0 = 0.0, ad1 = 0.0, ad2 = 0.0, am0 = 0.0, am1 = 0.0, am2 = 0.0, an = 0.0;
    for (int r = 0; r < nrange; ++r) {
      const int2 q = rg[r];
      float dx0 = 0.f, dy0 = 0.f, dz0 = 0.f, mx0 = 0.f, my0 = 0.f, mz0 = 0.f, n0 = 0.f;
      if (r < nrb.x)
        force_range<1, true>(rec, q.x, q.y, tx, ty, tz, nx, ny, nz, nx5, ny5, nz5, hc, tab, dx0,
                             dy0, dz0, mx0, my0, mz0, n0);
      else
        force_range<1, false>(rec, q.x, q.y, tx, ty, tz, nx, ny, nz, nx5, ny5, nz5, hc, tab, dx0,
                              dy0, dz0, mx0, my0, mz0, n0);
      ad0 += dx0;
      ad1 += dy0;
      ad2 += dz0;
      am0 += mx0;
      am1 += my0;
      am2 += mz0;
      an += n0;
    }
    // the m-sums hold -5 times their terms (force_range)
    if (t < te)
      out.put(perm[s0 + t], ad0 - 0.2 * am0 + static_cast<double>(nx) * an,
              ad1 - 0.2 * am1 + static_cast<double>(ny) * an,
              ad2 - 0.2 * am2 + static_cast<double>(nz) * an);
  }
}

template <int R>
__device__ __forceinline__ void near_force_remainder(const SrcRec *__restrict__ rec,
                                                     const int32_t *__restrict__ perm, int s0,
                                                     int cnt, const int *s_pre, const int *s_beg,
                                                     int nrange, int nmix, const HessConsts &hc,
                                                     const float4 *tab, int lane,
                                                     const ForceOut &out) {
  int tb, te;
  remainder_span(cnt, tb, te);
  for (int t0 = tb; t0 < te; t0 += R) {
    const int nt = min(R, te - t0);  // uniform
    float tx[R], ty[R], tz[R], nx[R], ny[R], nz[R];
    float ax[R], ay[R], az[R], an[R];
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const SrcRec me = rec[s0 + t0 + min(t, nt - 1)];
      tx[t] = me.x;
      ty[t] = me.y;
      tz[t] = me.z;
      nx[t] = me.mx;
      ny[t] = me.my;
      nz[t] = me.mz;
      ax[t] = ay[t] = az[t] = an[t] = 0.f;
    }
    LaneStream st{reinterpret_cast<const float4 *>(rec), s_pre, s_beg, s_pre[nrange], nmix};
    st.start(lane, nrange);
    for (int i0 = lane; i0 < st.total; i0 += 64 * kLaneBlock) {
      st.fetch(i0);
#pragma unroll
      for (int k = 0; k < kLaneBlock; ++k) {
        const float4 pa = st.pa[k];
        const float mx = st.pb[k].x, my = st.pb[k].y, mz = st.pb[k].z;
        const bool mixed = st.mixed[k];
#pragma unroll
        for (int t = 0; t < R; ++t) {
          if (t < nt) {
            const float dx = tx[t] - pa.x, dy = ty[t] - pa.y, dz = tz[t] - pa.z;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            const float ir = __builtin_amdgcn_rsqf(r2);
            // far curves r^-4 by default, the table's within 4h (force_range)
            const float irc = __builtin_amdgcn_fmed3f(ir, 0.f, 1e12f);
            const float ir2 = irc * irc;
            float c1 = ir2 * ir2, dq = c1;
            if (mixed && !(r2 > hc.near2)) hess_near_curves_tab(tab, hc.tab_scale, r2, irc, c1, dq);
            const float Pc = c1 * irc;
            const float Rc = -5.f * dq * (ir2 * irc);
            const float dm = fmaf(dx, mx, fmaf(dy, my, dz * mz));
            const float dn = fmaf(dx, nx[t], fmaf(dy, ny[t], dz * nz[t]));
            const float mn = fmaf(mx, nx[t], fmaf(my, ny[t], mz * nz[t]));
            const float e = fmaf(Rc, dm * dn, Pc * mn);
            const float f = Pc * dn;
            ax[t] = fmaf(dx, e, fmaf(mx, f, ax[t]));
            ay[t] = fmaf(dy, e, fmaf(my, f, ay[t]));
            az[t] = fmaf(dz, e, fmaf(mz, f, az[t]));
            an[t] = fmaf(Pc, dm, an[t]);
          }
        }
      }
    }
#pragma unroll
    for (int t = 0; t < R; ++t) {
      if (t < nt) {
        const double sx = wave_sum(static_cast<double>(ax[t]));
        const double sy = wave_sum(static_cast<double>(ay[t]));
        const double sz = wave_sum(static_cast<double>(az[t]));
        const double sn = wave_sum(static_cast<double>(an[t]));
        if (lane == 0)
          out.put(perm[s0 + t0 + t], sx + static_cast<double>(nx[t]) * sn,
                  sy + static_cast<double>(ny[t]) * sn, sz + static_cast<double>(nz[t]) * sn);
      }
    }
  }
}

// P2P, force: F[perm] (+)= (sum over each leaf's near ranges of B[m_j]) . m_i;
// the two leaf lists as in p2p_field_kernel.
template <int T>
__global__ __launch_bounds__(64) void p2p_force_kernel(
    const int32_t *__restrict__ full, int nfull, const int32_t *__restrict__ remain, int nrem,
    const int32_t *__restrict__ start, const int32_t *__restrict__ count,
    const SrcRec *__restrict__ rec, const int32_t *__restrict__ perm,
    const int2 *__restrict__ rng, const int2 *__restrict__ nr, int nslot, HessConsts hc,
    const float4 *__restrict__ htab, ForceOut out) {
  static_assert(T == 64, "one wave per leaf");
  bool rem;
  const int li = xcd_two_lists(nrem, nfull, rem);
  if (li < 0) return;  // uniform over the workgroup, before the barrier
  const int b = rem ? remain[li] : full[li];
  const int lane = threadIdx.x;
  __shared__ float4 s_tab[2 * kHessTabEntries];
  __shared__ int s_pre[kNearH + 1], s_beg[kNearH];
  for (int i = lane; i < 2 * kHessTabEntries; i += 64) s_tab[i] = htab[i];
  const float4 *tab = s_tab;
  const int2 nrb = nr[b];
  const int2 *__restrict__ rg = rng + static_cast<int64_t>(b) * nslot;
  if (rem) range_prefix(s_pre, s_beg, rg, nrb.x + nrb.y, lane);
  __syncthreads();
  if (rem)
    near_force_remainder<4>(rec, perm, start[b], count[b], s_pre, s_beg, nrb.x + nrb.y, nrb.x, hc,
                            tab, lane, out);
  else
    near_force_full(rec, perm, start[b], count[b], rg, nrb, hc, tab, lane, out);
}

// Moments into the sorted records: rec[i].m = scale * weight[perm[i]] * m[perm[i]]
// (weight: optional per-particle factor in caller order, e.g. kappa per system).
__global__ void pack_sorted_moments_kernel(int64_t n, const int32_t *__restrict__ perm,
                                           const float *__restrict__ mx,
                                           const float *__restrict__ my,
                                           const float *__restrict__ mz, float scale,
                                           const float *__restrict__ weight,
                                           SrcRec *__restrict__ rec) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = perm[i];
  const float w = weight ? scale * weight[o] : scale;
  reinterpret_cast<float4 *>(rec + i)[1] = make_float4(w * mx[o], w * my[o], w * mz[o], 0.f);
}

// Work counts: near pairs of the leaves and valid M2L slots of a box range.
__global__ void count_p2p_kernel(int nleaf, const int32_t *__restrict__ count,
                                 const int32_t *__restrict__ near, int nslot,
                                 unsigned long long *__restrict__ total) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nleaf) return;
  unsigned long long s = 0;
  for (int k = 0; k < nslot; ++k) {
    const int j = near[static_cast<int64_t>(b) * nslot + k];
    if (j >= 0) s += static_cast<unsigned long long>(count[j]);
  }
  atomicAdd(total, s * static_cast<unsigned long long>(count[b]));
}
__global__ void count_m2l_kernel(int nbox, int base, const int32_t *__restrict__ src, int nslot,
                                 unsigned long long *__restrict__ total) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbox) return;
  unsigned long long s = 0;
  for (int k = 0; k < nslot; ++k) s += src[static_cast<int64_t>(base + b) * nslot + k] >= 0 ? 1 : 0;
  atomicAdd(total, s);
}

// ---------------------------------------------------------------------------
// Host-side geometry tables.

std::vector<int3> near_offsets_field() {
  std::vector<int3> inner, outer;
  for (int x = -3; x <= 3; ++x)
    for (int y = -3; y <= 3; ++y)
      for (int z = -3; z <= 3; ++z) {
        if (x * x + y * y + z * z > 6) continue;
        const bool adj = std::abs(x) <= 1 && std::abs(y) <= 1 && std::abs(z) <= 1;
        (adj ? inner : outer).push_back(make_int3(x, y, z));
      }
  inner.insert(inner.end(), outer.begin(), outer.end());
  return inner;  // 27 + 54
}

int excess2(int x, int y, int z) {
  auto e = [](int v) {
    const int a = std::abs(v) - 1;
    return a > 0 ? a * a : 0;
  };
  return e(x) + e(y) + e(z);
}

// Boxes whose nearest points lie at most two leaf sides apart: every pair
// within 4h (leaf side 2h).  Ordered by that separation (excess2 is its square
// in leaf units), so the boxes that may hold pairs within 4h form a prefix.
std::vector<int3> near_offsets_hess() {
  std::vector<int3> v;
  for (int e = 0; e <= 4; ++e)
    for (int x = -3; x <= 3; ++x)
      for (int y = -3; y <= 3; ++y)
        for (int z = -3; z <= 3; ++z)
          if (excess2(x, y, z) == e) v.push_back(make_int3(x, y, z));
  return v;  // 179: 125 with excess2 <= 3, then 54 with excess2 == 4
}

bool in_near_field(int x, int y, int z) { return x * x + y * y + z * z <= 6; }
bool in_near_hess(int x, int y, int z) { return excess2(x, y, z) <= 4; }

// Interaction offsets (target - source) for a target of parity beta:
// children of the parent's N-neighbours, minus the near set of the level.
std::vector<int3> m2l_offsets(int beta, bool hess_leaf) {
  const int bx = beta & 1, by = (beta >> 1) & 1, bz = (beta >> 2) & 1;
  std::vector<int3> out;
  for (int dx = -2; dx <= 2; ++dx)
    for (int dy = -2; dy <= 2; ++dy)
      for (int dz = -2; dz <= 2; ++dz) {
        if (!in_near_field(dx, dy, dz)) continue;
        for (int g = 0; g < 8; ++g) {
          const int gx = g & 1, gy = (g >> 1) & 1, gz = (g >> 2) & 1;
          const int ox = bx - gx - 2 * dx, oy = by - gy - 2 * dy, oz = bz - gz - 2 * dz;
          const bool near = hess_leaf ? in_near_hess(ox, oy, oz) : in_near_field(ox, oy, oz);
          if (!near) out.push_back(make_int3(ox, oy, oz));
        }
      }
  std::sort(out.begin(), out.end(), [](const int3 &a, const int3 &b) {
    if (a.x != b.x) return a.x < b.x;
    if (a.y != b.y) return a.y < b.y;
    return a.z < b.z;
  });
  return out;
}

int off_code(int x, int y, int z) { return ((x + 8) * 17 + (y + 8)) * 17 + (z + 8); }

}  // namespace

// Operator tables of one expansion order on one device (built once per context).
struct FmmTables {
  int P = 0, KP = 0;
  int nslotF = 0, nslotH = 0;
  int noffF = 0;          // distinct offsets of the field M2L (all parities)
  std::vector<int> nearH_excess;  // excess2 of each force near slot (ascending)
  DeviceBuffer T;         // [noff][KP x KP in FP32 fragment order]
  DeviceBuffer Tb;        // BF16 bits [noff][3 parts][KP x KP in fragment order]
  DeviceBuffer W_m2m;     // [8][KP x KP in FP32 fragment order]
  DeviceBuffer W_l2l;     // [8][KP x KP in FP32 fragment order]
  DeviceBuffer W_m2m_b;   // BF16 bits [8][3 parts][KP x KP in fragment order]
  DeviceBuffer W_l2l_b;   // BF16 bits [8][3 parts][KP x KP in fragment order]
  DeviceBuffer slot_tab;  // [kSlotRows][slot_ld] table index per slot (-1 unused)
  DeviceBuffer offsF;     // [8][nslotF] int3 (padding 0x7fffffff)
  DeviceBuffer offsH;     // [8][nslotH]
  DeviceBuffer nearF;     // [81] int3
  DeviceBuffer nearH;     // [179] int3
  int slot_ld = 0;
};

// Near-pair coefficient tables of one smoothing length, built once per
// context and h: the field's (make_field_table) and the force's
// (make_hess_table).
struct FieldTable {
  DeviceBuffer buf;  // kFieldTabEntries or kHessTabEntries x 8 floats, split layout
};

struct FmmState {
  int P = 10, KP = 128;
  int D = 0;
  int nsys = 1;
  DeviceBuffer sys_lo;  // double [nsys][3]: root corner of each system
  DeviceBuffer chunks;  // int64 [nchunk][2]: record ranges of the bounding-box pass
  double L = 0.0, s_leaf = 0.0;
  int nmixH = 0;  // leading force near slots that may hold pairs within 4h
  int64_t n = 0;
  std::shared_ptr<FmmTables> tab;
  std::shared_ptr<FieldTable> ftab, htab;
  // host copies of the tile descriptors (sources of asynchronous uploads)
  std::vector<M2LTile> h_tilesF, h_tilesH, h_tilesM, h_tilesL;
  DeviceBuffer keys, keys_sorted, idx, perm, flags, scan, rec;  // particle arrays
  // boxes: level l occupies [base[l], base[l] + count[l]) of the global arrays
  std::vector<int> base, count;
  int nbox = 0;
  DeviceBuffer box_keys;     // uint64 [nbox] (level-local keys)
  DeviceBuffer box_first;    // int32 [nbox]: first particle (leaf) / first child (level-local, inner)
  DeviceBuffer box_count;    // int32 [nbox]: particles (leaf) / children (inner)
  DeviceBuffer box_parent;   // int32 [nbox] global parent index
  DeviceBuffer box_sys;      // int32 [nbox] system of each box (batch plans)
  DeviceBuffer nearF, nearH; // leaf lists
  DeviceBuffer cell_grid;    // int32 [8^l]: cell -> box of one level (list building)
  DeviceBuffer rngF, rngH;   // int2 merged near ranges [nleaf][nslot]
  DeviceBuffer nrF, nrH;     // int2 range counts per leaf (class 0, class 1)
  DeviceBuffer leaf_lists;   // int32: the four leaf-shape lists back to back, each in leaf order
  int leaf_list_n[kLeafLists] = {};
  int leaf_list_off[kLeafLists] = {};
  // lists 1 .. kLeafLists - 1 (the remainders) lie back to back
  int leaf_lists_remainder() const {
    int n = 0;
    for (int k = 1; k < kLeafLists; ++k) n += leaf_list_n[k];
    return n;
  }
  DeviceBuffer srcF, srcH;   // M2L source slots [nbox][nslotF], [nleaf][nslotH]
  DeviceBuffer class_ids, class_keys, class_keys2, class_ids2, class_counts;
  DeviceBuffer tilesF, tilesH;
  DeviceBuffer tlistF, tlistH, tnF, tnH;
  DeviceBuffer orderF, orderH;  // launch order of the tiles (tile_order_kernel)  // non-empty offsets per tile [ntiles][slot_ld], counts
  int ntilesF = 0, ntilesH = 0;
  int leaf_tile_start = 0;  // first tile of the leaf level (tiles are in level order)
  // Upward and downward translations on the M2L kernel: M2M tiles (parents,
  // slots = octants of the children) and L2L tiles (children, one slot: the
  // parent), in level order with per-level start indices.
  DeviceBuffer child_tab;  // int32 [nbox][8]: global child index per octant, -1 none
  DeviceBuffer box_ids;    // int32 [nbox]: identity list (M2M tile targets)
  DeviceBuffer tilesM, tilesL, tlistM, tnM, tlistL, tnL;
  std::vector<int> tileM_start, tileL_start;  // [level], size D + 2
  DeviceBuffer Mt, Lt;
  DeviceBuffer Mb;  // BF16 parts of Mt [nbox][3][KP], for the BF16-split M2L
  int64_t p2p_field_pairs = 0;  // sum over leaves of targets x near sources
  int64_t p2p_force_pairs = 0;
  int64_t m2l_pairs_field = 0;
  int64_t m2l_pairs_force = 0;
  int64_t m2l_slots_field = 0;  // sum over M2L tiles of their non-empty offsets
  int64_t m2l_slots_force = 0;
  bool counts_valid = false;
};

void destroy_fmm(FmmState *s) { delete s; }

namespace {

uint16_t bf16_bits_rne_host(float x) {
  uint32_t u;
  std::memcpy(&u, &x, sizeof(u));
  return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Position of element (row, col) of BF16 part p within one offset's table.
// The order is the MFMA fragment order of the M2L kernel: per (part, 32-wide
// k block, 16-row block), the 64 lanes' 8-element row pieces (lane = 16 kq +
// r16 holds row 16 rb + r16, columns 32 blk + 8 kq .. + 7) lie in lane order,
// so that one wave's fragment load reads 1 KB of consecutive memory.
size_t bf16_frag_index(int KP, int p, int row, int col) {
  const int blk = col / 32, kq = (col % 32) / 8, e = col % 8;
  const int rb = row / 16, r16 = row % 16;
  const size_t piece = ((static_cast<size_t>(p) * (KP / 32) + blk) * (KP / 16) + rb) * 64 + 16 * kq + r16;
  return piece * 8 + e;
}

// Position of element (row, col) within one FP32 operator table (M2L, M2M,
// L2L).  MFMA step j of 16-wide k block k16 takes k = 16 k16 + 4 kq + j from
// lane 16 kq + r16, so a lane's four steps are four consecutive k of its row:
// one 16-byte load per 16-row block and k block, 1 KB per wave in lane order.
size_t fp32_frag_index(int KP, int row, int col) {
  const int k16 = col / 16, kq = (col % 16) / 4, j = col % 4;
  const int rb = row / 16, r16 = row % 16;
  return ((static_cast<size_t>(k16) * (KP / 16) + rb) * 64 + 16 * kq + r16) * 4 + j;
}

int kp_for(int P) {
  const int nd = (P + 1) * (P + 1);
  return ((nd + 31) / 32) * 32;
}

Status build_tables(mf_ctx *ctx, int P, std::shared_ptr<FmmTables> *out) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  const std::string key = "fmm_tables_P" + std::to_string(P);
  auto it = ctx->cache.find(key);
  if (it != ctx->cache.end()) {
    *out = std::static_pointer_cast<FmmTables>(it->second);
    return Status::Ok();
  }
  auto t = std::make_shared<FmmTables>();
  t->P = P;
  t->KP = kp_for(P);
  const int KP = t->KP;
  // offsets per parity, F (all levels) and H (leaf level, force)
  std::vector<std::vector<int3>> of(8), oh(8);
  for (int b = 0; b < 8; ++b) {
    of[b] = m2l_offsets(b, false);
    oh[b] = m2l_offsets(b, true);
    t->nslotF = std::max<int>(t->nslotF, static_cast<int>(of[b].size()));
    t->nslotH = std::max<int>(t->nslotH, static_cast<int>(oh[b].size()));
  }
  // distinct offsets -> table index
  std::map<int, int> code2tab;
  std::vector<int3> uniq;
  for (int b = 0; b < 8; ++b)
    for (auto *lst : {&of[b], &oh[b]})
      for (const int3 &o : *lst) {
        const int c = off_code(o.x, o.y, o.z);
        if (!code2tab.count(c)) {
          code2tab[c] = static_cast<int>(uniq.size());
          uniq.push_back(o);
        }
      }
  {
    std::map<int, int> field_codes;
    for (int b = 0; b < 8; ++b)
      for (const int3 &o : of[b]) field_codes[off_code(o.x, o.y, o.z)] = 1;
    t->noffF = static_cast<int>(field_codes.size());
  }
  t->slot_ld = std::max(t->nslotF, t->nslotH);
  // rows: 0..7 field M2L per parity, 8..15 force leaf M2L, 16 M2M (slot k ->
  // octant k), 17..24 L2L of a child of octant k (slot 0 -> octant k)
  std::vector<int32_t> slots(kSlotRows * t->slot_ld, -1);
  for (int k = 0; k < 8; ++k) {
    slots[kRowM2M * t->slot_ld + k] = k;
    slots[(kRowL2L + k) * t->slot_ld] = k;
  }
  std::vector<int3> hof(8 * t->nslotF, make_int3(0x7fffffff, 0, 0));
  std::vector<int3> hoh(8 * t->nslotH, make_int3(0x7fffffff, 0, 0));
  for (int b = 0; b < 8; ++b) {
    for (size_t k = 0; k < of[b].size(); ++k) {
      slots[(0 * 8 + b) * t->slot_ld + k] = code2tab[off_code(of[b][k].x, of[b][k].y, of[b][k].z)];
      hof[b * t->nslotF + k] = of[b][k];
    }
    for (size_t k = 0; k < oh[b].size(); ++k) {
      slots[(1 * 8 + b) * t->slot_ld + k] = code2tab[off_code(oh[b][k].x, oh[b][k].y, oh[b][k].z)];
      hoh[b * t->nslotH + k] = oh[b][k];
    }
  }
  // BF16 parts of a KP x KP matrix (column-major FP64) at out[base ..], in
  // fragment order: part p is the BF16 of what parts 0..p-1 left over.
  auto split_bf16_parts = [KP](const std::vector<double> &m, uint16_t *out) {
    for (int row = 0; row < KP; ++row)
      for (int col = 0; col < KP; ++col) {
        double r = m[static_cast<size_t>(col) * KP + row];
        for (int p = 0; p < 3; ++p) {
          const uint16_t bits = bf16_bits_rne_host(static_cast<float>(r));
          out[bf16_frag_index(KP, p, row, col)] = bits;
          uint32_t w = static_cast<uint32_t>(bits) << 16;
          float f;
          std::memcpy(&f, &w, sizeof(f));
          r -= static_cast<double>(f);
        }
      }
  };
  // M2L tables: FP32 and BF16 parts, both in fragment order
  std::vector<float> T(static_cast<size_t>(uniq.size()) * KP * KP);
  std::vector<uint16_t> Tb(static_cast<size_t>(uniq.size()) * 3 * KP * KP);
  for (size_t u = 0; u < uniq.size(); ++u) {
    const std::vector<double> m = m2l_matrix(P, KP, uniq[u].x, uniq[u].y, uniq[u].z);
    for (int col = 0; col < KP; ++col)
      for (int row = 0; row < KP; ++row)
        T[u * KP * KP + fp32_frag_index(KP, row, col)] = static_cast<float>(m[static_cast<size_t>(col) * KP + row]);
    split_bf16_parts(m, Tb.data() + u * 3 * KP * KP);
  }
  std::vector<float> W1(8 * KP * KP), W2(8 * KP * KP);
  std::vector<uint16_t> W1b(8 * 3 * KP * KP), W2b(8 * 3 * KP * KP);
  for (int oct = 0; oct < 8; ++oct) {
    // child octant bits (x, y, z) = (oct & 1, oct >> 1 & 1, oct >> 2 & 1); d in parent units
    const double dx = ((oct & 1) ? 0.25 : -0.25), dy = (((oct >> 1) & 1) ? 0.25 : -0.25),
                 dz = (((oct >> 2) & 1) ? 0.25 : -0.25);
    const std::vector<double> a = m2m_matrix(P, KP, dx, dy, dz);
    const std::vector<double> b = l2l_matrix(P, KP, dx, dy, dz);
    for (int col = 0; col < KP; ++col)
      for (int row = 0; row < KP; ++row) {
        const size_t e = static_cast<size_t>(col) * KP + row, f = oct * KP * KP + fp32_frag_index(KP, row, col);
        W1[f] = static_cast<float>(a[e]);
        W2[f] = static_cast<float>(b[e]);
      }
    split_bf16_parts(a, W1b.data() + static_cast<size_t>(oct) * 3 * KP * KP);
    split_bf16_parts(b, W2b.data() + static_cast<size_t>(oct) * 3 * KP * KP);
  }
  const std::vector<int3> nf = near_offsets_field(), nh = near_offsets_hess();
  if (static_cast<int>(nf.size()) != kNearF || static_cast<int>(nh.size()) != kNearH)
    return Status::Err(MF_ERR_UNSUPPORTED, "near-set size mismatch");
  for (const int3 &o : nh) t->nearH_excess.push_back(excess2(o.x, o.y, o.z));
  auto upload = [&](DeviceBuffer &buf, const void *src, size_t bytes) -> Status {
    MF_TRY(buf.reserve(bytes));
    MF_HIP(hipMemcpy(buf.get(), src, bytes, hipMemcpyHostToDevice));
    return Status::Ok();
  };
  MF_TRY(upload(t->T, T.data(), T.size() * sizeof(float)));
  MF_TRY(upload(t->Tb, Tb.data(), Tb.size() * sizeof(uint16_t)));
  MF_TRY(upload(t->W_m2m, W1.data(), W1.size() * sizeof(float)));
  MF_TRY(upload(t->W_l2l, W2.data(), W2.size() * sizeof(float)));
  MF_TRY(upload(t->W_m2m_b, W1b.data(), W1b.size() * sizeof(uint16_t)));
  MF_TRY(upload(t->W_l2l_b, W2b.data(), W2b.size() * sizeof(uint16_t)));
  MF_TRY(upload(t->slot_tab, slots.data(), slots.size() * sizeof(int32_t)));
  MF_TRY(upload(t->offsF, hof.data(), hof.size() * sizeof(int3)));
  MF_TRY(upload(t->offsH, hoh.data(), hoh.size() * sizeof(int3)));
  MF_TRY(upload(t->nearF, nf.data(), nf.size() * sizeof(int3)));
  MF_TRY(upload(t->nearH, nh.data(), nh.size() * sizeof(int3)));
  ctx->cache[key] = t;
  *out = t;
  return Status::Ok();
}

Status build_field_table(mf_ctx *ctx, double h, bool hess, std::shared_ptr<FieldTable> *out) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  char key[64];
  std::snprintf(key, sizeof(key), hess ? "hess_table_h%a" : "field_table_h%a", h);
  auto it = ctx->cache.find(key);
  if (it != ctx->cache.end()) {
    *out = std::static_pointer_cast<FieldTable>(it->second);
    return Status::Ok();
  }
  auto t = std::make_shared<FieldTable>();
  const int entries = hess ? kHessTabEntries : kFieldTabEntries;
  std::vector<float> rows(8 * entries), tab(8 * entries);
  if (hess) make_hess_table(h, rows.data());
  else make_field_table(h, rows.data());
  split_near_table(rows.data(), entries, tab.data());
  MF_TRY(t->buf.reserve(tab.size() * sizeof(float)));
  MF_HIP(hipMemcpy(t->buf.get(), tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
  ctx->cache[key] = t;
  *out = t;
  return Status::Ok();
}

template <class T>
Status need(DeviceBuffer &b, int64_t count) {
  return b.reserve(sizeof(T) * static_cast<size_t>(count > 0 ? count : 1));
}

Status read_ints(mf_ctx *ctx, const int32_t *d, int count, std::vector<int32_t> &h) {
  h.resize(count);
  if (count == 0) return Status::Ok();
  MF_HIP(hipMemcpyAsync(h.data(), d, sizeof(int32_t) * count, hipMemcpyDeviceToHost, ctx->stream));
  MF_HIP(hipStreamSynchronize(ctx->stream));
  return Status::Ok();
}

}  // namespace

namespace {

// Builds the tree and every list with leaves of side s_leaf (>= 2h).
Status fmm_build_at(mf_plan *plan, double s_leaf) {
  mf_ctx *ctx = plan->ctx;
  if (!plan->fmm) plan->fmm.reset(new FmmState());
  FmmState &F = *plan->fmm;
  F.P = plan->opts.order > 0 ? plan->opts.order : 10;
  if (F.P != 6 && F.P != 8 && F.P != 10)
    return Status::Err(MF_ERR_INVALID_ARG, "FMM expansion order must be 6, 8 or 10");
  F.KP = kp_for(F.P);
  MF_TRY(build_tables(ctx, F.P, &F.tab));
  MF_TRY(build_field_table(ctx, plan->h, false, &F.ftab));
  MF_TRY(build_field_table(ctx, plan->h, true, &F.htab));
  const FmmTables &tab = *F.tab;
  const int64_t n = plan->n;
  F.n = n;
  const double h = plan->h;
  const SrcRec *rec_in = plan->rec.as<SrcRec>();
  if (n == 0) {
    F.D = 0;
    F.nbox = 0;
    F.base.assign(1, 0);
    F.count.assign(1, 0);
    return Status::Ok();
  }
  // 1. bounding box of every system, from per-chunk partials (a chunk never
  //    spans two systems)
  const int nsys = plan->nsys;
  F.nsys = nsys;
  std::vector<int64_t> off = plan->sys_offsets;
  if (nsys == 1) off = {0, n};
  std::vector<int64_t> ch;
  std::vector<int> ch_sys;
  constexpr int64_t kChunk = 4096;  // 282 workgroups on 1.15 M particles
  for (int sy = 0; sy < nsys; ++sy)
    for (int64_t b0 = off[sy]; b0 < off[sy + 1]; b0 += kChunk) {
      ch.push_back(b0);
      ch.push_back(std::min(off[sy + 1], b0 + kChunk));
      ch_sys.push_back(sy);
    }
  const int nch = static_cast<int>(ch_sys.size());
  MF_TRY(need<int64_t>(F.chunks, 2 * nch));
  MF_HIP(hipMemcpyAsync(F.chunks.get(), ch.data(), sizeof(int64_t) * ch.size(), hipMemcpyHostToDevice,
                        ctx->stream));
  void *pw = nullptr;
  MF_TRY(workspace(ctx, "fmm_bbox", sizeof(float) * 6 * nch, &pw));
  hipLaunchKernelGGL(bbox_partial_kernel, dim3(nch), dim3(256), 0, ctx->stream,
                     F.chunks.as<int64_t>(), rec_in, static_cast<float *>(pw));
  MF_LAUNCH_CHECK();
  std::vector<float> part(6 * nch);
  MF_HIP(hipMemcpyAsync(part.data(), pw, sizeof(float) * 6 * nch, hipMemcpyDeviceToHost, ctx->stream));
  MF_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<double> lo(3 * nsys, INFINITY), hi(3 * nsys, -INFINITY);
  for (int c = 0; c < nch; ++c)
    for (int a = 0; a < 3; ++a) {
      if (!std::isfinite(part[c * 6 + a]) || !std::isfinite(part[c * 6 + 3 + a]))
        return Status::Err(MF_ERR_INVALID_ARG, "particle positions must be finite");
      const int sy = ch_sys[c];
      lo[3 * sy + a] = std::min<double>(lo[3 * sy + a], part[c * 6 + a]);
      hi[3 * sy + a] = std::max<double>(hi[3 * sy + a], part[c * 6 + 3 + a]);
    }
  double ext = 0.0;
  for (int sy = 0; sy < nsys; ++sy) {
    if (off[sy + 1] == off[sy]) {  // empty system: any finite frame
      for (int a = 0; a < 3; ++a) lo[3 * sy + a] = hi[3 * sy + a] = 0.0;
      continue;
    }
    for (int a = 0; a < 3; ++a) {
      if (!std::isfinite(lo[3 * sy + a]) || !std::isfinite(hi[3 * sy + a]))
        return Status::Err(MF_ERR_INVALID_ARG, "particle positions must be finite");
      ext = std::max(ext, hi[3 * sy + a] - lo[3 * sy + a]);
    }
  }
  ext = ext * (1.0 + 1e-6) + 1e-6;
  int D = 0;
  while (s_leaf * std::ldexp(1.0, D) < ext) ++D;
  int sys_bits = 0;
  while ((1LL << sys_bits) < nsys) ++sys_bits;
  if (D > kMaxLevels - 1 || 3 * D + sys_bits > 63)
    return Status::Err(MF_ERR_UNSUPPORTED, "particle set too wide for the octree depth");
  F.D = D;
  F.L = s_leaf * std::ldexp(1.0, D);
  F.s_leaf = s_leaf;
  {
    // force near slots with separation (sqrt(excess2) leaf sides) below 4h
    const double t = 4.0 * h / s_leaf;
    F.nmixH = 0;
    for (int e : tab.nearH_excess)
      if (static_cast<double>(e) < t * t) ++F.nmixH;
  }
  std::vector<double> root(3 * nsys);
  for (int sy = 0; sy < nsys; ++sy)
    for (int a = 0; a < 3; ++a)
      root[3 * sy + a] = 0.5 * (lo[3 * sy + a] + hi[3 * sy + a]) - 0.5 * F.L;
  MF_TRY(need<double>(F.sys_lo, 3 * nsys));
  MF_HIP(hipMemcpyAsync(F.sys_lo.get(), root.data(), sizeof(double) * root.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  MF_HIP(hipStreamSynchronize(ctx->stream));  // host vectors go out of scope
  const int cells = 1 << D;
  // 2. keys and sort
  MF_TRY(need<uint64_t>(F.keys, n));
  MF_TRY(need<uint64_t>(F.keys_sorted, n));
  MF_TRY(need<int32_t>(F.idx, n));
  MF_TRY(need<int32_t>(F.perm, n));
  MF_TRY(need<int32_t>(F.flags, n + 1));
  MF_TRY(need<int32_t>(F.scan, n + 1));
  MF_TRY(need<SrcRec>(F.rec, n));
  hipLaunchKernelGGL(keys_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, rec_in,
                     nsys > 1 ? plan->sys_of.as<int32_t>() : static_cast<const int32_t *>(nullptr),
                     F.sys_lo.as<double>(), 1.0 / F.s_leaf, cells, 3 * D, F.keys.as<uint64_t>(),
                     F.idx.as<int32_t>());
  MF_LAUNCH_CHECK();
  const int key_bits = 3 * D + sys_bits;
  MF_TRY(sort_pairs_u64_i32(ctx, F.keys.as<uint64_t>(), F.keys_sorted.as<uint64_t>(),
                            F.idx.as<int32_t>(), F.perm.as<int32_t>(), n, 0,
                            key_bits > 0 ? key_bits : 1));
  hipLaunchKernelGGL(gather_sorted_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n,
                     rec_in, F.perm.as<int32_t>(), F.rec.as<SrcRec>());
  MF_LAUNCH_CHECK();
  // 3. leaves (level D) and coarser levels by key prefix
  F.base.assign(D + 1, 0);
  F.count.assign(D + 1, 0);
  // leaf runs over particles
  std::vector<int32_t> tmp;
  hipLaunchKernelGGL(run_flags_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n,
                     F.keys_sorted.as<uint64_t>(), 0, F.flags.as<int32_t>());
  MF_LAUNCH_CHECK();
  MF_TRY(exclusive_scan_i32(ctx, F.flags.as<int32_t>(), F.scan.as<int32_t>(), n));
  int32_t last[2];
  MF_HIP(hipMemcpyAsync(&last[0], F.scan.as<int32_t>() + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  MF_HIP(hipMemcpyAsync(&last[1], F.flags.as<int32_t>() + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  MF_HIP(hipStreamSynchronize(ctx->stream));
  F.count[D] = last[0] + last[1];
  // every level holds at most as many boxes as the leaf level
  const int64_t cap = static_cast<int64_t>(F.count[D]) * (D + 1) + 8;
  MF_TRY(need<uint64_t>(F.box_keys, cap));
  MF_TRY(need<int32_t>(F.box_first, cap));
  MF_TRY(need<int32_t>(F.box_count, cap));
  MF_TRY(need<int32_t>(F.box_parent, cap));
  // Box arrays: the leaves (level D) at base 0, then levels D - 1 .. 0:
  // base[D - 1] = cnt[D], base[D - 2] = cnt[D] + cnt[D - 1], ...
  std::vector<int> cnt(D + 1);
  cnt[D] = F.count[D];
  int cursor = cnt[D];
  F.base[D] = 0;
  hipLaunchKernelGGL(run_starts_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n,
                     F.keys_sorted.as<uint64_t>(), 0, F.flags.as<int32_t>(), F.scan.as<int32_t>(),
                     F.box_keys.as<uint64_t>(), F.box_first.as<int32_t>(),
                     static_cast<int32_t *>(nullptr));
  MF_LAUNCH_CHECK();
  hipLaunchKernelGGL(run_counts_kernel, dim3(grid_for(cnt[D], 256)), dim3(256), 0, ctx->stream,
                     cnt[D], static_cast<int>(n), F.box_first.as<int32_t>(), F.box_count.as<int32_t>());
  MF_LAUNCH_CHECK();
  if (D > 0) {
    // the box counts of all levels above the leaves, in one readback
    void *pc = nullptr;
    MF_TRY(workspace(ctx, "fmm_level_counts", sizeof(int32_t) * D, &pc));
    MF_HIP(hipMemsetAsync(pc, 0, sizeof(int32_t) * D, ctx->stream));
    hipLaunchKernelGGL(level_counts_kernel, dim3(grid_for(cnt[D], 256)), dim3(256), 0, ctx->stream,
                       cnt[D], F.box_keys.as<uint64_t>(), D, static_cast<int32_t *>(pc));
    MF_LAUNCH_CHECK();
    std::vector<int32_t> lc;
    MF_TRY(read_ints(ctx, static_cast<const int32_t *>(pc), D, lc));
    for (int l = 0; l < D; ++l) cnt[l] = lc[l] + 1;
  }
  // every level above the leaves in one pass (level_flags / scan /
  // level_boxes / level_child_counts)
  if (D > 0) {
    LevelTable lt{};
    lt.D = D;
    for (int l = D; l >= 0; --l) {
      lt.cnt[l] = cnt[l];
      lt.base[l] = l == D ? 0 : cursor;
      if (l < D) cursor += cnt[l];
    }
    for (int l = 0; l < D; ++l) F.base[l] = lt.base[l];
    const int64_t ne = static_cast<int64_t>(D) * cnt[D];
    MF_TRY(need<int32_t>(F.flags, ne));
    MF_TRY(need<int32_t>(F.scan, ne));
    const uint64_t *lk = F.box_keys.as<uint64_t>();
    hipLaunchKernelGGL(level_flags_kernel, dim3(grid_for(ne, 256)), dim3(256), 0, ctx->stream,
                       cnt[D], lk, D, F.flags.as<int32_t>());
    MF_LAUNCH_CHECK();
    MF_TRY(exclusive_scan_i32(ctx, F.flags.as<int32_t>(), F.scan.as<int32_t>(), ne));
    hipLaunchKernelGGL(level_boxes_kernel, dim3(grid_for(ne, 256)), dim3(256), 0, ctx->stream,
                       cnt[D], lk, lt, F.flags.as<int32_t>(), F.scan.as<int32_t>(),
                       F.box_keys.as<uint64_t>(), F.box_first.as<int32_t>(),
                       F.box_parent.as<int32_t>());
    MF_LAUNCH_CHECK();
    const int nupper = cursor - cnt[D];
    hipLaunchKernelGGL(level_child_counts_kernel, dim3(grid_for(nupper, 256)), dim3(256), 0,
                       ctx->stream, nupper, lt, F.box_first.as<int32_t>(), F.box_count.as<int32_t>());
    MF_LAUNCH_CHECK();
  }
  F.count = cnt;
  F.nbox = cursor;
  // Cell grid of level l (see grid_scatter_kernel) for single-system plans of
  // at most 2^24 cells per level; nullptr (binary search) otherwise.
  auto level_grid = [&](int l, const int32_t **out) -> Status {
    *out = nullptr;
    // at most 2^24 cells (64 MB); wider levels take the binary search
    if (3 * l + sys_bits > 24 || cnt[l] == 0) return Status::Ok();
    const int64_t ncell = 1LL << (3 * l + sys_bits);
    MF_TRY(need<int32_t>(F.cell_grid, ncell));
    hipLaunchKernelGGL(grid_clear_kernel, dim3(grid_for(ncell, 256)), dim3(256), 0, ctx->stream,
                       ncell, F.cell_grid.as<int32_t>());
    MF_LAUNCH_CHECK();
    hipLaunchKernelGGL(grid_scatter_kernel, dim3(grid_for(cnt[l], 256)), dim3(256), 0, ctx->stream,
                       cnt[l], F.box_keys.as<uint64_t>() + F.base[l], F.cell_grid.as<int32_t>());
    MF_LAUNCH_CHECK();
    *out = F.cell_grid.as<int32_t>();
    return Status::Ok();
  };
  // 4. near lists of the leaves
  const int nleaf = cnt[D];
  const int32_t *leaf_grid = nullptr;
  MF_TRY(level_grid(D, &leaf_grid));
  MF_TRY(need<int32_t>(F.nearF, static_cast<int64_t>(nleaf) * kNearF));
  MF_TRY(need<int32_t>(F.nearH, static_cast<int64_t>(nleaf) * kNearH));
  hipLaunchKernelGGL(near_list_kernel, dim3(grid_for(static_cast<int64_t>(nleaf) * kNearF, 256)),
                     dim3(256), 0, ctx->stream, nleaf, F.box_keys.as<uint64_t>(), cells, 3 * D,
                     tab.nearF.as<int3>(), kNearF, leaf_grid, F.nearF.as<int32_t>());
  MF_LAUNCH_CHECK();
  hipLaunchKernelGGL(near_list_kernel, dim3(grid_for(static_cast<int64_t>(nleaf) * kNearH, 256)),
                     dim3(256), 0, ctx->stream, nleaf, F.box_keys.as<uint64_t>(), cells, 3 * D,
                     tab.nearH.as<int3>(), kNearH, leaf_grid, F.nearH.as<int32_t>());
  MF_LAUNCH_CHECK();
  MF_TRY(need<int2>(F.rngF, static_cast<int64_t>(nleaf) * kNearF));
  MF_TRY(need<int2>(F.rngH, static_cast<int64_t>(nleaf) * kNearH));
  MF_TRY(need<int2>(F.nrF, nleaf));
  MF_TRY(need<int2>(F.nrH, nleaf));
  static_assert(kNearF <= 128 && kNearH <= 256, "near-range sort sizes");
  hipLaunchKernelGGL(near_ranges_kernel<128>, dim3(nleaf), dim3(64), 0, ctx->stream, nleaf,
                     F.nearF.as<int32_t>(), kNearF, kNearFMixed, F.box_first.as<int32_t>(),
                     F.box_count.as<int32_t>(), F.rngF.as<int2>(), F.nrF.as<int2>());
  MF_LAUNCH_CHECK();
  hipLaunchKernelGGL(near_ranges_kernel<256>, dim3(nleaf), dim3(64), 0, ctx->stream, nleaf,
                     F.nearH.as<int32_t>(), kNearH, F.nmixH, F.box_first.as<int32_t>(),
                     F.box_count.as<int32_t>(), F.rngH.as<int2>(), F.nrH.as<int2>());
  MF_LAUNCH_CHECK();
  {
    // leaf shapes of the near-field kernels, each list in leaf order
    const int64_t ne = 2 * static_cast<int64_t>(nleaf);
    MF_TRY(need<int32_t>(F.leaf_lists, ne));
    void *pk = nullptr, *pz = nullptr;
    MF_TRY(workspace(ctx, "fmm_leaf_keys", (sizeof(uint64_t) * 2 + sizeof(int32_t)) * ne, &pk));
    MF_TRY(workspace(ctx, "fmm_leaf_sizes", sizeof(int32_t) * kLeafLists, &pz));
    uint64_t *k_in = static_cast<uint64_t *>(pk), *k_out = k_in + ne;
    int32_t *v_in = reinterpret_cast<int32_t *>(k_out + ne);
    MF_HIP(hipMemsetAsync(pz, 0, sizeof(int32_t) * kLeafLists, ctx->stream));
    hipLaunchKernelGGL(classify_leaves_kernel, dim3(grid_for(nleaf, 256)), dim3(256), 0,
                       ctx->stream, nleaf, F.box_count.as<int32_t>(), k_in, v_in,
                       static_cast<int32_t *>(pz));
    MF_LAUNCH_CHECK();
    MF_TRY(sort_pairs_u64_i32(ctx, k_in, k_out, v_in, F.leaf_lists.as<int32_t>(), ne, 0, 3));
    MF_HIP(hipMemcpyAsync(F.leaf_list_n, pz, sizeof(int32_t) * kLeafLists, hipMemcpyDeviceToHost, ctx->stream));
    MF_HIP(hipStreamSynchronize(ctx->stream));
    F.leaf_list_off[0] = 0;
    for (int k = 1; k < kLeafLists; ++k) F.leaf_list_off[k] = F.leaf_list_off[k - 1] + F.leaf_list_n[k - 1];
  }
  // 5. M2L source slots (levels >= 2) and the class-sorted tile lists
  // one row past the last box of each slot array holds -1 throughout: the
  // M2L kernel's "no source" row
  MF_TRY(need<int32_t>(F.srcF, static_cast<int64_t>(F.nbox + 1) * tab.nslotF));
  MF_TRY(need<int32_t>(F.srcH, static_cast<int64_t>(nleaf + 1) * tab.nslotH));
  hipLaunchKernelGGL(grid_clear_kernel, dim3(grid_for(tab.nslotF, 256)), dim3(256), 0, ctx->stream,
                     static_cast<int64_t>(tab.nslotF),
                     F.srcF.as<int32_t>() + static_cast<int64_t>(F.nbox) * tab.nslotF);
  MF_LAUNCH_CHECK();
  hipLaunchKernelGGL(grid_clear_kernel, dim3(grid_for(tab.nslotH, 256)), dim3(256), 0, ctx->stream,
                     static_cast<int64_t>(tab.nslotH),
                     F.srcH.as<int32_t>() + static_cast<int64_t>(nleaf) * tab.nslotH);
  MF_LAUNCH_CHECK();
  for (int l = 2; l <= D; ++l) {
    if (cnt[l] == 0) continue;
    const int32_t *g = nullptr;
    MF_TRY(level_grid(l, &g));
    hipLaunchKernelGGL(m2l_list_kernel, dim3(grid_for(static_cast<int64_t>(cnt[l]) * tab.nslotF, 256)),
                       dim3(256), 0, ctx->stream, cnt[l], F.base[l],
                       F.box_keys.as<uint64_t>() + F.base[l], 1 << l, 3 * l, tab.offsF.as<int3>(),
                       tab.nslotF, g, F.srcF.as<int32_t>());
    MF_LAUNCH_CHECK();
  }
  if (D >= 2) {
    // leaf level Hessian slots: indexed by leaf (global index == leaf index, base[D] = 0)
    const int32_t *g = nullptr;
    MF_TRY(level_grid(D, &g));
    hipLaunchKernelGGL(m2l_list_kernel, dim3(grid_for(static_cast<int64_t>(nleaf) * tab.nslotH, 256)),
                       dim3(256), 0, ctx->stream, nleaf, 0, F.box_keys.as<uint64_t>(), cells, 3 * D,
                       tab.offsH.as<int3>(), tab.nslotH, g, F.srcH.as<int32_t>());
    MF_LAUNCH_CHECK();
  }
  int nm2l = 0;
  for (int l = 2; l <= D; ++l) nm2l += cnt[l];
  F.ntilesF = F.ntilesH = 0;
  if (nm2l > 0) {
    MF_TRY(need<uint64_t>(F.class_keys, nm2l));
    MF_TRY(need<uint64_t>(F.class_keys2, nm2l));
    MF_TRY(need<int32_t>(F.class_ids, nm2l));
    MF_TRY(need<int32_t>(F.class_ids2, nm2l));
    MF_TRY(need<int32_t>(F.class_counts, 8 * (kMaxLevels + 1)));
    MF_HIP(hipMemsetAsync(F.class_counts.get(), 0, sizeof(int32_t) * 8 * (kMaxLevels + 1), ctx->stream));
    int off = 0;
    for (int l = 2; l <= D; ++l) {
      hipLaunchKernelGGL(class_key_kernel, dim3(grid_for(cnt[l], 256)), dim3(256), 0, ctx->stream,
                         cnt[l], F.base[l], l, F.box_keys.as<uint64_t>() + F.base[l],
                         F.class_keys.as<uint64_t>() + off, F.class_ids.as<int32_t>() + off,
                         F.class_counts.as<int32_t>());
      MF_LAUNCH_CHECK();
      off += cnt[l];
    }
    MF_TRY(sort_pairs_u64_i32(ctx, F.class_keys.as<uint64_t>(), F.class_keys2.as<uint64_t>(),
                              F.class_ids.as<int32_t>(), F.class_ids2.as<int32_t>(), nm2l, 0, 8));
    std::vector<int32_t> cc;
    MF_TRY(read_ints(ctx, F.class_counts.as<int32_t>(), 8 * (kMaxLevels + 1), cc));
    // the host tile descriptors live in the plan until its next build, so the
    // asynchronous copies below need no wait
    std::vector<M2LTile> &tf = F.h_tilesF, &th = F.h_tilesH;
    tf.clear();
    th.clear();
    int pos = 0;
    for (int l = 2; l <= D; ++l) {
      if (l == D) F.leaf_tile_start = static_cast<int>(th.size());
      for (int par = 0; par < 8; ++par) {
        const int c = cc[l * 8 + par];
        for (int s = 0; s < c; s += kTileT) {
          M2LTile t{pos + s, std::min(kTileT, c - s), par, tab.nslotF};
          tf.push_back(t);
          M2LTile u = t;
          if (l == D) {
            u.slots = 8 + par;
            u.nslot = tab.nslotH;
          }
          th.push_back(u);
        }
        pos += c;
      }
    }
    F.ntilesF = static_cast<int>(tf.size());
    F.ntilesH = static_cast<int>(th.size());
    MF_TRY(need<M2LTile>(F.tilesF, F.ntilesF));
    MF_TRY(need<M2LTile>(F.tilesH, F.ntilesH));
    MF_HIP(hipMemcpyAsync(F.tilesF.get(), tf.data(), sizeof(M2LTile) * tf.size(), hipMemcpyHostToDevice, ctx->stream));
    MF_HIP(hipMemcpyAsync(F.tilesH.get(), th.data(), sizeof(M2LTile) * th.size(), hipMemcpyHostToDevice, ctx->stream));
    // non-empty offsets of every tile
    const int ld = tab.slot_ld;
    if (ld > 1024) return Status::Err(MF_ERR_UNSUPPORTED, "offset list longer than 1024");
    MF_TRY(need<int32_t>(F.tlistF, static_cast<int64_t>(F.ntilesF) * ld));
    MF_TRY(need<int32_t>(F.tlistH, static_cast<int64_t>(F.ntilesH) * ld));
    MF_TRY(need<int32_t>(F.tnF, F.ntilesF));
    MF_TRY(need<int32_t>(F.tnH, F.ntilesH));
    auto tile_lists = [&](const M2LTile *tiles, int nt, const int32_t *src, int src_ld,
                          int32_t *list, int32_t *nlist) -> Status {
      if (nt <= 0) return Status::Ok();
      hipLaunchKernelGGL(tile_slots_kernel, dim3(nt), dim3(256), 0, ctx->stream, tiles,
                         F.class_ids2.as<int32_t>(), src, src_ld, tab.slot_tab.as<int32_t>(), ld,
                         ld, list, nlist);
      MF_LAUNCH_CHECK();
      return Status::Ok();
    };
    MF_TRY(tile_lists(F.tilesF.as<M2LTile>(), F.ntilesF, F.srcF.as<int32_t>(), tab.nslotF,
                      F.tlistF.as<int32_t>(), F.tnF.as<int32_t>()));
    const int lt = F.leaf_tile_start;
    MF_TRY(tile_lists(F.tilesH.as<M2LTile>(), lt, F.srcF.as<int32_t>(), tab.nslotF,
                      F.tlistH.as<int32_t>(), F.tnH.as<int32_t>()));
    MF_TRY(tile_lists(F.tilesH.as<M2LTile>() + lt, F.ntilesH - lt, F.srcH.as<int32_t>(), tab.nslotH,
                      F.tlistH.as<int32_t>() + static_cast<int64_t>(lt) * ld,
                      F.tnH.as<int32_t>() + lt));
    MF_TRY(need<int32_t>(F.orderF, std::max(F.ntilesF, 1)));
    MF_TRY(need<int32_t>(F.orderH, std::max(F.ntilesH, 1)));
    if (F.ntilesF > 0) {
      hipLaunchKernelGGL(tile_order_kernel, dim3(grid_for(F.ntilesF, 256)), dim3(256), 0, ctx->stream,
                         F.ntilesF, F.tnF.as<int32_t>(), F.orderF.as<int32_t>());
      MF_LAUNCH_CHECK();
    }
    if (F.ntilesH > 0) {
      hipLaunchKernelGGL(tile_order_kernel, dim3(grid_for(F.ntilesH, 256)), dim3(256), 0, ctx->stream,
                         F.ntilesH, F.tnH.as<int32_t>(), F.orderH.as<int32_t>());
      MF_LAUNCH_CHECK();
    }
    // M2M / L2L tiles
    MF_TRY(need<int32_t>(F.child_tab, static_cast<int64_t>(F.nbox + 1) * 8));
    MF_TRY(need<int32_t>(F.box_ids, F.nbox));
    MF_HIP(hipMemsetAsync(F.child_tab.get(), 0xff,
                          sizeof(int32_t) * 8 * static_cast<size_t>(F.nbox + 1), ctx->stream));
    // box_parent[nbox] = -1: the L2L's "no source" row (the array has room:
    // its capacity bounds the boxes of every level)
    hipLaunchKernelGGL(grid_clear_kernel, dim3(1), dim3(64), 0, ctx->stream, int64_t{1},
                       F.box_parent.as<int32_t>() + F.nbox);
    MF_LAUNCH_CHECK();
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(F.nbox, 256)), dim3(256), 0, ctx->stream, F.nbox,
                       F.box_ids.as<int32_t>());
    MF_LAUNCH_CHECK();
    for (int l = 3; l <= D; ++l) {
      hipLaunchKernelGGL(child_table_kernel, dim3(grid_for(cnt[l], 256)), dim3(256), 0, ctx->stream,
                         cnt[l], F.base[l], F.box_parent.as<int32_t>() + F.base[l],
                         F.box_keys.as<uint64_t>() + F.base[l], F.child_tab.as<int32_t>());
      MF_LAUNCH_CHECK();
    }
    std::vector<M2LTile> &tm = F.h_tilesM, &tlv = F.h_tilesL;
    tm.clear();
    tlv.clear();
    F.tileM_start.assign(D + 2, 0);
    F.tileL_start.assign(D + 2, 0);
    for (int l = 2; l <= D; ++l) {  // M2M targets: parents at levels 2 .. D - 1
      F.tileM_start[l] = static_cast<int>(tm.size());
      if (l < D)
        for (int s = 0; s < cnt[l]; s += kTileT)
          tm.push_back(M2LTile{F.base[l] + s, std::min(kTileT, cnt[l] - s), kRowM2M, 8});
    }
    F.tileM_start[D + 1] = static_cast<int>(tm.size());
    {
      int level = 2, pos_in = 0;  // walk the field tiles (level order) to find level bounds
      std::vector<int> tile_level(tf.size());
      int acc_boxes = 0;
      for (size_t k = 0; k < tf.size(); ++k) {
        while (level <= D && tf[k].first >= acc_boxes + cnt[level]) {
          acc_boxes += cnt[level];
          ++level;
        }
        tile_level[k] = level;
      }
      (void)pos_in;
      for (int l = 2; l <= D + 1; ++l) F.tileL_start[l] = 0;
      for (int l = 3; l <= D; ++l) {  // L2L targets: children at levels 3 .. D
        F.tileL_start[l] = static_cast<int>(tlv.size());
        for (size_t k = 0; k < tf.size(); ++k)
          if (tile_level[k] == l)
            tlv.push_back(M2LTile{tf[k].first, tf[k].count, kRowL2L + tf[k].slots, 1});
      }
      F.tileL_start[D + 1] = static_cast<int>(tlv.size());
    }
    MF_TRY(need<M2LTile>(F.tilesM, static_cast<int64_t>(tm.size())));
    MF_TRY(need<M2LTile>(F.tilesL, static_cast<int64_t>(tlv.size())));
    MF_TRY(need<int32_t>(F.tlistM, static_cast<int64_t>(tm.size()) * ld));
    MF_TRY(need<int32_t>(F.tnM, static_cast<int64_t>(tm.size())));
    MF_TRY(need<int32_t>(F.tlistL, static_cast<int64_t>(tlv.size()) * ld));
    MF_TRY(need<int32_t>(F.tnL, static_cast<int64_t>(tlv.size())));
    if (!tm.empty())
      MF_HIP(hipMemcpyAsync(F.tilesM.get(), tm.data(), sizeof(M2LTile) * tm.size(), hipMemcpyHostToDevice,
                            ctx->stream));
    if (!tlv.empty())
      MF_HIP(hipMemcpyAsync(F.tilesL.get(), tlv.data(), sizeof(M2LTile) * tlv.size(),
                            hipMemcpyHostToDevice, ctx->stream));
    if (!tm.empty()) {
      hipLaunchKernelGGL(tile_slots_kernel, dim3(static_cast<int>(tm.size())), dim3(256), 0, ctx->stream,
                         F.tilesM.as<M2LTile>(), F.box_ids.as<int32_t>(), F.child_tab.as<int32_t>(),
                         8, tab.slot_tab.as<int32_t>(), ld, ld, F.tlistM.as<int32_t>(),
                         F.tnM.as<int32_t>());
      MF_LAUNCH_CHECK();
    }
    if (!tlv.empty()) {
      hipLaunchKernelGGL(tile_slots_kernel, dim3(static_cast<int>(tlv.size())), dim3(256), 0,
                         ctx->stream, F.tilesL.as<M2LTile>(), F.class_ids2.as<int32_t>(),
                         F.box_parent.as<int32_t>(), 1, tab.slot_tab.as<int32_t>(), ld, ld,
                         F.tlistL.as<int32_t>(), F.tnL.as<int32_t>());
      MF_LAUNCH_CHECK();
    }
  }
  if (nsys > 1) {
    MF_TRY(need<int32_t>(F.box_sys, F.nbox));
    for (int l = 0; l <= D; ++l)
      if (cnt[l] > 0) {
        hipLaunchKernelGGL(box_sys_kernel, dim3(grid_for(cnt[l], 256)), dim3(256), 0, ctx->stream,
                           cnt[l], F.box_keys.as<uint64_t>() + F.base[l], 3 * l,
                           F.box_sys.as<int32_t>() + F.base[l]);
        MF_LAUNCH_CHECK();
      }
  }
  MF_TRY(need<float>(F.Mt, static_cast<int64_t>(F.nbox) * F.KP));
  MF_TRY(need<float>(F.Lt, static_cast<int64_t>(F.nbox) * F.KP));
  MF_TRY(need<uint16_t>(F.Mb, static_cast<int64_t>(F.nbox) * 3 * F.KP));
  F.counts_valid = false;  // work counts are computed on the first fmm_info
  return Status::Ok();
}

}  // namespace

// The leaf side: leaf_side * h when set; otherwise 2h, or with a particle
// budget leaf_size (the role of the reference's ncrit, NDTree.hpp:413-496,
// magnetic_force.cu:647-648) the largest 2h * 2^j whose non-empty leaves hold
// at most leaf_size particles on average.  Any side >= 2h keeps every pair
// with q < 2 (field) and q < 4 (force) inside the near sets.
Status fmm_build(mf_plan *plan) {
  const double h = plan->h;
  const float side = plan->opts.leaf_side;
  MF_ARG(side == 0.f || (side >= 2.f && side <= 1024.f), "leaf_side must be 0 or in [2, 1024] (units of h)");
  MF_ARG(plan->opts.leaf_size >= 0, "leaf_size must be >= 0");
  if (side > 0.f) return fmm_build_at(plan, static_cast<double>(side) * h);
  MF_TRY(fmm_build_at(plan, 2.0 * h));
  const FmmState &F = *plan->fmm;
  if (plan->opts.leaf_size <= 0 || plan->n == 0) return Status::Ok();
  int j = 0;
  for (int jj = 1; jj <= F.D; ++jj) {
    if (static_cast<double>(plan->n) / F.count[F.D - jj] > plan->opts.leaf_size) break;
    j = jj;
  }
  if (j == 0) return Status::Ok();
  return fmm_build_at(plan, 2.0 * h * std::ldexp(1.0, j));
}

Status fmm_counts(mf_plan *plan) {
  mf_ctx *ctx = plan->ctx;
  FmmState &F = *plan->fmm;
  const FmmTables &tab = *F.tab;
  const int D = F.D;
  const std::vector<int> &cnt = F.count;
  const int nleaf = F.nbox > 0 ? cnt[D] : 0;
  if (nleaf == 0) {
    F.p2p_field_pairs = F.p2p_force_pairs = F.m2l_pairs_field = F.m2l_pairs_force = 0;
    F.counts_valid = true;
    return Status::Ok();
  }
  {
    void *pc = nullptr;
    MF_TRY(workspace(ctx, "fmm_counts", sizeof(unsigned long long) * 4, &pc));
    unsigned long long *c = static_cast<unsigned long long *>(pc);
    MF_HIP(hipMemsetAsync(c, 0, sizeof(unsigned long long) * 4, ctx->stream));
    hipLaunchKernelGGL(count_p2p_kernel, dim3(grid_for(nleaf, 256)), dim3(256), 0, ctx->stream, nleaf,
                       F.box_count.as<int32_t>(), F.nearF.as<int32_t>(), kNearF, c);
    MF_LAUNCH_CHECK();
    hipLaunchKernelGGL(count_p2p_kernel, dim3(grid_for(nleaf, 256)), dim3(256), 0, ctx->stream, nleaf,
                       F.box_count.as<int32_t>(), F.nearH.as<int32_t>(), kNearH, c + 1);
    MF_LAUNCH_CHECK();
    for (int l = 2; l <= D; ++l) {
      hipLaunchKernelGGL(count_m2l_kernel, dim3(grid_for(cnt[l], 256)), dim3(256), 0, ctx->stream,
                         cnt[l], F.base[l], F.srcF.as<int32_t>(), tab.nslotF, c + 2);
      MF_LAUNCH_CHECK();
      if (l < D) {
        hipLaunchKernelGGL(count_m2l_kernel, dim3(grid_for(cnt[l], 256)), dim3(256), 0, ctx->stream,
                           cnt[l], F.base[l], F.srcF.as<int32_t>(), tab.nslotF, c + 3);
      } else {
        hipLaunchKernelGGL(count_m2l_kernel, dim3(grid_for(cnt[l], 256)), dim3(256), 0, ctx->stream,
                           cnt[l], 0, F.srcH.as<int32_t>(), tab.nslotH, c + 3);
      }
      MF_LAUNCH_CHECK();
    }
    unsigned long long hc[4];
    MF_HIP(hipMemcpyAsync(hc, c, sizeof(hc), hipMemcpyDeviceToHost, ctx->stream));
    MF_HIP(hipStreamSynchronize(ctx->stream));
    F.p2p_field_pairs = static_cast<int64_t>(hc[0]);
    F.p2p_force_pairs = static_cast<int64_t>(hc[1]);
    F.m2l_pairs_field = static_cast<int64_t>(hc[2]);
    F.m2l_pairs_force = static_cast<int64_t>(hc[3]);
  }
  {
    std::vector<int32_t> nf, nh;
    MF_TRY(read_ints(ctx, F.tnF.as<int32_t>(), F.ntilesF, nf));
    MF_TRY(read_ints(ctx, F.tnH.as<int32_t>(), F.ntilesH, nh));
    F.m2l_slots_field = F.m2l_slots_force = 0;
    for (int32_t v : nf) F.m2l_slots_field += v;
    for (int32_t v : nh) F.m2l_slots_force += v;
  }
  F.counts_valid = true;
  return Status::Ok();
}

Status fmm_info(mf_plan *plan, mf_plan_info *info) {
  FmmState &F = *plan->fmm;
  if (!F.counts_valid) MF_TRY(fmm_counts(plan));
  info->order = F.P;
  info->depth = F.D;
  info->leaves = F.count.empty() ? 0 : F.count[F.D];
  info->boxes = F.nbox;
  info->p2p_field_pairs = F.p2p_field_pairs;
  info->p2p_force_pairs = F.p2p_force_pairs;
  info->m2l_field_pairs = F.m2l_pairs_field;
  info->m2l_force_pairs = F.m2l_pairs_force;
  info->m2l_field_tile_slots = F.m2l_slots_field;
  info->m2l_force_tile_slots = F.m2l_slots_force;
  info->m2l_field_operators = F.tab ? F.tab->noffF : 0;
  return Status::Ok();
}

namespace {

template <int P, int KP>
Status launch_m2l(mf_ctx *ctx, FmmState &F, const FmmTables &tab, int prec, const M2LTile *tiles,
                  int ntiles, const int32_t *src, int src_rows, int nslot, const int32_t *tlist,
                  const int32_t *tn, SysMask mask, bool fence,
                  M2LSrc2 src2 = M2LSrc2{nullptr, 0, 0, INT_MAX, nullptr}) {
  if (ntiles <= 0) return Status::Ok();
  const dim3 grid(8 * ((ntiles + 7) / 8)), block(2 * KP);
  const int npart = prec == MF_M2L_FP32 ? 0 : 3;
  if (!m2l_gather_range_ok(F.nbox, KP, npart))
    return Status::Err(MF_ERR_UNSUPPORTED, "too many boxes for the M2L gather's 32-bit range");
  const uint32_t msrc_bytes = static_cast<uint32_t>(static_cast<int64_t>(F.nbox) * (npart ? 6 * KP : 4 * KP));
#define MF_M2L_ARGS                                                                           \
  ctx->stream, msrc, msrc_bytes, F.Lt.as<float>(), tables, tiles, ntiles,                     \
      F.class_ids2.as<int32_t>(), src, nslot, src_rows, tab.slot_tab.as<int32_t>(),            \
      tab.slot_ld, tlist,                                                                      \
      tn, tab.slot_ld, 0, mask, src2
#define MF_DEG_LAUNCH                                                                         \
  ctx->stream, F.Mb.get(), msrc_bytes, F.Lt.as<float>(), tab.Tb.as<uint16_t>(), tiles, ntiles, \
      F.class_ids2.as<int32_t>(), src, nslot, src_rows, tab.slot_tab.as<int32_t>(),            \
      tab.slot_ld, tlist, tn, tab.slot_ld, mask, src2
  switch (prec) {
    case MF_M2L_FP32: {
      const void *tables = tab.T.get();
      const void *msrc = F.Mt.get();
      if (fence) hipLaunchKernelGGL((m2l_kernel<KP, 0, false, true>), grid, block, 0, MF_M2L_ARGS);
      else hipLaunchKernelGGL((m2l_kernel<KP, 0>), grid, block, 0, MF_M2L_ARGS);
      break;
    }
    case MF_M2L_BF16X6: {
      const void *tables = tab.Tb.get();
      const void *msrc = F.Mb.get();
      if (fence) hipLaunchKernelGGL((m2l_kernel<KP, 3, false, true>), grid, block, 0, MF_M2L_ARGS);
      else hipLaunchKernelGGL((m2l_kernel<KP, 3>), grid, block, 0, MF_M2L_ARGS);
      break;
    }
    case MF_M2L_BF16X3:
      hipLaunchKernelGGL((m2l_deg_kernel<KP, P, true>), grid, block, 0, MF_DEG_LAUNCH);
      break;
    case MF_M2L_BF16X1: {
      const void *tables = tab.Tb.get();
      const void *msrc = F.Mb.get();
      if (fence) hipLaunchKernelGGL((m2l_kernel<KP, 1, false, true>), grid, block, 0, MF_M2L_ARGS);
      else hipLaunchKernelGGL((m2l_kernel<KP, 1>), grid, block, 0, MF_M2L_ARGS);
      break;
    }
    case MF_M2L_BF16X1_CUT:
      hipLaunchKernelGGL((m2l_deg_kernel<KP, P, false>), grid, block, 0, MF_DEG_LAUNCH);
      break;
    default:
      return Status::Err(MF_ERR_INVALID_ARG, "unknown M2L precision");
  }
#undef MF_M2L_ARGS
#undef MF_DEG_LAUNCH
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

// Upward pass, M2L and L2L on the context stream.  `launch_near` runs first:
// the products launch their near field there, on the auxiliary stream, so
// that the whole far-field chain (P2M, the M2M levels, whose launches are
// short and latency-bound, M2L and L2L) runs beside it.  Measured 1.8 ms per
// solve + force step faster on the 1.15 M pool than launching it just before
// the M2L.
template <int P, int KP>
Status upward_and_m2l(mf_plan *plan, bool hess, const std::function<Status()> &launch_near,
                      SysMask mask) {
  mf_ctx *ctx = plan->ctx;
  FmmState &F = *plan->fmm;
  const FmmTables &tab = *F.tab;
  const int D = F.D;
  const int nleaf = F.count[D];
  MF_TRY(launch_near());
  if (D < 2) return Status::Ok();
  {
    KernelTimer timer(ctx, "p2m");
    hipLaunchKernelGGL((p2m_kernel<P, KP>), dim3(nleaf), dim3(64), 0, ctx->stream, nleaf,
                       F.box_first.as<int32_t>(), F.box_count.as<int32_t>(),
                       F.box_keys.as<uint64_t>(), F.rec.as<SrcRec>(),
                       Frame{F.sys_lo.as<double>(), 3 * D}, F.s_leaf, F.Mt.as<float>(), mask);
    MF_LAUNCH_CHECK();
  }
  // M2M and L2L run on the M2L kernel, in the M2L's precision: FP32 MFMA, or
  // the BF16 parts of the octant matrices against BF16 parts of the source
  // expansions, which are split level by level just before they are read.
  const int prec = plan->m2l_precision();
  const int npart = prec == MF_M2L_FP32 ? 0
                    : prec == MF_M2L_BF16X6 ? 3
                    : prec == MF_M2L_BF16X3 ? 2
                                            : 1;
  auto split_level = [&](const float *src, int l) -> Status {
    const int64_t ne = static_cast<int64_t>(F.count[l]) * KP;
    if (ne == 0) return Status::Ok();
    const int64_t off = static_cast<int64_t>(F.base[l]) * KP;
    hipLaunchKernelGGL(split_multipoles_kernel, dim3(grid_for(ne, 256)), dim3(256), 0,
                       ctx->stream, ne, KP, src + off, F.Mb.as<uint16_t>() + 3 * off);
    MF_LAUNCH_CHECK();
    return Status::Ok();
  };
  auto octant_pass = [&](const void *msrc, float *out, const void *tables, const M2LTile *tiles, int nt,
                         const int32_t *boxes, const int32_t *src, int src_ld,
                         const int32_t *tlist, const int32_t *tn, int accumulate) -> Status {
    const dim3 grid(8 * ((nt + 7) / 8)), block(2 * KP);
    if (!m2l_gather_range_ok(F.nbox, KP, npart))
      return Status::Err(MF_ERR_UNSUPPORTED, "too many boxes for the M2L gather's 32-bit range");
    const uint32_t msrc_bytes =
        static_cast<uint32_t>(static_cast<int64_t>(F.nbox) * (npart ? 6 * KP : 4 * KP));
#define MF_OCTANT_ARGS                                                                          \
  ctx->stream, msrc, msrc_bytes, out, tables, tiles, nt, boxes, src, src_ld, F.nbox,          \
      tab.slot_tab.as<int32_t>(), tab.slot_ld, tlist, tn, tab.slot_ld, accumulate, mask,       \
      M2LSrc2{nullptr, 0, 0, INT_MAX, nullptr}
    if (npart == 0) hipLaunchKernelGGL((m2l_kernel<KP, 0, true>), grid, block, 0, MF_OCTANT_ARGS);
    else if (npart == 1) hipLaunchKernelGGL((m2l_kernel<KP, 1, true>), grid, block, 0, MF_OCTANT_ARGS);
    else if (npart == 2) hipLaunchKernelGGL((m2l_kernel<KP, 2, true>), grid, block, 0, MF_OCTANT_ARGS);
    else hipLaunchKernelGGL((m2l_kernel<KP, 3, true>), grid, block, 0, MF_OCTANT_ARGS);
#undef MF_OCTANT_ARGS
    MF_LAUNCH_CHECK();
    return Status::Ok();
  };
  {
    // M2M: parents of level l gather their children's multipoles per octant
    // and apply that octant's matrix.
    KernelTimer timer(ctx, "m2m");
    for (int l = D - 1; l >= 2; --l) {
      if (npart) MF_TRY(split_level(F.Mt.as<float>(), l + 1));
      const int t0 = F.tileM_start[l], nt = F.tileM_start[l + 1] - t0;
      if (nt <= 0) continue;
      MF_TRY(octant_pass(npart ? F.Mb.get() : F.Mt.get(), F.Mt.as<float>(),
                         npart ? tab.W_m2m_b.get() : tab.W_m2m.get(),
                         F.tilesM.as<M2LTile>() + t0, nt, F.box_ids.as<int32_t>(),
                         F.child_tab.as<int32_t>(), 8,
                         F.tlistM.as<int32_t>() + static_cast<int64_t>(t0) * tab.slot_ld,
                         F.tnM.as<int32_t>() + t0, 0));
    }
  }
  if (npart) {
    // levels D .. 3 were split for the M2M; the M2L also reads level 2
    KernelTimer timer(ctx, "split");
    MF_TRY(split_level(F.Mt.as<float>(), 2));
  }
  {
    // field M2L launches made at a precision other than the plan's own (the
    // relaxed products of a solve) are timed apart
    const bool own = plan->m2l_precision() == plan->opts.m2l_precision;
    KernelTimer timer(ctx, hess ? "m2l_force" : own ? "m2l" : "m2l_relaxed");
    if (!hess) {
      MF_TRY((launch_m2l<P, KP>(ctx, F, tab, prec, F.tilesF.as<M2LTile>(), F.ntilesF, F.srcF.as<int32_t>(),
                            F.nbox, tab.nslotF, F.tlistF.as<int32_t>(), F.tnF.as<int32_t>(), mask,
                            false, M2LSrc2{nullptr, 0, 0, INT_MAX, F.orderF.as<int32_t>()})));
    } else {
      // Tiles are in level order: the leaf level's (trailing) tiles read the
      // force lists srcH (indexed by leaf = global index), the others srcF;
      // one launch, so the upper levels' few tiles run beside the leaves'
      const int lt = F.leaf_tile_start;
      MF_TRY((launch_m2l<P, KP>(ctx, F, tab, prec, F.tilesH.as<M2LTile>(), F.ntilesH,
                            F.srcF.as<int32_t>(), F.nbox, tab.nslotF, F.tlistH.as<int32_t>(),
                            F.tnH.as<int32_t>(), mask, true,
                            M2LSrc2{F.srcH.as<int32_t>(), tab.nslotH, F.count[D], lt,
                                    F.orderH.as<int32_t>()})));
    }
  }
  {
    // L2L (accumulating): the children of level l (tiles of one octant class)
    // add their parent's local expansion.  With BF16 parts, the parents'
    // locals are split into the multipole-part buffer, which the M2L no
    // longer needs.
    KernelTimer timer(ctx, "l2l");
    for (int l = 3; l <= D; ++l) {
      const int t0 = F.tileL_start[l], nt = F.tileL_start[l + 1] - t0;
      if (nt <= 0) continue;
      if (npart) MF_TRY(split_level(F.Lt.as<float>(), l - 1));
      MF_TRY(octant_pass(npart ? F.Mb.get() : F.Lt.get(), F.Lt.as<float>(),
                         npart ? tab.W_l2l_b.get() : tab.W_l2l.get(), F.tilesL.as<M2LTile>() + t0,
                         nt, F.class_ids2.as<int32_t>(), F.box_parent.as<int32_t>(), 1,
                         F.tlistL.as<int32_t>() + static_cast<int64_t>(t0) * tab.slot_ld,
                         F.tnL.as<int32_t>() + t0, 1));
    }
  }
  return Status::Ok();
}

}  // namespace

}  // namespace mf

namespace mf {
namespace {

template <int P, int KP>
Status apply_field_impl(mf_plan *plan, float scale, const float *px, const float *py,
                        const float *pz, const float *hconst, float *ox, float *oy, float *oz) {
  mf_ctx *ctx = plan->ctx;
  FmmState &F = *plan->fmm;
  const int64_t n = F.n;
  if (n == 0) return Status::Ok();
  hipLaunchKernelGGL(pack_sorted_moments_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream,
                     n, F.perm.as<int32_t>(), px, py, pz, scale, plan->weights, F.rec.as<SrcRec>());
  MF_LAUNCH_CHECK();
  const int D = F.D, nleaf = F.count[D];
  const float h0 = hconst ? hconst[0] : 0.f, h1 = hconst ? hconst[1] : 0.f,
              h2 = hconst ? hconst[2] : 0.f;
  const int use_far = D >= 2 ? 1 : 0;
  SysMask mask{};
  if (plan->active_sys && F.nsys > 1) mask = SysMask{F.box_sys.as<int32_t>(), plan->active_sys};
  double *far3 = nullptr;
  if (use_far) {
    void *pf = nullptr;
    MF_TRY(workspace(ctx, "fmm_far3", sizeof(double) * 3 * static_cast<size_t>(n), &pf));
    far3 = static_cast<double *>(pf);
  }
  {
    // The near field (VALU) runs on the auxiliary stream, concurrently with
    // the far-field chain (P2M .. L2L, L2P); it writes H_const + near, the
    // L2P writes the far field into FP64 scratch, and add_far_kernel sums
    // them once both are done.
    NearOverlap ov(ctx);
    const bool serial = plan->opts.serial_near != 0;
    auto near = [&]() -> Status {
      hipStream_t ns = ctx->stream;
      if (!serial) {
        MF_TRY(ov.fork());
        ns = ctx->aux_stream;
      }
      KernelTimer timer(ctx, "p2p_field", ns);
      const int32_t *lists = F.leaf_lists.as<int32_t>();
      {
        const int nfull = F.leaf_list_n[0], nrem = F.leaf_lists_remainder();
        const int grid = 8 * ((nrem + 7) / 8 + (nfull + 7) / 8);
        if (grid > 0) {
          hipLaunchKernelGGL((p2p_field_kernel<64>), dim3(grid), dim3(64), 0, ns, lists, nfull,
                             lists + F.leaf_list_off[1], nrem, F.box_first.as<int32_t>(),
                             F.box_count.as<int32_t>(), F.rec.as<SrcRec>(), F.perm.as<int32_t>(),
                             F.rngF.as<int2>(), F.nrF.as<int2>(), kNearF, plan->fc,
                             F.ftab->buf.as<float4>(), FieldOut{0, h0, h1, h2, ox, oy, oz}, mask);
          MF_LAUNCH_CHECK();
        }
      }
      timer.stop();
      return serial ? Status::Ok() : ov.record();
    };
    MF_TRY((upward_and_m2l<P, KP>(plan, false, near, mask)));
    if (use_far) {
      // the far part into FP64 scratch on the context stream, beside the
      // near field; added once both are done
      KernelTimer timer(ctx, "l2p_field");
      hipLaunchKernelGGL((l2p_field_kernel<P, KP>), dim3(nleaf), dim3(64), 0, ctx->stream, nleaf,
                         F.box_first.as<int32_t>(), F.box_count.as<int32_t>(),
                         F.box_keys.as<uint64_t>(), F.rec.as<SrcRec>(), F.perm.as<int32_t>(),
                         F.Lt.as<float>(), Frame{F.sys_lo.as<double>(), 3 * D}, F.s_leaf, n, far3,
                         mask);
      MF_LAUNCH_CHECK();
    }
    MF_TRY(ov.join());
  }
  if (use_far) {
    hipLaunchKernelGGL(add_far_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, far3,
                       ox, oy, oz,
                       mask.active ? plan->sys_of.as<int32_t>() : static_cast<const int32_t *>(nullptr),
                       mask.active);
    MF_LAUNCH_CHECK();
  }
  return Status::Ok();
}

template <int P, int KP>
Status apply_force_impl(mf_plan *plan, const float *mx, const float *my, const float *mz,
                        const float *const *text9, float *fx, float *fy, float *fz) {
  mf_ctx *ctx = plan->ctx;
  FmmState &F = *plan->fmm;
  const int64_t n = F.n;
  if (n == 0) return Status::Ok();
  hipLaunchKernelGGL(pack_sorted_moments_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream,
                     n, F.perm.as<int32_t>(), mx, my, mz, 1.f, static_cast<const float *>(nullptr),
                     F.rec.as<SrcRec>());
  MF_LAUNCH_CHECK();
  Text9 t{};
  int has_text = 0;
  if (text9) {
    has_text = 1;
    for (int k = 0; k < 9; ++k) {
      MF_ARG(text9[k] != nullptr, "text9 entries must be non-null device pointers");
      t.p[k] = text9[k];
    }
  }
  double *far3 = nullptr;
  {
    void *pf = nullptr;
    MF_TRY(workspace(ctx, "fmm_far3", sizeof(double) * 3 * static_cast<size_t>(n), &pf));
    far3 = static_cast<double *>(pf);
  }
  {
    // near force on the auxiliary stream (writes), concurrently with the
    // far-field chain; the L2P writes the far part and T_ext . m into FP64
    // scratch, and add_far_kernel sums them once both are done
    NearOverlap ov(ctx);
    const bool serial = plan->opts.serial_near != 0;
    auto near = [&]() -> Status {
      hipStream_t ns = ctx->stream;
      if (!serial) {
        MF_TRY(ov.fork());
        ns = ctx->aux_stream;
      }
      KernelTimer timer(ctx, "p2p_force", ns);
      const int32_t *lists = F.leaf_lists.as<int32_t>();
      {
        const int nfull = F.leaf_list_n[0], nrem = F.leaf_lists_remainder();
        const int grid = 8 * ((nrem + 7) / 8 + (nfull + 7) / 8);
        if (grid > 0) {
          hipLaunchKernelGGL((p2p_force_kernel<64>), dim3(grid), dim3(64), 0, ns, lists, nfull,
                             lists + F.leaf_list_off[1], nrem, F.box_first.as<int32_t>(),
                             F.box_count.as<int32_t>(), F.rec.as<SrcRec>(), F.perm.as<int32_t>(),
                             F.rngH.as<int2>(), F.nrH.as<int2>(), kNearH, plan->hc,
                             F.htab->buf.as<float4>(), ForceOut{0, fx, fy, fz});
          MF_LAUNCH_CHECK();
        }
      }
      timer.stop();
      return serial ? Status::Ok() : ov.record();
    };
    MF_TRY((upward_and_m2l<P, KP>(plan, true, near, SysMask{})));
    {
      // far part + T_ext . m into FP64 scratch, beside the near field
      KernelTimer timer(ctx, "l2p_force");
      const int D = F.D, nleaf = F.count[D];
      hipLaunchKernelGGL((l2p_force_kernel<P, KP>), dim3(nleaf), dim3(64), 0, ctx->stream, nleaf,
                         F.box_first.as<int32_t>(), F.box_count.as<int32_t>(),
                         F.box_keys.as<uint64_t>(), F.rec.as<SrcRec>(), F.perm.as<int32_t>(),
                         F.Lt.as<float>(), D >= 2 ? 1 : 0, Frame{F.sys_lo.as<double>(), 3 * D},
                         F.s_leaf, t, has_text, n, far3);
      MF_LAUNCH_CHECK();
    }
    MF_TRY(ov.join());
  }
  hipLaunchKernelGGL(add_far_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, far3,
                     fx, fy, fz, static_cast<const int32_t *>(nullptr),
                     static_cast<const int32_t *>(nullptr));
  MF_LAUNCH_CHECK();
  return Status::Ok();
}

}  // namespace

Status fmm_apply_G(mf_plan *plan, float scale, const float *px, const float *py, const float *pz,
                   const float *hconst, float *ox, float *oy, float *oz) {
  if (!plan->fmm) return Status::Err(MF_ERR_INVALID_ARG, "FMM plan not built");
  switch (plan->fmm->P) {
    case 6: return apply_field_impl<6, 64>(plan, scale, px, py, pz, hconst, ox, oy, oz);
    case 8: return apply_field_impl<8, 96>(plan, scale, px, py, pz, hconst, ox, oy, oz);
    case 10: return apply_field_impl<10, 128>(plan, scale, px, py, pz, hconst, ox, oy, oz);
    default: return Status::Err(MF_ERR_INVALID_ARG, "unsupported FMM order");
  }
}

Status fmm_force(mf_plan *plan, const float *mx, const float *my, const float *mz,
                 const float *const *text9, float *fx, float *fy, float *fz) {
  if (!plan->fmm) return Status::Err(MF_ERR_INVALID_ARG, "FMM plan not built");
  switch (plan->fmm->P) {
    case 6: return apply_force_impl<6, 64>(plan, mx, my, mz, text9, fx, fy, fz);
    case 8: return apply_force_impl<8, 96>(plan, mx, my, mz, text9, fx, fy, fz);
    case 10: return apply_force_impl<10, 128>(plan, mx, my, mz, text9, fx, fy, fz);
    default: return Status::Err(MF_ERR_INVALID_ARG, "unsupported FMM order");
  }
}

}  // namespace mf
