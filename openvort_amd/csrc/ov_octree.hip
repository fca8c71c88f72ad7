// ov_octree.hip — Barnes-Hut octree on the device, cell for cell the tree of
// the reference (octree.c:79-343).
//
// Build:
//  1. k_oct_keys: every point descends from the domain box, computing each
//     child box with the reference's arithmetic (octree.c:158-235) and taking
//     the first child, in child order, that contains it strictly
//     (vec3_maths.c:67-77); the child indices form a path of 4-bit digits
//     (8 = contained in no child: dropped at that level, octree.c:118-127).
//  2. Paths are sorted stably (points enter in ascending tangle index), so
//     every cell is a contiguous range.
//  3. Cells are built level by level: a cell with more than one point whose
//     shortest edge is >= BH_grain * dl_max (octree.c:99-108) gets one child
//     per non-empty digit group.
//  4. Moments per cell (octree.c:237-266): centre = (sum r) / n, circulation
//     Gamma = sum ds, tensor M_ij = sum ds_i (r - centre)_j (one wavefront per
//     cell; summation order differs from the reference's by rounding only).
//  5. Cells are renumbered in preorder; `skip` = subtree size.
// Traversal (octree.c:275-343), one lane per (target, image) pair, a few
// consecutive targets times a group of images per wavefront: the wave walks
// nodes in preorder; a lane that takes a node (empty, leaf or accepted
// multipole) waits until the wave leaves that subtree; the wave descends
// when any active lane opens the node and otherwise jumps over it.
#include <hip/hip_runtime.h>

#include "ov_device.h"

namespace ov {

constexpr int OCT_MAXD = 48;       // path digits (3 x 16 nibbles)

struct OctCell {                   // build-time cell, breadth-first order
  double lo[3], hi[3];
  int start, count;                // range in the sorted point order
  int level, parent;
  int child_first, n_child;
};

// Traversal node in preorder, split in two: what the walk reads at every
// visit (one 64-byte line, so one scalar load) and the moments, read only
// where a node is taken as a multipole.
struct alignas(64) OctNode {
  double com[3];
  double lmax, lmin;
  int n, first, skip;
  float q;   // (lmax / theta)^2 rounded to FP32: the FP32 opening test's threshold on d^2
};
struct OctMom {
  double gam[3];
  double M[9];
};
// The same moments rounded to FP32 (the FP32-multipole traversal), written
// over the FP64 array's storage (48 of its 96 bytes per node) when the tree is
// built for that traversal.
struct alignas(16) OctMom32 {
  float gam[3];
  float M[9];
};

struct LeafSeg {                   // segment of a sorted point
  double a[3], b[3];
  int m, fwd;
};

struct OctParams {
  BoxDev box;
  double grain;                    // BH_grain (leaf test, octree.c:289)
  double grain_dl;                 // BH_grain * dl_max (split test, octree.c:104)
  double theta;                    // BH_resolution
  int quad;
  int dipole;                      // 1: true dipole instead of the reference form
};

// Child-box faces (octree.c:158-235): per child (BLF, BLB, BRF, BRB, TLF,
// TLB, TRF, TRB) and axis, the low face is one of {P.lo, P.lo + L/2, c,
// c - L/2} and the high face one of {c, c + L/2, P.hi, P.hi - L/2}.
__constant__ unsigned char oct_lo_code[8][3] = {
    {1, 0, 0}, {0, 0, 0}, {2, 2, 3}, {0, 1, 0}, {2, 3, 2}, {0, 0, 1}, {2, 2, 2}, {3, 2, 2}};
__constant__ unsigned char oct_hi_code[8][3] = {
    {1, 0, 0}, {0, 0, 0}, {2, 2, 3}, {0, 1, 0}, {2, 3, 2}, {0, 0, 1}, {2, 2, 2}, {3, 2, 2}};

// (lo, hi may alias plo, phi.)
__device__ inline void oct_child_box(const double plo[3], const double phi[3], int ci, int quad,
                                     double lo[3], double hi[3]) {
#pragma clang fp contract(off)
  double nlo[3], nhi[3];
  for (int i = 0; i < 3; ++i) {
    double c = plo[i] + phi[i];
    c = c * 0.5;
    double L = phi[i] - plo[i];
    switch (oct_lo_code[ci][i]) {
      case 0: nlo[i] = plo[i]; break;
      case 1: nlo[i] = plo[i] + L / 2; break;
      case 2: nlo[i] = c; break;
      default: nlo[i] = c - L / 2; break;
    }
    switch (oct_hi_code[ci][i]) {
      case 0: nhi[i] = c; break;
      case 1: nhi[i] = c + L / 2; break;
      case 2: nhi[i] = phi[i]; break;
      default: nhi[i] = phi[i] - L / 2; break;
    }
  }
  if (quad) nhi[2] = phi[2];
  for (int i = 0; i < 3; ++i) {
    lo[i] = nlo[i];
    hi[i] = nhi[i];
  }
}

__device__ inline double oct_min_len(const double lo[3], const double hi[3]) {
  double Lx = hi[0] - lo[0], Ly = hi[1] - lo[1], Lz = hi[2] - lo[2];
  if (Lx < Ly) return Lx < Lz ? Lx : Lz;
  return Ly < Lz ? Ly : Lz;
}
__device__ inline double oct_max_len(const double lo[3], const double hi[3]) {
  double Lx = hi[0] - lo[0], Ly = hi[1] - lo[1], Lz = hi[2] - lo[2];
  if (Lx > Ly) return Lx > Lz ? Lx : Lz;
  return Ly > Lz ? Ly : Lz;
}

__device__ inline int oct_digit(const uint64_t* __restrict__ k0, const uint64_t* __restrict__ k1,
                                const uint64_t* __restrict__ k2, int j, int level0) {
  // digit of level level0 + 1 (0-based nibble index level0)
  const int w = level0 >> 4, sh = 60 - 4 * (level0 & 15);
  const uint64_t v = w == 0 ? k0[j] : (w == 1 ? k1[j] : k2[j]);
  return (int)((v >> sh) & 15);
}

// 1. Path of every point.  Point i of the ascending-tangle-index order is
// stream point order[i].
__global__ void __launch_bounds__(256) k_oct_keys(const double* __restrict__ P, long cap,
                                                  const int* __restrict__ order, int n,
                                                  OctParams op, uint64_t* __restrict__ k0,
                                                  uint64_t* __restrict__ k1,
                                                  uint64_t* __restrict__ k2,
                                                  int* __restrict__ vals, int* __restrict__ maxdepth) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int s = order[i];
  const double r[3] = {P[s], P[cap + s], P[2 * cap + s]};
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = op.box.lo[a];
    hi[a] = op.box.hi[a];
  }
  uint64_t w[3] = {0, 0, 0};
  const int nch = op.quad ? 4 : 8;
  int depth = 0;
  while (depth < OCT_MAXD && !(oct_min_len(lo, hi) < op.grain_dl)) {
    int found = 8;
    double clo[3], chi[3];
    for (int ci = 0; ci < nch; ++ci) {
      oct_child_box(lo, hi, ci, op.quad, clo, chi);
      if (r[0] > clo[0] && r[0] < chi[0] && r[1] > clo[1] && r[1] < chi[1] && r[2] > clo[2] &&
          r[2] < chi[2]) {
        found = ci;
        break;
      }
    }
    w[depth >> 4] |= (uint64_t)found << (60 - 4 * (depth & 15));
    ++depth;
    if (found == 8) break;
    oct_child_box(lo, hi, found, op.quad, lo, hi);
  }
  k0[i] = w[0];
  k1[i] = w[1];
  k2[i] = w[2];
  vals[i] = i;
  atomicMax(maxdepth, depth);
}

__global__ void __launch_bounds__(256) k_gather_u64(const uint64_t* __restrict__ src,
                                                    const int* __restrict__ idx, int n,
                                                    uint64_t* __restrict__ dst) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// Sorted position j -> stream point, and the point's segment (the segment
// of octree cells, octree.c:253-258 and tangle.c:519-528).  A point without
// a forward link (the pinned end of an open chain) has none: its entry is
// the zero-length segment a -> a, which adds exactly nothing to the cell
// moments (the reference skips it, octree.c:254-255) and whose field is
// exactly zero by the collinearity rule (segment_field returns 0 for it).
__global__ void __launch_bounds__(256) k_oct_leafsegs(const double* __restrict__ P, long cap,
                                                      const int* __restrict__ fwd,
                                                      const int* __restrict__ order,
                                                      const int* __restrict__ perm, int n,
                                                      BoxDev box, int* __restrict__ sorted_s,
                                                      LeafSeg* __restrict__ segs) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const int s = order[perm[j]];
  sorted_s[j] = s;
  LeafSeg q;
  const int f = fwd[s];
  q.a[0] = P[s];
  q.a[1] = P[cap + s];
  q.a[2] = P[2 * cap + s];
  if (f >= 0) {
    const double pf[3] = {P[f], P[cap + f], P[2 * cap + f]};
    pwrap(q.a, pf, box, q.b);
  } else {
    q.b[0] = q.a[0];
    q.b[1] = q.a[1];
    q.b[2] = q.a[2];
  }
  q.m = s;
  q.fwd = f;
  segs[j] = q;
}

// First position in [lo, hi) whose digit (level0) is >= d.
__device__ inline int oct_lower_bound(const uint64_t* k0, const uint64_t* k1, const uint64_t* k2,
                                      int lo, int hi, int level0, int d) {
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (oct_digit(k0, k1, k2, mid, level0) < d)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ inline bool oct_splits(const OctCell& c, const OctParams& op) {
  return c.count > 1 && !(oct_min_len(c.lo, c.hi) < op.grain_dl) && c.level < OCT_MAXD;
}

// 3a. Children of each cell of one level: the number of non-empty digit
// groups below `nch` (the dropped group makes no cell).
__global__ void __launch_bounds__(256) k_oct_count(const OctCell* __restrict__ cells, int ls,
                                                   int le, const uint64_t* __restrict__ k0,
                                                   const uint64_t* __restrict__ k1,
                                                   const uint64_t* __restrict__ k2, OctParams op,
                                                   int* __restrict__ nchild) {
  const int c = ls + blockIdx.x * 256 + threadIdx.x;
  if (c >= le) return;
  const OctCell cell = cells[c];
  int cnt = 0;
  if (oct_splits(cell, op)) {
    const int nch = op.quad ? 4 : 8;
    const int e = cell.start + cell.count;
    int b = cell.start;
    for (int d = 0; d < nch; ++d) {
      int nb = oct_lower_bound(k0, k1, k2, b, e, cell.level, d + 1);
      if (nb > b) ++cnt;
      b = nb;
    }
  }
  nchild[c - ls] = cnt;
}

// 3b. Write the children of one level at next + offset (offsets = exclusive
// scan of the counts).
__global__ void __launch_bounds__(256) k_oct_emit(OctCell* __restrict__ cells, int ls, int le,
                                                  int next, const uint64_t* __restrict__ k0,
                                                  const uint64_t* __restrict__ k1,
                                                  const uint64_t* __restrict__ k2, OctParams op,
                                                  const int* __restrict__ offs) {
  const int c = ls + blockIdx.x * 256 + threadIdx.x;
  if (c >= le) return;
  OctCell cell = cells[c];
  cell.child_first = -1;
  cell.n_child = 0;
  if (oct_splits(cell, op)) {
    const int nch = op.quad ? 4 : 8;
    const int e = cell.start + cell.count;
    int b = cell.start;
    int out = next + offs[c - ls];
    cell.child_first = out;
    for (int d = 0; d < nch; ++d) {
      int nb = oct_lower_bound(k0, k1, k2, b, e, cell.level, d + 1);
      if (nb > b) {
        OctCell ch;
        oct_child_box(cell.lo, cell.hi, d, op.quad, ch.lo, ch.hi);
        ch.start = b;
        ch.count = nb - b;
        ch.level = cell.level + 1;
        ch.parent = c;
        ch.child_first = -1;
        ch.n_child = 0;
        cells[out++] = ch;
        cell.n_child++;
      }
      b = nb;
    }
  }
  cells[c].child_first = cell.child_first;
  cells[c].n_child = cell.n_child;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// 4. Moments of every cell over its points (octree.c:237-266), one
// wavefront per cell: lane sums over a stride of 64, then a fixed butterfly
// (deterministic; the order differs from the reference's sequential sum by
// rounding only).  M_ij = sum ds_i (r - centre)_j with r the segment's start
// (the reference) or, for the dipole expansion (midpoint != 0), its midpoint:
// the first-order term of a straight segment's field is taken about the
// midpoint, and start points would leave a first-order error.
// Cells of more than OCT_BIG points are only listed here (big, n_big) and
// summed by the k_oct_big_* kernels, OCT_SLICES blocks per cell.
constexpr int OCT_BIG = 2048;
constexpr int OCT_SLICES = 64;

__global__ void __launch_bounds__(256) k_oct_moments(const OctCell* __restrict__ cells, int ncell,
                                                     const LeafSeg* __restrict__ segs,
                                                     double* __restrict__ mom, int* __restrict__ big,
                                                     int* __restrict__ n_big, int midpoint) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ncell) return;
  const OctCell cell = cells[c];
  if (cell.count > OCT_BIG) {
    if (lane == 0) big[atomicAdd(n_big, 1)] = c;
    return;
  }
  const int e = cell.start + cell.count;
  double s[3] = {0, 0, 0};
  for (int j = cell.start + lane; j < e; j += 64)
    for (int a = 0; a < 3; ++a) s[a] += segs[j].a[a];
  double com[3];
  for (int a = 0; a < 3; ++a) com[a] = wave_sum(s[a]) * (1.0 / cell.count);
  double g[3] = {0, 0, 0}, M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = cell.start + lane; j < e; j += 64) {
    const LeafSeg q = segs[j];
    double ds[3], sl[3];
    for (int a = 0; a < 3; ++a) {
      ds[a] = q.b[a] - q.a[a];
      sl[a] = q.a[a] - com[a];
      g[a] += ds[a];
    }
    if (midpoint)
      for (int a = 0; a < 3; ++a) sl[a] = sl[a] + 0.5 * ds[a];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) M[3 * a + b] += ds[a] * sl[b];
  }
  for (int a = 0; a < 3; ++a) g[a] = wave_sum(g[a]);
  for (int a = 0; a < 9; ++a) M[a] = wave_sum(M[a]);
  if (lane == 0) {
    double* o = mom + 15L * c;
    for (int a = 0; a < 3; ++a) {
      o[a] = com[a];
      o[3 + a] = g[a];
    }
    for (int a = 0; a < 9; ++a) o[6 + a] = M[a];
  }
}

// Block-wide sum of one value per thread (256 threads): the four wave sums
// added in wave order (deterministic).
__device__ inline double block_sum256(double v, double* lds) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  return ((lds[0] + lds[1]) + lds[2]) + lds[3];
}

// Big cells, pass 1: per slice the sums of the points and of the segment
// vectors (part[cell][slice][6]); slice x covers [start + x count / S,
// start + (x + 1) count / S).
__global__ void __launch_bounds__(256) k_oct_big_sum1(const OctCell* __restrict__ cells,
                                                      const int* __restrict__ big,
                                                      const int* __restrict__ n_big,
                                                      const LeafSeg* __restrict__ segs,
                                                      double* __restrict__ part) {
  __shared__ double lds[4];
  const int y = blockIdx.y;
  if (y >= *n_big) return;
  const OctCell cell = cells[big[y]];
  const long b = cell.start + (long)cell.count * blockIdx.x / OCT_SLICES;
  const long e = cell.start + (long)cell.count * (blockIdx.x + 1) / OCT_SLICES;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (long j = b + threadIdx.x; j < e; j += 256) {
    const LeafSeg q = segs[j];
    for (int a = 0; a < 3; ++a) {
      s[a] += q.a[a];
      s[3 + a] += q.b[a] - q.a[a];
    }
  }
  double* o = part + ((long)y * OCT_SLICES + blockIdx.x) * 15;
  for (int a = 0; a < 6; ++a) {
    const double t = block_sum256(s[a], lds);
    if (threadIdx.x == 0) o[a] = t;
  }
}

// Big cells, pass 2: per slice the moment about the cell's centre of mass,
// which every block derives from pass 1's slices summed in slice order.
__global__ void __launch_bounds__(256) k_oct_big_sum2(const OctCell* __restrict__ cells,
                                                      const int* __restrict__ big,
                                                      const int* __restrict__ n_big,
                                                      const LeafSeg* __restrict__ segs,
                                                      double* __restrict__ part, int midpoint) {
  __shared__ double lds[4];
  const int y = blockIdx.y;
  if (y >= *n_big) return;
  const OctCell cell = cells[big[y]];
  const double* p = part + (long)y * OCT_SLICES * 15;
  double com[3];
  for (int a = 0; a < 3; ++a) {
    double t = 0;
    for (int x = 0; x < OCT_SLICES; ++x) t += p[15 * x + a];
    com[a] = t * (1.0 / cell.count);
  }
  const long b = cell.start + (long)cell.count * blockIdx.x / OCT_SLICES;
  const long e = cell.start + (long)cell.count * (blockIdx.x + 1) / OCT_SLICES;
  double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (long j = b + threadIdx.x; j < e; j += 256) {
    const LeafSeg q = segs[j];
    double ds[3], sl[3];
    for (int a = 0; a < 3; ++a) {
      ds[a] = q.b[a] - q.a[a];
      sl[a] = q.a[a] - com[a];
    }
    if (midpoint)
      for (int a = 0; a < 3; ++a) sl[a] = sl[a] + 0.5 * ds[a];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) M[3 * a + c] += ds[a] * sl[c];
  }
  // pass 1 wrote fields 0..5 of every slice, this pass writes 6..14 of its own
  double* o = part + ((long)y * OCT_SLICES + blockIdx.x) * 15;
  for (int a = 0; a < 9; ++a) {
    const double t = block_sum256(M[a], lds);
    if (threadIdx.x == 0) o[6 + a] = t;
  }
}

// Big cells: slices summed in order into the cell's moments (one thread each).
__global__ void __launch_bounds__(64) k_oct_big_fin(const OctCell* __restrict__ cells,
                                                    const int* __restrict__ big,
                                                    const int* __restrict__ n_big,
                                                    const double* __restrict__ part,
                                                    double* __restrict__ mom) {
  const int y = blockIdx.x * 64 + threadIdx.x;
  if (y >= *n_big) return;
  const int c = big[y];
  const double* p = part + (long)y * OCT_SLICES * 15;
  double v[15];
  for (int a = 0; a < 15; ++a) {
    double t = 0;
    for (int x = 0; x < OCT_SLICES; ++x) t += p[15 * x + a];
    v[a] = t;
  }
  double* o = mom + 15L * c;
  for (int a = 0; a < 3; ++a) {
    o[a] = v[a] * (1.0 / cells[c].count);
    o[3 + a] = v[3 + a];
  }
  for (int a = 0; a < 9; ++a) o[6 + a] = v[6 + a];
}

// 5a. Subtree sizes, one level at a time from the deepest.
__global__ void __launch_bounds__(256) k_oct_size(const OctCell* __restrict__ cells, int ls, int le,
                                                  int* __restrict__ size) {
  const int c = ls + blockIdx.x * 256 + threadIdx.x;
  if (c >= le) return;
  const OctCell cell = cells[c];
  int sz = 1;
  for (int k = 0; k < cell.n_child; ++k) sz += size[cell.child_first + k];
  size[c] = sz;
}

// 5b. Preorder index of the children of one level's cells.
__global__ void __launch_bounds__(256) k_oct_pre(const OctCell* __restrict__ cells, int ls, int le,
                                                 const int* __restrict__ size,
                                                 int* __restrict__ pre) {
  const int c = ls + blockIdx.x * 256 + threadIdx.x;
  if (c >= le) return;
  const OctCell cell = cells[c];
  int p = pre[c] + 1;
  for (int k = 0; k < cell.n_child; ++k) {
    pre[cell.child_first + k] = p;
    p += size[cell.child_first + k];
  }
}

// 5c. Traversal nodes in preorder.
__global__ void __launch_bounds__(256) k_oct_nodes(const OctCell* __restrict__ cells, int ncell,
                                                   const double* __restrict__ mom,
                                                   const int* __restrict__ size,
                                                   const int* __restrict__ pre,
                                                   OctNode* __restrict__ nodes,
                                                   OctMom* __restrict__ nmom, double theta,
                                                   int fp32_moments) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c > ncell) return;
  OctNode nd;
  if (c == ncell) {   // past the last node: read by the walk's look-ahead only
    nd.com[0] = nd.com[1] = nd.com[2] = 0;
    nd.lmax = nd.lmin = 0;
    nd.n = nd.first = 0;
    nd.q = 0.f;
    nd.skip = 1;
    nodes[ncell] = nd;
    return;
  }
  const OctCell cell = cells[c];
  OctMom mm;
  const double* m = mom + 15L * c;
  for (int a = 0; a < 3; ++a) {
    nd.com[a] = m[a];
    mm.gam[a] = m[3 + a];
  }
  for (int a = 0; a < 9; ++a) mm.M[a] = m[6 + a];
  nd.lmax = oct_max_len(cell.lo, cell.hi);
  nd.lmin = oct_min_len(cell.lo, cell.hi);
  nd.n = cell.count;
  nd.first = cell.start;
  nd.skip = size[c];
  // theta = 0 gives +inf: never accepted, as lmax / d < 0 never holds
  nd.q = (float)(nd.lmax * nd.lmax / (theta * theta));
  nodes[pre[c]] = nd;
  if (fp32_moments) {
    OctMom32 m32;
    for (int a = 0; a < 3; ++a) m32.gam[a] = (float)mm.gam[a];
    for (int a = 0; a < 9; ++a) m32.M[a] = (float)mm.M[a];
    reinterpret_cast<OctMom32*>(nmom)[pre[c]] = m32;
  } else {
    nmom[pre[c]] = mm;
  }
}

// ---------------------------------------------------------------------------
// traversal
// ---------------------------------------------------------------------------

// One segment at r in FP64 with the reference's zero rules (tangle.c:486-514):
// the estimate decides unless it lies within 1e-13 of a threshold, where the
// reference's own expression does.
__device__ inline void oct_segment(const double a[3], const double b[3], const double r[3],
                                   double acc[3]) {
  const double R[3] = {a[0] - r[0], a[1] - r[1], a[2] - r[2]};
  const double Rp[3] = {b[0] - r[0], b[1] - r[1], b[2] - r[2]};
  const double l2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
  const double lp2 = Rp[0] * Rp[0] + Rp[1] * Rp[1] + Rp[2] * Rp[2];
  const double u = rsq64(l2), up = rsq64(lp2);
  const double dab = R[0] * Rp[0] + R[1] * Rp[1] + R[2] * Rp[2];
  const double w = u * up;
  const double tt = fma(w, dab, 1.0);
  const double nd1 = fabs(tt - 2.0);
  bool zero = l2 < 1e-16 || lp2 < 1e-16 || nd1 < 1e-8;
  const bool band = fabs(l2 - 1e-16) < 1e-22 || fabs(lp2 - 1e-16) < 1e-22 || fabs(nd1 - 1e-8) < 1e-13;
  if (band) zero = guard_exact(a, b, r);
  if (zero) return;
  const double f = (u + up) * w * rcp64(tt);
  acc[0] = fma(f, R[1] * Rp[2] - R[2] * Rp[1], acc[0]);
  acc[1] = fma(f, R[2] * Rp[0] - R[0] * Rp[2], acc[1]);
  acc[2] = fma(f, R[0] * Rp[1] - R[1] * Rp[0], acc[2]);
}

struct OctArgs {
  const double* P;   // targets (SoA)
  long cap;
  int n;
  const OctNode* nodes;   // n_nodes + 1 (the last one past the end)
  const OctMom* nmom;
  int n_nodes;
  const LeafSeg* segs;
  const ImgDev* img;
  int n_img;
  int img_per_walk;  // images per wavefront (1..64); 64 / img_per_walk targets
  int n_groups;      // ceil(n_img / img_per_walk)
  int n_items;       // (groups of OCT_WAVES wavefronts) x n_groups
  int per_xcd;       // ceil(n_items / 8)
  OctParams op;
  double kappa4pi;
  double* partial;   // [n_img][3][cap]
  long* stats;       // [0] leaf pairs, [1] multipoles, [2] node visits (wave)
};

// Far-field value of a node at r (octree.c:296-330), kappa/4pi applied by the
// caller.  Reference form: Gamma x R/|R|^3 + (M_21, 0, 0)/|R|^3 + (u x R)/|R|^5,
// u_i = M_i2 R_2; dipole form: Gamma x R/|R|^3 - eps:M/|R|^3 + 3 (M R) x R/|R|^5.
// R = r - centre, not minimum-imaged.
__device__ inline void oct_multipole(const double com[3], const OctMom& nd, const double r[3],
                                     bool dipole, double acc[3]) {
  const double R[3] = {r[0] - com[0], r[1] - com[1], r[2] - com[2]};
  const double R2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
  const double ir = rsq64(R2);
  const double i3 = ir * ir * ir;
  const double i5 = i3 * ir * ir;
  const double* M = nd.M;
  double v[3] = {(nd.gam[1] * R[2] - nd.gam[2] * R[1]) * i3, (nd.gam[2] * R[0] - nd.gam[0] * R[2]) * i3,
                 (nd.gam[0] * R[1] - nd.gam[1] * R[0]) * i3};
  double u[3];
  if (dipole) {
    v[0] -= (M[5] - M[7]) * i3;
    v[1] -= (M[6] - M[2]) * i3;
    v[2] -= (M[1] - M[3]) * i3;
    for (int i = 0; i < 3; ++i) u[i] = 3.0 * (M[3 * i] * R[0] + M[3 * i + 1] * R[1] + M[3 * i + 2] * R[2]);
  } else {
    v[0] += M[7] * i3;
    for (int i = 0; i < 3; ++i) u[i] = M[3 * i + 2] * R[2];
  }
  v[0] += (u[1] * R[2] - u[2] * R[1]) * i5;
  v[1] += (u[2] * R[0] - u[0] * R[2]) * i5;
  v[2] += (u[0] * R[1] - u[1] * R[0]) * i5;
  acc[0] += v[0];
  acc[1] += v[1];
  acc[2] += v[2];
}

// FP32 form of oct_multipole: R = r - centre in FP64 rounded to FP32 (given
// as its negation dn, which the opening test formed), everything after in
// FP32 (the node's moments rounded once).
__device__ inline void oct_multipole32(const float dn[3], const OctMom32& nd, bool dipole,
                                       double acc[3]) {
  const float R[3] = {-dn[0], -dn[1], -dn[2]};
  const float R2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
  const float ir = __builtin_amdgcn_rsqf(R2);
  const float i3 = ir * ir * ir;
  const float i5 = i3 * ir * ir;
  const float* M = nd.M;
  const float* G = nd.gam;
  float v[3] = {(G[1] * R[2] - G[2] * R[1]) * i3, (G[2] * R[0] - G[0] * R[2]) * i3,
                (G[0] * R[1] - G[1] * R[0]) * i3};
  float u[3];
  if (dipole) {
    v[0] -= (M[5] - M[7]) * i3;
    v[1] -= (M[6] - M[2]) * i3;
    v[2] -= (M[1] - M[3]) * i3;
    for (int i = 0; i < 3; ++i) u[i] = 3.f * (M[3 * i] * R[0] + M[3 * i + 1] * R[1] + M[3 * i + 2] * R[2]);
  } else {
    v[0] += M[7] * i3;
    for (int i = 0; i < 3; ++i) u[i] = M[3 * i + 2] * R[2];
  }
  v[0] += (u[1] * R[2] - u[2] * R[1]) * i5;
  v[1] += (u[2] * R[0] - u[0] * R[2]) * i5;
  v[2] += (u[0] * R[1] - u[1] * R[0]) * i5;
  acc[0] += (double)v[0];
  acc[1] += (double)v[1];
  acc[2] += (double)v[2];
}

// Opening test of octree.c:287-295: accept when d > 0 and Lmax / d < theta
// (d = minimum-image distance to the centre of mass).  Decided on squares
// (Lmax^2 against theta^2 d^2) outside a 1e-12 relative band, where the
// rounding of either form cannot change the answer; inside it, the
// reference's square root and division decide.
__device__ __forceinline__ bool oct_accept(const double dv[3], double lmax, double theta) {
  const double d2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
  const double t2 = theta * theta * d2, l2 = lmax * lmax;
  if (l2 < t2 * (1.0 - 1e-12)) return true;   // t2 > 0 here, so d > 0
  if (l2 > t2 * (1.0 + 1e-12)) return false;
  const double d = sqrt(d2);
  return d > 0 && lmax / d < theta;
}

// The opening test in FP32, decided only away from its threshold.  With
// d = c - rs rounded once from FP64 and wrapped by +-L on periodic axes where
// |d| > L/2 (d^2 is continuous across that switch, so an FP32 wrap decision
// that differs from the FP64 one near |d| = L/2 moves d^2 by rounding errors
// only), each wrapped component e_a is within 2^-23 (|e_a| + L_a) of the
// FP64 value, so d^2 = sum e_a^2 is within B = 1e-5 d^2 + 5e-7 sum |e_a| L_a
// (both terms with a margin over the rounding bounds): d^2 - B > q means
// lmax / d < theta (accept) and d^2 + B < q means it does not (open); in
// between the FP64 test of oct_accept decides.  q = (lmax / theta)^2 per
// node.  The test's constants sit in VGPRs (taken from the lane's own
// position, so the compiler keeps them there) and its FP32 instructions
// dual-issue, where the FP64 test's do not.
struct OctFast {
  float h[3], nh[3], L[3], nL[3];   // +-L/2 and +-L on periodic axes; inf, -inf and 0 elsewhere
  float eps;                        // 1e-5
  float ceps;                       // 5e-7
};
__device__ __forceinline__ OctFast oct_fast_consts(const BoxDev& b, float lane_zero) {
  OctFast F;
  for (int a = 0; a < 3; ++a) {
    const bool p = b.periodic[a] != 0;
    F.h[a] = (p ? (float)(b.L[a] / 2) : __builtin_inff()) + lane_zero;
    F.nh[a] = -F.h[a];
    F.L[a] = (p ? (float)b.L[a] : 0.f) + lane_zero;
    F.nL[a] = -F.L[a];
  }
  F.eps = 1e-5f + lane_zero;
  F.ceps = 5e-7f + lane_zero;
  return F;
}
// 1: accept, 0: open, -1: within the band (the FP64 test decides).  dn:
// c - rs rounded to FP32 before the wrap; its negation is rs - c rounded to
// FP32 (the FP64 difference and the rounding are both symmetric under
// negation), the FP32 multipole's R, which so costs no FP64 operation.
__device__ __forceinline__ int oct_accept32(const double rs[3], const double c[3], float q,
                                            const OctFast& F, float dn[3]) {
  float e[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float d = (float)(c[a] - rs[a]);
    dn[a] = d;
    float k = d > F.h[a] ? F.nL[a] : 0.f;
    k = d < F.nh[a] ? F.L[a] : k;
    e[a] = d + k;
  }
  const float d2 = fmaf(e[2], e[2], fmaf(e[1], e[1], e[0] * e[0]));
  const float s = fmaf(fabsf(e[2]), F.L[2], fmaf(fabsf(e[1]), F.L[1], fabsf(e[0]) * F.L[0]));
  const float B = fmaf(d2, F.eps, s * F.ceps);
  if (d2 - B > q) return 1;
  if (d2 + B < q) return 0;
  return -1;
}

constexpr int OCT_WAVES = 4;   // wavefronts per block

// Target of image (sh, w2, ax) for the point r (reference shifted(),
// tangle.c:39-79): r - shift, or the reflection 2 wall - r on axis ax.
__device__ __forceinline__ void oct_image_target(const double r[3], const double sh[3], double w2, int ax,
                                                 double rs[3]) {
  rs[0] = r[0] - sh[0];
  rs[1] = r[1] - sh[1];
  rs[2] = r[2] - sh[2];
  if (__builtin_expect(ax >= 0, 0)) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (k == ax) rs[k] = w2 - r[k];
  }
}

// rs - (minimum image of c seen from rs), the displacement of octree.c:
// 286-287 (seg_pwrap moves c by one box length on each periodic axis where
// |c - rs| > L/2).  Branch-free: fma(k, L, c) with k in {-1, 0, 1} rounds
// exactly as c - L, c or c + L does.
__device__ __forceinline__ void oct_min_image(const double rs[3], const double c[3], const BoxDev& b,
                                              double dv[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double d = c[a] - rs[a];
    const double h = b.L[a] / 2;
    const double k = b.periodic[a] ? (d > h ? -1.0 : (d < -h ? 1.0 : 0.0)) : 0.0;
    dv[a] = rs[a] - fma(k, b.L[a], c[a]);
  }
}

// One lane = one (target, image) pair.  A wavefront holds tpw consecutive
// stream targets x ipw images of one image group (tpw ipw <= 64): the
// minimum-image opening test (octree.c:286-295) makes a target's images
// decide alike except near the far faces, and consecutive targets sit next
// to each other on a filament, so the wave's lanes walk almost the same
// nodes and most lanes work at every node.  Each lane walks the preorder as
// octree.c:275-343 does (accept: multipole; leaf: segments; else open): it
// resumes at node >= next; the wave steps to node + 1 when some lane opened
// the node and over its subtree otherwise.  Every lane accumulates its own
// walk's terms in its own order, into the partial plane of its image.
//
// Grid: one dimension of n_items = (group of OCT_WAVES waves) x (image
// group) work items, XCD-major: block b runs on XCD b % 8 (blocks are
// dispatched round-robin over the 8 XCDs) and takes item (b % 8) per_xcd +
// b / 8, so each XCD's L2 serves a contiguous range of targets, whose walks
// touch the same nodes.  MIXED: multipoles in FP32 (leaf pairs and all
// decisions in FP64).
template <bool MIXED>
__device__ __forceinline__ void octree_body(const OctArgs& A) {
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.x;
  const int item = (b & 7) * A.per_xcd + (b >> 3);
  if (item >= A.n_items) return;
  const int wgrp = item / A.n_groups;
  const int g = item - wgrp * A.n_groups;
  const int W = wgrp * OCT_WAVES + wid;            // target group of this wave
  const int ipw = A.img_per_walk;
  const int tpw = TILE / ipw;
  if (W * tpw >= A.n) return;
  const int tl = lane / ipw;
  const int I = g * ipw + (lane - tl * ipw);       // this lane's image
  const int s = W * tpw + tl;                      // this lane's target
  const bool valid = tl < tpw && s < A.n && I < A.n_img;
  const int st = valid ? s : W * tpw;
  const ImgDev im = A.img[valid ? I : 0];
  const double r[3] = {A.P[st], A.P[A.cap + st], A.P[2 * A.cap + st]};
  double rs[3];
  oct_image_target(r, im.shift, im.wall2, im.axis, rs);
  const int skip_id = (valid && im.primary) ? s : -1;
  const bool dipole = A.op.dipole != 0;
  // 0 in every lane, but a per-lane value to the compiler
  const int lane_zero_i = lane >> 6;
  const float lane_zero = (float)lane_zero_i;
  const OctFast F = oct_fast_consts(A.op.box, lane_zero);
  double acc[3] = {0, 0, 0};
  unsigned leaf_pairs = 0, mps = 0, visits = 0;
  int next = valid ? 0 : 0x7fffffff;   // this lane resumes at node >= next
  int node = 0;
  while (node < A.n_nodes) {
    const OctNode nd = A.nodes[node];
    ++visits;
    const int after = node + nd.skip;
    bool open = false;
    if (node >= next) {
      if (nd.n == 0) {
        next = after;
      } else if (nd.n == 1 || nd.lmin < A.op.grain) {
        // segs holds one entry past the last sorted point, so the look-ahead
        // load of segment j + 1 stays in bounds
        LeafSeg q = A.segs[nd.first];
        for (int j = nd.first; j < nd.first + nd.n; ++j) {
          const LeafSeg qn = A.segs[j + 1];
          // the primary sum skips the target's own two segments (octree.c:291)
          if (q.m != skip_id && q.fwd != skip_id) {
            oct_segment(q.a, q.b, rs, acc);
            ++leaf_pairs;
          }
          q = qn;
        }
        next = after;
      } else {
        float dn[3];
        int dec = oct_accept32(rs, nd.com, nd.q, F, dn);
        if (__builtin_expect(dec < 0, 0)) {
          double dv[3];
          oct_min_image(rs, nd.com, A.op.box, dv);
          dec = oct_accept(dv, nd.lmax, A.op.theta) ? 1 : 0;
        }
        if (dec) {
          if (MIXED) {
            // a vector load of the FP32 moments (the index looks per-lane to
            // the compiler): they arrive in VGPRs, so the multipole's FP32
            // instructions dual-issue, where SGPR operands would not
            const OctMom32 mm = reinterpret_cast<const OctMom32*>(A.nmom)[node + lane_zero_i];
            oct_multipole32(dn, mm, dipole, acc);
          } else {
            const OctMom mm = A.nmom[node];
            oct_multipole(nd.com, mm, rs, dipole, acc);
          }
          ++mps;
          next = after;
        } else {
          open = true;
          next = node + 1;
        }
      }
    }
    node = __builtin_amdgcn_ballot_w64(open) != 0 ? node + 1 : after;
  }
  // mirror image: the velocity is reflected (reference mirror_dir_reflect)
  if (im.axis >= 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (k == im.axis) acc[k] = -acc[k];
  }
  if (valid) {
    double* out = A.partial + (long)I * 3 * A.cap;
    out[s] = acc[0] * A.kappa4pi;
    out[A.cap + s] = acc[1] * A.kappa4pi;
    out[2 * A.cap + s] = acc[2] * A.kappa4pi;
    if (A.stats) {
      atomicAdd((unsigned long long*)&A.stats[0], (unsigned long long)leaf_pairs);
      atomicAdd((unsigned long long*)&A.stats[1], (unsigned long long)mps);
    }
  }
  if (A.stats && lane == 0) atomicAdd((unsigned long long*)&A.stats[2], (unsigned long long)visits);
}

__global__ void __launch_bounds__(256) k_octree_fp64(OctArgs A) { octree_body<false>(A); }
__global__ void __launch_bounds__(256) k_octree_mixed(OctArgs A) { octree_body<true>(A); }

}  // namespace ov
