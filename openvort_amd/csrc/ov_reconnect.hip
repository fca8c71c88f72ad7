// ov_reconnect.hip — reconnection candidate search (reference reconnect,
// vortex_dynamics.c:140-315).
//
// The reference walks every pair k < l of points (caller indices) and, for
// those that survive its tests in order, rewires the first acceptable l of
// each k.  The tests that cannot change during the pass are found here:
//   * both points free (EMPTY points are not in the stream; points pinned on
//     walls, which the reference's pair loop skips, are skipped by their pin
//     codes; its wall reconnections, vortex_dynamics.c:152-208, stay with
//     the host);
//   * minimum-image distance |v_k - pwrap(v_k, v_l)| <= rec_dist
//     (vortex_dynamics.c:236-242);
//   * |t_k| >= 0.5 and |t_l| >= 0.5 (vortex_dynamics.c:256-257; tangents are
//     not recomputed during the pass);
//   * t_k.t_l / |t_k| / |t_l| <= cos(rec_angle) (vortex_dynamics.c:259-265).
// All in FP64 with the reference's operation order and no fused multiply-add,
// so every decision is the reference's.  The neighbour test, the
// "recalculate" taint flags, the velocity test and the rewiring depend on
// reconnections made earlier in the pass and stay with the host, which visits
// the sorted candidates in the reference's (k, l) order.
//
// Search: points are binned into cells of edge >= rec_dist (a periodic axis
// of length L gets floor(L / rec_dist) cells, wrapped), sorted by cell key,
// and each point scans the cells around its own (binary search for each
// cell's run), emitting a pair only from the point with the smaller caller
// index.  Pairs are sorted by (k, l) afterwards.
#include <hip/hip_runtime.h>

#include "ov_device.h"

namespace ov {

// Cell grid: per axis the cell edge, the cell count (0: open axis, unbounded,
// coordinates offset by 2^20 and clamped to 21 bits) and the low corner.
struct CellGrid {
  double lo[3];
  double h[3];
  int dims[3];   // periodic: cells per axis (>= 1); open: 0
  int pad;
};

__device__ __forceinline__ int cell_coord(double x, const CellGrid& g, int ax) {
  const double c = floor((x - g.lo[ax]) / g.h[ax]);
  if (g.dims[ax] > 0) {
    long m = (long)c % g.dims[ax];
    if (m < 0) m += g.dims[ax];
    return (int)m;
  }
  const double o = c + 1048576.0;
  return (int)fmin(fmax(o, 0.0), 2097151.0);
}

__device__ __forceinline__ uint64_t cell_key(int cx, int cy, int cz) {
  return ((uint64_t)cx << 42) | ((uint64_t)cy << 21) | (uint64_t)cz;
}

__global__ void __launch_bounds__(256) kr_keys(const double* __restrict__ P, int n, long cap,
                                               CellGrid g, uint64_t* __restrict__ keys,
                                               int* __restrict__ vals) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  keys[s] = cell_key(cell_coord(P[s], g, 0), cell_coord(P[cap + s], g, 1),
                     cell_coord(P[2 * cap + s], g, 2));
  vals[s] = s;
}

// first index in sorted keys[0, n) with keys[i] >= key
__device__ __forceinline__ int key_lower_bound(const uint64_t* __restrict__ keys, int n, uint64_t key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// The reference's arithmetic of the static tests (vec3_maths.c:105-138,
// 286-301, 341-350): sums from zero left to right, no contraction.
__device__ __forceinline__ bool rec_static(const double a[3], const double b[3], const double ta[3],
                                           const double tb[3], const BoxDev& box, double rec_dist,
                                           double cos_angle) {
#pragma clang fp contract(off)
  double w[3];
  for (int i = 0; i < 3; ++i) {
    double o = b[i];
    if (box.periodic[i]) {
      const double d = b[i] - a[i];
      if (d > box.L[i] / 2)
        o -= box.L[i];
      else if (d < -box.L[i] / 2)
        o += box.L[i];
    }
    w[i] = o;
  }
  double d2 = 0;
  for (int i = 0; i < 3; ++i) {
    const double v = a[i] - w[i];
    d2 += v * v;
  }
  if (sqrt(d2) > rec_dist) return false;
  double na = 0, nb = 0, dot = 0;
  for (int i = 0; i < 3; ++i) na += ta[i] * ta[i];
  for (int i = 0; i < 3; ++i) nb += tb[i] * tb[i];
  na = sqrt(na);
  nb = sqrt(nb);
  if (na < 0.5 || nb < 0.5) return false;
  for (int i = 0; i < 3; ++i) dot += ta[i] * tb[i];
  const double c = dot / na / nb;
  return !(c > cos_angle);
}

struct RecArgs {
  const double* P;        // positions, stream order (SoA, stride cap)
  const double* T;        // caller tangents, stream order
  const int* act;         // stream index -> caller index
  const uint64_t* keys;   // sorted cell keys
  const int* order;       // stream index of each sorted key
  const int* pin;         // pin codes (stream order), nullptr without pinned points
  long cap;
  int n;
  CellGrid g;
  BoxDev box;
  double rec_dist, cos_angle;
  unsigned long long* pairs;   // (k << 32) | l
  int max_pairs;
  int* count;
};

// One thread per point (in cell order): scan the neighbouring cells.
__global__ void __launch_bounds__(256) kr_pairs(RecArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const int s = A.order[i];
  if (A.pin && A.pin[s] >= 0) return;
  const int ks = A.act[s];
  double a[3], ta[3];
  load3(A.P, A.cap, s, a);
  load3(A.T, A.cap, s, ta);
  int c[3];
  for (int ax = 0; ax < 3; ++ax) c[ax] = cell_coord(a[ax], A.g, ax);
  // neighbour offsets per axis: -1, 0, +1, without repeating a cell on a
  // periodic axis of one or two cells
  int lo[3], hi[3];
  for (int ax = 0; ax < 3; ++ax) {
    const int d = A.g.dims[ax];
    lo[ax] = (d == 1) ? 0 : -1;
    hi[ax] = (d == 1 || d == 2) ? 0 : 1;
    if (d == 2) lo[ax] = -1;   // cells c and c-1 (= c+1)
  }
  for (int ox = lo[0]; ox <= hi[0]; ++ox)
    for (int oy = lo[1]; oy <= hi[1]; ++oy)
      for (int oz = lo[2]; oz <= hi[2]; ++oz) {
        int q[3] = {c[0] + ox, c[1] + oy, c[2] + oz};
        bool out = false;
        for (int ax = 0; ax < 3; ++ax) {
          const int d = A.g.dims[ax];
          if (d > 0) {
            q[ax] = (q[ax] % d + d) % d;
          } else if (q[ax] < 0 || q[ax] > 2097151) {
            out = true;
          }
        }
        if (out) continue;
        const uint64_t key = cell_key(q[0], q[1], q[2]);
        int j = key_lower_bound(A.keys, A.n, key);
        for (; j < A.n && A.keys[j] == key; ++j) {
          const int sl = A.order[j];
          const int kl = A.act[sl];
          if (kl <= ks || (A.pin && A.pin[sl] >= 0)) continue;
          double b[3], tb[3];
          load3(A.P, A.cap, sl, b);
          load3(A.T, A.cap, sl, tb);
          if (!rec_static(a, b, ta, tb, A.box, A.rec_dist, A.cos_angle)) continue;
          const int slot = atomicAdd(A.count, 1);
          if (slot < A.max_pairs)
            A.pairs[slot] = ((unsigned long long)(unsigned)ks << 32) | (unsigned)kl;
        }
      }
}

}  // namespace ov
