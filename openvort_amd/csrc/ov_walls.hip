// ov_walls.hip — geometry of tangles with points pinned on walls (reference
// src/tangle.c:165-243 step_node / arc_length, 245-480 update_tangent_normal,
// 530-551 lia_velocity, 704-721 initialize_dxi).
//
// Stream points carry a pin code: -1 for a free point, else
//   wall (0..5, X_L X_H Y_L Y_H Z_L Z_H) | slip << 3 | fwd_pos << 4
// with slip = 1 for PINNED_SLIP and fwd_pos = 1 when the caller's forward
// index is > 0 (the reference tests `forward > 0`, not `>= 0`, to choose the
// side of a PINNED point's one-sided stencil, tangle.c:285).  A pinned point
// may lack its forward or reverse link (an open chain ends on it); walking
// off such an end turns back along the chain and mirrors the result in the
// point's wall (step_node, tangle.c:176-191).

#pragma once
#include <hip/hip_runtime.h>

#include "ov_device.h"

namespace ov {

constexpr int WALK_MAX = 16;   // turns + steps of one walk (|where| <= 4, chains of >= 2 points)

__device__ __forceinline__ bool pin_is_pinned(int pc) { return pc >= 0 && !((pc >> 3) & 1); }

// 2 wall - x on the wall's axis (mirror_shift, vec3_maths.c:204-258).
__device__ __forceinline__ void wall_mirror(double p[3], int wall, const BoxDev& box) {
  const int ax = wall >> 1;
  const double w = (wall & 1) ? box.hi[ax] : box.lo[ax];
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (a == ax) p[a] = 2 * w - p[a];
}

// step_node(s, where): the point `where` links away, turning back at chain
// ends and mirroring in their walls (outermost turn applied last).  Returns
// false when the walk does not end (a point without either link).
__device__ inline bool wall_walk(int s, int where, const double* __restrict__ P, long cap,
                                 const int* __restrict__ fwd, const int* __restrict__ rev,
                                 const int* __restrict__ pin, const BoxDev& box, double out[3]) {
  int walls[WALK_MAX];
  int nm = 0;
  int node = s, w = where;
  for (int it = 0; w != 0; ++it) {
    if (it >= WALK_MAX) return false;
    const int nxt = w > 0 ? fwd[node] : rev[node];
    if (nxt < 0) {           // only pinned points lack a link
      walls[nm++] = pin[node] & 7;
      w = -w;
      continue;
    }
    node = nxt;
    w += w > 0 ? -1 : 1;
  }
  load3(P, cap, node, out);
  for (int i = nm - 1; i >= 0; --i) wall_mirror(out, walls[i], box);
  return true;
}

// arc_length(s, dist) (tangle.c:200-243): dxi of the segments walked,
// summed right to left as the reference's recursion does; a chain end turns
// the walk back without a term.
__device__ inline double wall_arc(int s, int dist, const double* __restrict__ dxi,
                                  const int* __restrict__ fwd, const int* __restrict__ rev) {
  double terms[WALK_MAX];
  int nt = 0;
  int node = s, d = dist;
  for (int it = 0; d != 0 && it < WALK_MAX; ++it) {
    if (d > 0) {
      if (fwd[node] < 0) {
        d = -d;
        continue;
      }
      terms[nt++] = dxi[node];
      node = fwd[node];
      --d;
    } else {
      const int r = rev[node];
      if (r < 0) {
        d = -d;
        continue;
      }
      terms[nt++] = dxi[r];
      node = r;
      ++d;
    }
  }
  double sum = 0;
  for (int i = nt - 1; i >= 0; --i) sum = terms[i] + sum;
  return sum;
}

// Five-node derivative stencil at node 0 with the other nodes at x[0..3]
// (chord coordinates) and points pj[0..3]: s' and s'' (the same Lagrange
// weights as the free stencil, tangle.c:432-479 and 312-357).
__device__ inline void stencil5(const double x[4], const double pj[4][3], const double p0[3],
                                const BoxDev& box, double t[3], double nm[3]) {
  t[0] = t[1] = t[2] = 0;
  nm[0] = nm[1] = nm[2] = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double o[3];
    int q = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m)
      if (m != j) o[q++] = x[m];
    const double den = x[j] * (x[j] - o[0]) * (x[j] - o[1]) * (x[j] - o[2]);
    const double w1 = -(o[0] * o[1] * o[2]) / den;
    const double w2 = 2 * (o[0] * o[1] + o[0] * o[2] + o[1] * o[2]) / den;
    double s0w[3];
    pwrap(pj[j], p0, box, s0w);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double d = pj[j][i] - s0w[i];
      t[i] += w1 * d;
      nm[i] += w2 * d;
    }
  }
}

// Whether point s takes the reference's neighbour-tangent stencil
// (tangle.c:364-430): it is not PINNED and a linked neighbour is.
__device__ __forceinline__ bool wall_hermite(int s, const int* __restrict__ fwd,
                                             const int* __restrict__ rev,
                                             const int* __restrict__ pin) {
  if (pin_is_pinned(pin[s])) return false;
  const int f = fwd[s], r = rev[s];
  return (f >= 0 && pin_is_pinned(pin[f])) || (r >= 0 && pin_is_pinned(pin[r]));
}

// dxi[s] = |s - minimum image of step_node(s, +1)| (initialize_dxi); false
// when the walk does not end.
__device__ inline bool wall_dxi_point(int s, const double* __restrict__ P, long cap,
                                      const int* __restrict__ fwd, const int* __restrict__ rev,
                                      const int* __restrict__ pin, const BoxDev& box,
                                      double* __restrict__ dxi) {
  double p0[3], p1[3];
  load3(P, cap, s, p0);
  if (!wall_walk(s, 1, P, cap, fwd, rev, pin, box, p1)) {
    dxi[s] = __builtin_nan("");
    return false;
  }
  dxi[s] = chord(p0, p1, box);
  return true;
}

// s', s'' of point s (update_tangent_normal with walls): the neighbour-
// tangent stencil when herm (reading the neighbours' s' from T), the
// one-sided stencil of a PINNED point, the free stencil through step_node /
// arc_length otherwise.  p0, pn, pp: s and its walked neighbours (+1, -1).
// false when a walk does not end.
__device__ inline bool wall_tn_point(int s, bool herm, const double p0[3], const double pn[3],
                                     const double pp[3], const double* __restrict__ P, long cap,
                                     const int* __restrict__ fwd, const int* __restrict__ rev,
                                     const int* __restrict__ pin, const BoxDev& box,
                                     const double* __restrict__ T, const double* __restrict__ dxi,
                                     double t[3], double nm[3]) {
  bool ok = true;
  if (herm) {
    // tangle.c:364-430: Hermite-type stencil from the neighbours' tangents t1,
    // t-1 and ds = (s0 seen from s_j) - s_j for j = +-1, a = arc(s, 1),
    // b = -arc(s, -1); the coefficients are the reference's (its geometry at
    // walls is defined by them)
    const int f = fwd[s], r = rev[s];
    double t1[3], tm1[3], w1[3], wm1[3], ds1[3], dsm1[3];
    load3(T, cap, f, t1);
    load3(T, cap, r, tm1);
    pwrap(pn, p0, box, w1);
    pwrap(pp, p0, box, wm1);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      ds1[i] = w1[i] - pn[i];
      dsm1[i] = wm1[i] - pp[i];
    }
    const double a = wall_arc(s, 1, dxi, fwd, rev);
    const double b = -wall_arc(s, -1, dxi, fwd, rev);
    const double D = a * a - 2 * a * b + b * b;
    const double D3 = a * a * a - 3 * a * a * b + 3 * a * b * b - b * b * b;
    const double ct1 = -b * b / D, ctm1 = -a * a / D;
    const double ct_ds1 = 2 * b * b * (-2 * a + b) / (a * D3);
    const double ct_dsm1 = 2 * a * a * (-a + 2 * b) / (b * D3);
    const double cn_ds1 = 2 * b * (8 * a * a - a * b - b * b) / (a * a * D3);
    const double cn_dsm1 = 2 * a * (a * a + a * b - 8 * b * b) / (b * b * D3);
    const double cn_t1 = 2 * b * (2 * a + b) / (a * D);
    const double cn_tm1 = 2 * a * (a + 2 * b) / (b * D);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      t[i] = ct1 * t1[i] + ctm1 * tm1[i] + ct_ds1 * ds1[i] + ct_dsm1 * dsm1[i];
      nm[i] = cn_t1 * t1[i] + cn_tm1 * tm1[i] + cn_ds1 * ds1[i] + cn_dsm1 * dsm1[i];
    }
    return true;
  }
  const int pc = pin[s];
  double x[4], pj[4][3];
  if (pin_is_pinned(pc)) {
    // one-sided stencil on the nodes 1..4 links away (tangle.c:275-362),
    // forward when the caller's forward index is > 0
    const int dir = ((pc >> 4) & 1) ? 1 : -1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      ok = wall_walk(s, dir * (j + 1), P, cap, fwd, rev, pin, box, pj[j]) && ok;
      x[j] = dir * wall_arc(s, dir * (j + 1), dxi, fwd, rev);
    }
  } else {
    // the free stencil (tangle.c:432-479) through step_node / arc_length
    ok = wall_walk(s, 2, P, cap, fwd, rev, pin, box, pj[0]) && ok;
    pj[1][0] = pn[0], pj[1][1] = pn[1], pj[1][2] = pn[2];
    pj[2][0] = pp[0], pj[2][1] = pp[1], pj[2][2] = pp[2];
    ok = wall_walk(s, -2, P, cap, fwd, rev, pin, box, pj[3]) && ok;
    x[0] = wall_arc(s, 2, dxi, fwd, rev);
    x[1] = wall_arc(s, 1, dxi, fwd, rev);
    x[2] = -wall_arc(s, -1, dxi, fwd, rev);
    x[3] = -wall_arc(s, -2, dxi, fwd, rev);
  }
  if (!ok) return false;
  stencil5(x, pj, p0, box, t, nm);
  return true;
}

// lia_velocity (tangle.c:530-551) of point p0 with its walked neighbours pn,
// pp and its s', s''.
__device__ inline void wall_lia_point(const double p0[3], const double pn[3], const double pp[3],
                                      const double t[3], const double nm[3], const BoxDev& box,
                                      const PhysDev& ph, double v[3]) {
  const double lf = chord(p0, pn, box), lr = chord(pp, p0, box);
  const double pi = 3.14159265358979323846;
  const double beta = ph.kappa * log(2 * sqrt(lf * lr) / sqrt(sqrt(2.718281828459045)) / ph.a0) / 4 / pi;
  v[0] = (t[1] * nm[2] - t[2] * nm[1]) * beta;
  v[1] = (t[2] * nm[0] - t[0] * nm[2]) * beta;
  v[2] = (t[0] * nm[1] - t[1] * nm[0]) * beta;
}

__global__ void __launch_bounds__(256) k_wall_dxi(const double* __restrict__ P, const int* __restrict__ fwd,
                                                  const int* __restrict__ rev, const int* __restrict__ pin,
                                                  int n, long cap, BoxDev box, double* __restrict__ dxi,
                                                  int* __restrict__ flag) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  if (!wall_dxi_point(s, P, cap, fwd, rev, pin, box, dxi)) atomicOr(flag, 1);
}

// s', s'' (when compute_geom) and the local induction of point s.  Phase 1:
// every point except those next to a PINNED point; phase 2: those, from
// their neighbours' tangents of phase 1.  With compute_geom = 0 (caller's
// s', s'') phase 1 does the local induction of every point.
__global__ void __launch_bounds__(256) k_wall_geometry(const double* __restrict__ P,
                                                       const int* __restrict__ fwd,
                                                       const int* __restrict__ rev,
                                                       const int* __restrict__ pin, int n, long cap,
                                                       BoxDev box, PhysDev ph, int compute_geom,
                                                       int phase, double* __restrict__ T,
                                                       double* __restrict__ Nn,
                                                       const double* __restrict__ dxi,
                                                       double* __restrict__ lia,
                                                       int* __restrict__ flag) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const bool herm = compute_geom && wall_hermite(s, fwd, rev, pin);
  if (herm != (phase == 2)) return;
  double p0[3], pn[3], pp[3];
  load3(P, cap, s, p0);
  bool ok = wall_walk(s, 1, P, cap, fwd, rev, pin, box, pn);
  ok = wall_walk(s, -1, P, cap, fwd, rev, pin, box, pp) && ok;
  if (!ok) {
    atomicOr(flag, 1);
    return;
  }
  double t[3], nm[3];
  if (!compute_geom) {
    load3(T, cap, s, t);
    load3(Nn, cap, s, nm);
  } else if (!wall_tn_point(s, herm, p0, pn, pp, P, cap, fwd, rev, pin, box, T, dxi, t, nm)) {
    atomicOr(flag, 1);
    return;
  }
  if (compute_geom) {
    store3(T, cap, s, t);
    store3(Nn, cap, s, nm);
  }
  double v[3];
  wall_lia_point(p0, pn, pp, t, nm, box, ph, v);
  store3(lia, cap, s, v);
}

}  // namespace ov
