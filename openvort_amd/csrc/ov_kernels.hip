// ov_kernels.hip — gfx950 kernels of the vortex-filament velocity / RK4 path.
//
// Layout (see ov_device.h): SoA FP64 arrays in stream order with stride
// `cap` between the x, y and z planes; one wavefront per 64-entry tile.
#include <hip/hip_runtime.h>
#include <math.h>

#include "ov_device.h"
#include "ov_flow.h"

namespace ov {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

// Minimum image of r2 seen from r1 (reference seg_pwrap, vec3_maths.c:105-138):
// only r2 moves, by one box length per periodic axis beyond half a length.
__device__ __forceinline__ void pwrap(const double r1[3], const double r2[3], const BoxDev& b,
                                      double out[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double o = r2[i];
    if (b.periodic[i]) {
      double d = r2[i] - r1[i];
      if (d > b.L[i] / 2)
        o -= b.L[i];
      else if (d < -b.L[i] / 2)
        o += b.L[i];
    }
    out[i] = o;
  }
}

__device__ __forceinline__ double norm3(const double v[3]) {
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

// |s - minimum image of s'| (reference segment_len of seg_pwrap, tangle.c:536-540)
__device__ __forceinline__ double chord(const double s[3], const double sp[3], const BoxDev& b) {
  double w[3], d[3];
  pwrap(s, sp, b, w);
  d[0] = s[0] - w[0];
  d[1] = s[1] - w[1];
  d[2] = s[2] - w[2];
  return norm3(d);
}

__device__ __forceinline__ void load3(const double* __restrict__ P, long cap, int i, double r[3]) {
  r[0] = P[i];
  r[1] = P[cap + i];
  r[2] = P[2 * cap + i];
}
__device__ __forceinline__ void store3(double* __restrict__ P, long cap, int i, const double r[3]) {
  P[i] = r[0];
  P[cap + i] = r[1];
  P[2 * cap + i] = r[2];
}

// FP64 1/sqrt(x) and 1/x: hardware estimate refined by two Newton steps.
__device__ __forceinline__ double rsq64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double h = 0.5 * x;
  double e = fma(-h * y, y, 0.5);
  y = fma(y, e, y);
  e = fma(-h * y, y, 0.5);
  y = fma(y, e, y);
  return y;
}
__device__ __forceinline__ double rcp64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  return y;
}

// ---------------------------------------------------------------------------
// geometry: dxi, s', s'' and the LIA velocity of every point
// ---------------------------------------------------------------------------

// Five-node Lagrange stencil in cumulative chord length (reference
// update_tangent_normal free branch, tangle.c:245-273 and 432-479): with
// nodes x_j (the other three nonzero nodes a, b, c)
//   w'_j  = -x_a x_b x_c / (x_j (x_j-x_a)(x_j-x_b)(x_j-x_c))
//   w''_j = 2 (x_a x_b + x_a x_c + x_b x_c) / (x_j (x_j-x_a)(x_j-x_b)(x_j-x_c))
// applied to the minimum-image differences s_j - s_0.  LIA (tangle.c:530-551):
//   v_LIA = kappa/4pi ln(2 sqrt(l+ l-) / (e^(1/4) a0)) s' x s''.
// With compute_geom = 0 the caller's s', s'' (T, Nn) are kept.
__device__ __forceinline__ void geometry_point(int s, const double* __restrict__ P,
                                               const int* __restrict__ fwd,
                                               const int* __restrict__ rev, long cap,
                                               const BoxDev& box, const PhysDev& ph,
                                               int compute_geom, double* __restrict__ T,
                                               double* __restrict__ Nn, double* __restrict__ dxi,
                                               double* __restrict__ lia) {
  int f = fwd[s], r = rev[s];
  double p0[3], pf[3], pr[3];
  load3(P, cap, s, p0);
  load3(P, cap, f, pf);
  load3(P, cap, r, pr);
  double dk = chord(p0, pf, box);   // dxi[s]
  double dr = chord(pr, p0, box);   // dxi[rev s]
  double t[3], nm[3];
  if (compute_geom) {
    int ff = fwd[f], rr = rev[r];
    double pff[3], prr[3];
    load3(P, cap, ff, pff);
    load3(P, cap, rr, prr);
    double df = chord(pf, pff, box);
    double drr = chord(prr, pr, box);
    double x[4] = {dk + df, dk, -dr, -(dr + drr)};
    const double* pj[4] = {pff, pf, pr, prr};
    t[0] = t[1] = t[2] = 0;
    nm[0] = nm[1] = nm[2] = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double o[3];
      int q = 0;
#pragma unroll
      for (int m = 0; m < 4; ++m)
        if (m != j) o[q++] = x[m];
      double den = x[j] * (x[j] - o[0]) * (x[j] - o[1]) * (x[j] - o[2]);
      double w1 = -(o[0] * o[1] * o[2]) / den;
      double w2 = 2 * (o[0] * o[1] + o[0] * o[2] + o[1] * o[2]) / den;
      double s0w[3];
      pwrap(pj[j], p0, box, s0w);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double d = pj[j][i] - s0w[i];
        t[i] += w1 * d;
        nm[i] += w2 * d;
      }
    }
    store3(T, cap, s, t);
    store3(Nn, cap, s, nm);
    dxi[s] = dk;
  } else {
    load3(T, cap, s, t);
    load3(Nn, cap, s, nm);
  }
  const double pi = 3.14159265358979323846;
  double beta = ph.kappa * log(2 * sqrt(dk * dr) / sqrt(sqrt(2.718281828459045)) / ph.a0) / 4 / pi;
  double v[3] = {(t[1] * nm[2] - t[2] * nm[1]) * beta, (t[2] * nm[0] - t[0] * nm[2]) * beta,
                 (t[0] * nm[1] - t[1] * nm[0]) * beta};
  store3(lia, cap, s, v);
}

__global__ void __launch_bounds__(256) k_geometry(const double* __restrict__ P,
                                                  const int* __restrict__ fwd,
                                                  const int* __restrict__ rev, int n, long cap,
                                                  BoxDev box, PhysDev ph, int compute_geom,
                                                  double* __restrict__ T, double* __restrict__ Nn,
                                                  double* __restrict__ dxi,
                                                  double* __restrict__ lia) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  geometry_point(s, P, fwd, rev, cap, box, ph, compute_geom, T, Nn, dxi, lia);
}

// ---------------------------------------------------------------------------
// record stream: per-point segment ends, scan, records, tile frames
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Block-wide exclusive scan of one int per thread (256 threads).
__device__ __forceinline__ int block_exscan256(int v, int* lds, int* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wid] = x;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    int t = lds[w];
    if (w < wid) base += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

// Segment s -> fwd[s] ends exactly at stream point s+1 ("continues",
// brk = 0) unless it closes a loop or its end is the periodic image of the
// forward point (reference seg_pwrap, vec3_maths.c:105-138): brk = 1, and
// the point emits a second record.  A point without a forward link (the
// pinned end of an open chain) has no segment: brk = 2, no second record.
// Per block: exclusive scan of the second-record flags (rank[s]) and the
// block total.  Each wavefront also writes the bounding box of its 64-point
// target tile, and with fill_adj the adjacency entries of its points start
// at -1 (open chains leave some of them unset).
// Bounding box of each 64-point target tile.
__device__ __forceinline__ void target_tile(const double* __restrict__ P, int n, long cap, int T, int lane,
                                            TgtTile* __restrict__ tiles) {
  const int s = T * TILE + lane;
  double r[3];
  if (s < n) {
    load3(P, cap, s, r);
  } else {
    load3(P, cap, T * TILE, r);
  }
  TgtTile t;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    t.lo[i] = wave_min(r[i]);
    t.hi[i] = wave_max(r[i]);
  }
  if (lane == 0) tiles[T] = t;
}

__device__ __forceinline__ int cont_flag(const double* __restrict__ P, const int* __restrict__ fwd,
                                         int s, long cap, const BoxDev& box) {
  const int f = fwd[s];
  if (f < 0) return 2;
  double a[3], pf[3], b[3];
  load3(P, cap, s, a);
  load3(P, cap, f, pf);
  pwrap(a, pf, box, b);
  const bool cont = (f == s + 1) && b[0] == pf[0] && b[1] == pf[1] && b[2] == pf[2];
  return cont ? 0 : 1;
}

__global__ void __launch_bounds__(256) k_cont(const double* __restrict__ P,
                                              const int* __restrict__ fwd, int n, long cap,
                                              BoxDev box, int* __restrict__ brk,
                                              int* __restrict__ rank, int* __restrict__ bsum,
                                              TgtTile* __restrict__ tgt, int fill_adj,
                                              int* __restrict__ adj) {
  __shared__ int lds[4];
  const int s = blockIdx.x * 256 + threadIdx.x;
  const int T = __builtin_amdgcn_readfirstlane(s >> 6);
  if (T * TILE < n) target_tile(P, n, cap, T, threadIdx.x & 63, tgt);
  if (fill_adj && s < n) {
    adj[2 * s] = -1;
    adj[2 * s + 1] = -1;
  }
  int nc = 0;
  if (s < n) {
    const int b = cont_flag(P, fwd, s, cap, box);
    brk[s] = b;
    nc = b == 1 ? 1 : 0;
  }
  int tot;
  int r = block_exscan256(nc, lds, &tot);
  if (s < n) rank[s] = r;
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Exclusive scan of the block totals in place (one block of 1024 threads);
// writes the record count M = n + sum and the source tile count.
__global__ void __launch_bounds__(1024) k_scan_blocks(int* __restrict__ bsum, int nb, int n,
                                                      int* __restrict__ counts) {
  __shared__ int lds[16];
  __shared__ int carry_s;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    int i = base + threadIdx.x;
    int v = i < nb ? bsum[i] : 0;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    int pre = 0, tot = 0;
    for (int w = 0; w < 16; ++w) {
      int t = lds[w];
      if (w < wid) pre += t;
      tot += t;
    }
    int carry = carry_s;
    if (i < nb) bsum[i] = carry + pre + x - v;
    __syncthreads();
    if (threadIdx.x == 0) carry_s = carry + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int M = n + carry_s;
    counts[0] = M;
    counts[1] = (M + TILE - 1) / TILE;
  }
}

// Records of point s at offset off = s + (breaks before s): the point
// itself, closing segment s-1 when that one continues, then, for a break,
// the end of segment s (the forward point's minimum image).
__device__ __forceinline__ void point_records(const double* __restrict__ P, const int* __restrict__ fwd,
                                              int s, long cap, const BoxDev& box,
                                              const int* __restrict__ brk, int off,
                                              NearRec* __restrict__ near) {
  double a[3];
  load3(P, cap, s, a);
  NearRec r0;
  r0.p[0] = a[0];
  r0.p[1] = a[1];
  r0.p[2] = a[2];
  if (s > 0 && brk[s - 1] == 0) {
    r0.seg = s - 1;
    r0.seg_fwd = s;
  } else {
    r0.seg = -1;
    r0.seg_fwd = -1;
  }
  near[off] = r0;
  if (brk[s] == 1) {
    int f = fwd[s];
    double pf[3], b[3];
    load3(P, cap, f, pf);
    pwrap(a, pf, box, b);
    NearRec r1;
    r1.p[0] = b[0];
    r1.p[1] = b[1];
    r1.p[2] = b[2];
    r1.seg = s;
    r1.seg_fwd = f;
    near[off + 1] = r1;
  }
}

__global__ void __launch_bounds__(256) k_records(const double* __restrict__ P,
                                                 const int* __restrict__ fwd, int n, long cap,
                                                 BoxDev box, const int* __restrict__ brk,
                                                 const int* __restrict__ rank,
                                                 const int* __restrict__ bsum,
                                                 NearRec* __restrict__ near) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  point_records(P, fwd, s, cap, box, brk, s + bsum[blockIdx.x] + rank[s], near);
}

// One 64-thread block per source tile: frame centre = centre of the bounding
// box of the tile's records and of the record before the tile; FP32 records
// in that frame with the FP64 segment vector rounded once; padding records
// beyond M (unlinked, at the centre) complete the last tile.
__device__ __forceinline__ void source_tile(NearRec* __restrict__ near, int M, int J, int lane,
                                            SrcTile* __restrict__ tiles, FarRec* __restrict__ far) {
  const int e = J * TILE + lane;
  const bool valid = e < M;
  NearRec r;
  if (valid) r = near[e];
  double prev[3];
  {
    int pe = J > 0 ? J * TILE - 1 : 0;
    NearRec q = near[pe];
    prev[0] = q.p[0];
    prev[1] = q.p[1];
    prev[2] = q.p[2];
  }
  double lo[3], hi[3], c[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double v = valid ? r.p[i] : prev[i];
    double mn = fmin(v, prev[i]), mx = fmax(v, prev[i]);
    lo[i] = wave_min(mn);
    hi[i] = wave_max(mx);
    c[i] = 0.5 * (lo[i] + hi[i]);
  }
  // previous record of each lane (lane 0: the record before the tile)
  double pp[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double v = valid ? r.p[i] : c[i];
    double up = __shfl_up(v, 1, 64);
    pp[i] = lane == 0 ? prev[i] : up;
  }
  FarRec fr;
  if (valid) {
    const bool link = r.seg >= 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      fr.p[i] = (float)(r.p[i] - c[i]);
      fr.d[i] = link ? 0.5f * (float)(r.p[i] - pp[i]) : 0.f;
    }
  } else {
    NearRec pad;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      fr.p[i] = 0.f;
      fr.d[i] = 0.f;
      pad.p[i] = c[i];
    }
    pad.seg = -1;
    pad.seg_fwd = -1;
    near[e] = pad;
  }
  far[e] = fr;
  const double dmax2 =
      4.0 * wave_max((double)(fr.d[0] * fr.d[0] + fr.d[1] * fr.d[1] + fr.d[2] * fr.d[2]));
  if (lane == 0) {
    SrcTile t;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      t.c[i] = c[i];
      t.lo[i] = lo[i];
      t.hi[i] = hi[i];
      t.prev[i] = prev[i];
      t.prev_loc[i] = (float)(prev[i] - c[i]);
    }
    t.dmax2 = (float)dmax2;
    tiles[J] = t;
  }
}

// One 64-thread block per source tile (source_tile), and for each target
// the records closing its two adjacent segments (the one that starts at it
// and the one that ends at it, reference tangle.c:612-615): adj[2 s] and
// adj[2 s + 1].  A wrapped adjacent segment lies a box length away in the
// primary frame, so the sum must find it wherever it is.
__global__ void __launch_bounds__(64) k_src_tiles(NearRec* __restrict__ near,
                                                  const int* __restrict__ counts,
                                                  SrcTile* __restrict__ tiles,
                                                  FarRec* __restrict__ far, int* __restrict__ adj) {
  const int M = counts[0], nst = counts[1];
  const int J = blockIdx.x;
  if (J >= nst) return;
  const int e = J * TILE + threadIdx.x;
  if (e < M) {
    const NearRec r = near[e];
    if (r.seg >= 0) {
      adj[2 * r.seg] = e;
      adj[2 * r.seg_fwd + 1] = e;
    }
  }
  source_tile(near, M, J, threadIdx.x, tiles, far);
}



// ---------------------------------------------------------------------------
// direct Biot-Savart sum over all segments and images
// ---------------------------------------------------------------------------

// Per pair the straight-segment law (reference segment_field1,
// tangle.c:486-514),
//   v = kappa/4pi (l + l') / (l l' (l l' + R.R')) R x R',  R = a - r, R' = b - r,
// is evaluated as (u + u') u u' / (1 + u u' R.R') R x R' with u = 1/l,
// u' = 1/l'; a record's distance serves the segment it closes and the one it
// starts.  The reference returns exactly 0 when l or l' < 1e-8 or
// |R.R'/(l l') - 1| < 1e-8; that rule is kept: pairs whose cheap estimate
// falls in a narrow band around a threshold are decided by the reference's
// own FP64 expression (guard_exact).

// Reference decision of the zero-result rules for one pair, evaluated as the
// reference does (left-to-right sums from 0, no fused multiply-add).
__device__ __forceinline__ bool guard_exact(const double a[3], const double b[3],
                                            const double r[3]) {
#pragma clang fp contract(off)
  double R[3], Rp[3];
  for (int i = 0; i < 3; ++i) {
    R[i] = a[i] - r[i];
    Rp[i] = b[i] - r[i];
  }
  double l2 = 0, lp2 = 0, dd = 0;
  for (int i = 0; i < 3; ++i) {
    l2 += R[i] * R[i];
    lp2 += Rp[i] * Rp[i];
    dd += R[i] * Rp[i];
  }
  double l = sqrt(l2), lp = sqrt(lp2);
  if (l < 1e-8 || lp < 1e-8) return true;
  double nd = dd / l / lp;
  return fabs(nd - 1) < 1e-8;
}

struct DirectArgs {
  const double* P;  // target positions (SoA, stride cap)
  long cap;
  int n;
  int n_tgt_tiles;
  const TgtTile* tgt;
  const SrcTile* src;
  const int* counts;  // [M, n_src_tiles]
  const FarRec* far;
  const NearRec* near;
  const ImgDev* img;
  const int* adj;   // [2 n] records of each target's adjacent segments
  int n_img;
  double rnear2;    // squared box distance of the FP64 near field (mixed)
  double kappa4pi;
  double* partial;  // [n_chunks][3][cap]
  int n_chunks;
};

// One FP64 pair: segment from the previous record (R_prev, |R_prev|^2,
// 1/|R_prev|) to the record at p, for the image target rs; returns the
// distance data of p for the next pair.  nd1 = |cos - 1| estimates the
// reference's |R.R'/(l l') - 1|.
struct Pair64 {
  double R[3], l2, u;
  double c[3], f, nd1;
};
__device__ __forceinline__ Pair64 pair64(const double Rp[3], double l2p, double up,
                                         const double p[3], const double rs[3]) {
  Pair64 o;
  o.R[0] = p[0] - rs[0];
  o.R[1] = p[1] - rs[1];
  o.R[2] = p[2] - rs[2];
  o.l2 = o.R[0] * o.R[0] + o.R[1] * o.R[1] + o.R[2] * o.R[2];
  o.u = rsq64(o.l2);
  o.c[0] = Rp[1] * o.R[2] - Rp[2] * o.R[1];
  o.c[1] = Rp[2] * o.R[0] - Rp[0] * o.R[2];
  o.c[2] = Rp[0] * o.R[1] - Rp[1] * o.R[0];
  double dab = Rp[0] * o.R[0] + Rp[1] * o.R[1] + Rp[2] * o.R[2];
  double w = up * o.u;
  double tt = fma(w, dab, 1.0);
  o.nd1 = fabs(tt - 2.0);
  o.f = (up + o.u) * w * rcp64(tt);
  return o;
}

// A source tile's FP64 records staged in LDS, one copy per wavefront, read
// back by broadcast LDS reads (the small-tangle kernels): in small tangles
// every tile is near, each wavefront's work is a few tiles, and reading the
// records one by one through the scalar cache leaves the pass waiting on
// memory.  lane e writes record e;
// only this wavefront reads or writes the copy, and its LDS operations
// complete in order.
struct NearLds {
  NearRec rec[TILE];
};
__device__ __forceinline__ void stage_near(const NearRec* __restrict__ near, int base, int lane,
                                           NearLds* lds) {
  lds->rec[lane] = near[base + lane];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FP64 pass over one source tile for the image target rs of lane target s.
// Zero rules (l, l' < 1e-8; |cos - 1| < 1e-8) are decided on the estimate,
// which carries ~1e-15 error: records where some lane's estimate lies within
// 1e-13 of the cosine threshold or 1e-6 relative of l^2 = 1e-16 are noted
// in a wave-uniform mask (and, per lane, the band flag and the decision
// taken) and re-decided after the pass with the reference's own expression,
// correcting the sum where the decisions differ.
__device__ __forceinline__ void near_tile(const NearRec* __restrict__ lrec, const NearRec* __restrict__ rec,
                                          const SrcTile& st, int base, const double rs[3], int s,
                                          bool primary, double acc[3]) {
  const double R0[3] = {st.prev[0] - rs[0], st.prev[1] - rs[1], st.prev[2] - rs[2]};
  double Rp[3] = {R0[0], R0[1], R0[2]};
  double l2p = Rp[0] * Rp[0] + Rp[1] * Rp[1] + Rp[2] * Rp[2];
  double up = rsq64(l2p);
  uint64_t bandmask = 0, lane_band = 0, lane_zero = 0;
#pragma unroll 2
  for (int e = 0; e < TILE; ++e) {
    const NearRec q = lrec[e];
    Pair64 o = pair64(Rp, l2p, up, q.p, rs);
    const bool skip = q.seg < 0 || (primary && (q.seg == s || q.seg_fwd == s));
    const bool zero = l2p < 1e-16 || o.l2 < 1e-16 || o.nd1 < 1e-8;
    const bool band = fabs(l2p - 1e-16) < 1e-22 || fabs(o.l2 - 1e-16) < 1e-22 ||
                      fabs(o.nd1 - 1e-8) < 1e-13;
    const double f = (skip || zero) ? 0.0 : o.f;
    acc[0] = fma(f, o.c[0], acc[0]);
    acc[1] = fma(f, o.c[1], acc[1]);
    acc[2] = fma(f, o.c[2], acc[2]);
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(band && !skip) != 0, 0)) {
      bandmask |= 1ull << e;
      if (band && !skip) lane_band |= 1ull << e;
      if (band && !skip && zero) lane_zero |= 1ull << e;
    }
    Rp[0] = o.R[0];
    Rp[1] = o.R[1];
    Rp[2] = o.R[2];
    l2p = o.l2;
    up = o.u;
  }
  while (__builtin_expect(bandmask != 0, 0)) {
    const int e = __builtin_ctzll(bandmask);
    bandmask &= bandmask - 1;
    double a[3], P[3];
    const NearRec q = rec[base + e];
    if (e == 0) {
      a[0] = st.prev[0];
      a[1] = st.prev[1];
      a[2] = st.prev[2];
    } else {
      const NearRec qp = rec[base + e - 1];
      a[0] = qp.p[0];
      a[1] = qp.p[1];
      a[2] = qp.p[2];
    }
    P[0] = a[0] - rs[0];
    P[1] = a[1] - rs[1];
    P[2] = a[2] - rs[2];
    const double pl2 = P[0] * P[0] + P[1] * P[1] + P[2] * P[2];
    const double pu = rsq64(pl2);
    Pair64 o = pair64(P, pl2, pu, q.p, rs);
    if ((lane_band >> e) & 1) {
      const bool zero_est = (lane_zero >> e) & 1;
      const bool zero = guard_exact(a, q.p, rs);
      if (zero != zero_est) {
        const double f = zero ? -o.f : o.f;
        acc[0] = fma(f, o.c[0], acc[0]);
        acc[1] = fma(f, o.c[1], acc[1]);
        acc[2] = fma(f, o.c[2], acc[2]);
      }
    }
  }
}

// One FP32 pair: segment from the previous record (R_prev, |R_prev|^2,
// 1/|R_prev|) to record q; returns the point distance data of q for the next
// pair.  g = sin^2 - 2e-8 is the zero-rule estimate.
struct Pair32 {
  float R[3], l2, u;
  float c[3], f, g;
};
__device__ __forceinline__ Pair32 pair32(const float Rp[3], float l2p, float up, const FarRec& qh,
                                         const float off[3]) {
  const FarRec q = {{qh.p[0], qh.p[1], qh.p[2]}, {2.f * qh.d[0], 2.f * qh.d[1], 2.f * qh.d[2]}};
  Pair32 o;
  o.R[0] = q.p[0] + off[0];
  o.R[1] = q.p[1] + off[1];
  o.R[2] = q.p[2] + off[2];
  o.l2 = o.R[0] * o.R[0] + o.R[1] * o.R[1] + o.R[2] * o.R[2];
  o.u = __builtin_amdgcn_rsqf(o.l2);
  o.c[0] = Rp[1] * q.d[2] - Rp[2] * q.d[1];
  o.c[1] = Rp[2] * q.d[0] - Rp[0] * q.d[2];
  o.c[2] = Rp[0] * q.d[1] - Rp[1] * q.d[0];
  float dab = l2p + (Rp[0] * q.d[0] + Rp[1] * q.d[1] + Rp[2] * q.d[2]);
  float w = up * o.u;
  float tt = fmaf(w, dab, 1.f);
  float c2 = o.c[0] * o.c[0] + o.c[1] * o.c[1] + o.c[2] * o.c[2];
  o.g = fmaf(c2, w * w, -2e-8f);
  o.f = (up + o.u) * w * __builtin_amdgcn_rcpf(tt);
  return o;
}

// FP32 pass over one source tile in the tile's frame: R = p_loc + off with
// off = c_tile - rs rounded once per tile; FP32 partial sums of 8 records are
// added into the lane's FP64 sum; the cross product is taken with
// the FP64-rounded segment vector d (R_prev x d = R_prev x R), free of the
// cancellation of R_prev x R.  Zero rule: 1 - cos < 1e-8 <=> sin^2 < 2e-8
// here (cos > 0 in the far field), sin^2 = |R_prev x d|^2 u_prev^2 u^2.
// Records where some lane's estimate lies within a relative 5e-4 of the
// threshold are noted in a wave-uniform mask (and, per lane, whether it was
// in the band and what it decided) and re-decided after the pass with the
// reference's FP64 expression, correcting the sum where the decisions differ.
__device__ __forceinline__ void far_tile(const FarRec* __restrict__ rec,
                                         const NearRec* __restrict__ nrec, const SrcTile& st,
                                         int base, const float off[3], const double rs[3],
                                         double acc[3]) {
  const float R0[3] = {st.prev_loc[0] + off[0], st.prev_loc[1] + off[1], st.prev_loc[2] + off[2]};
  float Rp[3] = {R0[0], R0[1], R0[2]};
  float l2p = Rp[0] * Rp[0] + Rp[1] * Rp[1] + Rp[2] * Rp[2];
  float up = __builtin_amdgcn_rsqf(l2p);
  uint64_t bandmask = 0, lane_band = 0, lane_zero = 0;
  for (int e0 = 0; e0 < TILE; e0 += 8) {
    FarRec q[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) q[k] = rec[base + e0 + k];
    float a8[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      Pair32 o = pair32(Rp, l2p, up, q[k], off);
      float f = o.g < 0.f ? 0.f : o.f;
      a8[0] = fmaf(f, o.c[0], a8[0]);
      a8[1] = fmaf(f, o.c[1], a8[1]);
      a8[2] = fmaf(f, o.c[2], a8[2]);
      const bool band = fabsf(o.g) < 1e-11f;
      if (__builtin_expect(__builtin_amdgcn_ballot_w64(band) != 0, 0)) {
        bandmask |= 1ull << (e0 + k);
        if (band) lane_band |= 1ull << (e0 + k);
        if (band && o.g < 0.f) lane_zero |= 1ull << (e0 + k);
      }
      Rp[0] = o.R[0];
      Rp[1] = o.R[1];
      Rp[2] = o.R[2];
      l2p = o.l2;
      up = o.u;
    }
    acc[0] += (double)a8[0];
    acc[1] += (double)a8[1];
    acc[2] += (double)a8[2];
  }
  while (__builtin_expect(bandmask != 0, 0)) {
    const int e = __builtin_ctzll(bandmask);
    bandmask &= bandmask - 1;
    float P[3];
    if (e == 0) {
      P[0] = R0[0];
      P[1] = R0[1];
      P[2] = R0[2];
    } else {
      const FarRec& qp = rec[base + e - 1];
      P[0] = qp.p[0] + off[0];
      P[1] = qp.p[1] + off[1];
      P[2] = qp.p[2] + off[2];
    }
    float pl2 = P[0] * P[0] + P[1] * P[1] + P[2] * P[2];
    float pu = __builtin_amdgcn_rsqf(pl2);
    Pair32 o = pair32(P, pl2, pu, rec[base + e], off);
    if ((lane_band >> e) & 1) {
      const bool zero32 = (lane_zero >> e) & 1;
      const double* a = e > 0 ? nrec[base + e - 1].p : st.prev;
      const bool zero = guard_exact(a, nrec[base + e].p, rs);
      if (zero != zero32) {
        const double f = zero ? -(double)o.f : (double)o.f;
        acc[0] += f * (double)o.c[0];
        acc[1] += f * (double)o.c[1];
        acc[2] += f * (double)o.c[2];
      }
    }
  }
}

// Far field at small angles.  Every pair of a source tile whose longest
// segment subtends at most theta (sin^2 theta <= SMALL_ANGLE2) from all
// targets has s = sin^2 of the angle between R and R' below SMALL_ANGLE2, and
//   2 / (1 + cos) = 1 + s/4 + s^2/8 + ...,
// so 2f = 2 (u + u') u u' / (1 + cos) needs no reciprocal: 1 + s/4 is within
// SMALL_ANGLE2^2 / 8 = 2e-8 relative (within 1e-10 for the 1e-4 cm spacing at
// the 0.02 cm near radius).  With the half segment vector of FarRec the cross
// product is c/2 = R x d/2 and |c/2|^2 u^2 u'^2 = s/4, the quantity the zero
// rule tests (s < 2e-8), so the per-pair reciprocal and R.R' drop out and
// (2f)(c/2) = f c.  2f = b + b (s/4) with b = (u + u') u u' is rounded once:
// rounding 1/2 + s/8 instead would truncate the small term to 0 for most far
// pairs (it lies below half an ulp of 1/2) and bias every pair low.
constexpr float SMALL_ANGLE2 = 4e-4f;

// v_rsq_f32 is within 1 ulp but not unbiased: its mean relative error over
// operands spread across binades is -7.4e-9 (measured on gfx950 against the
// correctly rounded value, whose mean is -5e-12).  f carries three of these
// reciprocal square roots ((u + u') u u') and, unlike the reciprocal form,
// nothing compensates them, so the per-pair mean error of 2f (c/2) is
// -2.654e-8 (measured over 2^20 far pairs at 0.02-0.3 cm).  Summed over the
// ~1e13 pairs of an evaluation that bias is coherent, so each tile's sum is
// scaled back by this factor (per-pair rms error, 1.2e-7, is unchanged).
constexpr double RSQ_BIAS_FIX = 1.0 + 2.654e-8;

// One FP32 small-angle pair (the FP32 arithmetic of far_tile_fast, written
// with explicit fusions so that a recomputation gives the same bits):
// c = R_prev x d/2 and g = s/4 - ZT, with the zero rule's threshold ZT =
// 5e-9 (s < 2e-8) folded into the fused multiply-add that forms s/4, so that
// the rule is g < 0 and its band |g| < 2.5e-12.  f = b (1 + g) is 2f scaled by
// 1 - ZT / (1 + s/4) (to within 5e-13 of the constant 1 - ZT, which the
// tile's sum takes back).
struct PairS {
  float R[3], l2, u;
  float c[3], f, g;
};
constexpr float SMALL_ZT = 5e-9f;
__device__ __forceinline__ PairS pair_small(const float Rp[3], float up, const FarRec& q,
                                            const float off[3]) {
#pragma clang fp contract(off)
  PairS o;
  o.R[0] = q.p[0] + off[0];
  o.R[1] = q.p[1] + off[1];
  o.R[2] = q.p[2] + off[2];
  o.l2 = fmaf(o.R[2], o.R[2], fmaf(o.R[1], o.R[1], o.R[0] * o.R[0]));
  o.u = __builtin_amdgcn_rsqf(o.l2);
  o.c[0] = fmaf(Rp[1], q.d[2], -(Rp[2] * q.d[1]));
  o.c[1] = fmaf(Rp[2], q.d[0], -(Rp[0] * q.d[2]));
  o.c[2] = fmaf(Rp[0], q.d[1], -(Rp[1] * q.d[0]));
  const float c2 = fmaf(o.c[2], o.c[2], fmaf(o.c[1], o.c[1], o.c[0] * o.c[0]));
  const float w = up * o.u;
  o.g = fmaf(c2, w * w, -SMALL_ZT);
  const float b = (up + o.u) * w;
  o.f = fmaf(b, o.g, b);
  return o;
}
// zero rule s < 2e-8 and the band re-decided exactly, in units of s/4
__device__ __forceinline__ bool small_zero(float g) { return g < 0.f; }
// small_zero(g) ? 0 : f through VCC: a compare into VCC and v_cndmask_b32_e32
// both dual-issue on gfx950, where the compiler's own choice (compares into
// SGPR pairs, kept live across the unrolled group, and v_cndmask_b32_e64
// reading them) puts an SGPR operand on the select, which never dual-issues.
__device__ __forceinline__ float zero_rule_select(float g, float f) {
  float r;
  asm("v_cmp_ngt_f32_e32 vcc, 0, %1\n\tv_cndmask_b32_e32 %0, 0, %2, vcc"
      : "=v"(r)
      : "v"(g), "v"(f)
      : "vcc");
  return r;
}
constexpr float SMALL_BAND = 2.5e-12f;
__device__ __forceinline__ bool small_band(float g) { return fabsf(g) < SMALL_BAND; }
// per-tile scale of the fast pass: the mean v_rsq_f32 bias and the folded
// threshold (1 + ZT)
constexpr double SMALL_TILE_SCALE = RSQ_BIAS_FIX * (1.0 + (double)SMALL_ZT);

// A source tile's FP32 far records staged in LDS, one copy per wavefront:
// point and half segment vector of record e as two float4 (w unused), so the
// pass reads each with one broadcast LDS read into VGPRs.  On gfx950 a
// VALU instruction with an SGPR operand is never dual-issued, while
// instructions on VGPRs (and inline or literal constants) pair with another
// wave's (SQ_ACTIVE_INST_VALU2); with the records in SGPRs (scalar loads)
// nine of the pair's operations read an SGPR.
struct FarLds {
  float4 v[TILE][2];
};

// Copy source tile J's records into the wavefront's LDS copy (lane e writes
// record e).  Only this wavefront reads or writes the copy, and its LDS
// operations complete in order.
__device__ __forceinline__ void stage_far(const FarRec* __restrict__ far, int base, int lane,
                                          FarLds* lds) {
  const FarRec r = far[base + lane];
  lds->v[lane][0] = make_float4(r.p[0], r.p[1], r.p[2], 0.f);
  lds->v[lane][1] = make_float4(r.d[0], r.d[1], r.d[2], 0.f);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FP32 pass over one source tile at small angles (far_tile without the
// reciprocal), reading the records from the wavefront's LDS copy (lds; rec
// holds the same records for the rare band re-decision).  Zero-rule band
// detection: each record's band test (|g| < band, the threshold held in a
// VGPR) becomes a wave mask; one branch per 8-record group, and only when
// some lane of the group is in a band are that group's records re-evaluated
// (pair_small gives the same bits) to note which lanes were in a band and
// what they decided.  acc receives this tile's sum only (it starts at zero
// and is scaled by SMALL_TILE_SCALE at the end).
__device__ __forceinline__ void far_tile_fast(const FarLds* lds, const FarRec* __restrict__ rec,
                                              const NearRec* __restrict__ nrec, const SrcTile& st,
                                              int base, const float off[3], const double rs[3],
                                              float band, double acc[3]) {
  const float R0[3] = {st.prev_loc[0] + off[0], st.prev_loc[1] + off[1], st.prev_loc[2] + off[2]};
  float Rp[3] = {R0[0], R0[1], R0[2]};
  float up = __builtin_amdgcn_rsqf(fmaf(Rp[2], Rp[2], fmaf(Rp[1], Rp[1], Rp[0] * Rp[0])));
  uint64_t bandmask = 0, lane_band = 0, lane_zero = 0;
  for (int e0 = 0; e0 < TILE; e0 += 8) {
    FarRec q[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float4 a = lds->v[e0 + k][0], b = lds->v[e0 + k][1];
      q[k] = {{a.x, a.y, a.z}, {b.x, b.y, b.z}};
    }
    const float Rg[3] = {Rp[0], Rp[1], Rp[2]};
    const float ug = up;
    float a8[3] = {0.f, 0.f, 0.f};
    uint64_t any = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const PairS o = pair_small(Rp, up, q[k], off);
      const float f = zero_rule_select(o.g, o.f);
      a8[0] = fmaf(f, o.c[0], a8[0]);
      a8[1] = fmaf(f, o.c[1], a8[1]);
      a8[2] = fmaf(f, o.c[2], a8[2]);
      any |= __builtin_amdgcn_ballot_w64(fabsf(o.g) < band);
      Rp[0] = o.R[0];
      Rp[1] = o.R[1];
      Rp[2] = o.R[2];
      up = o.u;
    }
    acc[0] += (double)a8[0];
    acc[1] += (double)a8[1];
    acc[2] += (double)a8[2];
    if (__builtin_expect(any != 0, 0)) {
      float Rb[3] = {Rg[0], Rg[1], Rg[2]};
      float ub = ug;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const PairS o = pair_small(Rb, ub, q[k], off);
        const bool inb = fabsf(o.g) < band;
        if (__builtin_amdgcn_ballot_w64(inb) != 0) bandmask |= 1ull << (e0 + k);
        if (inb) lane_band |= 1ull << (e0 + k);
        if (inb && small_zero(o.g)) lane_zero |= 1ull << (e0 + k);
        Rb[0] = o.R[0];
        Rb[1] = o.R[1];
        Rb[2] = o.R[2];
        ub = o.u;
      }
    }
  }
  while (__builtin_expect(bandmask != 0, 0)) {
    const int e = __builtin_ctzll(bandmask);
    bandmask &= bandmask - 1;
    float P[3];
    if (e == 0) {
      P[0] = R0[0];
      P[1] = R0[1];
      P[2] = R0[2];
    } else {
      const FarRec& qp = rec[base + e - 1];
      P[0] = qp.p[0] + off[0];
      P[1] = qp.p[1] + off[1];
      P[2] = qp.p[2] + off[2];
    }
    const float pu = __builtin_amdgcn_rsqf(fmaf(P[2], P[2], fmaf(P[1], P[1], P[0] * P[0])));
    const PairS o = pair_small(P, pu, rec[base + e], off);
    if ((lane_band >> e) & 1) {
      const bool zero32 = (lane_zero >> e) & 1;
      const double* a = e > 0 ? nrec[base + e - 1].p : st.prev;
      const bool zero = guard_exact(a, nrec[base + e].p, rs);
      if (zero != zero32) {
        const double f = zero ? -(double)o.f : (double)o.f;
        acc[0] += f * (double)o.c[0];
        acc[1] += f * (double)o.c[1];
        acc[2] += f * (double)o.c[2];
      }
    }
  }
  acc[0] *= SMALL_TILE_SCALE;
  acc[1] *= SMALL_TILE_SCALE;
  acc[2] *= SMALL_TILE_SCALE;
}

// A lane's target: position, stream index, whether it is a real point, and
// the source tiles holding the records of its two adjacent segments.
struct LaneTarget {
  double r[3];
  int s;
  bool valid;
  int adjA, adjB;
};
// Source tile of the record that closes a target's adjacent segment; -1
// where the target has no such segment (a chain end) or is not a point.
__device__ __forceinline__ int adj_tile(int rec) { return rec >= 0 ? rec / TILE : -1; }

// Work items [w0, w1) of one target tile: item w = (source tile J_base +
// w / n_img, image w % n_img), walked with an incremental (tile, image)
// counter; the mirror axis is wave-uniform, so it selects by scalar branches.  MIXED: FP32 pass for source tiles beyond the
// near radius; otherwise all FP64.  Adds the raw sums (without kappa/4pi).
template <bool MIXED, bool STAGE_NEAR>
__device__ __forceinline__ void direct_items(const SrcTile* __restrict__ src,
                                             const FarRec* __restrict__ far,
                                             const NearRec* __restrict__ near,
                                             const ImgDev* __restrict__ img, int n_img, int J_base,
                                             const TgtTile& tt, const LaneTarget& L, long w0,
                                             long w1, double rnear2, FarLds* lds, NearLds* nlds,
                                             double acc[3]) {
  const int lane = threadIdx.x & 63;
  // the band threshold as a VGPR operand (a VOP3 compare cannot take a
  // literal, and an SGPR operand would keep it from dual issue)
  float band = SMALL_BAND;
  asm volatile("" : "+v"(band));
  int staged = -1;   // source tile held in lds
  int staged_near = -1;   // source tile held in nlds
  int Jl = (int)(w0 / n_img);
  int I = (int)(w0 - (long)Jl * n_img);
  for (long w = w0; w < w1; ++w, ++I) {
    if (I == n_img) {
      I = 0;
      ++Jl;
    }
    const int J = J_base + Jl;
    const ImgDev im = img[I];
    const SrcTile sj = src[J];
    // image target and the target tile's box under the same map
    double rs[3], tlo[3], thi[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      rs[i] = L.r[i] - im.shift[i];
      tlo[i] = tt.lo[i] - im.shift[i];
      thi[i] = tt.hi[i] - im.shift[i];
    }
    if (__builtin_expect(im.axis >= 0, 0)) {
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
        if (ax == im.axis) {
          rs[ax] = im.wall2 - L.r[ax];
          double l = im.wall2 - tt.hi[ax], h = im.wall2 - tt.lo[ax];
          tlo[ax] = l;
          thi[ax] = h;
        }
    }
    double d2 = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double g = fmax(0.0, fmax(sj.lo[i] - thi[i], tlo[i] - sj.hi[i]));
      d2 += g * g;
    }
    double v[3] = {0, 0, 0};
    // the primary sum skips each target's adjacent segments: a tile holding
    // one (possibly wrapped, hence far away) goes through the FP64 pass
    const bool adjacent = im.primary && __builtin_amdgcn_ballot_w64(J == L.adjA || J == L.adjB) != 0;
    if (MIXED && d2 >= rnear2 && !adjacent) {
      float off[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) off[i] = (float)(sj.c[i] - rs[i]);
      // d2 bounds every pair's l^2 from below, so dmax2 / d2 bounds sin^2
      if ((double)sj.dmax2 <= (double)SMALL_ANGLE2 * d2) {
        if (J != staged) {
          stage_far(far, J * TILE, lane, lds);
          staged = J;
        }
        far_tile_fast(lds, far, near, sj, J * TILE, off, rs, band, v);
      } else
        far_tile(far, near, sj, J * TILE, off, rs, v);
    } else {
      if (STAGE_NEAR) {
        if (J != staged_near) {
          stage_near(near, J * TILE, lane, nlds);
          staged_near = J;
        }
        near_tile(nlds->rec, near, sj, J * TILE, rs, L.s, im.primary != 0, v);
      } else {
        near_tile(near + J * TILE, near, sj, J * TILE, rs, L.s, im.primary != 0, v);
      }
    }
    if (__builtin_expect(im.axis >= 0, 0)) {
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
        if (ax == im.axis) v[ax] = -v[ax];
    }
    acc[0] += v[0];
    acc[1] += v[1];
    acc[2] += v[2];
  }
}

// Grid: blockIdx.x = group of 4 target tiles (one wavefront each),
// blockIdx.y = chunk of the (source tile, image) work list.  Each wave writes
// its chunk's sum to partial[chunk]; k_reduce adds the chunks in order.
template <bool MIXED, bool STAGE_NEAR>
__device__ __forceinline__ void direct_body(const DirectArgs& A) {
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int T = blockIdx.x * 4 + wid;
  if (T >= A.n_tgt_tiles) return;
  const int chunk = blockIdx.y;
  const long n_work = (long)A.counts[1] * A.n_img;
  const long w0 = (n_work * chunk) / A.n_chunks;
  const long w1 = (n_work * (chunk + 1)) / A.n_chunks;
  LaneTarget L;
  L.s = T * TILE + lane;
  L.valid = L.s < A.n;
  load3(A.P, A.cap, L.valid ? L.s : T * TILE, L.r);
  L.adjA = L.valid ? adj_tile(A.adj[2 * L.s]) : -1;
  L.adjB = L.valid ? adj_tile(A.adj[2 * L.s + 1]) : -1;
  const TgtTile tt = A.tgt[T];
  double acc[3] = {0, 0, 0};
  __shared__ FarLds lds[MIXED ? 4 : 1];
  __shared__ NearLds nlds[STAGE_NEAR ? 4 : 1];
  direct_items<MIXED, STAGE_NEAR>(A.src, A.far, A.near, A.img, A.n_img, 0, tt, L, w0, w1, A.rnear2,
                                  &lds[MIXED ? wid : 0], &nlds[STAGE_NEAR ? wid : 0], acc);
  if (L.valid) {
    double* out = A.partial + (long)chunk * 3 * A.cap;
    out[L.s] = acc[0] * A.kappa4pi;
    out[A.cap + L.s] = acc[1] * A.kappa4pi;
    out[2 * A.cap + L.s] = acc[2] * A.kappa4pi;
  }
}

// Mixed form: at most 128 VGPRs (4 waves per SIMD); the fast FP32 loop keeps
// its registers (spills stay in the rare band/correction paths) and the extra
// wave per SIMD hides the scalar-load latency of the record groups (6 % faster
// than the unconstrained 148-VGPR build at 200k points).
__global__ void __launch_bounds__(256, 4) k_direct_mixed(DirectArgs A) { direct_body<true, false>(A); }
__global__ void __launch_bounds__(256) k_direct_fp64(DirectArgs A) { direct_body<false, false>(A); }
// Small tangles (at most SMALL_TANGLE_TILES target tiles): each wave's work
// is one or a few source tiles, and staging their FP64 records in LDS takes
// the scalar-load latency out of the near pass (4096-point ring: 68 -> 57 us
// FP64, 61 -> 54 us mixed); at large sizes, where a wave walks thousands of
// tiles, it does not pay.
constexpr int SMALL_TANGLE_TILES = 128;
__global__ void __launch_bounds__(256, 4) k_direct_mixed_small(DirectArgs A) { direct_body<true, true>(A); }
__global__ void __launch_bounds__(256) k_direct_fp64_small(DirectArgs A) { direct_body<false, true>(A); }

// Grid (points / 256, 3): one component per blockIdx.y, chunks added in order.
__global__ void __launch_bounds__(256) k_reduce(const double* __restrict__ partial, int n_chunks,
                                                int n, long cap, double* __restrict__ bs) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (s >= n) return;
  const double* p = partial + i * cap + s;
  double a = 0;
#pragma unroll 8
  for (int c = 0; c < n_chunks; ++c) a += p[(long)c * 3 * cap];
  bs[i * cap + s] = a;
}
// ---------------------------------------------------------------------------
// per-point epilogue: vs and vels
// ---------------------------------------------------------------------------

// Reference update_velocity tail (tangle.c:596-683):
//   vs   = v_LIA + vs_ext(r, t) + v_BS
//   vels = vs - [|s''| > c/dl_max] alpha_h s' x v_LIA
//             + alpha s' x (vn(r, t) - vs) - alpha' s' x (s' x (vn(r, t) - vs))
__device__ __forceinline__ void epilogue_point(int s, const double* __restrict__ P,
                                               const double* __restrict__ T,
                                               const double* __restrict__ Nn,
                                               const double* __restrict__ lia,
                                               const double* __restrict__ bs, int have_bs,
                                               long cap, double time, const PhysDev& ph,
                                               int pc, double* __restrict__ vs_out,
                                               double* __restrict__ vel_out,
                                               int* __restrict__ flag) {
  double r[3], t[3], nm[3], l[3], vs[3], v[3], ext[3];
  load3(P, cap, s, r);
  if (pc >= 0 && !((pc >> 3) & 1)) {
    // PINNED: v_s = 0, the point moves with the boundary (tangle.c:589-594)
    const double z[3] = {0, 0, 0};
    flow_eval(ph.vb, r, time, v);
    store3(vs_out, cap, s, z);
    store3(vel_out, cap, s, v);
    if (!(isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]))) atomicOr(flag, 1);
    return;
  }
  load3(T, cap, s, t);
  load3(Nn, cap, s, nm);
  load3(lia, cap, s, l);
  flow_eval(ph.vs, r, time, ext);
#pragma unroll
  for (int i = 0; i < 3; ++i) vs[i] = l[i] + ext[i];
  if (have_bs) {
    double b[3];
    load3(bs, cap, s, b);
#pragma unroll
    for (int i = 0; i < 3; ++i) vs[i] += b[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] = vs[i];
  if (ph.hyperfriction) {
    double spp = norm3(nm);
    if (spp > ph.max_curv / ph.dl_max) {
      double h[3] = {t[1] * l[2] - t[2] * l[1], t[2] * l[0] - t[0] * l[2], t[0] * l[1] - t[1] * l[0]};
#pragma unroll
      for (int i = 0; i < 3; ++i) v[i] += h[i] * -ph.hyperalpha;
    }
  }
  if (ph.use_mf) {
    double vn[3];
    flow_eval(ph.vn, r, time, vn);
    double dv[3] = {vn[0] - vs[0], vn[1] - vs[1], vn[2] - vs[2]};
    double a1[3] = {t[1] * dv[2] - t[2] * dv[1], t[2] * dv[0] - t[0] * dv[2],
                    t[0] * dv[1] - t[1] * dv[0]};
    double a2[3] = {t[1] * a1[2] - t[2] * a1[1], t[2] * a1[0] - t[0] * a1[2],
                    t[0] * a1[1] - t[1] * a1[0]};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      v[i] += a1[i] * ph.alpha;
      v[i] += a2[i] * -ph.alpha_p;
    }
  }
  if (pc >= 0) {
    // PINNED_SLIP: the velocity normal to the wall is removed (tangle.c:677-683)
    const int ax = (pc & 7) >> 1;
    const double sg = (pc & 1) ? -1.0 : 1.0;   // boundary_normals: +e on the low face
    double nrm[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (a == ax) nrm[a] = sg;
    const double nv = nrm[0] * v[0] + nrm[1] * v[1] + nrm[2] * v[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) v[i] -= nrm[i] * nv;
  }
  store3(vs_out, cap, s, vs);
  store3(vel_out, cap, s, v);
  bool bad = !(isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]) && isfinite(vs[0]) &&
               isfinite(vs[1]) && isfinite(vs[2]));
  if (bad) atomicOr(flag, 1);
}

__global__ void __launch_bounds__(256) k_epilogue(const double* __restrict__ P,
                                                  const double* __restrict__ T,
                                                  const double* __restrict__ Nn,
                                                  const double* __restrict__ lia,
                                                  const double* __restrict__ bs, int have_bs, int n,
                                                  long cap, double time, PhysDev ph,
                                                  const int* __restrict__ pin,
                                                  double* __restrict__ vs_out,
                                                  double* __restrict__ vel_out,
                                                  int* __restrict__ flag) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  epilogue_point(s, P, T, Nn, lia, bs, have_bs, cap, time, ph, pin ? pin[s] : -1, vs_out, vel_out, flag);
}

// ---------------------------------------------------------------------------
// RK4 stages (reference vortex_dynamics.c:36-52, 58-117)
// ---------------------------------------------------------------------------

// dst = src for the first n entries of three planes of stride cap (a kernel,
// so that captured graphs hold kernels only and copy what is used, whatever
// the buffers' capacity).
__global__ void __launch_bounds__(256) k_copy3(const double* __restrict__ src, double* __restrict__ dst,
                                               int n, long cap) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  dst[s] = src[s];
  dst[cap + s] = src[cap + s];
  dst[2 * cap + s] = src[2 * cap + s];
}

// Xs = X0 + h K (euler_step2: move = v dt, then x + move, each rounded as the
// reference rounds it: no fused multiply-add)
__global__ void __launch_bounds__(256) k_stage(const double* __restrict__ X0,
                                               const double* __restrict__ K, double h, int n,
                                               long cap, double* __restrict__ Xs) {
#pragma clang fp contract(off)
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double mv = K[i * cap + s] * h;
    Xs[i * cap + s] = X0[i * cap + s] + mv;
  }
}

// X0 += dt/6 k1 + dt/3 k2 + dt/3 k3 + dt/6 k4, accumulated in that order
// with every product and sum rounded separately (vortex_dynamics.c:99-110).
__global__ void __launch_bounds__(256) k_combine(double* __restrict__ X0,
                                                 const double* __restrict__ K1,
                                                 const double* __restrict__ K2,
                                                 const double* __restrict__ K3,
                                                 const double* __restrict__ K4, double dt, int n,
                                                 long cap, int* __restrict__ flag) {
#pragma clang fp contract(off)
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  bool bad = false;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    long o = i * cap + s;
    double mv = K1[o] * (dt / 6);
    mv = mv + K2[o] * (dt / 3);
    mv = mv + K3[o] * (dt / 3);
    mv = mv + K4[o] * (dt / 6);
    double x = X0[o] + mv;
    X0[o] = x;
    bad |= !isfinite(x);
  }
  if (bad) atomicOr(flag, 2);
}

// Periodic wrap (reference enforce_boundaries, tangle.c:766-819): while the
// point is beyond a periodic face (the last offending face in x, y, z order),
// move it by one box length.  Non-finite points are left alone.
__device__ __forceinline__ void wrap_point(int s, double* __restrict__ X0, long cap,
                                           const BoxDev& box, int wall_periodic_mask) {
  double x[3];
  load3(X0, cap, s, x);
  if (!(isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]))) return;
  for (int it = 0; it < 64; ++it) {
    int face = -1;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      if (x[ax] < box.lo[ax]) face = 2 * ax;
      if (x[ax] > box.hi[ax]) face = 2 * ax + 1;
    }
    if (face < 0 || !((wall_periodic_mask >> face) & 1)) break;
    int ax = face / 2;
    if (face & 1)
      x[ax] -= box.L[ax];
    else
      x[ax] += box.L[ax];
  }
  store3(X0, cap, s, x);
}

// Image list passed by value (kernel argument), written to device memory by
// k_set_images: a host-to-device copy without a staging buffer.
struct ImgList {
  ImgDev im[MAX_IMAGES];
};
__global__ void __launch_bounds__(64) k_set_images(ImgList L, int n, ImgDev* __restrict__ dst) {
  const int i = threadIdx.x;
  if (i < n) dst[i] = L.im[i];
}

// Zero one int (a fill kernel, not a memset node, so it may be captured).
__global__ void __launch_bounds__(64) k_zero_int(int* __restrict__ p) {
  if (threadIdx.x == 0) *p = 0;
}

__global__ void __launch_bounds__(256) k_wrap(double* __restrict__ X0, int n, long cap, BoxDev box,
                                              int wall_periodic_mask) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  wrap_point(s, X0, cap, box, wall_periodic_mask);
}

}  // namespace ov
