// ov_batch.hip — ensembles of independent tangles in one launch per stage.
//
// Layout: the tangles' stream-ordered points are concatenated, each tangle
// starting at a multiple of 64 (slots between tangles are padding, tid = -1);
// each tangle owns a record region of ceil(2 n / 64) tiles (at most two
// records per point) of which the first S_k = ceil(M_k / 64) are used.
// The direct sum is one persistent launch: waves take work units from a
// queue; a unit is up to U (source tile, image) items of one target tile of
// one tangle.  Unit partial sums are reduced per target in unit order.
// Tangles with points pinned on walls carry open chains (a chain's last
// point starts no segment) and pin codes (ov_walls.hip) in the batch's pin
// array.
#include <hip/hip_runtime.h>

#include "ov_device.h"
#include "ov_walls.hip"

namespace ov {

struct TangleDev {
  BoxDev box;
  PhysDev ph;
  double kappa4pi;
  int pt_base, n;        // points [pt_base, pt_base + n)
  int rec_base;          // record region base (multiple of 64)
  int tgt_base, n_tgt;   // target tiles
  int img_base, n_img;   // images (primary first) in the batch image list
  int lia_only;
  int wall_mask;
  int pins;              // 1: points pinned on walls (the walls' geometry)
};

__global__ void __launch_bounds__(256) kb_geometry(const double* __restrict__ P,
                                                   const int* __restrict__ fwd,
                                                   const int* __restrict__ rev, long cap,
                                                   const int* __restrict__ tid,
                                                   const TangleDev* __restrict__ tg,
                                                   int compute_geom, double* __restrict__ T,
                                                   double* __restrict__ Nn,
                                                   double* __restrict__ dxi,
                                                   double* __restrict__ lia, int ntot) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= ntot) return;
  const int k = tid[s];
  if (k < 0) return;
  geometry_point(s, P, fwd, rev, cap, tg[k].box, tg[k].ph, compute_geom, T, Nn, dxi, lia);
}

// Geometry of a batch with points pinned on walls, as geometry() does it for
// one tangle: phase 0 dxi of the tangles with pins; phase 1 the free
// tangles' geometry and, in the others, every point but those next to a
// PINNED point; phase 2 those (compute_geom: s', s'' as well as the local
// induction; otherwise the caller's s', s'' in T, Nn and phase 1 only).
__global__ void __launch_bounds__(256) kb_wall_geometry(const double* __restrict__ P,
                                                        const int* __restrict__ fwd,
                                                        const int* __restrict__ rev,
                                                        const int* __restrict__ pin, long cap,
                                                        const int* __restrict__ tid,
                                                        const TangleDev* __restrict__ tg,
                                                        int compute_geom, int phase,
                                                        double* __restrict__ T,
                                                        double* __restrict__ Nn,
                                                        double* __restrict__ dxi,
                                                        double* __restrict__ lia,
                                                        int* __restrict__ flag, int ntot) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= ntot) return;
  const int k = tid[s];
  if (k < 0) return;
  const TangleDev& t = tg[k];
  if (!t.pins) {
    if (phase == 1) geometry_point(s, P, fwd, rev, cap, t.box, t.ph, compute_geom, T, Nn, dxi, lia);
    return;
  }
  if (phase == 0) {
    if (!wall_dxi_point(s, P, cap, fwd, rev, pin, t.box, dxi)) atomicOr(flag, 1);
    return;
  }
  const bool herm = compute_geom && wall_hermite(s, fwd, rev, pin);
  if (herm != (phase == 2)) return;
  double p0[3], pn[3], pp[3];
  load3(P, cap, s, p0);
  bool ok = wall_walk(s, 1, P, cap, fwd, rev, pin, t.box, pn);
  ok = wall_walk(s, -1, P, cap, fwd, rev, pin, t.box, pp) && ok;
  double tv[3], nm[3];
  if (ok && !compute_geom) {
    load3(T, cap, s, tv);
    load3(Nn, cap, s, nm);
  } else if (ok) {
    ok = wall_tn_point(s, herm, p0, pn, pp, P, cap, fwd, rev, pin, t.box, T, dxi, tv, nm);
  }
  if (!ok) {
    atomicOr(flag, 1);
    return;
  }
  if (compute_geom) {
    store3(T, cap, s, tv);
    store3(Nn, cap, s, nm);
  }
  double v[3];
  wall_lia_point(p0, pn, pp, tv, nm, t.box, t.ph, v);
  store3(lia, cap, s, v);
}

// Segment breaks (as k_cont) with the tangle's own box; padding slots have
// none; a chain's last point (no forward link) is 2: no segment, no record.
__global__ void __launch_bounds__(256) kb_cont(const double* __restrict__ P,
                                               const int* __restrict__ fwd, long cap,
                                               const int* __restrict__ tid,
                                               const TangleDev* __restrict__ tg,
                                               int* __restrict__ brk, int* __restrict__ rank,
                                               int* __restrict__ bsum, int ntot) {
  __shared__ int lds[4];
  const int s = blockIdx.x * 256 + threadIdx.x;
  int nc = 0;
  if (s < ntot) {
    const int k = tid[s];
    int b = 0;
    if (k >= 0) b = cont_flag(P, fwd, s, cap, tg[k].box);
    brk[s] = b;
    nc = b == 1 ? 1 : 0;
  }
  int tot;
  const int r = block_exscan256(nc, lds, &tot);
  if (s < ntot) rank[s] = r;
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__device__ __forceinline__ int scanned(const int* __restrict__ rank, const int* __restrict__ bsum,
                                       int i) {
  return bsum[i >> 8] + rank[i];
}

// Records of each tangle in its region; the last point of a tangle stores
// the tangle's record count M and source tile count S.
__global__ void __launch_bounds__(256) kb_records(const double* __restrict__ P,
                                                  const int* __restrict__ fwd, long cap,
                                                  const int* __restrict__ tid,
                                                  const TangleDev* __restrict__ tg,
                                                  const int* __restrict__ brk,
                                                  const int* __restrict__ rank,
                                                  const int* __restrict__ bsum,
                                                  NearRec* __restrict__ near, int* __restrict__ tM,
                                                  int* __restrict__ tS, int ntot) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= ntot) return;
  const int k = tid[s];
  if (k < 0) return;
  const TangleDev& t = tg[k];
  const int pb = t.pt_base;
  const int E0 = scanned(rank, bsum, pb);
  const int Es = scanned(rank, bsum, s);
  const int off = t.rec_base + (s - pb) + Es - E0;
  double a[3];
  load3(P, cap, s, a);
  NearRec r0;
  r0.p[0] = a[0];
  r0.p[1] = a[1];
  r0.p[2] = a[2];
  if (s > pb && brk[s - 1] == 0) {
    r0.seg = s - 1;
    r0.seg_fwd = s;
  } else {
    r0.seg = -1;
    r0.seg_fwd = -1;
  }
  near[off] = r0;
  if (brk[s] == 1) {
    const int f = fwd[s];
    double pf[3], b[3];
    load3(P, cap, f, pf);
    pwrap(a, pf, t.box, b);
    NearRec r1;
    r1.p[0] = b[0];
    r1.p[1] = b[1];
    r1.p[2] = b[2];
    r1.seg = s;
    r1.seg_fwd = f;
    near[off + 1] = r1;
  }
  if (s == pb + t.n - 1) {
    const int M = t.n + (Es + (brk[s] == 1 ? 1 : 0) - E0);
    tM[k] = M;
    tS[k] = (M + TILE - 1) / TILE;
  }
}

// Tile frames of each tangle's used source tiles (as k_src_tiles).
__global__ void __launch_bounds__(64) kb_src_tiles(NearRec* __restrict__ near,
                                                   const int* __restrict__ tile_tangle,
                                                   const TangleDev* __restrict__ tg,
                                                   const int* __restrict__ tM,
                                                   const int* __restrict__ tS,
                                                   SrcTile* __restrict__ tiles,
                                                   FarRec* __restrict__ far) {
  const int j = blockIdx.x;
  const int k = tile_tangle[j];
  const int rb = tg[k].rec_base;
  const int jl = j - rb / TILE;
  if (jl >= tS[k]) return;
  const int lane = threadIdx.x;
  const int e = j * TILE + lane;
  const bool valid = (e - rb) < tM[k];
  NearRec r;
  if (valid) r = near[e];
  double prev[3];
  {
    const NearRec q = near[jl > 0 ? j * TILE - 1 : rb];
    prev[0] = q.p[0];
    prev[1] = q.p[1];
    prev[2] = q.p[2];
  }
  double lo[3], hi[3], c[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double v = valid ? r.p[i] : prev[i];
    lo[i] = wave_min(fmin(v, prev[i]));
    hi[i] = wave_max(fmax(v, prev[i]));
    c[i] = 0.5 * (lo[i] + hi[i]);
  }
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
    tiles[j] = t;
  }
}

__global__ void __launch_bounds__(64) kb_tgt_tiles(const double* __restrict__ P, long cap,
                                                   const int* __restrict__ tgt_tangle,
                                                   const TangleDev* __restrict__ tg,
                                                   TgtTile* __restrict__ tiles) {
  const int T = blockIdx.x;
  const TangleDev& t = tg[tgt_tangle[T]];
  const int s = T * TILE + threadIdx.x;
  double r[3];
  load3(P, cap, (s - t.pt_base) < t.n ? s : T * TILE, r);
  TgtTile o;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    o.lo[i] = wave_min(r[i]);
    o.hi[i] = wave_max(r[i]);
  }
  if (threadIdx.x == 0) tiles[T] = o;
}

__global__ void __launch_bounds__(256) kb_adjacent(const NearRec* __restrict__ near,
                                                   const int* __restrict__ tile_tangle,
                                                   const TangleDev* __restrict__ tg,
                                                   const int* __restrict__ tM,
                                                   int* __restrict__ adj, int nrec) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nrec) return;
  const int k = tile_tangle[e / TILE];
  if (e - tg[k].rec_base >= tM[k]) return;
  const NearRec r = near[e];
  if (r.seg < 0) return;
  adj[2 * r.seg] = e;
  adj[2 * r.seg_fwd + 1] = e;
}

// Work units per target tile: ceil(S_k n_img / U) (0 for LIA-only tangles);
// block-wise exclusive scan as kb_cont.
__global__ void __launch_bounds__(256) kb_units(const int* __restrict__ tgt_tangle,
                                                const TangleDev* __restrict__ tg,
                                                const int* __restrict__ tS, int U,
                                                int* __restrict__ ucnt, int* __restrict__ rank,
                                                int* __restrict__ bsum, int n_tgt) {
  __shared__ int lds[4];
  const int T = blockIdx.x * 256 + threadIdx.x;
  int c = 0;
  if (T < n_tgt) {
    const int k = tgt_tangle[T];
    if (!tg[k].lia_only) {
      const long items = (long)tS[k] * tg[k].n_img;
      c = (int)((items + U - 1) / U);
    }
    ucnt[T] = c;
  }
  int tot;
  const int r = block_exscan256(c, lds, &tot);
  if (T < n_tgt) rank[T] = r;
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Queue reset (a kernel, so that a captured graph holds no memset node).
__global__ void kb_reset(int* __restrict__ q) {
  if (threadIdx.x == 0) q[0] = 0;
}

struct BatchArgs {
  const double* P;
  long cap;
  const TangleDev* tg;
  const int* tgt_tangle;
  int n_tgt;
  const TgtTile* tgt;
  const SrcTile* src;
  const FarRec* far;
  const NearRec* near;
  const ImgDev* img;
  const int* adj;
  const int* tS;
  const int* urank;      // unit offsets: ubsum[T >> 8] + urank[T]
  const int* ubsum;
  const int* ucounts;    // [0] = total units
  int* queue;
  int U;
  double rnear2;
  double* upart;         // [unit][64 lanes][3]
};

// Direct sum over all tangles' work units.  PERSIST: a fixed grid of
// wavefronts takes units from an atomic queue until none are left;
// otherwise one unit per wavefront (unit = block * 4 + wave; the grid is a
// bound on the unit count and the surplus waves exit at once).
template <bool MIXED, bool PERSIST>
__device__ __forceinline__ void batch_body(const BatchArgs& A) {
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  __shared__ FarLds lds_far[MIXED ? 4 : 1];
  __shared__ NearLds lds_near[1];   // not staged: a unit walks many source tiles
  const int total = A.ucounts[0];
  for (;;) {
    int u = 0;
    if (PERSIST) {
      if (lane == 0) u = atomicAdd(A.queue, 1);
      u = __builtin_amdgcn_readfirstlane(u);
    } else {
      u = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    }
    if (u >= total) break;
    // target tile of unit u: the last T with offset(T) <= u
    int lo = 0, hi = A.n_tgt - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (A.ubsum[mid >> 8] + A.urank[mid] <= u)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int T = lo;
    const TangleDev& t = A.tg[A.tgt_tangle[T]];
    const long items = (long)A.tS[A.tgt_tangle[T]] * t.n_img;
    const long w0 = (long)(u - (A.ubsum[T >> 8] + A.urank[T])) * A.U;
    const long w1 = w0 + A.U < items ? w0 + A.U : items;
    LaneTarget L;
    L.s = T * TILE + lane;
    L.valid = (L.s - t.pt_base) < t.n;
    load3(A.P, A.cap, L.valid ? L.s : T * TILE, L.r);
    L.adjA = L.valid ? adj_tile(A.adj[2 * L.s]) : -1;
    L.adjB = L.valid ? adj_tile(A.adj[2 * L.s + 1]) : -1;
    const TgtTile tt = A.tgt[T];
    double acc[3] = {0, 0, 0};
    direct_items<MIXED, false>(A.src, A.far, A.near, A.img + t.img_base, t.n_img, t.rec_base / TILE,
                               tt, L, w0, w1, A.rnear2, &lds_far[MIXED ? wid : 0], &lds_near[0], acc);
    double* o = A.upart + ((long)u * TILE + lane) * 3;
    o[0] = acc[0] * t.kappa4pi;
    o[1] = acc[1] * t.kappa4pi;
    o[2] = acc[2] * t.kappa4pi;
    if (!PERSIST) break;
  }
}

__global__ void __launch_bounds__(256, 4) kb_direct_mixed(BatchArgs A) { batch_body<true, true>(A); }
__global__ void __launch_bounds__(256) kb_direct_fp64(BatchArgs A) { batch_body<false, true>(A); }
__global__ void __launch_bounds__(256, 4) kb_grid_mixed(BatchArgs A) { batch_body<true, false>(A); }
__global__ void __launch_bounds__(256) kb_grid_fp64(BatchArgs A) { batch_body<false, false>(A); }

// Sum of each target's unit partials in unit order.
__global__ void __launch_bounds__(256) kb_reduce(const double* __restrict__ upart,
                                                 const int* __restrict__ ucnt,
                                                 const int* __restrict__ urank,
                                                 const int* __restrict__ ubsum,
                                                 const int* __restrict__ tid, long cap,
                                                 double* __restrict__ bs, int ntot) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= ntot || tid[s] < 0) return;
  const int T = s / TILE, lane = s % TILE;
  const int u0 = ubsum[T >> 8] + urank[T];
  double a[3] = {0, 0, 0};
  for (int u = u0; u < u0 + ucnt[T]; ++u) {
    const double* o = upart + ((long)u * TILE + lane) * 3;
    a[0] += o[0];
    a[1] += o[1];
    a[2] += o[2];
  }
  store3(bs, cap, s, a);
}

__global__ void __launch_bounds__(256) kb_epilogue(const double* __restrict__ P,
                                                   const double* __restrict__ T,
                                                   const double* __restrict__ Nn,
                                                   const double* __restrict__ lia,
                                                   const double* __restrict__ bs, long cap,
                                                   double time, const int* __restrict__ tid,
                                                   const TangleDev* __restrict__ tg,
                                                   const int* __restrict__ pin,
                                                   double* __restrict__ vs_out,
                                                   double* __restrict__ vel_out,
                                                   int* __restrict__ flag, int ntot) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= ntot) return;
  const int k = tid[s];
  if (k < 0) return;
  epilogue_point(s, P, T, Nn, lia, bs, !tg[k].lia_only, cap, time, tg[k].ph, pin ? pin[s] : -1, vs_out,
                 vel_out, flag);
}

__global__ void __launch_bounds__(256) kb_wrap(double* __restrict__ X0, long cap,
                                               const int* __restrict__ tid,
                                               const TangleDev* __restrict__ tg, int ntot) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= ntot) return;
  const int k = tid[s];
  if (k < 0) return;
  wrap_point(s, X0, cap, tg[k].box, tg[k].wall_mask);
}

}  // namespace ov
