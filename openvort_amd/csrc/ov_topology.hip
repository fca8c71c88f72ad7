// ov_topology.hip — topology edits of a device-resident tangle (reference
// src/tangle.c:766-1070): remesh with point removal and insertion,
// elimination of small loops and of loops near the origin / the z axis,
// periodic wrap, and the stream order of the point stream rebuilt on the
// device.
//
// These kernels work in the caller's index space: arrays of the tangle's
// capacity N indexed like the caller's tangle (positions as three planes of
// stride `cap`), status[k] < 0 for EMPTY points, links fwd / rev as caller
// indices.  New points take the slot the reference's get_tangle_next_free
// would give them, so indices, links and positions follow the reference.
// Tangles with points pinned on walls carry open chains (a pinned point
// may lack one link) and a pin code per index (ov_walls.hip's codes, with
// the forward-index bit kept up to date as links change); their geometry
// is the walls' (step_node with mirrors, one-sided and neighbour-tangent
// stencils).
#include <hip/hip_runtime.h>

#include "ov_device.h"
#include "ov_walls.hip"

namespace ov {

struct TopoArrays {
  double* X;      // positions, 3 planes of stride cap
  double* T;      // s'
  double* Nn;     // s''
  double* dxi;
  int* fwd;
  int* rev;
  int* status;    // OV_EMPTY / OV_FREE / OV_PINNED / OV_PINNED_SLIP
  int* pin;       // pin codes (walls), nullptr when no point is pinned
  long cap;
  int N;          // the tangle's capacity (indices 0..N-1)
};

// |s0 - minimum image of s1| with the reference's rounding (seg_pwrap,
// vec3_dist: differences, then the sum of squares left to right, no fused
// multiply-add), so that dxi is the reference's bit for bit.
__device__ __forceinline__ double chord_ref(const double s0[3], const double s1[3], const BoxDev& b) {
#pragma clang fp contract(off)
  double w[3];
  pwrap(s0, s1, b, w);
  double d0 = s0[0] - w[0], d1 = s0[1] - w[1], d2 = s0[2] - w[2];
  double o = 0;
  o += d0 * d0;
  o += d1 * d1;
  o += d2 * d2;
  return sqrt(o);
}

// dxi of point k (reference initialize_dxi, tangle.c:704-721).
__device__ __forceinline__ void topo_dxi(const TopoArrays& A, const BoxDev& box, int k) {
  double a[3], b[3];
  load3(A.X, A.cap, k, a);
  load3(A.X, A.cap, A.fwd[k], b);
  A.dxi[k] = chord_ref(a, b, box);
}

// s', s'' of point k from the current positions and dxi (the five-node
// stencil of geometry_point, on the index-space arrays).
__device__ __forceinline__ void topo_tn(const TopoArrays& A, const BoxDev& box, int k) {
  const int f = A.fwd[k], r = A.rev[k], ff = A.fwd[f], rr = A.rev[r];
  double p0[3], pf[3], pr[3], pff[3], prr[3];
  load3(A.X, A.cap, k, p0);
  load3(A.X, A.cap, f, pf);
  load3(A.X, A.cap, r, pr);
  load3(A.X, A.cap, ff, pff);
  load3(A.X, A.cap, rr, prr);
  const double dk = A.dxi[k], df = A.dxi[f], dr = A.dxi[r], drr = A.dxi[rr];
  const double x[4] = {dk + df, dk, -dr, -(dr + drr)};
  const double* pj[4] = {pff, pf, pr, prr};
  double t[3] = {0, 0, 0}, nm[3] = {0, 0, 0};
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
  store3(A.T, A.cap, k, t);
  store3(A.Nn, A.cap, k, nm);
}

// Geometry of the points whose five-node windows a change touched: up to 8
// indices in aff[], refreshed by one lane each (dxi first, then s', s'').
__device__ __forceinline__ void refresh_geometry(const TopoArrays& A, const BoxDev& box,
                                                 const int* aff, int naff) {
  const int lane = threadIdx.x;
  if (lane < naff) topo_dxi(A, box, aff[lane]);
  __threadfence_block();
  __syncthreads();
  if (lane < naff) topo_tn(A, box, aff[lane]);
  __threadfence_block();
  __syncthreads();
}

// The same refresh with walls (the geometry of update_tangents_normals
// with pinned points: dxi, then every point but those next to a PINNED one,
// then those from their neighbours' s').  A point's s', s'' depend on the
// points and dxi up to four links away (the one-sided stencil of a PINNED
// point), and those next to a PINNED point on its s': the caller lists
// every point within six links of a change.  The forward-index bit of a
// pinned point's code follows its current forward link.  Returns false
// (on every lane) when a walk does not end.
__device__ __forceinline__ bool refresh_geometry_walls(const TopoArrays& A, const BoxDev& box,
                                                       const int* aff, int naff) {
  const int lane = threadIdx.x;
  const int s = lane < naff ? aff[lane] : -1;
  if (s >= 0 && A.pin[s] >= 0) A.pin[s] = (A.pin[s] & 15) | (A.fwd[s] > 0 ? 16 : 0);
  __threadfence_block();
  __syncthreads();
  bool ok = true;
  if (s >= 0) ok = wall_dxi_point(s, A.X, A.cap, A.fwd, A.rev, A.pin, box, A.dxi);
  __threadfence_block();
  __syncthreads();
  const bool herm = s >= 0 && wall_hermite(s, A.fwd, A.rev, A.pin);
  for (int phase = 1; phase <= 2; ++phase) {
    if (s >= 0 && herm == (phase == 2)) {
      double p0[3], pn[3], pp[3], t[3], nm[3];
      load3(A.X, A.cap, s, p0);
      bool w = wall_walk(s, 1, A.X, A.cap, A.fwd, A.rev, A.pin, box, pn);
      w = wall_walk(s, -1, A.X, A.cap, A.fwd, A.rev, A.pin, box, pp) && w;
      if (w && wall_tn_point(s, herm, p0, pn, pp, A.X, A.cap, A.fwd, A.rev, A.pin, box, A.T, A.dxi,
                             t, nm)) {
        store3(A.T, A.cap, s, t);
        store3(A.Nn, A.cap, s, nm);
      } else {
        ok = false;
      }
    }
    __threadfence_block();
    __syncthreads();
  }
  return __builtin_amdgcn_ballot_w64(!ok) == 0;
}

// Whether point s takes the neighbour-tangent stencil (not PINNED, a linked
// neighbour PINNED), on the index-space arrays.
__device__ __forceinline__ bool topo_hermite(const TopoArrays& A, int s) {
  if (A.status[s] == OV_PINNED) return false;
  const int f = A.fwd[s], r = A.rev[s];
  return (f >= 0 && A.status[f] == OV_PINNED) || (r >= 0 && A.status[r] == OV_PINNED);
}

// remove_point (tangle.c:1017-1066) with merge direction dir (+1: merge into
// the forward neighbour, -1: into the reverse one).  Scalar work of lane 0.
__device__ __forceinline__ void remove_point_ref(const TopoArrays& A, int k, int dir) {
#pragma clang fp contract(off)
  const int prev = A.rev[k], next = A.fwd[k];
  if (dir != 0) {
    int other;
    double l;
    if (dir > 0) {
      other = next;
      l = A.dxi[k];
    } else {
      other = prev;
      l = -A.dxi[prev];
    }
    if (A.status[other] == OV_FREE) {
      double s0[3], s0p[3], s0pp[3], s1[3], s1p[3], s1pp[3], nw[3], a[3];
      load3(A.X, A.cap, k, s0);
      load3(A.T, A.cap, k, s0p);
      load3(A.Nn, A.cap, k, s0pp);
      load3(A.X, A.cap, other, s1);
      load3(A.T, A.cap, other, s1p);
      load3(A.Nn, A.cap, other, s1pp);
      for (int i = 0; i < 3; ++i) {
        a[i] = s0[i] + s1[i];
        nw[i] = a[i] * 0.5;
      }
      const double lh = l / 2;
      for (int i = 0; i < 3; ++i) {
        a[i] = s0p[i] - s1p[i];
        a[i] = a[i] * lh;
        nw[i] = nw[i] + a[i];
      }
      const double lq = l * l / 16.0;
      for (int i = 0; i < 3; ++i) {
        a[i] = s0pp[i] + s1pp[i];
        a[i] = a[i] * lq;
        nw[i] = nw[i] + a[i];
      }
      store3(A.X, A.cap, other, nw);
    }
  }
  if (prev >= 0) A.fwd[prev] = next;
  if (next >= 0) A.rev[next] = prev;
  A.fwd[k] = -1;
  A.rev[k] = -1;
  A.status[k] = OV_EMPTY;
}

// The new point of add_point (tangle.c:1069-1160) between p and next: on the
// circle through the two points with the mean curvature of their s'' when
// that curvature is resolvable (below max_curv / dl_max and above 1e-8),
// else their midpoint.
__device__ __forceinline__ void add_point_position(const TopoArrays& A, const BoxDev& box, int p,
                                                   int next, double max_curv, double dl_max,
                                                   double out[3]) {
#pragma clang fp contract(off)
  double s0[3], s1r[3], s1[3], s0pp[3], s1pp[3], n[3];
  load3(A.X, A.cap, p, s0);
  load3(A.X, A.cap, next, s1r);
  pwrap(s0, s1r, box, s1);
  load3(A.Nn, A.cap, p, s0pp);
  load3(A.Nn, A.cap, next, s1pp);
  for (int i = 0; i < 3; ++i) {
    n[i] = s0pp[i] + s1pp[i];
    n[i] = n[i] * 0.5;
  }
  auto vd = [](const double v[3]) {
    double o = 0;
    o += v[0] * v[0];
    o += v[1] * v[1];
    o += v[2] * v[2];
    return sqrt(o);
  };
  const double nd = vd(n);
  const double lf = A.dxi[p];
  if (nd < max_curv / dl_max && nd > 1e-8) {
    const double R = 1 / vd(n);
    const double dR = R * R - lf * lf / 4;
    const double delta = dR > 0 ? R - sqrt(dR) : R;
    const double inv = 1 / vd(n);
    for (int i = 0; i < 3; ++i) {
      n[i] = n[i] * inv;
      n[i] = n[i] * -1;
    }
    for (int i = 0; i < 3; ++i) {
      double a = s0[i] + s1[i];
      a = a * 0.5;
      const double b = n[i] * delta;
      out[i] = a + b;
    }
  } else {
    for (int i = 0; i < 3; ++i) {
      out[i] = s0[i] + s1[i];
      out[i] = out[i] * 0.5;
    }
  }
}

// First EMPTY index at or after `from` (the wave scans 64 at a time);
// -1 when there is none below N.
__device__ __forceinline__ int first_empty(const TopoArrays& A, int from) {
  for (int b = from; b < A.N; b += 64) {
    const int i = b + (int)threadIdx.x;
    const bool e = i < A.N && A.status[i] == OV_EMPTY;
    const uint64_t m = __builtin_amdgcn_ballot_w64(e);
    if (m) return b + __builtin_ctzll(m);
  }
  return -1;
}

// remesh (tangle.c:824-873) with the reference's order of visits: point k
// in increasing index (points added at an index above the current one are
// visited later, as the reference's loop does), each change followed by the
// geometry refresh the reference's update_tangents_normals makes (only the
// points whose stencil window changed get different values, so only those
// are recomputed).  One wavefront: the 64 next indices are tested at once
// from the current state, the first one that changes anything is applied,
// and the scan resumes after it.
// state: [0] next_free (in/out), [1] points added, [2] 0 ok / 1 no EMPTY slot
// left (the reference would expand_tangle; nothing past that point is done)
// / 2 a removal left two neighbouring points between PINNED points, or a
// walk that does not end (the reference's update_tangent_normal would
// recurse without end; nothing past that point is done), [3] changes made.
__global__ void __launch_bounds__(64) k_remesh(TopoArrays A, BoxDev box, double min_dist,
                                               double max_dist, double max_curv, double dl_max,
                                               int* __restrict__ state) {
  const int lane = threadIdx.x;
  __shared__ int aff[16];
  __shared__ int n_aff;
  __shared__ int bad;
  if (lane == 0) bad = 0;
  int next_free = state[0];
  int low = 0;          // no EMPTY index below it
  int added = 0, changes = 0;
  int k = 0;
  while (k < A.N) {
    const int i = k + lane;
    int decision = 0;   // 1 remove merging forward, 2 remove merging backward, 3 add
    if (i < A.N && A.status[i] != OV_EMPTY) {
      const int nx = A.fwd[i], pv = A.rev[i];
      const double lf = nx >= 0 ? A.dxi[i] : 0.0;
      const double lr = pv >= 0 ? A.dxi[pv] : 0.0;
      if (nx >= 0 && pv >= 0 && !(A.rev[pv] < 0 && A.fwd[nx] < 0) &&
          ((lf < min_dist || lr < min_dist) && (lf + lr) < max_dist))
        decision = lf < min_dist ? 1 : 2;
      else if (nx >= 0 && lf > max_dist)
        decision = 3;
    }
    const uint64_t m = __builtin_amdgcn_ballot_w64(decision != 0);
    if (m == 0) {
      k += 64;
      continue;
    }
    const int j = __builtin_ctzll(m);
    const int c = k + j;
    const int dec = __shfl(decision, j, 64);
    int naff = 0;
    if (dec == 1 || dec == 2) {
      if (lane == 0) {
        const int prev = A.rev[c], next = A.fwd[c];
        remove_point_ref(A, c, dec == 1 ? 1 : -1);
        if (!A.pin) {
          // windows containing prev or next: two links either side
          const int r1 = A.rev[prev], r2 = A.rev[r1], f1 = A.fwd[next], f2 = A.fwd[f1];
          aff[0] = r2;
          aff[1] = r1;
          aff[2] = prev;
          aff[3] = next;
          aff[4] = f1;
          aff[5] = f2;
          naff = 6;
        } else {
          // six links either side, stopping at chain ends
          naff = 0;
          for (int x = prev, q = 0; q < 7 && x >= 0; ++q, x = A.rev[x]) aff[naff++] = x;
          for (int x = next, q = 0; q < 7 && x >= 0; ++q, x = A.fwd[x]) aff[naff++] = x;
          // two neighbouring points next to PINNED ones (or a chain end there)
          int b = 0;
          for (int q = 0; q < naff; ++q) {
            const int x = aff[q];
            if (!topo_hermite(A, x)) continue;
            const int f = A.fwd[x], r = A.rev[x];
            if (f < 0 || r < 0 || topo_hermite(A, f) || topo_hermite(A, r)) b = 1;
          }
          bad = b;
        }
      }
      low = min(low, c);
    } else {
      // get_tangle_next_free (tangle.c:729-757) without expansion
      int slot = -1;
      if (next_free < A.N && A.status[next_free] == OV_EMPTY) {
        slot = next_free++;
      } else {
        slot = first_empty(A, low);
        if (slot >= 0) {
          next_free = slot + 1;
          low = slot;
        }
      }
      if (slot < 0) {
        if (lane == 0) {
          state[0] = next_free;
          state[1] = added;
          state[2] = 1;
          state[3] = changes;
        }
        return;
      }
      if (lane == 0) {
        const int next = A.fwd[c];
        double nw[3];
        add_point_position(A, box, c, next, max_curv, dl_max, nw);
        store3(A.X, A.cap, slot, nw);
        A.status[slot] = OV_FREE;
        A.rev[slot] = c;
        A.fwd[slot] = next;
        A.fwd[c] = slot;
        A.rev[next] = slot;
        if (A.pin) A.pin[slot] = -1;
        if (!A.pin) {
          const int r1 = A.rev[c], r2 = A.rev[r1], f1 = A.fwd[next], f2 = A.fwd[f1];
          aff[0] = r2;
          aff[1] = r1;
          aff[2] = c;
          aff[3] = slot;
          aff[4] = next;
          aff[5] = f1;
          aff[6] = f2;
          naff = 7;
        } else {
          naff = 0;
          for (int x = c, q = 0; q < 7 && x >= 0; ++q, x = A.rev[x]) aff[naff++] = x;
          aff[naff++] = slot;
          for (int x = next, q = 0; q < 7 && x >= 0; ++q, x = A.fwd[x]) aff[naff++] = x;
        }
      }
      ++added;
    }
    ++changes;
    // a short loop lists a point more than once: refresh each point once
    if (lane == 0) {
      int w = 0;
      for (int q = 0; q < naff; ++q) {
        bool seen = false;
        for (int r = 0; r < w; ++r) seen = seen || aff[r] == aff[q];
        if (!seen) aff[w++] = aff[q];
      }
      n_aff = w;
    }
    __threadfence_block();
    __syncthreads();
    if (!A.pin) {
      refresh_geometry(A, box, aff, n_aff);
    } else if (bad || !refresh_geometry_walls(A, box, aff, n_aff)) {
      if (lane == 0) {
        state[0] = next_free;
        state[1] = added;
        state[2] = 2;
        state[3] = changes;
      }
      return;
    }
    k = c + 1;
  }
  if (lane == 0) {
    state[0] = next_free;
    state[1] = added;
    state[2] = 0;
    state[3] = changes;
  }
}

// enforce_boundaries (tangle.c:766-819) on the index-space arrays: EMPTY
// points are skipped.
__global__ void __launch_bounds__(256) k_wrap_index(double* __restrict__ X,
                                                    const int* __restrict__ status, int N,
                                                    long cap, BoxDev box, int wall_periodic_mask,
                                                    const int* __restrict__ only_if) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N || status[k] == OV_EMPTY) return;
  if (only_if && *only_if == 0) return;
  wrap_point(k, X, cap, box, wall_periodic_mask);
}

// ---- loops: labels, ranks, lengths --------------------------------------------
// Pointer jumping on the forward links of the closed loops: after enough
// rounds lab[k] is the lowest index of k's loop.  EMPTY points keep -1.
__global__ void __launch_bounds__(256) k_loop_init(const int* __restrict__ fwd,
                                                   const int* __restrict__ status, int N,
                                                   int* __restrict__ lab, int* __restrict__ jmp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const bool act = status[k] != OV_EMPTY;
  lab[k] = act ? k : -1;
  jmp[k] = act ? fwd[k] : -1;
}
__global__ void __launch_bounds__(256) k_loop_min(const int* __restrict__ lab_in,
                                                  const int* __restrict__ jmp_in, int N,
                                                  int* __restrict__ lab_out, int* __restrict__ jmp_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int j = jmp_in[k];
  if (j < 0) {
    lab_out[k] = lab_in[k];
    jmp_out[k] = j;
    return;
  }
  lab_out[k] = min(lab_in[k], lab_in[j]);
  jmp_out[k] = jmp_in[j];
}
// Ranks along the loop from its lowest index (the loop cut before it):
// d[k] = steps from the head, by pointer jumping on the reverse links.
__global__ void __launch_bounds__(256) k_rank_init(const int* __restrict__ rev,
                                                   const int* __restrict__ lab, int N,
                                                   int* __restrict__ d, int* __restrict__ jmp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  if (lab[k] < 0) {
    d[k] = 0;
    jmp[k] = -1;
    return;
  }
  const bool head = lab[k] == k;
  d[k] = head ? 0 : 1;
  jmp[k] = head ? -1 : rev[k];
}
__global__ void __launch_bounds__(256) k_rank_step(const int* __restrict__ d_in,
                                                   const int* __restrict__ jmp_in, int N,
                                                   int* __restrict__ d_out, int* __restrict__ jmp_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int j = jmp_in[k];
  if (j < 0) {
    d_out[k] = d_in[k];
    jmp_out[k] = -1;
    return;
  }
  d_out[k] = d_in[k] + d_in[j];
  jmp_out[k] = jmp_in[j];
}
// Loop length L at each loop head (rank of the point before it + 1), 0
// elsewhere: the counts whose exclusive scan places the loops in the point
// stream in order of their lowest index.
__global__ void __launch_bounds__(256) k_loop_len(const int* __restrict__ rev,
                                                  const int* __restrict__ lab,
                                                  const int* __restrict__ d, int N,
                                                  int* __restrict__ len) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  len[k] = (lab[k] == k) ? d[rev[k]] + 1 : 0;
}

// With open chains (points pinned on walls) the forward pass gives each
// point the lowest index of its chain's remainder; a second pass on the
// reverse links completes the chain's lowest index (for a loop it changes
// nothing).
__global__ void __launch_bounds__(256) k_jmp_rev(const int* __restrict__ rev,
                                                 const int* __restrict__ lab, int N,
                                                 int* __restrict__ jmp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < N) jmp[k] = lab[k] >= 0 ? rev[k] : -1;
}
// chain[lab] = 1 for a component with an end (a point without a reverse
// link, its head), counted in *n_chains.
__global__ void __launch_bounds__(256) k_chain_flag(const int* __restrict__ rev,
                                                    const int* __restrict__ lab, int N,
                                                    int* __restrict__ chain, int* __restrict__ n_chains) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N || lab[k] < 0 || rev[k] >= 0) return;
  chain[lab[k]] = 1;
  atomicAdd(n_chains, 1);
}
// Ranks with chains: a chain is ranked from its head (the stream walks it
// from there), a loop from its lowest index.
__global__ void __launch_bounds__(256) k_rank_init_ch(const int* __restrict__ rev,
                                                      const int* __restrict__ lab,
                                                      const int* __restrict__ chain, int N,
                                                      int* __restrict__ d, int* __restrict__ jmp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  if (lab[k] < 0) {
    d[k] = 0;
    jmp[k] = -1;
    return;
  }
  const bool head = chain[lab[k]] ? rev[k] < 0 : lab[k] == k;
  d[k] = head ? 0 : 1;
  jmp[k] = head ? -1 : rev[k];
}
// Component sizes at the label (len zeroed before): a loop's from its lowest
// index, a chain's from its last point.  eliminate_small_loops counts a
// chain of L points as L - 1 steps too (tangle.c:899-922: forward to the
// wall, then back from the start point to the other end).
__global__ void __launch_bounds__(256) k_loop_len_ch(const int* __restrict__ fwd,
                                                     const int* __restrict__ rev,
                                                     const int* __restrict__ lab,
                                                     const int* __restrict__ d,
                                                     const int* __restrict__ chain, int N,
                                                     int* __restrict__ len) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N || lab[k] < 0) return;
  if (chain[lab[k]]) {
    if (fwd[k] < 0) len[lab[k]] = d[k] + 1;
  } else if (lab[k] == k) {
    len[k] = d[rev[k]] + 1;
  }
}
// Stream pin codes: the index space's, with the forward-index bit of the
// caller's current forward link.
__global__ void __launch_bounds__(256) k_gather_pin(const int* __restrict__ pin,
                                                    const int* __restrict__ fwd,
                                                    const int* __restrict__ act, int n,
                                                    int* __restrict__ pin_s) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int k = act[s];
  pin_s[s] = pin[k] < 0 ? -1 : ((pin[k] & 15) | (fwd[k] > 0 ? 16 : 0));
}

// Loops to delete (eliminate_small_loops, tangle.c:878-945 and the
// origin / z-axis variants, 947-1015): flag[head] = 1 when the loop has fewer
// than loop_length + 1 points, or when every point lies within the origin
// cutoff, within the z-axis cutoff (inside) or beyond the outer cutoff.
// Per point: whether it satisfies each region test (bits 0..2).
__global__ void __launch_bounds__(256) k_loop_region(const double* __restrict__ X,
                                                     const int* __restrict__ lab, int N, long cap,
                                                     double r_origin, double r_zaxis, double r_outer,
                                                     int* __restrict__ bits) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  if (lab[k] < 0) {
    bits[k] = 0;
    return;
  }
  double p[3];
  load3(X, cap, k, p);
  double o = 0;
  o += p[0] * p[0];
  o += p[1] * p[1];
  o += p[2] * p[2];
  const double d = sqrt(o);
  const double zr = sqrt(p[0] * p[0] + p[1] * p[1]);
  bits[k] = (d < r_origin ? 1 : 0) | (zr < r_zaxis ? 2 : 0) | (zr > r_outer ? 4 : 0);
}
// AND of the region bits over each loop, at its head (one atomic per point).
__global__ void __launch_bounds__(256) k_loop_and(const int* __restrict__ lab,
                                                  const int* __restrict__ bits, int N,
                                                  int* __restrict__ all) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N || lab[k] < 0) return;
  atomicAnd(&all[lab[k]], bits[k]);
}
__global__ void __launch_bounds__(256) k_fill_int(int* __restrict__ p, int n, int v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}
// Delete every point of a flagged loop (remove_point with no merge on every
// point, or the EMPTY marking of eliminate_small_loops: the same result for
// a whole loop).  mask selects which region tests are on; a loop of L points
// goes when L - 1 < loop_length.
__global__ void __launch_bounds__(256) k_loop_delete(const int* __restrict__ lab,
                                                     const int* __restrict__ len,
                                                     const int* __restrict__ all, int N, int mask,
                                                     int loop_length, int* __restrict__ fwd,
                                                     int* __restrict__ rev, int* __restrict__ status) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N || lab[k] < 0) return;
  const int h = lab[k];
  const bool region = (all[h] & mask) != 0;
  const bool small = len[h] - 1 < loop_length;
  if (region || small) {
    fwd[k] = -1;
    rev[k] = -1;
    status[k] = OV_EMPTY;
  }
}

// Stream order of the active points: s(k) = off[lab[k]] + d[k] (loops in
// order of their lowest index, each walked forward from it), and the maps
// both ways.
__global__ void __launch_bounds__(256) k_stream_pos(const int* __restrict__ lab,
                                                    const int* __restrict__ d,
                                                    const int* __restrict__ off, int N,
                                                    int* __restrict__ pos, int* __restrict__ act) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  if (lab[k] < 0) {
    pos[k] = -1;
    return;
  }
  const int s = off[lab[k]] + d[k];
  pos[k] = s;
  act[s] = k;
}
// The octree's insertion order (points in ascending caller index, as the
// reference inserts them) as stream positions: flag[k] = 1 for an active
// index, then order[scan[k]] = pos[k] after an exclusive scan of the flags.
__global__ void __launch_bounds__(256) k_active_flags(const int* __restrict__ pos, int N,
                                                      int* __restrict__ flag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < N) flag[k] = pos[k] >= 0 ? 1 : 0;
}
__global__ void __launch_bounds__(256) k_tree_order(const int* __restrict__ pos,
                                                    const int* __restrict__ scan, int N,
                                                    int* __restrict__ order) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < N && pos[k] >= 0) order[scan[k]] = pos[k];
}

// Stream-order positions and links gathered from the index space.
__global__ void __launch_bounds__(256) k_gather_stream(const double* __restrict__ Xi, long icap,
                                                       const int* __restrict__ fwd,
                                                       const int* __restrict__ rev,
                                                       const int* __restrict__ act,
                                                       const int* __restrict__ pos, int n, long cap,
                                                       double* __restrict__ Xs, int* __restrict__ fwd_s,
                                                       int* __restrict__ rev_s) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int k = act[s];
  double p[3];
  load3(Xi, icap, k, p);
  store3(Xs, cap, s, p);
  fwd_s[s] = fwd[k] >= 0 ? pos[fwd[k]] : -1;
  rev_s[s] = rev[k] >= 0 ? pos[rev[k]] : -1;
}
// A stream-order plane set scattered to the index space.
__global__ void __launch_bounds__(256) k_scatter3(const double* __restrict__ Ps, long cap,
                                                  const int* __restrict__ act, int n, long icap,
                                                  double* __restrict__ Pi) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double p[3];
  load3(Ps, cap, s, p);
  store3(Pi, icap, act[s], p);
}
__global__ void __launch_bounds__(256) k_scatter1(const double* __restrict__ Ps,
                                                  const int* __restrict__ act, int n,
                                                  double* __restrict__ Pi) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  Pi[act[s]] = Ps[s];
}

}  // namespace ov
