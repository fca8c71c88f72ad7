// ov_adjoint.hip — vector-Jacobian products (VJPs) of the point velocity and
// of the RK4 step, FP64, direct sum, open / periodic / mirror images.
//
// The velocity of point k is vels_k = F(p_{k-2..k+2}, u_k) with u_k the
// Biot-Savart sum at p_k (reference update_velocity, tangle.c:583-684):
//   * local part F — stencil geometry (tangle.c:245-273, 432-479), LIA
//     (tangle.c:530-551), hyperfriction and mutual friction (tangle.c:641-675)
//     — differentiated per point with dual numbers (value + 3 seeded
//     directions; 6 seedings: the five stencil points and u_k); the five
//     transposed blocks are gathered back to the points they came from;
//   * the imposed flows vs, vn (external_velocity.c:176-512) at p_k and time
//     t — evaluated once per point on duals seeded along x, y, z, their 3x3
//     Jacobians entering the seeding of p_k only;
//   * the segment law (tangle.c:486-514) summed over segments and images:
//     for one pair, with R = a - r, R' = b - r, l = |R|, l' = |R'|,
//     d = R.R', q = l l' + d, g = l l' q, f = (l + l') / g, c = R x R',
//     B = f c, and for a cotangent w
//       w^T dB/dR  = f (R' x w) + (w.c) (f_l R / l  + f_d R')
//       w^T dB/dR' = f (w x R)  + (w.c) (f_l' R' / l' + f_d R)
//     with f_l = (1 - f l' (l l' + q)) / g, f_l' = (1 - f l (l l' + q)) / g,
//     f_d = -f l l' / g.  The target gets -(dB/dR + dB/dR'), the segment's
//     start dB/dR and its (minimum-image) end dB/dR'.  Pairs the reference
//     zeroes (l or l' < 1e-8, |R.R'/(l l') - 1| < 1e-8) contribute nothing.
// Branch points (the minimum image, the hyperfriction switch, the flows'
// cut-offs) are differentiated on the side the values select.
#include <hip/hip_runtime.h>

#include "ov_device.h"
#include "ov_flow.h"

namespace ov {

struct D3 {   // value and derivatives along 3 seeded directions
  double v, d[3];
};
__device__ inline D3 dc(double x) { return {x, {0, 0, 0}}; }
__device__ inline D3 operator+(const D3& a, const D3& b) {
  return {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
}
__device__ inline D3 operator-(const D3& a, const D3& b) {
  return {a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
}
__device__ inline D3 operator-(const D3& a) { return {-a.v, {-a.d[0], -a.d[1], -a.d[2]}}; }
__device__ inline D3 operator*(const D3& a, const D3& b) {
  return {a.v * b.v,
          {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]}};
}
__device__ inline D3 operator*(const D3& a, double s) {
  return {a.v * s, {a.d[0] * s, a.d[1] * s, a.d[2] * s}};
}
__device__ inline D3 operator*(double s, const D3& a) { return a * s; }
__device__ inline D3 operator/(const D3& a, const D3& b) {
  const double q = a.v / b.v;
  return {q, {(a.d[0] - q * b.d[0]) / b.v, (a.d[1] - q * b.d[1]) / b.v, (a.d[2] - q * b.d[2]) / b.v}};
}
__device__ inline D3 dsqrt(const D3& a) {
  const double r = sqrt(a.v), h = 0.5 / r;
  return {r, {a.d[0] * h, a.d[1] * h, a.d[2] * h}};
}
__device__ inline D3 dlog(const D3& a) { return {log(a.v), {a.d[0] / a.v, a.d[1] / a.v, a.d[2] / a.v}}; }
__device__ inline D3 operator+(const D3& a, double s) { return {a.v + s, {a.d[0], a.d[1], a.d[2]}}; }
__device__ inline D3 operator+(double s, const D3& a) { return a + s; }
__device__ inline D3 operator-(const D3& a, double s) { return {a.v - s, {a.d[0], a.d[1], a.d[2]}}; }
__device__ inline D3 operator-(double s, const D3& a) { return {s - a.v, {-a.d[0], -a.d[1], -a.d[2]}}; }
__device__ inline D3 operator/(const D3& a, double s) { return {a.v / s, {a.d[0] / s, a.d[1] / s, a.d[2] / s}}; }
__device__ inline D3 operator/(double s, const D3& a) { return dc(s) / a; }
__device__ inline D3 dscale(const D3& a, double f, double df) {   // g(a) with g = f, g' = df
  return {f, {a.d[0] * df, a.d[1] * df, a.d[2] * df}};
}
__device__ inline D3 dexp(const D3& a) {
  const double e = exp(a.v);
  return dscale(a, e, e);
}
__device__ inline D3 dcos(const D3& a) { return dscale(a, cos(a.v), -sin(a.v)); }
__device__ inline D3 dsin(const D3& a) { return dscale(a, sin(a.v), cos(a.v)); }

// The imposed flow of descriptor f at the dual position r and time t: the
// formulas of flow_eval (ov_flow.h, external_velocity.c:176-512) on dual
// numbers, so the point velocity's VJP carries the flows' spatial Jacobians.
// Branches (cut-offs) are taken on the values.
__device__ inline void dflow_eval(const ov_flow& f, const D3 r[3], double t, D3 out[3]) {
  const D3 x = r[0], y = r[1], z = r[2];
  const D3 zero = dc(0);
  auto spherical = [&](D3 o[3]) {
    const double strength = f.a[0], cutoff = f.a[1];
    const D3 d = dsqrt(x * x + y * y + z * z);
    const D3 attn = (cutoff / d) * (cutoff / d);
    if (d.v < 1e-8 || attn.v > 100) {
      o[0] = o[1] = o[2] = zero;
      return;
    }
    const D3 m = strength / (4 * FLOW_PI * d * d) * dexp(-attn);
    for (int i = 0; i < 3; ++i) o[i] = (r[i] / d) * m;
  };
  switch (f.type) {
    case OV_FLOW_SIMPLE:
      for (int i = 0; i < 3; ++i) out[i] = dc(f.v[i]);
      return;
    case OV_FLOW_SPHERICAL:
      spherical(out);
      return;
    case OV_FLOW_SPHERICAL_AND_SIMPLE:
      spherical(out);
      for (int i = 0; i < 3; ++i) out[i] = out[i] + f.v[i];
      return;
    case OV_FLOW_CYLINDRICAL: {
      const double strength = f.a[0], cutoff = f.a[1];
      const D3 d = dsqrt(x * x + y * y);
      const D3 attn = (cutoff / d) * (cutoff / d);
      if (d.v < 1e-8 || attn.v > 100) {
        out[0] = out[1] = out[2] = zero;
        return;
      }
      const D3 m = strength / (2 * FLOW_PI * d) * dexp(-attn);
      out[0] = (x / d) * m;
      out[1] = (y / d) * m;
      out[2] = zero;
      return;
    }
    case OV_FLOW_OSCILLATING: {
      const double strength = f.a[0], freq = f.a[1];
      const D3 kr = f.w[0] * x + f.w[1] * y + f.w[2] * z;
      const D3 m = dcos(2 * FLOW_PI * kr) * (strength * sin(2 * FLOW_PI * freq * t));
      for (int i = 0; i < 3; ++i) out[i] = f.q[i] * m;
      return;
    }
    case OV_FLOW_OSCILLATING_SLIT: {
      const double freq = f.a[1], cutoff = f.a[2];
      const D3 rr = dsqrt(x * x + z * z);
      const D3 m = dexp(-cutoff * cutoff / rr / rr) * cos(2 * FLOW_PI * freq * t);
      out[0] = (x / rr) * m;
      out[1] = zero;
      out[2] = (z / rr) * m;
      return;
    }
    case OV_FLOW_COSCOS: {
      const double k = f.a[0], v0 = f.a[1], l = f.a[2];
      out[0] = out[1] = zero;
      out[2] = v0 * (1 + dcos(k * x) * dcos(k * y) * dexp(-z / l));
      return;
    }
    case OV_FLOW_COSCOS_DIVFREE: {
      const double k = f.a[0], v0 = f.a[1], l = f.a[2];
      const D3 e = dexp(-z / l);
      out[0] = v0 / 2 / l / k * dsin(k * x) * dcos(k * y) * e;
      out[1] = v0 / 2 / l / k * dcos(k * x) * dsin(k * y) * e;
      out[2] = v0 * (1 + dcos(k * x) * dcos(k * y) * e);
      return;
    }
    case OV_FLOW_COS_DIVFREE: {
      const double k = f.a[0], v0 = f.a[1], l = f.a[2];
      const D3 e = dexp(-z / l);
      out[0] = v0 / l / k * dsin(k * x) * e;
      out[1] = zero;
      out[2] = v0 * (1 + dcos(k * x) * e);
      return;
    }
    case OV_FLOW_SIMPLE_SHEAR:
      out[0] = out[1] = zero;
      out[2] = f.a[1] + f.a[0] * x;
      return;
    case OV_FLOW_ROTOSC:
    case OV_FLOW_ROTOSC_BOUNDARY: {
      const double m = f.a[1] * cos(2 * FLOW_PI * f.a[0] * t);
      out[0] = -y * m;
      out[1] = x * m;
      out[2] = zero;
      return;
    }
    case OV_FLOW_DG:
      out[0] = -y * f.a[0];
      out[1] = x * f.a[0];
      out[2] = dc(f.a[1]);
      return;
    case OV_FLOW_CONSTANT_ACCELERATION:
      out[0] = dc(f.a[0] * t);
      out[1] = out[2] = zero;
      return;
    case OV_FLOW_PARABOLIC:
      out[0] = f.a[1] * z * (f.a[0] - z);
      out[1] = out[2] = zero;
      return;
    default:
      out[0] = out[1] = out[2] = zero;
      return;
  }
}

__device__ inline void dcross(const D3 a[3], const D3 b[3], D3 o[3]) {
  const D3 x = a[1] * b[2] - a[2] * b[1];
  const D3 y = a[2] * b[0] - a[0] * b[2];
  const D3 z = a[0] * b[1] - a[1] * b[0];
  o[0] = x;
  o[1] = y;
  o[2] = z;
}

// minimum image of r2 seen from r1 (the shift is decided on values)
__device__ inline void dpwrap(const D3 r1[3], const D3 r2[3], const BoxDev& b, D3 out[3]) {
  for (int i = 0; i < 3; ++i) {
    D3 o = r2[i];
    if (b.periodic[i]) {
      const double d = r2[i].v - r1[i].v;
      if (d > b.L[i] / 2)
        o.v -= b.L[i];
      else if (d < -b.L[i] / 2)
        o.v += b.L[i];
    }
    out[i] = o;
  }
}

__device__ inline D3 dchord(const D3 s[3], const D3 sp[3], const BoxDev& b) {
  D3 w[3];
  dpwrap(s, sp, b, w);
  const D3 d0 = s[0] - w[0], d1 = s[1] - w[1], d2 = s[2] - w[2];
  return dsqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

// Local part of the point velocity (the same arithmetic as geometry_point +
// epilogue_point): p[0..4] = positions of k-2, k-1, k, k+1, k+2, u = the
// Biot-Savart sum at k.
__device__ void local_velocity(const D3 p[5][3], const D3 u[3], const BoxDev& box, const PhysDev& ph,
                               const D3 ext[3], const D3 vn[3], D3 vel[3]) {
  const D3 dk = dchord(p[2], p[3], box), df = dchord(p[3], p[4], box);
  const D3 dr = dchord(p[1], p[2], box), drr = dchord(p[0], p[1], box);
  const D3 x[4] = {dk + df, dk, -dr, -(dr + drr)};
  const int nbi[4] = {4, 3, 1, 0};
  D3 t[3] = {dc(0), dc(0), dc(0)}, nm[3] = {dc(0), dc(0), dc(0)};
  for (int j = 0; j < 4; ++j) {
    D3 o[3];
    int q = 0;
    for (int m = 0; m < 4; ++m)
      if (m != j) o[q++] = x[m];
    const D3 den = x[j] * (x[j] - o[0]) * (x[j] - o[1]) * (x[j] - o[2]);
    const D3 w1 = -(o[0] * o[1] * o[2]) / den;
    const D3 w2 = 2.0 * (o[0] * o[1] + o[0] * o[2] + o[1] * o[2]) / den;
    D3 s0w[3];
    dpwrap(p[nbi[j]], p[2], box, s0w);
    for (int i = 0; i < 3; ++i) {
      const D3 d = p[nbi[j]][i] - s0w[i];
      t[i] = t[i] + w1 * d;
      nm[i] = nm[i] + w2 * d;
    }
  }
  const double pi = 3.14159265358979323846;
  const D3 beta =
      dlog(dsqrt(dk * dr) * (2.0 / sqrt(sqrt(2.718281828459045)) / ph.a0)) * (ph.kappa / 4 / pi);
  D3 lia[3];
  dcross(t, nm, lia);
  for (int i = 0; i < 3; ++i) lia[i] = lia[i] * beta;
  D3 vs[3];
  for (int i = 0; i < 3; ++i) vs[i] = lia[i] + ext[i] + u[i];
  for (int i = 0; i < 3; ++i) vel[i] = vs[i];
  if (ph.hyperfriction) {
    const double spp = sqrt(nm[0].v * nm[0].v + nm[1].v * nm[1].v + nm[2].v * nm[2].v);
    if (spp > ph.max_curv / ph.dl_max) {
      D3 h[3];
      dcross(t, lia, h);
      for (int i = 0; i < 3; ++i) vel[i] = vel[i] - h[i] * ph.hyperalpha;
    }
  }
  if (ph.use_mf) {
    D3 dv[3], a1[3], a2[3];
    for (int i = 0; i < 3; ++i) dv[i] = vn[i] - vs[i];
    dcross(t, dv, a1);
    dcross(t, a1, a2);
    for (int i = 0; i < 3; ++i) vel[i] = vel[i] + a1[i] * ph.alpha - a2[i] * ph.alpha_p;
  }
}

// Local VJP of every point k: gnb[j][k] = (dvel_k/dp_j)^T w_k for the five
// stencil slots j (k-2, k-1, k, k+1, k+2), gu[k] = (dvel_k/du_k)^T w_k.
// One thread per (point, seeding): blockIdx.y is the seeding (0..4 the
// stencil slots, 5 the Biot-Savart sum), so the six seedings of a point run
// side by side rather than in one thread's sequence.
__global__ void __launch_bounds__(128) ka_local(const double* __restrict__ P,
                                                const int* __restrict__ fwd,
                                                const int* __restrict__ rev,
                                                const double* __restrict__ bs, int have_bs,
                                                const double* __restrict__ w, int n, long cap,
                                                BoxDev box, PhysDev ph, double t,
                                                double* __restrict__ gnb, double* __restrict__ gu) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int seed = blockIdx.y;
  if (k >= n) return;
  int ids[5];
  ids[2] = k;
  ids[3] = fwd[k];
  ids[4] = fwd[ids[3]];
  ids[1] = rev[k];
  ids[0] = rev[ids[1]];
  double wk[3], uk[3] = {0, 0, 0};
  load3(w, cap, k, wk);
  if (have_bs) load3(bs, cap, k, uk);
  // The imposed flows depend on the point's own position only: evaluate them
  // and their 3x3 Jacobians seeded along x, y, z of point k, and hand the
  // derivative part only to the seeding of point k (slot 2).
  D3 fs[3], fn[3];
  {
    D3 r[3];
    for (int i = 0; i < 3; ++i) {
      r[i] = dc(P[i * cap + k]);
      r[i].d[i] = 1.0;
    }
    dflow_eval(ph.vs, r, t, fs);
    if (ph.use_mf)
      dflow_eval(ph.vn, r, t, fn);
    else
      fn[0] = fn[1] = fn[2] = dc(0);
  }
  {
    D3 p[5][3], u[3];
    for (int j = 0; j < 5; ++j)
      for (int i = 0; i < 3; ++i) {
        p[j][i] = dc(P[i * cap + ids[j]]);
        if (seed == j) p[j][i].d[i] = 1.0;
      }
    for (int i = 0; i < 3; ++i) {
      u[i] = dc(uk[i]);
      if (seed == 5) u[i].d[i] = 1.0;
    }
    D3 ext[3], vn[3];
    for (int i = 0; i < 3; ++i) {
      ext[i] = seed == 2 ? fs[i] : dc(fs[i].v);
      vn[i] = seed == 2 ? fn[i] : dc(fn[i].v);
    }
    D3 vel[3];
    local_velocity(p, u, box, ph, ext, vn, vel);
    double g[3];
    for (int c = 0; c < 3; ++c) g[c] = vel[0].d[c] * wk[0] + vel[1].d[c] * wk[1] + vel[2].d[c] * wk[2];
    if (seed < 5) {
      for (int c = 0; c < 3; ++c) gnb[(long)(seed * 3 + c) * cap + k] = g[c];
    } else {
      store3(gu, cap, k, g);
    }
  }
}

// gl[p] = the local VJP blocks naming p: p is slot 0 of k+2, slot 1 of k+1,
// slot 2 of itself, slot 3 of k-1 and slot 4 of k-2.
__global__ void __launch_bounds__(256) ka_gather(const double* __restrict__ gnb,
                                                 const int* __restrict__ fwd,
                                                 const int* __restrict__ rev, int n, long cap,
                                                 double* __restrict__ gl) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int f = fwd[p], r = rev[p];
  const int owner[5] = {fwd[f], f, p, r, rev[r]};
  double g[3] = {0, 0, 0};
  for (int j = 0; j < 5; ++j)
    for (int c = 0; c < 3; ++c) g[c] += gnb[(long)(j * 3 + c) * cap + owner[j]];
  store3(gl, cap, p, g);
}

// SB[m] = minimum image of the segment end fwd[m] seen from m
__global__ void __launch_bounds__(256) ka_seg_end(const double* __restrict__ P,
                                                  const int* __restrict__ fwd, int n, long cap,
                                                  BoxDev box, double* __restrict__ SB) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double a[3], f[3], b[3];
  load3(P, cap, m, a);
  load3(P, cap, fwd[m], f);
  pwrap(a, f, box, b);
  store3(SB, cap, m, b);
}

// VJP of one pair: gR = w^T dB/dR, gRp = w^T dB/dR' (B without kappa/4pi).
// Returns false (and nothing) for pairs the reference zeroes.
__device__ __forceinline__ bool pair_vjp(const double a[3], const double b[3], const double rs[3],
                                         const double w[3], double gR[3], double gRp[3]) {
  double R[3], Rp[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    R[i] = a[i] - rs[i];
    Rp[i] = b[i] - rs[i];
  }
  const double l2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
  const double lp2 = Rp[0] * Rp[0] + Rp[1] * Rp[1] + Rp[2] * Rp[2];
  const double d = R[0] * Rp[0] + R[1] * Rp[1] + R[2] * Rp[2];
  // zero rules: a cheap estimate; pairs near a threshold take the exact test.
  // Lengths are compared squared first: a segment end can coincide with an
  // image target (a wrapped segment seen from the image that undoes the wrap).
  if (l2 < 1.000002e-16 || lp2 < 1.000002e-16) {
    if (guard_exact(a, b, rs)) return false;
  }
  const double il = rsq64(l2), ilp = rsq64(lp2);
  const double l = l2 * il, lp = lp2 * ilp;
  if (fabs(d * il * ilp - 1.0) < 1.000001e-8) {
    if (guard_exact(a, b, rs)) return false;
  }
  const double ll = l * lp, q = ll + d, g = ll * q, ig = rcp64(g);
  const double f = (l + lp) * ig;
  const double fl = ig * (1.0 - f * lp * (ll + q)) * il;   // f_l / l
  const double flp = ig * (1.0 - f * l * (ll + q)) * ilp;  // f_l' / l'
  const double fd = -f * ll * ig;
  const double c[3] = {R[1] * Rp[2] - R[2] * Rp[1], R[2] * Rp[0] - R[0] * Rp[2],
                       R[0] * Rp[1] - R[1] * Rp[0]};
  const double wc = w[0] * c[0] + w[1] * c[1] + w[2] * c[2];
  const double pw[3] = {Rp[1] * w[2] - Rp[2] * w[1], Rp[2] * w[0] - Rp[0] * w[2],
                        Rp[0] * w[1] - Rp[1] * w[0]};   // R' x w
  const double wr[3] = {w[1] * R[2] - w[2] * R[1], w[2] * R[0] - w[0] * R[2],
                        w[0] * R[1] - w[1] * R[0]};     // w x R
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    gR[i] = f * pw[i] + wc * (fl * R[i] + fd * Rp[i]);
    gRp[i] = f * wr[i] + wc * (flp * Rp[i] + fd * R[i]);
  }
  return true;
}

// image target and image cotangent: rs = r - shift, mirrored on the image's
// axis (rs[ax] = 2 wall - r[ax]); the image velocity is mirrored on that axis,
// so the cotangent is too.
__device__ __forceinline__ void image_map(const ImgDev& im, const double r[3], const double w[3],
                                          double rs[3], double wi[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    rs[i] = r[i] - im.shift[i];
    wi[i] = w[i];
  }
  if (im.axis >= 0) {
    rs[im.axis] = im.wall2 - r[im.axis];
    wi[im.axis] = -w[im.axis];
  }
}

struct AdjArgs {
  const double* P;    // positions (SoA, stride cap)
  const double* SB;   // minimum-image segment ends
  const int* fwd;
  const double* gu;   // cotangent of each point's Biot-Savart sum (without kappa/4pi)
  const ImgDev* img;
  long cap;
  long ps;            // plane stride of the partial sums (the point count rounded up to 64)
  int n;
  int n_tiles;
  int n_img;
  int n_chunks;
  double* part;       // targets: [chunks][3][ps]; segments: [chunks][6][ps]
};

// Targets per lane, (segment, image) pairs wave-uniform: partial sums of
// sum_{m, images} (dB/dr)^T gu_k over a chunk of segments.
__global__ void __launch_bounds__(256) ka_bs_targets(AdjArgs A) {
  const int lane = threadIdx.x & 63;
  const int T = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (T >= A.n_tiles) return;
  const int chunk = blockIdx.y;
  const int m0 = (int)(((long)A.n * chunk) / A.n_chunks);
  const int m1 = (int)(((long)A.n * (chunk + 1)) / A.n_chunks);
  const int k = T * TILE + lane;
  const bool valid = k < A.n;
  const int kk = valid ? k : T * TILE;
  double r[3], w[3], g[3] = {0, 0, 0};
  load3(A.P, A.cap, kk, r);
  load3(A.gu, A.cap, kk, w);
  for (int m = m0; m < m1; ++m) {
    double a[3], b[3];
    load3(A.P, A.cap, m, a);
    load3(A.SB, A.cap, m, b);
    const int fm = A.fwd[m];
    const bool adjacent = m == kk || fm == kk;
    for (int I = 0; I < A.n_img; ++I) {
      const ImgDev im = A.img[I];
      if (im.primary && adjacent) continue;
      double rs[3], wi[3], gR[3], gRp[3];
      image_map(im, r, w, rs, wi);
      if (!pair_vjp(a, b, rs, wi, gR, gRp)) continue;
      double gr[3] = {-(gR[0] + gRp[0]), -(gR[1] + gRp[1]), -(gR[2] + gRp[2])};
      if (im.axis >= 0) gr[im.axis] = -gr[im.axis];
      g[0] += gr[0];
      g[1] += gr[1];
      g[2] += gr[2];
    }
  }
  if (valid) store3(A.part + (long)chunk * 3 * A.ps, A.ps, k, g);
}

// Segments per lane, (target, image) pairs wave-uniform: partial sums of
// sum_{k, images} (dB/da)^T gu_k and (dB/db)^T gu_k over a chunk of targets.
__global__ void __launch_bounds__(256) ka_bs_sources(AdjArgs A) {
  const int lane = threadIdx.x & 63;
  const int T = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (T >= A.n_tiles) return;
  const int chunk = blockIdx.y;
  const int k0 = (int)(((long)A.n * chunk) / A.n_chunks);
  const int k1 = (int)(((long)A.n * (chunk + 1)) / A.n_chunks);
  const int m = T * TILE + lane;
  const bool valid = m < A.n;
  const int mm = valid ? m : T * TILE;
  double a[3], b[3], ga[3] = {0, 0, 0}, gb[3] = {0, 0, 0};
  load3(A.P, A.cap, mm, a);
  load3(A.SB, A.cap, mm, b);
  const int fm = A.fwd[mm];
  for (int k = k0; k < k1; ++k) {
    double r[3], w[3];
    load3(A.P, A.cap, k, r);
    load3(A.gu, A.cap, k, w);
    const bool adjacent = mm == k || fm == k;
    for (int I = 0; I < A.n_img; ++I) {
      const ImgDev im = A.img[I];
      if (im.primary && adjacent) continue;
      double rs[3], wi[3], gR[3], gRp[3];
      image_map(im, r, w, rs, wi);
      if (!pair_vjp(a, b, rs, wi, gR, gRp)) continue;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        ga[i] += gR[i];
        gb[i] += gRp[i];
      }
    }
  }
  if (valid) {
    double* out = A.part + (long)chunk * 6 * A.ps;
    store3(out, A.ps, m, ga);
    store3(out + 3 * A.ps, A.ps, m, gb);
  }
}

// Both passes in one (tangles of few target tiles): targets per lane with
// the target pass's sums, and the terms of each (segment, image) pair for
// the segment's ends, (dB/da)^T gu and (dB/db)^T gu, summed over the
// wavefront's 64 targets by a butterfly that halves the values each lane
// carries (6 -> 3 -> 2 -> 1 over lane distances 32, 16, 8, then a full sum
// over 4, 2, 1), written per target tile: spart[T][6][ps].  Each pair is
// evaluated once instead of once per pass.
__global__ void __launch_bounds__(256) ka_bs_fused(AdjArgs A, double* __restrict__ spart) {
  const int lane = threadIdx.x & 63;
  const int T = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (T >= A.n_tiles) return;
  const int chunk = blockIdx.y;
  const int m0 = (int)(((long)A.n * chunk) / A.n_chunks);
  const int m1 = (int)(((long)A.n * (chunk + 1)) / A.n_chunks);
  const int k = T * TILE + lane;
  const bool valid = k < A.n;
  const int kk = valid ? k : T * TILE;
  const bool h5 = (lane & 32) != 0, h4 = (lane & 16) != 0, h3 = (lane & 8) != 0;
  const int vi = (h4 ? 2 : 0) + (h3 ? 1 : 0);   // which of this half's values the lane ends with
  double r[3], w[3], g[3] = {0, 0, 0};
  load3(A.P, A.cap, kk, r);
  load3(A.gu, A.cap, kk, w);
  for (int m = m0; m < m1; ++m) {
    double a[3], b[3];
    load3(A.P, A.cap, m, a);
    load3(A.SB, A.cap, m, b);
    const int fm = A.fwd[m];
    const bool adjacent = m == kk || fm == kk;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int I = 0; I < A.n_img; ++I) {
      const ImgDev im = A.img[I];
      if (im.primary && adjacent) continue;
      double rs[3], wi[3], gR[3], gRp[3];
      image_map(im, r, w, rs, wi);
      if (!pair_vjp(a, b, rs, wi, gR, gRp)) continue;
      double gr[3] = {-(gR[0] + gRp[0]), -(gR[1] + gRp[1]), -(gR[2] + gRp[2])};
      if (im.axis >= 0) gr[im.axis] = -gr[im.axis];
      g[0] += gr[0];
      g[1] += gr[1];
      g[2] += gr[2];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        v[i] += gR[i];
        v[3 + i] += gRp[i];
      }
    }
    if (!valid) {
#pragma unroll
      for (int i = 0; i < 6; ++i) v[i] = 0;
    }
    double u[4];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double send = h5 ? v[i] : v[3 + i], keep = h5 ? v[3 + i] : v[i];
      u[i] = keep + __shfl_xor(send, 32, 64);
    }
    u[3] = 0;
    double x2[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const double send = h4 ? u[j] : u[2 + j], keep = h4 ? u[2 + j] : u[j];
      x2[j] = keep + __shfl_xor(send, 16, 64);
    }
    double x;
    {
      const double send = h3 ? x2[0] : x2[1], keep = h3 ? x2[1] : x2[0];
      x = keep + __shfl_xor(send, 8, 64);
    }
    x += __shfl_xor(x, 4, 64);
    x += __shfl_xor(x, 2, 64);
    x += __shfl_xor(x, 1, 64);
    if ((lane & 7) == 0 && vi < 3) spart[((long)T * 6 + (h5 ? 3 : 0) + vi) * A.ps + m] = x;
  }
  if (valid) store3(A.part + (long)chunk * 3 * A.ps, A.ps, k, g);
}

// red[q][p] (9 planes, q = blockIdx.y): the sums of the target pass's t_terms
// chunks (q = 0..2) and of the segment pass's s_terms chunks or target tiles
// (q = 3..5: start, 6..8: end), in order.  red reuses the local pass's gnb
// planes, free once ka_gather ran.
__global__ void __launch_bounds__(256) ka_bs_reduce(const double* __restrict__ tpart, int t_terms,
                                                    const double* __restrict__ spart, int s_terms,
                                                    int n, long ps, long cap, double* __restrict__ red) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (p >= n) return;
  const double* src = q < 3 ? tpart + (long)q * ps + p : spart + (long)(q - 3) * ps + p;
  const long stride = (q < 3 ? 3 : 6) * ps;
  const int terms = q < 3 ? t_terms : s_terms;
  double a = 0;
#pragma unroll 8
  for (int c = 0; c < terms; ++c) a += src[c * stride];
  red[(long)q * cap + p] = a;
}

// G[p] = gl[p] + kappa/4pi (targets[p] + (starts[p] + ends[rev p]))
// (p starts segment p and ends segment rev p).
__global__ void __launch_bounds__(256) ka_bs_finish(const double* __restrict__ gl,
                                                    const double* __restrict__ red,
                                                    const int* __restrict__ rev, int n, long cap,
                                                    double kappa4pi, double* __restrict__ G) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int r = rev[p];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    G[i * cap + p] = gl[i * cap + p] +
                     kappa4pi * (red[i * cap + p] + (red[(3 + i) * cap + p] + red[(6 + i) * cap + r]));
}

// y = x + a z (y may be x)
__global__ void __launch_bounds__(256) ka_axpy(const double* x, double a,
                                               const double* __restrict__ z, double* y, int n,
                                               long cap) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) y[c * cap + s] = x[c * cap + s] + a * z[c * cap + s];
}

// y = a x
__global__ void __launch_bounds__(256) ka_scale(const double* __restrict__ x, double a,
                                                double* __restrict__ y, int n, long cap) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) y[c * cap + s] = a * x[c * cap + s];
}

// flag bit 4: a non-finite gradient entry
__global__ void __launch_bounds__(256) ka_finite(const double* __restrict__ G, int n, long cap,
                                                 int* __restrict__ flag) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  if (!(isfinite(G[s]) && isfinite(G[cap + s]) && isfinite(G[2 * cap + s]))) atomicOr(flag, 4);
}

}  // namespace ov
