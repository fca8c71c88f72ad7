/*
 * ov_oracle.c — FP64 CPU restatement of OpenVort's velocity / RK4 path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the GPU library.  Only tests/,
 * the smoke check and bench.py's CPU-baseline leg load it; the product path
 * never does.
 *
 * Written from the mathematics of the reference (emil-varga/openvort); each
 * function names the reference lines whose behaviour it restates.  Where the
 * reference's evaluation order fixes rounding that matters for parity (guard
 * thresholds, accumulation order), the same order is kept.  Compiled with
 * -ffp-contract=off so that no multiply-add is fused (the reference's gcc
 * build targets plain x86-64, where none is).
 *
 * Parity pins (tests/test_oracle.py): the compiled reference (oracle/_ref)
 * on the same inputs for every function here; closed forms for the segment
 * law, the FD stencil, the ring speed and the RK amplification polynomial;
 * periodic images against an explicitly replicated open tangle; octree at
 * theta=0 against the direct sum.
 */
#include "ov_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static int g_threads = 0;
void ovo_set_threads(int n) { g_threads = n; }
static int nthreads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---- small vector helpers ------------------------------------------------ */

/* |v| summed x, y, z in that order from 0 (vec3_maths.c:338-345). */
static double norm3(const double v[3]) {
  double s = 0;
  s += v[0] * v[0];
  s += v[1] * v[1];
  s += v[2] * v[2];
  return sqrt(s);
}
static double dot3(const double u[3], const double v[3]) {
  double s = 0;
  s += u[0] * v[0];
  s += u[1] * v[1];
  s += u[2] * v[2];
  return s;
}
static void cross3(const double u[3], const double v[3], double o[3]) {
  double a0 = u[1] * v[2] - u[2] * v[1];
  double a1 = u[2] * v[0] - u[0] * v[2];
  double a2 = u[0] * v[1] - u[1] * v[0];
  o[0] = a0; o[1] = a1; o[2] = a2;
}

static void box_len(const ov_domain_box *b, double L[3]) {
  for (int i = 0; i < 3; ++i) L[i] = b->top_right_front.p[i] - b->bottom_left_back.p[i];
}

/* Minimum-image copy of r2 relative to r1: only r2 moves, by one box length
 * per periodic axis when |r2 - r1| exceeds half of it (vec3_maths.c:105-138;
 * the periodic flag of an axis is its low wall). */
static void pwrap(const double r1[3], const double r2[3], const ov_domain_box *b,
                  double out[3]) {
  double L[3];
  box_len(b, L);
  for (int i = 0; i < 3; ++i) {
    out[i] = r2[i];
    if (b->wall[2 * i] == OV_WALL_PERIODIC) {
      double d = r2[i] - r1[i];
      if (d > L[i] / 2)
        out[i] -= L[i];
      else if (d < -L[i] / 2)
        out[i] += L[i];
    }
  }
}

/* Image target of r for one image entry (tangle.c:39-79): shift by whole box
 * lengths, then, for a mirror image, reflect the unshifted coordinate about
 * the wall. */
static void image_target(const ov_image_tangle *im, const ov_domain_box *b,
                         const double r[3], double rs[3]) {
  double L[3];
  box_len(b, L);
  for (int i = 0; i < 3; ++i) rs[i] = r[i] - L[i] * im->shift[i];
  if (im->reflect >= 0 && im->reflect < 6) {
    int ax = im->reflect / 2;
    double wall = (im->reflect & 1) ? b->top_right_front.p[ax] : b->bottom_left_back.p[ax];
    rs[ax] = 2 * wall - r[ax];
  }
}

/* Reflect a direction in the mirror of face w: the component along the face
 * normal changes sign (vec3_maths.c:260-273). */
static void mirror_reflect(double v[3], int w) {
  int ax = w / 2;
  double n[3] = {0, 0, 0};
  n[ax] = (w & 1) ? -1.0 : 1.0;
  double nc = dot3(v, n);
  for (int i = 0; i < 3; ++i) v[i] = v[i] + n[i] * (-2 * nc);
}

/* ---- Biot-Savart of a straight segment ---------------------------------- */

/* v = kappa/4pi (l + l') / (l l' (l l' + R.R')) R x R', R = a - r, R' = b - r;
 * exactly zero when either endpoint is within 1e-8 of r or when R, R' are
 * parallel to 1e-8 in their normalised dot product (tangle.c:486-514). */
void ovo_segment_field(const double a[3], const double b[3], const double r[3],
                       double kappa, double out[3]) {
  double R[3], Rp[3];
  for (int i = 0; i < 3; ++i) { R[i] = a[i] - r[i]; Rp[i] = b[i] - r[i]; }
  double l = norm3(R), lp = norm3(Rp);
  double dd = dot3(R, Rp);
  double denom = l * lp * (l * lp + dd);
  double f = kappa / 4 / M_PI;
  out[0] = out[1] = out[2] = 0;
  if (l < 1e-8 || lp < 1e-8) return;
  double nd = dd / l / lp;
  if (fabs(nd - 1) < 1e-8) return;
  double c[3];
  cross3(R, Rp, c);
  double m = f * (l + lp) / denom;
  for (int i = 0; i < 3; ++i) out[i] = c[i] * m;
}

/* Segment of point m: m -> minimum image of its forward neighbour
 * (tangle.c:519-528). */
static int segment_of(const ov_tangle_state *t, int m, double a[3], double b[3]) {
  int f = t->connections[m].forward;
  if (f == -1) return 0;
  for (int i = 0; i < 3; ++i) a[i] = t->vnodes[m].p[i];
  pwrap(a, t->vnodes[f].p, &t->box, b);
  return 1;
}

/* ---- geometry ------------------------------------------------------------ */

/* Chord length from s to the minimum image of s' (tangle.c:536-540, 710-719). */
static double chord(const double s[3], const double sp[3], const ov_domain_box *b) {
  double w[3], d[3];
  pwrap(s, sp, b, w);
  for (int i = 0; i < 3; ++i) d[i] = s[i] - w[i];
  return norm3(d);
}

static int check_point(const ov_tangle_state *t, int k) {
  int st = t->status[k].status;
  if (st == OV_PINNED || st == OV_PINNED_SLIP) return OV_E_UNSUPPORTED;
  if (st != OV_FREE) return OV_E_INVALID;
  int f = t->connections[k].forward, r = t->connections[k].reverse;
  if (f < 0 || r < 0 || f >= t->N || r >= t->N) return OV_E_INVALID;
  if (t->status[f].status != OV_FREE || t->status[r].status != OV_FREE) {
    if (t->status[f].status == OV_PINNED || t->status[r].status == OV_PINNED ||
        t->status[f].status == OV_PINNED_SLIP || t->status[r].status == OV_PINNED_SLIP)
      return OV_E_UNSUPPORTED;
    return OV_E_INVALID;
  }
  return OV_OK;
}

/* s' and s'' at k from the 5-node Lagrange stencil in cumulative chord
 * length {-(dxi_pp+dxi_p), -dxi_p, 0, dxi_k, dxi_k+dxi_f}.  With nodes x_j
 * (j over the four nonzero nodes, the other three being a,b,c):
 *   w'_j  = -x_a x_b x_c / (x_j (x_j-x_a)(x_j-x_b)(x_j-x_c))
 *   w''_j = 2 (x_a x_b + x_a x_c + x_b x_c) / (x_j (x_j-x_a)(x_j-x_b)(x_j-x_c))
 *   s' = sum_j w'_j (s_j - s_0),  s'' = sum_j w''_j (s_j - s_0)
 * where s_j - s_0 is the minimum-image difference (tangle.c:245-273,
 * 432-479; 4point.wxmx).  Only the free-point branch exists here: pinned
 * neighbours are OV_E_UNSUPPORTED (tangle.c:275-430). */
static void fd_weights(const double x[4], double w1[4], double w2[4]) {
  for (int j = 0; j < 4; ++j) {
    double o[3];
    int q = 0;
    for (int m = 0; m < 4; ++m)
      if (m != j) o[q++] = x[m];
    double den = x[j] * (x[j] - o[0]) * (x[j] - o[1]) * (x[j] - o[2]);
    w1[j] = -(o[0] * o[1] * o[2]) / den;
    w2[j] = 2 * (o[0] * o[1] + o[0] * o[2] + o[1] * o[2]) / den;
  }
}

int ovo_update_tangents_normals(ov_tangle_state *t) {
  const int N = t->N;
  for (int k = 0; k < N; ++k) {
    if (t->status[k].status == OV_EMPTY) continue;
    int rc = check_point(t, k);
    if (rc) return rc;
  }
  for (int k = 0; k < N; ++k) {
    if (t->status[k].status == OV_EMPTY) continue;
    int f = t->connections[k].forward;
    t->dxi[k] = chord(t->vnodes[k].p, t->vnodes[f].p, &t->box);
  }
  for (int k = 0; k < N; ++k) {
    if (t->status[k].status == OV_EMPTY) continue;
    int f = t->connections[k].forward, ff = t->connections[f].forward;
    int r = t->connections[k].reverse, rr = t->connections[r].reverse;
    const double *s0 = t->vnodes[k].p;
    const int nb[4] = {ff, f, r, rr};
    double x[4];
    x[0] = t->dxi[k] + t->dxi[f];
    x[1] = t->dxi[k];
    x[2] = -t->dxi[r];
    x[3] = -(t->dxi[r] + t->dxi[rr]);
    double w1[4], w2[4];
    fd_weights(x, w1, w2);
    double tg[3] = {0, 0, 0}, nm[3] = {0, 0, 0};
    for (int j = 0; j < 4; ++j) {
      double s0w[3], d[3];
      /* s_j - s_0 with s_0 taken as the minimum image seen from s_j */
      pwrap(t->vnodes[nb[j]].p, s0, &t->box, s0w);
      for (int i = 0; i < 3; ++i) d[i] = t->vnodes[nb[j]].p[i] - s0w[i];
      for (int i = 0; i < 3; ++i) {
        tg[i] += w1[j] * d[i];
        nm[i] += w2[j] * d[i];
      }
    }
    for (int i = 0; i < 3; ++i) {
      t->tangents[k].p[i] = tg[i];
      t->normals[k].p[i] = nm[i];
    }
  }
  return OV_OK;
}

/* ---- external flows ------------------------------------------------------ */

/* |u| as vec3_d (vec3_maths.c:338-345): squares summed from zero. */
static double vnorm(double x, double y, double z) {
  double o = 0;
  o += x * x;
  o += y * y;
  o += z * z;
  return sqrt(o);
}

/* Radial source/sink field, strength/(4 pi r^2) or, about the z axis,
 * strength/(2 pi r), damped by exp(-(cutoff/r)^2) and cut to zero below
 * r = 1e-8 or beyond an exponent of 100 (external_velocity.c:197-223,
 * 243-270).  cyl selects the axial form. */
static void radial_flow(const ov_flow *fl, const double r[3], int cyl, double out[3]) {
  const double strength = fl->a[0], cutoff = fl->a[1];
  const double d = cyl ? sqrt(r[0] * r[0] + r[1] * r[1]) : vnorm(r[0], r[1], r[2]);
  const double attn = (cutoff / d) * (cutoff / d);
  if (d < 1e-8 || attn > 100) {
    out[0] = out[1] = out[2] = 0;
    return;
  }
  const double factor = exp(-attn);
  const double u[3] = {r[0], r[1], cyl ? 0.0 : r[2]};
  const double inv = 1 / vnorm(u[0], u[1], u[2]); /* vec3_normalize */
  const double m = cyl ? strength / (2 * M_PI * d) * factor : strength / (4 * M_PI * d * d) * factor;
  for (int i = 0; i < 3; ++i) out[i] = (u[i] * inv) * m;
}

/* Every flow of the reference table (external_velocity.c:85-174) at r, t;
 * the parameter slots are those of ov_flow (include/openvort_gpu.h). */
static int flow_eval(const ov_flow *fl, const double r[3], double t, double out[3]) {
  const double x = r[0], y = r[1], z = r[2];
  const double *a = fl->a;
  switch (fl->type) {
  case OV_FLOW_NONE: /* :176-182 */
    out[0] = out[1] = out[2] = 0;
    return 0;
  case OV_FLOW_SIMPLE: /* :184-188 */
    for (int i = 0; i < 3; ++i) out[i] = fl->v[i];
    return 0;
  case OV_FLOW_SPHERICAL:
    radial_flow(fl, r, 0, out);
    return 0;
  case OV_FLOW_SPHERICAL_AND_SIMPLE: { /* :225-241, the sum it computes */
    double sph[3];
    radial_flow(fl, r, 0, sph);
    for (int i = 0; i < 3; ++i) out[i] = sph[i] + fl->v[i];
    return 0;
  }
  case OV_FLOW_CYLINDRICAL:
    radial_flow(fl, r, 1, out);
    return 0;
  case OV_FLOW_OSCILLATING: { /* :272-294: pol strength cos(2 pi k.r) sin(2 pi f t) */
    double kr = 0;
    kr += fl->w[0] * x;
    kr += fl->w[1] * y;
    kr += fl->w[2] * z;
    const double m = a[0] * cos(2 * M_PI * kr) * sin(2 * M_PI * a[1] * t);
    for (int i = 0; i < 3; ++i) out[i] = fl->q[i] * m;
    return 0;
  }
  case OV_FLOW_OSCILLATING_SLIT: { /* :296-322, strength unused */
    const double rr = sqrt(x * x + z * z);
    const double dir[3] = {x / rr, 0.0, z / rr};
    const double m = cos(2 * M_PI * a[1] * t) * exp(-a[2] * a[2] / rr / rr);
    for (int i = 0; i < 3; ++i) out[i] = dir[i] * m;
    return 0;
  }
  case OV_FLOW_COSCOS: /* :324-344 */
    out[0] = out[1] = 0;
    out[2] = a[1] * (1 + cos(a[0] * x) * cos(a[0] * y) * exp(-z / a[2]));
    return 0;
  case OV_FLOW_COSCOS_DIVFREE: /* :347-370 */
    out[0] = a[1] / 2 / a[2] / a[0] * sin(a[0] * x) * cos(a[0] * y) * exp(-z / a[2]);
    out[1] = a[1] / 2 / a[2] / a[0] * cos(a[0] * x) * sin(a[0] * y) * exp(-z / a[2]);
    out[2] = a[1] * (1 + cos(a[0] * x) * cos(a[0] * y) * exp(-z / a[2]));
    return 0;
  case OV_FLOW_COS_DIVFREE: /* :372-394 */
    out[0] = a[1] / a[2] / a[0] * sin(a[0] * x) * exp(-z / a[2]);
    out[1] = 0;
    out[2] = a[1] * (1 + cos(a[0] * x) * exp(-z / a[2]));
    return 0;
  case OV_FLOW_SIMPLE_SHEAR: /* :396-413 */
    out[0] = out[1] = 0;
    out[2] = a[1] + a[0] * x;
    return 0;
  case OV_FLOW_ROTOSC: /* :415-433 */
  case OV_FLOW_ROTOSC_BOUNDARY: /* the table runs get_v_rotosc for it, :156-157 */
  case OV_FLOW_DG: { /* :438-457 */
    /* z x r written out as vec3_cross (vec3_maths.c:304-312) with u = (0,0,1) */
    const double v[3] = {0.0 * z - 1.0 * y, 1.0 * x - 0.0 * z, 0.0 * y - 0.0 * x};
    const double m = fl->type == OV_FLOW_DG ? a[0] : a[1] * cos(2 * M_PI * a[0] * t);
    for (int i = 0; i < 3; ++i) out[i] = v[i] * m;
    if (fl->type == OV_FLOW_DG) out[2] += a[1];
    return 0;
  }
  case OV_FLOW_CONSTANT_ACCELERATION: /* :483-495 */
    out[0] = a[0] * t;
    out[1] = out[2] = 0;
    return 0;
  case OV_FLOW_PARABOLIC: /* :497-512 */
    out[0] = a[1] * z * (a[0] - z);
    out[1] = out[2] = 0;
    return 0;
  default:
    return OV_E_INVALID;
  }
}

int ovo_flow(const ov_flow *fl, const double *r, double t, double *out) {
  return flow_eval(fl, r, t, out);
}

/* ---- octree (reference geometry and multipole) --------------------------- */

typedef struct otnode {
  int n;          /* points in the cell */
  int *ids;       /* their tangle indices, parent order */
  double lo[3], hi[3];
  double com[3], gam[3], M[3][3];
  struct otnode *child[8];
} otnode;

static long ot_cells, ot_leaf_pairs, ot_multipoles;

/* Child-box faces (octree.c:158-235).  Per child and axis, the low face is
 * one of {P.lo, P.lo + L/2, c, c - L/2} and the high face one of
 * {c, c + L/2, P.hi, P.hi - L/2}, with c = (P.lo + P.hi) * 0.5 and
 * L = P.hi - P.lo; the codes below give, for children 0..7 in the
 * reference's order (BLF, BLB, BRF, BRB, TLF, TLB, TRF, TRB), which one. */
enum { LO_P = 0, LO_PH = 1, LO_C = 2, LO_CM = 3 };
enum { HI_C = 0, HI_CH = 1, HI_P = 2, HI_PM = 3 };
static const unsigned char child_lo[8][3] = {
    {LO_PH, LO_P, LO_P}, {LO_P, LO_P, LO_P},  {LO_C, LO_C, LO_CM}, {LO_P, LO_PH, LO_P},
    {LO_C, LO_CM, LO_C}, {LO_P, LO_P, LO_PH}, {LO_C, LO_C, LO_C},  {LO_CM, LO_C, LO_C}};
static const unsigned char child_hi[8][3] = {
    {HI_CH, HI_C, HI_C}, {HI_C, HI_C, HI_C},  {HI_P, HI_P, HI_PM}, {HI_C, HI_CH, HI_C},
    {HI_P, HI_PM, HI_P}, {HI_C, HI_C, HI_CH}, {HI_P, HI_P, HI_P},  {HI_PM, HI_P, HI_P}};

static void child_box(const otnode *P, int ci, int quad, double lo[3], double hi[3]) {
  double c[3], L[3];
  for (int i = 0; i < 3; ++i) {
    c[i] = P->lo[i] + P->hi[i];
    c[i] = c[i] * 0.5;
    L[i] = P->hi[i] - P->lo[i];
  }
  for (int i = 0; i < 3; ++i) {
    switch (child_lo[ci][i]) {
      case LO_P: lo[i] = P->lo[i]; break;
      case LO_PH: lo[i] = P->lo[i] + L[i] / 2; break;
      case LO_C: lo[i] = c[i]; break;
      default: lo[i] = c[i] - L[i] / 2; break;
    }
    switch (child_hi[ci][i]) {
      case HI_C: hi[i] = c[i]; break;
      case HI_CH: hi[i] = c[i] + L[i] / 2; break;
      case HI_P: hi[i] = P->hi[i]; break;
      default: hi[i] = P->hi[i] - L[i] / 2; break;
    }
  }
  if (quad) hi[2] = P->hi[2];
}

static double box_min_len(const double lo[3], const double hi[3]) {
  double Lx = hi[0] - lo[0], Ly = hi[1] - lo[1], Lz = hi[2] - lo[2];
  if (Lx < Ly) return Lx < Lz ? Lx : Lz;
  return Ly < Lz ? Ly : Lz;
}
static double box_max_len(const double lo[3], const double hi[3]) {
  double Lx = hi[0] - lo[0], Ly = hi[1] - lo[1], Lz = hi[2] - lo[2];
  if (Lx > Ly) return Lx > Lz ? Lx : Lz;
  return Ly > Lz ? Ly : Lz;
}

/* Cell moments (octree.c:237-266): centre = (sum r) * (1/n); circulation
 * Gamma = sum ds; tensor M_ij = sum ds_i (r - centre)_j, with ds the
 * minimum-image segment of each point that has a forward neighbour and r its
 * start (the reference's form).  The dipole mode takes r at the segment's
 * midpoint instead: the first-order term of a straight segment's field is
 * ds x [-(mid - centre)/|R|^3 + 3 ((mid - centre).R) R/|R|^5], so start
 * points would leave a first-order error of 3/2 sum (ds.R) ds x R/|R|^5. */
static void ot_moments(otnode *c, const ov_tangle_state *t, int midpoint) {
  double s[3] = {0, 0, 0};
  for (int k = 0; k < c->n; ++k)
    for (int i = 0; i < 3; ++i) s[i] += t->vnodes[c->ids[k]].p[i];
  for (int i = 0; i < 3; ++i) c->com[i] = s[i] * (1.0 / c->n);
  memset(c->gam, 0, sizeof c->gam);
  memset(c->M, 0, sizeof c->M);
  for (int k = 0; k < c->n; ++k) {
    double a[3], b[3];
    if (!segment_of(t, c->ids[k], a, b)) continue;
    double ds[3], sl[3];
    for (int i = 0; i < 3; ++i) { ds[i] = b[i] - a[i]; sl[i] = a[i] - c->com[i]; }
    if (midpoint)
      for (int i = 0; i < 3; ++i) sl[i] = sl[i] + 0.5 * ds[i];
    for (int i = 0; i < 3; ++i) c->gam[i] += ds[i];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) c->M[i][j] += ds[i] * sl[j];
  }
}

/* Recursive split (octree.c:99-136): a cell with more than one point whose
 * shortest edge is at least BH_grain * dl_max gets 8 (quadtree: 4) child
 * boxes; each point goes to the first child, in child order, that contains
 * it strictly; points on shared faces or outside every child are dropped. */
static void ot_split(otnode *c, const ov_tangle_state *t, const ov_params *p) {
  if (c->n > 0) ot_cells++;
  ot_moments(c, t, p->multipole == OV_MP_DIPOLE);
  if (c->n <= 1) return;
  if (box_min_len(c->lo, c->hi) < p->BH_grain * p->dl_max) return;
  const int nch = p->BH_quadtree ? 4 : 8;
  for (int ci = 0; ci < nch; ++ci) {
    otnode *ch = (otnode *)calloc(1, sizeof(otnode));
    child_box(c, ci, p->BH_quadtree, ch->lo, ch->hi);
    ch->ids = (int *)malloc(sizeof(int) * (size_t)c->n);
    c->child[ci] = ch;
  }
  for (int k = 0; k < c->n; ++k) {
    const double *r = t->vnodes[c->ids[k]].p;
    for (int ci = 0; ci < nch; ++ci) {
      otnode *ch = c->child[ci];
      if (r[0] > ch->lo[0] && r[0] < ch->hi[0] && r[1] > ch->lo[1] && r[1] < ch->hi[1] &&
          r[2] > ch->lo[2] && r[2] < ch->hi[2]) {
        ch->ids[ch->n++] = c->ids[k];
        break;
      }
    }
  }
  for (int ci = 0; ci < nch; ++ci) ot_split(c->child[ci], t, p);
}

static void ot_free(otnode *c) {
  if (!c) return;
  for (int i = 0; i < 8; ++i) ot_free(c->child[i]);
  free(c->ids);
  free(c);
}

static otnode *ot_build(const ov_tangle_state *t, const ov_params *p) {
  otnode *root = (otnode *)calloc(1, sizeof(otnode));
  root->ids = (int *)malloc(sizeof(int) * (size_t)(t->N > 0 ? t->N : 1));
  for (int k = 0; k < t->N; ++k)
    if (t->status[k].status != OV_EMPTY) root->ids[root->n++] = k;
  for (int i = 0; i < 3; ++i) {
    root->lo[i] = t->box.bottom_left_back.p[i];
    root->hi[i] = t->box.top_right_front.p[i];
  }
  ot_cells = 0;
  ot_split(root, t, p);
  return root;
}

/* Sum over the listed segments at r, skipping segments that start or end at
 * `skip` (tangle.c:553-576). */
static long sum_segments(const ov_tangle_state *t, const int *ids, int n, const double r[3],
                         int skip, double kappa, double out[3]) {
  long cnt = 0;
  out[0] = out[1] = out[2] = 0;
  for (int q = 0; q < n; ++q) {
    int m = ids ? ids[q] : q;
    int f = t->connections[m].forward;
    if (f == -1 || m == skip || f == skip) continue;
    double a[3], b[3], v[3];
    segment_of(t, m, a, b);
    ovo_segment_field(a, b, r, kappa, v);
    for (int i = 0; i < 3; ++i) out[i] += v[i];
    ++cnt;
  }
  return cnt;
}

/* Far-field value of a cell at r.  Reference form (octree.c:296-330):
 *   kappa/4pi [ Gamma x R/|R|^3 + (M_21, 0, 0)/|R|^3 + (u x R)/|R|^5 ],
 *   u_i = M_i2 R_2, R = r - centre (not minimum-imaged);
 * dipole form: kappa/4pi [ Gamma x R/|R|^3 - eps:M/|R|^3 + 3 (M R) x R/|R|^5 ]. */
static void ot_multipole(const otnode *c, const double r[3], const ov_params *p, double out[3]) {
  double R[3];
  for (int i = 0; i < 3; ++i) R[i] = r[i] - c->com[i];
  double Rm = norm3(R);
  double i3 = 1 / Rm / Rm / Rm;
  double i5 = 1 / Rm / Rm / Rm / Rm / Rm;
  double v1[3], res[3] = {0, 0, 0};
  cross3(c->gam, R, v1);
  for (int i = 0; i < 3; ++i) res[i] += v1[i] * i3;
  if (p->multipole == OV_MP_DIPOLE) {
    double e[3] = {c->M[1][2] - c->M[2][1], c->M[2][0] - c->M[0][2], c->M[0][1] - c->M[1][0]};
    for (int i = 0; i < 3; ++i) res[i] += -e[i] * i3;
    double u[3];
    for (int i = 0; i < 3; ++i) u[i] = c->M[i][0] * R[0] + c->M[i][1] * R[1] + c->M[i][2] * R[2];
    cross3(u, R, v1);
    for (int i = 0; i < 3; ++i) res[i] += 3 * v1[i] * i5;
  } else {
    double e[3] = {c->M[2][1], 0, 0};
    for (int i = 0; i < 3; ++i) res[i] += e[i] * i3;
    double u[3];
    for (int i = 0; i < 3; ++i) u[i] = c->M[i][2] * R[2];
    cross3(u, R, v1);
    for (int i = 0; i < 3; ++i) res[i] += v1[i] * i5;
  }
  double f = p->kappa / 4 / M_PI;
  for (int i = 0; i < 3; ++i) out[i] = res[i] * f;
}

/* Traversal (octree.c:275-343): empty cell -> 0; one point or an edge below
 * BH_grain -> direct sum over the cell's points with `skip`; accepted when
 * d > 0 and Lmax/d < theta, d the minimum-image distance to the centre;
 * otherwise the children's sum in child order. */
static void ot_vs(const otnode *c, const ov_tangle_state *t, const double r[3], const ov_params *p,
                  int skip, double out[3], long *pairs, long *mps) {
  out[0] = out[1] = out[2] = 0;
  if (!c || c->n == 0) return;
  double Lm = box_max_len(c->lo, c->hi), Lmin = box_min_len(c->lo, c->hi);
  double d = chord(r, c->com, &t->box);
  if (c->n == 1 || Lmin < p->BH_grain) {
    *pairs += sum_segments(t, c->ids, c->n, r, skip, p->kappa, out);
    return;
  }
  if (d > 0 && Lm / d < p->BH_resolution) {
    ot_multipole(c, r, p, out);
    *mps += 1;
    return;
  }
  const int nch = p->BH_quadtree ? 4 : 8;
  for (int ci = 0; ci < nch; ++ci) {
    double part[3];
    ot_vs(c->child[ci], t, r, p, skip, part, pairs, mps);
    for (int i = 0; i < 3; ++i) out[i] += part[i];
  }
}

void ovo_octree_stats(long *n_cells, long *n_leaf_pairs, long *n_multipoles) {
  if (n_cells) *n_cells = ot_cells;
  if (n_leaf_pairs) *n_leaf_pairs = ot_leaf_pairs;
  if (n_multipoles) *n_multipoles = ot_multipoles;
}

/* Far-field value at r of one cell holding the listed points: the moments
 * and the multipole of the octree path above (the cell's box plays no part). */
int ovo_cell_far_field(const ov_tangle_state *t, const int *ids, int n, const double *r,
                       const ov_params *p, double *out) {
  if (n <= 0 || !ids) return OV_E_INVALID;
  otnode c;
  memset(&c, 0, sizeof c);
  c.ids = (int *)ids;
  c.n = n;
  ot_moments(&c, t, p->multipole == OV_MP_DIPOLE);
  ot_multipole(&c, r, p, out);
  return OV_OK;
}

/* ---- per-point velocity -------------------------------------------------- */

/* Point assembly (tangle.c:583-684):
 *   vs   = beta s' x s'' + vs_ext + sum_{m not adjacent} B(m, r) + sum_img P_j sum_m B(m, r_j)
 *   vels = vs - [|s''| > c/dl_max] alpha_h s' x (beta s' x s'')
 *             + alpha s' x (vn - vs) - alpha' s' x (s' x (vn - vs)),
 * beta = kappa/4pi ln(2 sqrt(l+ l-) / (e^(1/4) a0)) with l+- the minimum-image
 * chords to the two neighbours (tangle.c:530-551). */
static int point_velocity(const ov_tangle_state *t, int k, double time, const ov_params *p,
                          const otnode *tree, double vs[3], double vel[3], long *pairs, long *mps) {
  const double *r = t->vnodes[k].p;
  const double *tg = t->tangents[k].p, *nm = t->normals[k].p;
  int f = t->connections[k].forward, rv = t->connections[k].reverse;
  double lnext = chord(r, t->vnodes[f].p, &t->box);
  double lprev = chord(t->vnodes[rv].p, r, &t->box);
  double beta = p->kappa * log(2 * sqrt(lnext * lprev) / sqrt(sqrt(M_E)) / p->vortex_width) / 4 / M_PI;
  double lia[3];
  cross3(tg, nm, lia);
  for (int i = 0; i < 3; ++i) lia[i] = lia[i] * beta;
  for (int i = 0; i < 3; ++i) vs[i] = lia[i];
  double e[3];
  int rc = flow_eval(&p->vs, r, time, e);
  if (rc) return rc;
  for (int i = 0; i < 3; ++i) vs[i] = vs[i] + e[i];

  if (!p->LIA_only) {
    if (tree) {
      double v[3];
      ot_vs(tree, t, r, p, k, v, pairs, mps);
      for (int i = 0; i < 3; ++i) vs[i] = vs[i] + v[i];
    } else {
      for (int m = 0; m < t->N; ++m) {
        int fm = t->connections[m].forward;
        if (fm == -1 || m == k || fm == k) continue;
        double a[3], b[3], v[3];
        segment_of(t, m, a, b);
        ovo_segment_field(a, b, r, p->kappa, v);
        for (int i = 0; i < 3; ++i) vs[i] = vs[i] + v[i];
      }
    }
    double tot[3] = {0, 0, 0};
    for (int j = 0; j < t->bimg.n; ++j) {
      double rs[3], v[3];
      image_target(&t->bimg.images[j], &t->box, r, rs);
      if (tree)
        ot_vs(tree, t, rs, p, -1, v, pairs, mps);
      else
        sum_segments(t, NULL, t->N, rs, -1, p->kappa, v);
      if (t->bimg.images[j].reflect > -1) mirror_reflect(v, t->bimg.images[j].reflect);
      for (int i = 0; i < 3; ++i) tot[i] = tot[i] + v[i];
    }
    for (int i = 0; i < 3; ++i) vs[i] = vs[i] + tot[i];
  }

  for (int i = 0; i < 3; ++i) vel[i] = vs[i];
  if (p->hyperfriction) {
    double spp = norm3(nm);
    if (spp > p->max_curvature_scale / p->dl_max) {
      double h[3];
      cross3(tg, lia, h);
      for (int i = 0; i < 3; ++i) vel[i] = vel[i] + h[i] * -p->hyperalpha;
    }
  }
  if (p->use_mutual_friction) {
    double vn[3], dv[3], a1[3], a2[3];
    rc = flow_eval(&p->vn, r, time, vn);
    if (rc) return rc;
    for (int i = 0; i < 3; ++i) dv[i] = vn[i] - vs[i];
    cross3(tg, dv, a1);
    for (int i = 0; i < 3; ++i) vel[i] = vel[i] + a1[i] * p->alpha;
    cross3(tg, a1, a2);
    for (int i = 0; i < 3; ++i) vel[i] = vel[i] + a2[i] * -p->alpha_p;
  }
  return OV_OK;
}

static int validate(const ov_tangle_state *t, const ov_params *p) {
  if (!t || !p || t->N < 0) return OV_E_INVALID;
  if (p->vs.type < 0 || p->vs.type >= OV_FLOW_COUNT || p->vn.type < 0 || p->vn.type >= OV_FLOW_COUNT)
    return OV_E_INVALID;
  for (int k = 0; k < t->N; ++k) {
    if (t->status[k].status == OV_EMPTY) {
      if (t->connections[k].forward != -1) return OV_E_INVALID;
      continue;
    }
    int rc = check_point(t, k);
    if (rc) return rc;
  }
  return OV_OK;
}

int ovo_velocities_sample(const ov_tangle_state *t, double time, const ov_params *p,
                          const int *idx, int n, double *vs_out, double *vels_out) {
  int rc = validate(t, p);
  if (rc) return rc;
  otnode *tree = p->use_BH ? ot_build(t, p) : NULL;
  int err = 0;
  long pairs = 0, mps = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads()) reduction(| : err) \
    reduction(+ : pairs, mps)
  for (int q = 0; q < n; ++q) {
    int k = idx[q];
    double vs[3] = {0, 0, 0}, vel[3] = {0, 0, 0};
    if (k >= 0 && k < t->N && t->status[k].status != OV_EMPTY)
      err |= -point_velocity(t, k, time, p, tree, vs, vel, &pairs, &mps);
    for (int i = 0; i < 3; ++i) {
      vs_out[3 * q + i] = vs[i];
      vels_out[3 * q + i] = vel[i];
    }
  }
  ot_leaf_pairs = pairs;
  ot_multipoles = mps;
  ot_free(tree);
  return err ? OV_E_UNSUPPORTED : OV_OK;
}

int ovo_update_velocities(ov_tangle_state *t, double time, const ov_params *p) {
  int rc = validate(t, p);
  if (rc) return rc;
  otnode *tree = p->use_BH ? ot_build(t, p) : NULL;
  int err = 0;
  long pairs = 0, mps = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads()) reduction(| : err) \
    reduction(+ : pairs, mps)
  for (int k = 0; k < t->N; ++k) {
    if (t->status[k].status == OV_EMPTY) continue;
    err |= -point_velocity(t, k, time, p, tree, t->vs[k].p, t->vels[k].p, &pairs, &mps);
  }
  ot_leaf_pairs = pairs;
  ot_multipoles = mps;
  ot_free(tree);
  return err ? OV_E_UNSUPPORTED : OV_OK;
}

/* ---- time stepping ------------------------------------------------------- */

/* rk4_step (vortex_dynamics.c:58-121; polynomial in Appendix form
 * 1 + z + 5z^2/12 + z^3/8 + z^4/48 for the reference variant):
 *   k2 = v(x + dt/2 k1, t + dt/2), k3 = v(x + dt/2 k2, t + dt/2),
 *   k4 = v(x + h4 k3, t + dt) with h4 = dt/2 (reference) or dt (classic),
 *   x += dt/6 k1 + dt/3 k2 + dt/3 k3 + dt/6 k4, accumulated in that order.
 * Geometry is recomputed at each stage position; k1 is t->vels. */
int ovo_rk4_step(ov_tangle_state *t, double time, double dt, const ov_params *p) {
  const int N = t->N;
  int rc = validate(t, p);
  if (rc) return rc;
  ov_tangle_state s = *t;
  size_t v3 = sizeof(ov_vec3d) * (size_t)(N > 0 ? N : 1);
  s.vnodes = (ov_vec3d *)malloc(v3);
  s.vels = (ov_vec3d *)malloc(v3);
  s.vs = (ov_vec3d *)malloc(v3);
  s.tangents = (ov_vec3d *)malloc(v3);
  s.normals = (ov_vec3d *)malloc(v3);
  s.dxi = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
  ov_vec3d *k[3];
  for (int q = 0; q < 3; ++q) k[q] = (ov_vec3d *)malloc(v3);
  const double h[3] = {dt / 2, dt / 2, p->rk_variant == OV_RK_CLASSIC ? dt : dt / 2};
  const double ts[3] = {time + dt / 2, time + dt / 2, time + dt};
  for (int st = 0; st < 3 && !rc; ++st) {
    const ov_vec3d *kin = st == 0 ? t->vels : k[st - 1];
    for (int q = 0; q < N; ++q) {
      if (t->status[q].status == OV_EMPTY) continue;
      for (int i = 0; i < 3; ++i) s.vnodes[q].p[i] = t->vnodes[q].p[i] + kin[q].p[i] * h[st];
    }
    rc = ovo_update_tangents_normals(&s);
    if (!rc) rc = ovo_update_velocities(&s, ts[st], p);
    memcpy(k[st], s.vels, v3);
  }
  if (!rc) {
    for (int q = 0; q < N; ++q) {
      if (t->status[q].status == OV_EMPTY) continue;
      for (int i = 0; i < 3; ++i) {
        double mv = t->vels[q].p[i] * (dt / 6);
        mv = mv + k[0][q].p[i] * (dt / 3);
        mv = mv + k[1][q].p[i] * (dt / 3);
        mv = mv + k[2][q].p[i] * (dt / 6);
        t->vnodes[q].p[i] = t->vnodes[q].p[i] + mv;
      }
    }
  }
  free(s.vnodes); free(s.vels); free(s.vs); free(s.tangents); free(s.normals); free(s.dxi);
  for (int q = 0; q < 3; ++q) free(k[q]);
  return rc;
}

/* While a point lies beyond a periodic face, move it by one box length
 * (tangle.c:766-819; the last offending face in x, y, z order is fixed
 * first). */
void ovo_enforce_boundaries(ov_tangle_state *t) {
  double L[3];
  box_len(&t->box, L);
  for (int k = 0; k < t->N; ++k) {
    if (t->status[k].status == OV_EMPTY) continue;
    for (;;) {
      int face = -1;
      for (int ax = 0; ax < 3; ++ax) {
        if (t->vnodes[k].p[ax] < t->box.bottom_left_back.p[ax]) face = 2 * ax;
        if (t->vnodes[k].p[ax] > t->box.top_right_front.p[ax]) face = 2 * ax + 1;
      }
      if (face < 0 || t->box.wall[face] != OV_WALL_PERIODIC) break;
      int ax = face / 2;
      if (face & 1)
        t->vnodes[k].p[ax] -= L[ax];
      else
        t->vnodes[k].p[ax] += L[ax];
    }
  }
}

/* Pass-invariant tests of the reconnection pair loop (vortex_dynamics.c:
 * 221-265), every pair k < l in the loop's order: both points free, the
 * minimum-image distance |v_k - pwrap(v_k, v_l)| <= rec_dist (vec3_dist,
 * vec3_maths.c:347-351), |t_k|, |t_l| >= 0.5 and the tangents' normalised
 * dot (vec3_ndot, vec3_maths.c:295-301) <= cos(rec_angle). */
long ovo_reconnection_candidates(const ov_tangle_state *t, double rec_dist, double rec_angle,
                                 int *pairs, long max_pairs) {
  const double ca = cos(rec_angle);
  long count = 0;
  for (int k = 0; k < t->N; ++k) {
    if (t->status[k].status != OV_FREE) continue;
    const double *a = t->vnodes[k].p, *ta = t->tangents[k].p;
    for (int l = k + 1; l < t->N; ++l) {
      if (t->status[l].status != OV_FREE) continue;
      double w[3], d[3];
      pwrap(a, t->vnodes[l].p, &t->box, w);
      for (int i = 0; i < 3; ++i) d[i] = a[i] - w[i];
      if (norm3(d) > rec_dist) continue;
      const double *tb = t->tangents[l].p;
      const double na = norm3(ta), nb = norm3(tb);
      if (na < 0.5 || nb < 0.5) continue;
      if (dot3(ta, tb) / na / nb > ca) continue;
      if (count < max_pairs) {
        pairs[2 * count] = k;
        pairs[2 * count + 1] = l;
      }
      ++count;
    }
  }
  return count;
}
