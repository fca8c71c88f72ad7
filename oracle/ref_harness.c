/*
 * ref_harness.c — drives the compiled reference (emil-varga/openvort) from
 * ctypes.  TEST INFRASTRUCTURE ONLY: linked with the reference's own
 * src/*.c (oracle/Makefile) into oracle/_ref/libovref.so, which pins the
 * restatement in ov_oracle.c and can serve as the timed CPU baseline.
 *
 * The reference's globals (src/vortex_constants.c:24-95) and flow configs
 * (src/external_velocity.c:76-78) are set from an ov_params; the tangle is
 * passed straight through, ov_tangle_state being layout-identical to
 * struct tangle_state (include/tangle.h:82-106; checked below).
 */
#include <stddef.h>
#include <string.h>

#include "external_velocity.h"
#include "octree.h"
#include "tangle.h"
#include "vortex_constants.h"
#include "vortex_dynamics.h"

#include "../include/openvort_gpu.h"
#include "../include/openvort_host.h"

_Static_assert(sizeof(ov_tangle_state) == sizeof(struct tangle_state), "layout");
_Static_assert(offsetof(ov_tangle_state, connections) == offsetof(struct tangle_state, connections), "layout");
_Static_assert(offsetof(ov_tangle_state, box) == offsetof(struct tangle_state, box), "layout");
_Static_assert(offsetof(ov_tangle_state, bimg) == offsetof(struct tangle_state, bimg), "layout");
_Static_assert(offsetof(ov_tangle_state, N) == offsetof(struct tangle_state, N), "layout");
_Static_assert(sizeof(ov_node_status) == sizeof(node_status), "layout");
_Static_assert(sizeof(ov_image_tangle) == sizeof(struct image_tangle), "layout");
_Static_assert(sizeof(ov_v_conf) == sizeof(struct v_conf_t), "layout");
_Static_assert(offsetof(ov_v_conf, v_params) == offsetof(struct v_conf_t, v_params), "layout");
_Static_assert(sizeof(ov_v_param) == sizeof(struct v_param_t), "layout");

extern int global_num_threads;

static int find_conf(const char *name) {
  for (int k = 0; v_confs[k].fun; ++k)
    if (strcmp(v_confs[k].name, name) == 0) return k;
  return -1;
}

/* Table names in the order of the OV_FLOW_* codes. */
static const char *const flow_names[OV_FLOW_COUNT] = {
    "no flow", "simple", "spherical", "spherical and simple", "cylindrical", "oscillating",
    "oscillating slit", "coscos", "coscos-divfree", "cos-divfree", "simple-shear", "rotosc",
    "DG", "rotosc-boundary", "constant acceleration", "parabolic"};

/* The table entry of f's type with f's parameter values filled in by name:
 * "external_v" <- v, "wave" <- w, "polarization" <- q, the scalars <- a[] in
 * table order (entries listed past n_params included, as rotosc-boundary's
 * delta). */
static int set_flow(struct v_conf_t *dst, const ov_flow *f) {
  if (f->type < 0 || f->type >= OV_FLOW_COUNT) return -1;
  const int k = find_conf(flow_names[f->type]);
  if (k < 0) return -1;
  *dst = v_confs[k];
  int scalar = 0;
  for (int j = 0; j < 32 && dst->v_params[j].name; ++j) {
    struct v_param_t *q = &dst->v_params[j];
    if (strcmp(q->name, "external_v") == 0)
      q->value.vector = vec3(f->v[0], f->v[1], f->v[2]);
    else if (strcmp(q->name, "wave") == 0)
      q->value.vector = vec3(f->w[0], f->w[1], f->w[2]);
    else if (strcmp(q->name, "polarization") == 0)
      q->value.vector = vec3(f->q[0], f->q[1], f->q[2]);
    else if (scalar < 4)
      q->value.scalar = f->a[scalar++];
  }
  return 0;
}

/* The reference's flow function of f at r, t (its return value ignored, as
 * get_vn/get_vs do). */
int ovref_flow(const ov_flow *f, const double *r, double t, double *out) {
  struct v_conf_t c;
  if (set_flow(&c, f)) return -1;
  struct vec3d w = vec3(r[0], r[1], r[2]), res = vec3(0, 0, 0);
  c.fun(&w, t, &res, &c);
  for (int i = 0; i < 3; ++i) out[i] = res.p[i];
  return 0;
}

/* Apply an ov_params to the reference globals. */
int ovref_set_params(const ov_params *p, int threads) {
  KAPPA = p->kappa;
  VORTEX_WIDTH = p->vortex_width;
  global_alpha = p->alpha;
  global_alpha_p = p->alpha_p;
  global_dl_max = p->dl_max;
  global_LIA_only = p->LIA_only;
  global_use_mutual_friction = p->use_mutual_friction;
  global_hyperfriction = p->hyperfriction;
  global_max_curvature_scale = p->max_curvature_scale;
  global_hyperalpha = p->hyperalpha;
  global_use_BH = p->use_BH;
  global_BH_resolution = p->BH_resolution;
  global_BH_grain = p->BH_grain;
  global_BH_quadtree = p->BH_quadtree;
  global_num_threads = threads > 0 ? threads : 1;
  if (set_flow(&vn_conf, &p->vn) || set_flow(&vs_conf, &p->vs) || set_flow(&vb_conf, &p->vb))
    return -2;
  return 0;
}

/* update_tangents_normals (tangle.c:723-727). */
void ovref_update_tangents_normals(ov_tangle_state *t) {
  update_tangents_normals((struct tangle_state *)t);
}

/* update_velocities (tangle.c:686-702), octree built inside when use_BH. */
void ovref_update_velocities(ov_tangle_state *t, double time) {
  update_velocities((struct tangle_state *)t, time, NULL);
}

/* update_velocity for the listed points only, OpenMP-parallel like
 * update_velocities; results copied to [n][3] arrays. */
void ovref_velocities_sample(ov_tangle_state *t, double time, const int *idx, int n,
                             double *vs_out, double *vels_out) {
  struct tangle_state *ts = (struct tangle_state *)t;
  struct octree *tree = global_use_BH ? octree_build(ts, global_BH_quadtree) : NULL;
#pragma omp parallel for schedule(dynamic, 4) num_threads(global_num_threads)
  for (int q = 0; q < n; ++q) {
    int k = idx[q];
    struct vec3d vs, vel;
    if (ts->status[k].status == EMPTY) {
      vs = vec3(0, 0, 0);
      vel = vs;
    } else {
      /* update_velocity writes into the tangle arrays at k; the caller gives
       * distinct k, so the OpenMP loop is race-free as in tangle.c:693-699 */
      update_velocity(ts, k, time, tree);
      vs = ts->vs[k];
      vel = ts->vels[k];
    }
    memcpy(vs_out + 3 * q, vs.p, sizeof vs.p);
    memcpy(vels_out + 3 * q, vel.p, sizeof vel.p);
  }
  if (tree) octree_destroy(tree);
}

/* rk4_step (vortex_dynamics.c:119-121): k1 = t->vels. */
void ovref_rk4_step(ov_tangle_state *t, double time, double dt) {
  rk4_step((struct tangle_state *)t, time, dt);
}

/* enforce_boundaries (tangle.c:796-819). */
void ovref_enforce_boundaries(ov_tangle_state *t) {
  enforce_boundaries((struct tangle_state *)t);
}

/* The reference's reconnect (vortex_dynamics.c:140-315) with the reconnection
 * globals it reads; returns the number of reconnections. */
int ovref_reconnect(ov_tangle_state *t, double time, double rdist, double rangle) {
  return reconnect((struct tangle_state *)t, time, rdist, rangle);
}

/* The vortex-vortex part of the same pass driven by a candidate list (pairs
 * k < l in increasing (k, l) order, the pair-loop tests that do not change
 * during the pass already applied): the tests that do change — the
 * "recalculate" flags, the neighbour test, the approach test on the two
 * points' velocities — and the rewiring are the reference's own calls, in the
 * order of its loop (vortex_dynamics.c:214-312).  For boxes without walls it
 * leaves the tangle as ovref_reconnect does; returns the reconnections. */
int do_reconnection(struct tangle_state *tangle, size_t k, size_t l);   /* vortex_dynamics.c:317 */

int ovref_reconnect_candidates(ov_tangle_state *t, double time, const int *pairs, long n_pairs) {
  struct tangle_state *ts = (struct tangle_state *)t;
  int recs = 0;
  for (int q = 0; q < ts->N; ++q) ts->recalculate[q] = 0;
  long i = 0;
  while (i < n_pairs) {
    const int k = pairs[2 * i];
    long next_k = i;
    while (next_k < n_pairs && pairs[2 * next_k] == k) ++next_k;   /* this k's run */
    if (!ts->recalculate[k]) {
      for (long j = i; j < next_k; ++j) {
        const int l = pairs[2 * j + 1];
        if (ts->connections[k].forward == l || ts->connections[k].reverse == l ||
            ts->recalculate[l])
          continue;
        struct segment seg = seg_pwrap(&ts->vnodes[k], &ts->vnodes[l], &ts->box);
        update_velocity(ts, k, time, NULL);
        update_velocity(ts, l, time, NULL);
        struct vec3d dx, dv;
        vec3_sub(&dx, &seg.r1, &seg.r2);
        vec3_sub(&dv, &ts->vels[k], &ts->vels[l]);
        if (vec3_dot(&dx, &dv) > 0) continue;
        const int knext = ts->connections[k].forward, kprev = ts->connections[k].reverse;
        const int lnext = ts->connections[l].forward, lprev = ts->connections[l].reverse;
        if (!do_reconnection(ts, k, l)) continue;
        ts->recalculate[k]++;
        ts->recalculate[ts->connections[k].forward]++;
        ts->recalculate[ts->connections[k].reverse]++;
        ts->recalculate[knext]++;
        ts->recalculate[kprev]++;
        ts->recalculate[l]++;
        ts->recalculate[ts->connections[l].forward]++;
        ts->recalculate[ts->connections[l].reverse]++;
        ts->recalculate[lnext]++;
        ts->recalculate[lprev]++;
        ++recs;
        break;
      }
    }
    i = next_k;
  }
  return recs;
}

/* ---- frames and flow configurations (host I/O checks) --------------------- */

#include "vortex_utils.h"

/* The reference's save_tangle (vortex_utils.c:306-352) on t. */
void ovref_save_tangle(ov_tangle_state *t, const char *path) {
  save_tangle(path, (struct tangle_state *)t);
}

/* The reference's load_tangle (vortex_utils.c:354-387) of path into a tangle
 * of capacity n it creates itself; the first n positions and links are
 * copied out ([n][3] doubles, [n][2] ints).  Returns load_tangle's value. */
int ovref_load_tangle(const char *path, int n, double *pos, int *links) {
  struct tangle_state t;
  create_tangle(&t, (size_t)n);
  int rc = load_tangle(path, &t);
  for (int k = 0; k < n && k < t.N; ++k) {
    for (int i = 0; i < 3; ++i) pos[3 * k + i] = t.vnodes[k].p[i];
    links[2 * k] = t.connections[k].forward;
    links[2 * k + 1] = t.connections[k].reverse;
  }
  free_tangle(&t);
  return rc;
}

/* sizeof(struct v_conf_t) */
int ovref_vconf_size(void) { return (int)sizeof(struct v_conf_t); }

/* The reference's v_conf_t for f (the table entry, parameters by name) into
 * out, which holds ovref_vconf_size() bytes. */
int ovref_vconf(const ov_flow *f, void *out) {
  struct v_conf_t c;
  if (set_flow(&c, f)) return -1;
  memcpy(out, &c, sizeof c);
  return 0;
}

/* ---- topology edits (tangle.c:796-1070) ------------------------------------ */

extern double global_dl_min;
extern int global_eliminate_origin_loops, global_eliminate_zaxis_loops, global_eliminate_outer_loops;
extern double global_eliminate_loops_origin_cutoff, global_eliminate_loops_zaxis_cutoff,
    global_eliminate_outer_loops_cutoff;

/* The reference's remesh (tangle.c:824-873).  The caller's tangle must have
 * EMPTY slots for every point remesh adds (expand_tangle would realloc the
 * caller's arrays); returns -1 without calling remesh when fewer than
 * `need_free` slots are EMPTY. */
int ovref_remesh(ov_tangle_state *t, double min_dist, double max_dist, int need_free) {
  int free_slots = 0;
  for (int k = 0; k < t->N; ++k) free_slots += t->status[k].status == OV_EMPTY;
  if (free_slots < need_free) return -1;
  remesh((struct tangle_state *)t, min_dist, max_dist);
  return 0;
}

/* The reference's eliminate_small_loops (tangle.c:878-945) with the
 * elimination globals set first. */
void ovref_eliminate_small_loops(ov_tangle_state *t, int loop_length, int origin, double origin_cut,
                                 int zaxis, double zaxis_cut, int outer, double outer_cut) {
  global_eliminate_origin_loops = origin;
  global_eliminate_loops_origin_cutoff = origin_cut;
  global_eliminate_zaxis_loops = zaxis;
  global_eliminate_loops_zaxis_cutoff = zaxis_cut;
  global_eliminate_outer_loops = outer;
  global_eliminate_outer_loops_cutoff = outer_cut;
  eliminate_small_loops((struct tangle_state *)t, loop_length);
}
