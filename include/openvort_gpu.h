/*
 * openvort_gpu.h — C ABI of the MI355X vortex-filament velocity / RK4 library.
 *
 * Drop-in boundary for OpenVort's hot path.  The host keeps its own
 * `struct tangle_state` (reference include/tangle.h:82-106), configuration,
 * reconnection and remeshing; it hands the tangle to these entry points in
 * place of
 *   update_velocities()        reference include/tangle.h:128, src/tangle.c:686-702
 *   update_tangle()            reference include/tangle.h:160-164
 *   update_tangents_normals()  reference include/tangle.h:120, src/tangle.c:723-727
 *   rk4_step()                 reference include/vortex_dynamics.h:49,
 *                              src/vortex_dynamics.c:58-121
 *
 * Every type below is layout-identical to its reference counterpart, so a
 * reference host passes `(ov_tangle_state *)&tangle` without copying.  The
 * physics globals of reference src/vortex_constants.c:24-95 and the flow
 * descriptors of include/external_velocity.h:80-85 travel in `ov_params`.
 *
 * Conventions (all entry points):
 *   - return OV_OK (0) or a negative OV_E_* code; never abort; the message of
 *     the last failure on a context is ov_last_error(ctx);
 *   - host arrays are caller-owned and re-read on every call (the reference
 *     reallocates them, src/tangle.c:114-150), nothing keeps host pointers;
 *   - calls are synchronous on return; one context per host thread;
 *   - a non-finite velocity or position (the reference would trap with
 *     SIGFPE, src/main.c:56) is reported as OV_E_NONFINITE after the outputs
 *     have been written.
 */
#ifndef OPENVORT_GPU_H
#define OPENVORT_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tangle types, layout-identical to the reference -------------------- */

typedef struct { double p[3]; } ov_vec3d;              /* vec3_maths.h:23-25 */
typedef struct { int forward, reverse; } ov_neighbour;  /* tangle.h:41-44 */
typedef struct { int status, pin_wall; } ov_node_status;/* tangle.h:36-39 */

enum { OV_EMPTY = -1, OV_FREE = 0, OV_PINNED = 1, OV_PINNED_SLIP = 2 }; /* tangle.h:28-33 */
enum { OV_WALL_OPEN = 0, OV_WALL_PERIODIC = 1, OV_WALL_MIRROR = 2 };    /* vec3_maths.h:38 */
enum { OV_X_L = 0, OV_X_H, OV_Y_L, OV_Y_H, OV_Z_L, OV_Z_H };            /* vec3_maths.h:46-51 */

typedef struct {                                        /* vec3_maths.h:61-68 */
  ov_vec3d bottom_left_back;
  ov_vec3d top_right_front;
  int wall[6];
} ov_domain_box;

typedef struct { int shift[3]; int reflect; } ov_image_tangle;   /* tangle.h:50-53 */
typedef struct { const ov_image_tangle *images; int n; } ov_boundary_images; /* tangle.h:55-58 */

typedef struct {                                        /* tangle.h:82-106 */
  ov_vec3d *vnodes;
  ov_vec3d *vnodes_new;
  ov_vec3d *vels;
  ov_vec3d *vs;
  ov_vec3d *tangents;
  ov_vec3d *normals;
  double *dxi;
  int *recalculate;
  ov_neighbour *connections;
  ov_node_status *status;
  ov_domain_box box;
  ov_boundary_images bimg;
  int N;
  int next_free;
  int total_free;
} ov_tangle_state;

/* ---- parameters: the reference globals + flow descriptors --------------- */

/* Imposed flows: every entry of the reference's flow table v_confs[]
 * (external_velocity.c:85-174; "no flow" and "simple" first, the others in
 * table order), evaluated per point at the
 * point's position and the evaluation's time (tangle.c:601, 662).
 * Parameters, by the names of the table's v_params:
 *   v     "external_v"                                     SIMPLE, SPHERICAL_AND_SIMPLE
 *   a[0..3] the scalar parameters in table order:
 *     SPHERICAL, SPHERICAL_AND_SIMPLE, CYLINDRICAL: strength, cutoff
 *     OSCILLATING:        strength, frequency      (w = "wave", q = "polarization")
 *     OSCILLATING_SLIT:   strength, frequency, cutoff
 *     COSCOS, COSCOS_DIVFREE, COS_DIVFREE:          k, v0, l
 *     SIMPLE_SHEAR:       rate, v0
 *     ROTOSC:             freq, amp
 *     DG:                 rot, vz
 *     ROTOSC_BOUNDARY:    freq, amp, delta
 *     CONSTANT_ACCELERATION: accel
 *     PARABOLIC:          H, amp
 * Each flow computes what the reference's function computes, quirks included:
 * "oscillating slit" ignores strength (external_velocity.c:296-322),
 * "rotosc-boundary" runs get_v_rotosc (the table entry, :156-157), so delta is
 * unused.  "spherical and simple" returns in the reference before writing its
 * result (get_v_simple returns 1, :225-241), leaving the caller's vector
 * uninitialised; here it is the sum of the two flows, what the function
 * computes before that return. */
enum { OV_FLOW_NONE = 0,                  /* "no flow" external_velocity.c:176-182 */
       OV_FLOW_SIMPLE = 1,                /* "simple"               :184-188 */
       OV_FLOW_SPHERICAL = 2,             /* "spherical"            :197-223 */
       OV_FLOW_SPHERICAL_AND_SIMPLE = 3,  /* "spherical and simple" :225-241 */
       OV_FLOW_CYLINDRICAL = 4,           /* "cylindrical"          :243-270 */
       OV_FLOW_OSCILLATING = 5,           /* "oscillating"          :272-294 */
       OV_FLOW_OSCILLATING_SLIT = 6,      /* "oscillating slit"     :296-322 */
       OV_FLOW_COSCOS = 7,                /* "coscos"               :324-344 */
       OV_FLOW_COSCOS_DIVFREE = 8,        /* "coscos-divfree"       :347-370 */
       OV_FLOW_COS_DIVFREE = 9,           /* "cos-divfree"          :372-394 */
       OV_FLOW_SIMPLE_SHEAR = 10,         /* "simple-shear"         :396-413 */
       OV_FLOW_ROTOSC = 11,               /* "rotosc"               :415-433 */
       OV_FLOW_DG = 12,                   /* "DG"                   :438-457 */
       OV_FLOW_ROTOSC_BOUNDARY = 13,      /* "rotosc-boundary" (runs get_v_rotosc) */
       OV_FLOW_CONSTANT_ACCELERATION = 14,/* "constant acceleration" :483-495 */
       OV_FLOW_PARABOLIC = 15,            /* "parabolic"            :497-512 */
       OV_FLOW_COUNT = 16 };
typedef struct {
  int type;          /* OV_FLOW_* */
  double v[3];       /* "external_v" */
  double a[4];       /* scalar parameters, table order */
  double w[3], q[3]; /* "wave", "polarization" */
} ov_flow;

enum { OV_PREC_FP64 = 0,     /* every pair term in FP64 */
       OV_PREC_MIXED = 1 };  /* FP32 far field, FP64 near field, LIA, RK */
enum { OV_RK_REFERENCE = 0,  /* k4 at x + dt/2 k3, vortex_dynamics.c:89 */
       OV_RK_CLASSIC = 1 };  /* k4 at x + dt k3 */
enum { OV_MP_REFERENCE = 0,  /* octree.c:303-328 as written */
       OV_MP_DIPOLE = 1 };   /* true first-order (dipole) expansion */

typedef struct {
  double kappa;               /* KAPPA                       vortex_constants.c:28 */
  double vortex_width;        /* VORTEX_WIDTH (a0)           vortex_constants.c:26 */
  double alpha, alpha_p;      /* global_alpha, _alpha_p      vortex_constants.c:34-35 */
  double dl_max;              /* global_dl_max               vortex_constants.c:44 */
  int LIA_only;               /* global_LIA_only             vortex_constants.c:31 */
  int use_mutual_friction;    /* global_use_mutual_friction  vortex_constants.c:53 */
  int hyperfriction;          /* global_hyperfriction        vortex_constants.c:55 */
  int use_BH;                 /* global_use_BH               vortex_constants.c:91 */
  double max_curvature_scale; /* global_max_curvature_scale  vortex_constants.c:56 */
  double hyperalpha;          /* global_hyperalpha           vortex_constants.c:57 */
  double BH_resolution;       /* global_BH_resolution (theta) vortex_constants.c:93 */
  double BH_grain;            /* global_BH_grain             vortex_constants.c:95 */
  int BH_quadtree;            /* global_BH_quadtree          vortex_constants.c:92 */
  int precision;              /* OV_PREC_*   (no reference counterpart) */
  int rk_variant;             /* OV_RK_*     (no reference counterpart) */
  int multipole;              /* OV_MP_*     (no reference counterpart) */
  double near_radius;         /* mixed: FP64 below this box distance (cm), <=0: 0.02 */
  ov_flow vn, vs, vb;         /* vn_conf, vs_conf, vb_conf   external_velocity.c:76-78 */
} ov_params;

/* Error codes. */
enum {
  OV_OK = 0,
  OV_E_INVALID = -1,      /* malformed tangle or arguments */
  OV_E_UNSUPPORTED = -2,  /* a feature outside an entry point's scope (pinned points in the adjoint, ...) */
  OV_E_NONFINITE = -3,    /* a result is NaN/Inf (outputs still written) */
  OV_E_DEVICE = -4,       /* HIP runtime failure */
  OV_E_NOMEM = -5         /* device or host allocation failed */
};

/* Flags of ov_gpu_velocities / ov_gpu_rk4_step / ov_gpu_rk4_run. */
enum {
  OV_UPDATE_GEOMETRY = 1,        /* recompute dxi/tangents/normals first (update_tangle) */
  OV_USE_CALLER_K1 = 2,          /* rk4: k1 = tangle->vels as given (rk4_step alone, vortex_dynamics.c:100) */
  OV_ENFORCE_BOUNDARIES = 4,     /* rk4_run: periodic wrap after each step (tangle.c:796-819) */
  OV_PROFILE = 8                 /* time each kernel with HIP events (ov_ctx_kernel_stats) */
};

typedef struct ov_ctx ov_ctx;

/* Reference defaults of vortex_constants.c:24-95 and the no-flow configs
 * (external_velocity.c:80-83); precision FP64, reference RK and multipole. */
void ov_params_default(ov_params *p);

/* Context: device buffers, stream and events for one host thread. */
ov_ctx *ov_ctx_create(int device);
void ov_ctx_destroy(ov_ctx *ctx);
const char *ov_last_error(const ov_ctx *ctx);
/* hipStream_t the context launches on (for callers that order their own work). */
void *ov_ctx_stream(ov_ctx *ctx);

/* update_tangents_normals (tangle.h:120): writes dxi, tangents, normals.
 *
 * Walls (this call, ov_gpu_velocities and ov_gpu_rk4_step / ov_gpu_rk4_run,
 * direct sum or octree, the ensembles, the topology edits and ov_gpu_run):
 * PINNED and PINNED_SLIP points with pin_wall 0..5
 * may end open chains (forward or reverse -1).  Neighbours past a chain end
 * are the mirror images of the chain in that point's wall (step_node,
 * tangle.c:165-198); PINNED points take the one-sided stencil (tangle.c:
 * 275-362, forward when the caller's forward index is > 0, as the reference
 * tests), their non-PINNED neighbours the neighbour-tangent stencil
 * (tangle.c:364-430); PINNED points get vs = 0 and vels = v_b
 * (tangle.c:589-594), PINNED_SLIP points lose the velocity normal to their
 * wall (tangle.c:677-683); in the octree a chain end has no segment
 * (octree.c:254-255).  Refused with OV_E_INVALID: a free point without
 * both links, a pinned point without any, two neighbouring points that both
 * border a PINNED point (the reference recurses without end there).  The
 * reference reads status[-1] at a PINNED_SLIP chain end; here that missing
 * neighbour counts as not PINNED. */
int ov_gpu_update_tangents_normals(ov_ctx *ctx, ov_tangle_state *tangle);

/* update_velocities (tangle.h:128) at time t: writes vs[] and vels[].
 * With OV_UPDATE_GEOMETRY it is update_tangle (tangle.h:160-164) and also
 * writes dxi/tangents/normals; otherwise it reads the caller's tangents and
 * normals (the reference precondition, tangle.h:124). */
int ov_gpu_velocities(ov_ctx *ctx, ov_tangle_state *tangle, double t,
                      const ov_params *params, int flags);

/* update_tangle(t) followed by rk4_step(t, dt) (main.c:112-113): writes the
 * new vnodes, vels = k1 at t, vs at t and the geometry at t.  With
 * OV_USE_CALLER_K1 the caller's vels are k1 (rk4_step alone). */
int ov_gpu_rk4_step(ov_ctx *ctx, ov_tangle_state *tangle, double t, double dt,
                    const ov_params *params, int flags);

/* nsteps of ov_gpu_rk4_step with the tangle resident on the device between
 * steps (the host loop of main.c:96-131 without reconnection or remeshing;
 * with OV_ENFORCE_BOUNDARIES also its periodic wrap, main.c:117). */
int ov_gpu_rk4_run(ov_ctx *ctx, ov_tangle_state *tangle, double t0, double dt,
                   int nsteps, const ov_params *params, int flags);

/* Ensembles: nsteps RK4 steps (update_tangle + rk4_step each, as
 * ov_gpu_rk4_run) of ntangles independent tangles, params[k] for tangle k
 * (each with its own box, images, flows and friction, and its own walls: a
 * tangle may hold points pinned on walls as ov_gpu_rk4_run takes them; all
 * with one precision and RK variant; direct sum only).  Every velocity
 * evaluation of the whole
 * ensemble is one persistent launch over a queue of work units; one RK4 step
 * of the ensemble is captured once as a HIP graph and replayed; unit size
 * and grid are picked by timing candidates on the first call for a layout.
 * No counterpart in the reference (it holds one tangle per process). */
int ov_gpu_rk4_run_batch(ov_ctx *ctx, ov_tangle_state *const *tangles, int ntangles,
                         double t0, double dt, int nsteps, const ov_params *params, int flags);
/* Unit size (work items per unit), persistent grid (blocks of 4 waves; 0:
 * one unit per wavefront, no queue) and the tuning time of the last batch
 * layout. */
void ov_ctx_batch_tuning(const ov_ctx *ctx, int *unit, int *blocks, double *tune_ms);

/* Adjoint of update_tangle + update_velocities (tangle.c:583-702, geometry
 * tangle.c:704-727): grad_vnodes = (d vels / d vnodes)^T cot_vels at the
 * tangle's positions, FP64, direct sum (use_BH must be 0; precision selects
 * the forward sum whose values enter the mutual-friction terms).  Both arrays
 * have the tangle's capacity N; EMPTY entries of grad_vnodes are set to 0.
 * Every imposed flow of the table is differentiated (its spatial Jacobian at
 * time t).  Flags: OV_PROFILE.  No counterpart in the reference (it has no
 * gradients). */
int ov_gpu_velocity_vjp(ov_ctx *ctx, const ov_tangle_state *tangle, double t,
                        const ov_params *params, const ov_vec3d *cot_vels,
                        ov_vec3d *grad_vnodes, int flags);

/* Adjoint of ov_gpu_rk4_run from the tangle's vnodes (each step
 * update_tangle + rk4_step, vortex_dynamics.c:58-121, in the params' RK
 * variant): grad_initial = (d final vnodes / d initial vnodes)^T cot_final.
 * The forward trajectory is recomputed on the device, keeping the positions
 * every checkpoint_every steps (<= 0: only the initial ones) and, for the
 * segment being differentiated, every step's stage data (8 N x 24 bytes per
 * step).  OV_ENFORCE_BOUNDARIES wraps as the forward did (the wrap's
 * derivative is the identity); OV_USE_CALLER_K1 is rejected.  EMPTY points
 * do not move, so their entries of grad_initial are cot_final's.  Stage k
 * of the step from t sees the flows at the forward's stage times.  The
 * tangle is not modified.  No counterpart in the reference. */
int ov_gpu_rk4_run_vjp(ov_ctx *ctx, const ov_tangle_state *tangle, double t0, double dt,
                       int nsteps, const ov_params *params, int flags, int checkpoint_every,
                       const ov_vec3d *cot_final, ov_vec3d *grad_initial);

/* Device memory the adjoint's segment pass may hold for its per-target-tile
 * partial sums (6 N ceil(N / 64) doubles): within the limit each (segment,
 * target) pair is evaluated once for both of its uses; beyond it (or with
 * bytes = 0) in two passes that need 6 N x (a few hundred) doubles.  Default
 * 128 MiB (tangles up to about 13,000 points).  No counterpart in the
 * reference. */
void ov_ctx_set_adjoint_partial_limit(ov_ctx *ctx, size_t bytes);

/* Reconnection candidates of reconnect (vortex_dynamics.c:140-315): every
 * pair k < l (caller indices) of free points whose minimum-image distance is
 * at most rec_dist and whose tangents (as the tangle holds them) both have
 * length >= 0.5 and an angle cosine <= cos(rec_angle) — the tests of the
 * reference's pair loop that do not change during a reconnection pass
 * (vortex_dynamics.c:236-265), with its arithmetic.  Pairs are written as
 * pairs[2i] = k, pairs[2i+1] = l in increasing (k, l) order, at most
 * max_pairs of them; *n_pairs is the total (call again with a larger buffer
 * when it exceeds max_pairs).  The host keeps the rest of the pass: visiting
 * the candidates in order with the neighbour test, the "recalculate" flags,
 * the velocity test and do_reconnection gives the reference's pass.  Points
 * must lie within half a box length of a periodic box (enforce_boundaries
 * keeps them inside).  Points pinned on walls are skipped as the pair loop
 * skips them; the reference's wall reconnections before that loop
 * (vortex_dynamics.c:152-208) stay with the host. */
int ov_gpu_reconnection_candidates(ov_ctx *ctx, const ov_tangle_state *tangle, double rec_dist,
                                   double rec_angle, int *pairs, int max_pairs, int *n_pairs);

/* ---- topology edits (reference src/tangle.c:766-1070) -------------------- */

/* Settings of remesh and eliminate_small_loops: the reference globals
 * global_dl_min, global_dl_max, global_small_loop_cutoff and
 * global_eliminate_* (vortex_constants.c:39-66). */
typedef struct {
  double dl_min, dl_max;                  /* remesh(min_dist, max_dist) */
  int small_loop_cutoff;                  /* eliminate_small_loops(loop_length) */
  int eliminate_origin_loops;             /* whole loops within the cutoff of the origin */
  double eliminate_loops_origin_cutoff;
  int eliminate_zaxis_loops;              /* whole loops within the cutoff of the z axis */
  double eliminate_loops_zaxis_cutoff;
  int eliminate_outer_loops;              /* whole loops beyond the cutoff from the z axis */
  double eliminate_outer_loops_cutoff;
} ov_topology;

/* The reference defaults (dl_min 1e-3, dl_max 5e-3, small_loop_cutoff 5,
 * eliminations off with cutoffs 3e-2, 2e-2, 1). */
void ov_topology_default(ov_topology *tp);

/* remesh(tangle, min_dist, max_dist) (tangle.c:824-873) on the GPU: points
 * visited in index order, a point removed (remove_point, tangle.c:1017-1066)
 * when a neighbouring segment is shorter than min_dist and the two together
 * shorter than max_dist, one added (add_point, tangle.c:1069-1160, at the
 * slot get_tangle_next_free gives) after a segment longer than max_dist, the
 * geometry refreshed after every change and the periodic wrap applied when
 * points were added.  Reads the caller's dxi / tangents / normals (current,
 * as the reference requires); writes vnodes, links, status, dxi, tangents,
 * normals and next_free.  Where the reference would expand_tangle (no EMPTY
 * slot left) it returns OV_E_NOMEM and leaves the tangle unchanged: pass a
 * tangle with spare capacity.  add_point's curvature test uses
 * params->max_curvature_scale / params->dl_max (global_max_curvature_scale,
 * global_dl_max).  With points pinned on walls the refresh is the walls'
 * geometry (above) and a pinned neighbour is not moved by remove_point;
 * a removal that leaves two neighbouring points between PINNED points
 * (where the reference's update_tangent_normal recurses without end) stops
 * the call with OV_E_INVALID. */
int ov_gpu_remesh(ov_ctx *ctx, ov_tangle_state *tangle, const ov_params *params,
                  double min_dist, double max_dist);

/* eliminate_small_loops(tangle, tp->small_loop_cutoff) (tangle.c:878-945)
 * with the origin / z-axis / outer eliminations it runs first when enabled
 * (tangle.c:947-1015): every loop of fewer than small_loop_cutoff + 1
 * points, or lying wholly inside (outside) the cutoff, is removed.  An open
 * chain of L points ending on walls counts L - 1 steps as a loop of L points
 * does and goes as a whole; the origin / z-axis / outer eliminations walk
 * closed loops only and return OV_E_UNSUPPORTED when points are pinned.
 * Writes links and status.  The reference's use of `recalculate` as scratch
 * marks is not reproduced. */
int ov_gpu_eliminate_small_loops(ov_ctx *ctx, ov_tangle_state *tangle, const ov_topology *tp);

/* enforce_boundaries (tangle.c:796-819): periodic wrap of every point. */
int ov_gpu_enforce_boundaries(ov_ctx *ctx, ov_tangle_state *tangle);

/* nsteps of the reference's main loop body (main.c:112-117) with the tangle
 * resident on the device: update_tangle(t); rk4_step(t, dt);
 * update_tangents_normals; remesh(dl_min, dl_max); eliminate_small_loops;
 * enforce_boundaries; t += dt.  The point stream is rebuilt on the device
 * every step (loops and open chains in order of their lowest index, a chain
 * from its head, as every other entry point orders them, with the pin codes
 * of points on walls), so each step equals ov_gpu_rk4_step followed by the
 * three edits above.  Writes vnodes, links, status, next_free, the geometry
 * after the last step's edits, and vels / vs of the last step's k1 (slots
 * created afterwards keep the values the slot held, as in the reference).
 * OV_E_NOMEM as ov_gpu_remesh when the capacity runs out (nothing written).
 * Direct sum or octree (use_BH: each evaluation rebuilds the tree, points
 * inserted in ascending index as the reference inserts them).  Flags:
 * OV_PROFILE. */
int ov_gpu_run(ov_ctx *ctx, ov_tangle_state *tangle, double t0, double dt, int nsteps,
               const ov_params *params, const ov_topology *tp, int flags);

/* Allocation-free variant of ov_gpu_velocities with OV_UPDATE_GEOMETRY
 * (update_tangle, direct sum).  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Every array is a
 * caller-owned DEVICE array: pos, tangents, normals, vs, vels are three planes
 * x[stride], y[stride], z[stride]; dxi one plane; fwd/rev the forward and
 * reverse links of the n points as indices into those planes (every point a
 * free point on a closed loop or a periodic line: no EMPTY or pinned entries).
 * box and bimg are host structs (reference tangle.h:55-58, vec3_maths.h:61-68).
 * Intermediate buffers (local induction, record stream, partial sums) live in
 * one file-scope device array of 14.8 MiB per GPU, held until the process
 * ends, so: stride must be at most 16384 points (larger is OV_E_INVALID),
 * use_BH must be 0 (OV_E_UNSUPPORTED), and calls that overlap on one GPU race
 * (they share that array).  Launches go on the context's stream; the call is
 * synchronous on return and calls no allocating routine. */
int ov_gpu_velocities_noalloc(ov_ctx *ctx, const double *pos, const int *fwd, const int *rev,
                              int n, long stride, const ov_domain_box *box,
                              const ov_boundary_images *bimg, double t, const ov_params *params,
                              double *tangents, double *normals, double *dxi, double *vs,
                              double *vels);

/* Per-kernel timings gathered while OV_PROFILE was set; kernel names are the
 * device function names.  Returns the number of launches, 0 if none. */
long ov_ctx_kernel_stats(ov_ctx *ctx, const char *kernel, double *total_ms);
void ov_ctx_reset_stats(ov_ctx *ctx);

/* Segment-point interactions of the last velocity evaluation (the
 * direct-sum count N_tgt (N_seg - 2) + n_img N_tgt N_seg). */
double ov_ctx_last_interactions(const ov_ctx *ctx);

/* Device time of the last velocities / rk4 / run / topology-edit call in
 * milliseconds: HIP events on the context stream around the kernels,
 * host<->device copies of the tangle excluded. */
double ov_ctx_last_device_ms(const ov_ctx *ctx);

/* Octree of the last evaluation with use_BH: cell count; leaf segment-point
 * interactions and multipole evaluations summed over the evaluations made
 * with OV_PROFILE since the previous call (then reset). */
void ov_ctx_tree_stats(ov_ctx *ctx, double *cells, double *leaf_pairs, double *multipoles);

#ifdef __cplusplus
}
#endif
#endif /* OPENVORT_GPU_H */
