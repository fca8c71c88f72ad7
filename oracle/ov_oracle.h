/*
 * ov_oracle.h — FP64 CPU restatement of OpenVort's velocity / RK4 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU library is compared
 * against; only tests/, the smoke check and the CPU-baseline leg of bench.py
 * may load it.  The product path (openvort_amd/) never calls it.
 *
 * Every function restates the reference algorithm (file:line cited at each
 * definition in ov_oracle.c) on the same ov_tangle_state / ov_params types as
 * the GPU ABI (include/openvort_gpu.h).  Return codes follow that ABI.
 */
#ifndef OV_ORACLE_H
#define OV_ORACLE_H

#include "../include/openvort_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* dxi, tangents, normals of every non-EMPTY point (tangle.c:704-727). */
int ovo_update_tangents_normals(ov_tangle_state *t);

/* vs/vels of every non-EMPTY point by the direct sum or the octree
 * (tangle.c:583-702).  Geometry must be current. */
int ovo_update_velocities(ov_tangle_state *t, double time, const ov_params *p);

/* vs/vels of the listed points only (the same per-point assembly as
 * ovo_update_velocities); out arrays are [n][3]. */
int ovo_velocities_sample(const ov_tangle_state *t, double time,
                          const ov_params *p, const int *idx, int n,
                          double *vs_out, double *vels_out);

/* rk4_step (vortex_dynamics.c:58-121): k1 is t->vels as given. */
int ovo_rk4_step(ov_tangle_state *t, double time, double dt, const ov_params *p);

/* Periodic wrap of every point into the box (tangle.c:796-819). */
void ovo_enforce_boundaries(ov_tangle_state *t);

/* Biot-Savart of one straight segment a->b at r (tangle.c:486-514), no
 * kappa/4pi factor applied twice: the result includes kappa/4pi. */
void ovo_segment_field(const double a[3], const double b[3], const double r[3],
                       double kappa, double out[3]);

/* Octree statistics of the last ovo_update_velocities with use_BH:
 * non-empty cells, evaluated leaf segment-point pairs, multipoles. */
void ovo_octree_stats(long *n_cells, long *n_leaf_pairs, long *n_multipoles);

/* Far-field value at r[3] of one octree cell holding the points ids[0..n-1]
 * (the cell moments and multipole of the octree path, mode p->multipole). */
int ovo_cell_far_field(const ov_tangle_state *t, const int *ids, int n, const double *r,
                       const ov_params *p, double *out);

/* Pass-invariant reconnection tests over all pairs k < l (vortex_dynamics.c:
 * 221-265): writes up to max_pairs (k, l) pairs in loop order, returns the
 * total count. */
long ovo_reconnection_candidates(const ov_tangle_state *t, double rec_dist, double rec_angle,
                                 int *pairs, long max_pairs);

/* The imposed flow fl at position r[3], time t (external_velocity.c:85-512);
 * OV_E_INVALID for an unknown type. */
int ovo_flow(const ov_flow *fl, const double *r, double t, double *out);

/* Threads used by the OpenMP loops (0: runtime default). */
void ovo_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
