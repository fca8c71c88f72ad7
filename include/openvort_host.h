/*
 * openvort_host.h — host-side configuration and frame I/O of the MI355X
 * library, for a reference host built without libconfig.
 *
 * The reference reads its configuration with libconfig (src/configuration.c:
 * 169-392 load_conf, 407-535 setup_init) and is unbuildable where libconfig is
 * absent; these entry points parse the same files (libconfig's syntax:
 * settings `name = value;` or `name: value`, optional terminators, groups
 * {...}, lists (...), arrays [...], ints / 64-bit ints / floats / booleans /
 * strings, `#`, `//` and C comments) with the same typing rules as libconfig
 * without auto-conversion, which load_conf relies on: an integer written for a
 * float key (`reconnection_angle_cutoff = 5;`) or a float for an integer key is
 * not applied, exactly as config_lookup_float / config_lookup_int ignore it.
 *
 * Frames: save_tangle (src/vortex_utils.c:306-352) writes 19 columns per point
 * but load_tangle (src/vortex_utils.c:287-304, 354-387) reads 13, so the
 * reference cannot restart from its own frames.  ov_frame_save writes
 * save_tangle's format byte for byte; ov_frame_load reads it (and 13-column
 * files), restoring positions, velocities, tangents, normals, dxi, status and
 * links.
 *
 * All functions return OV_OK or a negative OV_E_* code (include/openvort_gpu.h)
 * and never abort; parse errors carry the line in the message buffer.
 */
#ifndef OPENVORT_HOST_H
#define OPENVORT_HOST_H

#include "openvort_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The reference's flow configuration (include/external_velocity.h:54-85),
 * layout-identical, so a reference host passes &vn_conf unchanged. */
typedef struct {
  const char *name;
  int type;                       /* 1 scalar, 2 vector (scalar_param, vector_param) */
  union { ov_vec3d vector; double scalar; } value;
} ov_v_param;
typedef struct {
  const char *name;
  void *fun;
  int n_params;
  ov_v_param v_params[32];
} ov_v_conf;

/* Every setting load_conf and setup_init read, with the reference defaults of
 * src/vortex_constants.c:24-95 where a key is absent. */
typedef struct {
  int max_steps;                          /* "max_steps"                 int   */
  int frame_shots;                        /* "frame_shots"               int   */
  double kappa;                           /* "KAPPA"                     float */
  int num_threads;                        /* "num_threads"               int   */
  int pin_mode;                           /* "wall_type" "pin"/"slip" -> OV_PINNED / OV_PINNED_SLIP */
  int use_BH;                             /* "use_BH"                    bool  */
  double BH_resolution;                   /* "BH_resolution"             float */
  int BH_quadtree;                        /* "BH_quadtree"               bool  */
  double BH_grain;                        /* "BH_grain"                  float */
  int use_mutual_friction;                /* "use_mutual_friction"       bool  */
  double alpha, alpha_p;                  /* "alpha", "alpha_p"          float */
  int LIA_only;                           /* "LIA"                       bool  */
  int hyperfriction;                      /* "use_hyperfriction"         bool  */
  double hyperalpha;                      /* "hyperalpha"                float */
  double max_curvature_scale;             /* "max_curvature_scale"       float */
  double dt, dl_min, dl_max;              /* "dt", "dl_min", "dl_max"    float */
  int small_loop_cutoff;                  /* "small_loop_cutoff"         int   */
  double reconnection_angle_cutoff;       /* "reconnection_angle_cutoff" float, degrees in the file, radians here */
  double rec_dist;                        /* "reconnection_distance"     float */
  int eliminate_origin_loops;             /* "eliminate_origin_loops"    bool  */
  double eliminate_loops_origin_cutoff;   /* "eliminate_loops_origin_cutoff" float */
  int eliminate_zaxis_loops;              /* "eliminate_zaxis_loops"     bool  */
  double eliminate_loops_zaxis_cutoff;    /* "eliminate_loops_zaxis_cutoff"  float */
  int eliminate_outer_loops;              /* "eliminate_outer_loops"     bool  */
  double eliminate_outer_loops_cutoff;    /* "eliminate_outer_loops_cutoff"  float */
  int loop_injection;                     /* "loop_injection"            bool  */
  double loop_injection_frequency;        /* "loop_injection_frequency"  float */
  int line_injection;                     /* "line_injection"            bool  */
  double line_injection_frequency;        /* "line_injection_frequency"  float */
  int line_injection_n;                   /* "line_injection_n"          int   */
  int line_injection_polarized;           /* "line_injection_polarized"  bool  */
  ov_flow vn, vs, vb;                     /* groups "vn_conf", "vs_conf", "vb_conf" */
  int vn_set, vs_set, vb_set;             /* 1 when the group was present */
  ov_domain_box box;                      /* "domain" and the walls of "boundaries" */
  char boundaries[32];                    /* "boundaries" preset name */
  ov_image_tangle images[26];             /* the preset's image list (src/boundary_images.c) */
  int n_images;
  /* initial condition (setup_init): the mode and its parameters */
  char init_mode[64];                     /* "init_mode" */
  char init_file[256];                    /* "init_file" (restart) */
  int init_loops_n;                       /* "init_loops_n" */
  double loop_center[3], loop_dir[3];     /* "loop_center", "loop_dir" */
  double loop_r;                          /* "loop_r" */
  int loop_N;                             /* "loop_N" */
  double ring_r;                          /* "ring_r" */
  int ring_N;                             /* "ring_N" */
  int init_npairs, init_line_points, init_nvorts, init_direction, init_lattice_shells;
  double init_lattice_constant, init_separation;
  int dipole_direction, init_line_direction, init_k_KW;
  double init_r_KW;
} ov_config;

/* The reference defaults (vortex_constants.c:24-95), "no flow" flows, open box. */
void ov_config_default(ov_config *cfg);

/* Parse configuration text / a configuration file into cfg (which should hold
 * ov_config_default first); the keys and their effects follow load_conf and
 * setup_init.  A key that is missing keeps its value.  Errors a reference
 * run would abort on (unknown flow type or boundaries, missing domain,
 * boundaries, init_mode or a flow parameter, malformed syntax) return
 * OV_E_INVALID with a message in err[0..errlen). */
int ov_config_parse(const char *text, ov_config *cfg, char *err, int errlen);
int ov_config_load(const char *path, ov_config *cfg, char *err, int errlen);

/* The ov_params of a configuration: physics globals and the three flows.
 * VORTEX_WIDTH is not a configuration key of the reference
 * (vortex_constants.c:26), so vortex_width keeps p's value. */
void ov_params_from_config(const ov_config *cfg, ov_params *p);

/* A reference flow configuration (v_conf_t, found by name in the reference's
 * table, src/external_velocity.c:85-174) as an ov_flow; its parameters are
 * read by name over its first n_params entries, as get_v_param_scalar /
 * get_v_param_vector do.  OV_E_INVALID for an unknown name. */
int ov_flow_from_vconf(const ov_v_conf *conf, ov_flow *out);

/* save_tangle (vortex_utils.c:306-352): one line per point, loop by loop, in
 * the reference's column order and "%.15g" formatting; digits = 17 writes
 * every double exactly instead (a lossless variant of the same format). */
int ov_frame_save(const char *path, const ov_tangle_state *t, int digits);

/* Points in a frame file (non-empty lines). */
int ov_frame_count(const char *path, int *n_points);

/* Read a frame written by save_tangle or ov_frame_save into t, whose arrays
 * the caller allocated for t->N >= the file's point count: points in file
 * order, vels/tangents/normals as saved; with 19 columns also status,
 * pin_wall and dxi, and the links as saved (mapped through the saved point
 * indices); with 13 columns, each vortex is closed into a loop in file order
 * (load_tangle's linking).  Entries past the file's points become EMPTY.
 * next_free / total_free are set as the reference's tangle keeps them. */
int ov_frame_load(const char *path, ov_tangle_state *t, char *err, int errlen);

#ifdef __cplusplus
}
#endif
#endif /* OPENVORT_HOST_H */
