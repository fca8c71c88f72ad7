// ov_gpu.hip — host side of libopenvort_gpu: the C ABI of
// include/openvort_gpu.h, device buffers, stream ordering and launches.
//
// One call = validate the caller's tangle, renumber its active points in
// stream order (loop by loop along forward links), upload SoA FP64 planes,
// run the kernels of ov_kernels.hip on the context's stream, scatter the
// results back into the caller's AoS arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/openvort_gpu.h"
#include "ov_device.h"
#include "ov_kernels.hip"
#include "ov_octree.hip"
#include "ov_batch.hip"
#include "ov_adjoint.hip"
#include "ov_reconnect.hip"
#include "ov_topology.hip"
#include "ov_walls.hip"
#include "ov_host.hip"

#include <hipcub/hipcub.hpp>

using namespace ov;

namespace {

struct KStat {
  long launches = 0;
  double ms = 0;
};

struct Pending {
  const char* name;
  hipEvent_t a, b;
};

}  // namespace

struct ov_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  char err[512] = {0};

  // stream-order maps of the current call
  int n = 0;              // active points
  long cap = 0;           // plane stride of the SoA buffers
  std::vector<int> act;   // stream index -> caller index
  std::vector<int> fwd_s, rev_s;
  BoxDev box{};
  std::vector<ImgDev> imgs;
  PhysDev ph{};
  int mixed = 0;
  int n_img = 1;           // terms of the image sum (primary + boundary images)
  int max_chunks = 256;    // partial-sum planes of the direct sum at most
  unsigned char* scratch_ptr = nullptr;   // this device's `scratch` (no-alloc entry point)
  double rnear = 0.02;
  int wall_periodic_mask = 0;
  bool lia_only = false;
  bool flow_time_dep = false;   // an imposed flow depends on t (no graph replay)

  // device buffers (planes of `cap` doubles)
  double *X0 = nullptr, *Xs = nullptr, *K[4] = {nullptr, nullptr, nullptr, nullptr};
  double *T = nullptr, *Nn = nullptr, *dxi = nullptr, *lia = nullptr, *bs = nullptr;
  double *vs = nullptr, *vel_tmp = nullptr;
  double *T1 = nullptr, *N1 = nullptr, *dxi1 = nullptr, *vs1 = nullptr;
  int *fwd = nullptr, *rev = nullptr, *flag = nullptr;
  int *brk = nullptr, *rank = nullptr, *bsum = nullptr, *counts = nullptr, *adj = nullptr;
  TgtTile* tgt = nullptr;
  SrcTile* src = nullptr;
  FarRec* far = nullptr;
  NearRec* near = nullptr;
  ImgDev* img = nullptr;
  double* partial = nullptr;
  size_t partial_bytes = 0;
  long buf_cap = 0;

  // pinned host staging
  double* hst = nullptr;
  size_t hst_bytes = 0;

  // profiling
  bool profile = false;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> ev_pool;
  std::map<std::string, KStat> stats;
  double last_interactions = 0;
  // octree
  bool use_tree = false;
  OctParams op{};
  std::vector<int> pos_by_index;   // ascending tangle index -> stream index
  int* t_order = nullptr;
  uint64_t *t_k[3] = {nullptr, nullptr, nullptr}, *t_ks[3] = {nullptr, nullptr, nullptr};
  uint64_t *t_ktmp = nullptr, *t_kout = nullptr;
  int *t_vals = nullptr, *t_vals2 = nullptr, *t_sorted = nullptr, *t_maxd = nullptr;
  LeafSeg* t_segs = nullptr;
  long t_pcap = 0;
  OctCell* t_cells = nullptr;
  int *t_nchild = nullptr, *t_offs = nullptr, *t_size = nullptr, *t_pre = nullptr;
  double* t_mom = nullptr;
  OctNode* t_nodes = nullptr;   // tree_cells + 1
  OctMom* t_nmom = nullptr;
  long t_ccap = 0;
  void* t_cub = nullptr;
  size_t t_cub_bytes = 0;
  long* t_stats = nullptr;
  // points pinned on walls (stream order): pin code per point, see ov_walls.hip
  bool has_pins = false;
  int n_seg = 0;                  // stream points with a forward link
  std::vector<int> pin_s;
  int* pin = nullptr;
  long pin_cap = 0;
  bool tree_order_device = false;   // t_order already set on the device (ov_gpu_run)
  int oct_images_per_walk = 9;   // most images per traversal wavefront (64 / 9 = 7 targets)
  int* t_big = nullptr;
  long t_big_cap = 0;
  double* t_big_part = nullptr;
  size_t t_big_part_bytes = 0;
  long tree_cells = 0;
  double tree_leaf_pairs = 0, tree_multipoles = 0;
  hipEvent_t t_begin = nullptr, t_end = nullptr;
  double last_device_ms = 0;

  // ensembles
  int b_K = 0, b_ntot = 0, b_rec_tot = 0, b_n_tgt = 0, b_mixed = 0;
  std::vector<TangleDev> b_tg;
  std::vector<std::vector<int>> b_act;
  TangleDev* b_dtg = nullptr;
  int *b_tid = nullptr, *b_tile_tangle = nullptr, *b_tgt_tangle = nullptr, *b_tM = nullptr,
      *b_tS = nullptr, *b_ucnt = nullptr, *b_urank = nullptr, *b_ubsum = nullptr,
      *b_ucounts = nullptr, *b_queue = nullptr, *b_pin = nullptr;
  bool b_pins = false;   // a member has points pinned on walls (pin codes in b_pin)
  long b_cap_pin = 0;
  ImgDev* b_img = nullptr;
  double* b_upart = nullptr;
  size_t b_upart_bytes = 0;
  long b_cap_tid = 0, b_cap_rec = 0, b_cap_tgt = 0, b_cap_img = 0, b_cap_K = 0;
  std::map<std::pair<long, int>, std::pair<int, int>> b_tuned;   // (ntot, K) -> (U, blocks)
  int b_U = 256, b_blocks = 1024;
  double b_tune_ms = 0;

  // adjoint (planes of `cap` doubles: gnb 15, the others 3)
  double *a_gnb = nullptr, *a_gu = nullptr, *a_gl = nullptr, *a_SB = nullptr, *a_W = nullptr,
         *a_G = nullptr, *a_Gtmp = nullptr, *a_gK[4] = {nullptr, nullptr, nullptr, nullptr};
  long a_cap = 0;
  double *a_tpart = nullptr, *a_spart = nullptr, *a_ck = nullptr, *a_store = nullptr;
  size_t a_fused_limit = (size_t)128 << 20;   // per-target-tile partials of the fused segment pass
  size_t a_tpart_bytes = 0, a_spart_bytes = 0, a_ck_bytes = 0, a_store_bytes = 0;

  // reconnection candidates
  uint64_t *r_keys = nullptr, *r_keys_sorted = nullptr;
  int *r_vals = nullptr, *r_order = nullptr, *r_act = nullptr, *r_count = nullptr;
  long r_cap = 0;
  unsigned long long *r_pairs = nullptr, *r_pairs_sorted = nullptr;
  long r_pair_cap = 0;

  // topology edits: the tangle in the caller's index space (capacity ix_N)
  int ix_N = 0;
  long ix_alloc = 0;
  double *ix_X = nullptr, *ix_T = nullptr, *ix_N3 = nullptr, *ix_dxi = nullptr, *ix_vels = nullptr,
         *ix_vs = nullptr;
  int *ix_fwd = nullptr, *ix_rev = nullptr, *ix_status = nullptr;
  int *ix_lab[2] = {nullptr, nullptr}, *ix_jmp[2] = {nullptr, nullptr}, *ix_d[2] = {nullptr, nullptr};
  int *ix_len = nullptr, *ix_off = nullptr, *ix_pos = nullptr, *ix_act = nullptr, *ix_bits = nullptr,
      *ix_all = nullptr, *ix_state = nullptr, *ix_pin = nullptr, *ix_chain = nullptr;
  bool ix_has_pins = false;        // points pinned on walls (open chains, pin codes in ix_pin)
  std::vector<int> ix_status_in;   // caller status on entry (new points get pin_wall NOT_A_FACE)
};

// ---- device scratch of the allocation-free entry point ------------------------
// ov_gpu_velocities_noalloc keeps every buffer beyond its inputs and outputs
// in one file-scope device array (each GPU the library runs on holds its own
// copy until the process ends).  Layout for a plane stride of `cap` points:
// LIA planes, the Biot-Savart sum, NOALLOC_CHUNKS partial-sum planes, the
// record stream and its tiles, scan words, adjacency, images and a flag.
namespace {
constexpr long NOALLOC_MAX_POINTS = 16384;
constexpr int NOALLOC_CHUNKS = 32;

struct NoAllocLayout {
  size_t lia, bs, partial, near, far, src, tgt, brk, rank, bsum, counts, adj, img, flag, total;
};
constexpr size_t al256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr NoAllocLayout noalloc_layout(long cap) {
  NoAllocLayout L{};
  const size_t max_rec = 2 * (size_t)cap + TILE;
  size_t o = 0;
  L.lia = o;     o = al256(o + sizeof(double) * 3 * (size_t)cap);
  L.bs = o;      o = al256(o + sizeof(double) * 3 * (size_t)cap);
  L.partial = o; o = al256(o + sizeof(double) * 3 * (size_t)cap * NOALLOC_CHUNKS);
  L.near = o;    o = al256(o + sizeof(NearRec) * max_rec);
  L.far = o;     o = al256(o + sizeof(FarRec) * max_rec);
  L.src = o;     o = al256(o + sizeof(SrcTile) * (max_rec / TILE));
  L.tgt = o;     o = al256(o + sizeof(TgtTile) * ((size_t)cap / TILE + 1));
  L.brk = o;     o = al256(o + sizeof(int) * (size_t)cap);
  L.rank = o;    o = al256(o + sizeof(int) * (size_t)cap);
  L.bsum = o;    o = al256(o + sizeof(int) * ((size_t)cap / 256 + 1));
  L.counts = o;  o = al256(o + sizeof(int) * 4);
  L.adj = o;     o = al256(o + sizeof(int) * 2 * (size_t)cap);
  L.img = o;     o = al256(o + sizeof(ImgDev) * MAX_IMAGES);
  L.flag = o;    o = al256(o + sizeof(int));
  L.total = o;
  return L;
}
constexpr size_t NOALLOC_SCRATCH_BYTES = noalloc_layout(NOALLOC_MAX_POINTS).total;
static_assert(NOALLOC_SCRATCH_BYTES < (size_t)20 << 20, "no-alloc scratch above 20 MiB");
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

int fail(ov_ctx* c, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(c->err, sizeof c->err, fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(ctx, OV_E_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_));   \
  } while (0)

hipEvent_t get_event(ov_ctx* ctx) {
  if (!ctx->ev_pool.empty()) {
    hipEvent_t e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Launch on the context stream, bracketed by HIP events when profiling.
template <typename K, typename... Args>
int launch(ov_ctx* ctx, const char* name, K kernel, dim3 grid, dim3 block, Args... args) {
  hipEvent_t a = nullptr, b = nullptr;
  if (ctx->profile) {
    a = get_event(ctx);
    b = get_event(ctx);
    if (a) (void)hipEventRecord(a, ctx->stream);
  }
  hipLaunchKernelGGL(kernel, grid, block, 0, ctx->stream, args...);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, OV_E_DEVICE, "launch %s: %s", name, hipGetErrorString(e));
  if (ctx->profile && a && b) {
    (void)hipEventRecord(b, ctx->stream);
    ctx->pending.push_back({name, a, b});
  }
  return OV_OK;
}

void collect_profile(ov_ctx* ctx) {
  for (auto& p : ctx->pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      KStat& s = ctx->stats[p.name];
      s.launches += 1;
      s.ms += ms;
    }
    ctx->ev_pool.push_back(p.a);
    ctx->ev_pool.push_back(p.b);
  }
  ctx->pending.clear();
}

int dalloc(ov_ctx* ctx, void** p, size_t bytes) {
  if (*p) hipFree(*p);
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(ctx, OV_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  return OV_OK;
}

// Grow every per-point buffer to hold `need` points (x2 growth).
int ensure_capacity(ov_ctx* ctx, long need) {
  if (need <= ctx->buf_cap) return OV_OK;
  long cap = std::max<long>(need, 2 * ctx->buf_cap);
  cap = ((cap + TILE - 1) / TILE) * TILE;
  size_t v3 = sizeof(double) * 3 * (size_t)cap;
  size_t v1 = sizeof(double) * (size_t)cap;
  int rc = 0;
  double** v3b[] = {&ctx->X0, &ctx->Xs, &ctx->K[0], &ctx->K[1], &ctx->K[2], &ctx->K[3],
                    &ctx->T,  &ctx->Nn, &ctx->lia,  &ctx->bs,   &ctx->vs,   &ctx->vel_tmp,
                    &ctx->T1, &ctx->N1, &ctx->vs1};
  for (double** b : v3b)
    if ((rc = dalloc(ctx, (void**)b, v3))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->dxi, v1))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->dxi1, v1))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->fwd, sizeof(int) * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->rev, sizeof(int) * cap))) return rc;
  // record stream: at most two records per point, padded to whole tiles
  const long max_rec = 2 * cap + TILE;
  if ((rc = dalloc(ctx, (void**)&ctx->brk, sizeof(int) * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->rank, sizeof(int) * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->bsum, sizeof(int) * (cap / 256 + 1)))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->counts, sizeof(int) * 4))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->adj, sizeof(int) * 2 * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->tgt, sizeof(TgtTile) * (cap / TILE)))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->src, sizeof(SrcTile) * (max_rec / TILE)))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->far, sizeof(FarRec) * max_rec))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->near, sizeof(NearRec) * max_rec))) return rc;
  ctx->buf_cap = cap;
  return OV_OK;
}

int ensure_host(ov_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->hst_bytes) return OV_OK;
  if (ctx->hst) hipHostFree(ctx->hst);
  ctx->hst = nullptr;
  size_t b = std::max(bytes, 2 * ctx->hst_bytes);
  hipError_t e = hipHostMalloc((void**)&ctx->hst, b, hipHostMallocDefault);
  if (e != hipSuccess) {
    ctx->hst_bytes = 0;
    return fail(ctx, OV_E_NOMEM, "hipHostMalloc(%zu): %s", b, hipGetErrorString(e));
  }
  ctx->hst_bytes = b;
  return OV_OK;
}

// Box, image list and physics of one call (no host or device allocation):
// the image terms go to il[0..*nil) (primary first), the rest into ctx.
int setup_box_physics(ov_ctx* ctx, const ov_domain_box* bx, const ov_boundary_images* bimg,
                      const ov_params* p, ImgDev* il, int* nil) {
  for (const ov_flow* f : {&p->vs, &p->vn, &p->vb})
    if (f->type < OV_FLOW_NONE || f->type >= OV_FLOW_COUNT)
      return fail(ctx, OV_E_INVALID, "unknown flow type %d", f->type);
  if (bimg->n < 0 || bimg->n > MAX_IMAGES - 1 || (bimg->n > 0 && !bimg->images))
    return fail(ctx, OV_E_INVALID, "bad image list (n=%d)", bimg->n);
  for (int i = 0; i < 3; ++i) {
    ctx->box.lo[i] = bx->bottom_left_back.p[i];
    ctx->box.hi[i] = bx->top_right_front.p[i];
    ctx->box.L[i] = ctx->box.hi[i] - ctx->box.lo[i];
    ctx->box.periodic[i] = bx->wall[2 * i] == OV_WALL_PERIODIC;
  }
  ctx->wall_periodic_mask = 0;
  for (int f = 0; f < 6; ++f)
    if (bx->wall[f] == OV_WALL_PERIODIC) ctx->wall_periodic_mask |= 1 << f;
  ImgDev prim{};
  prim.axis = -1;
  prim.primary = 1;
  il[0] = prim;
  int m = 1;
  for (int j = 0; j < bimg->n; ++j) {
    const ov_image_tangle& it = bimg->images[j];
    if (it.reflect < -1 || it.reflect > 5)
      return fail(ctx, OV_E_INVALID, "image %d: bad reflect %d", j, it.reflect);
    ImgDev d{};
    for (int i = 0; i < 3; ++i) d.shift[i] = ctx->box.L[i] * it.shift[i];
    d.axis = -1;
    if (it.reflect >= 0) {
      int ax = it.reflect / 2;
      double wall = (it.reflect & 1) ? ctx->box.hi[ax] : ctx->box.lo[ax];
      d.axis = ax;
      d.wall2 = 2 * wall;
    }
    d.primary = 0;
    il[m++] = d;
  }
  *nil = m;
  ctx->n_img = m;

  PhysDev& ph = ctx->ph;
  ph.kappa = p->kappa;
  ph.a0 = p->vortex_width;
  ph.alpha = p->alpha;
  ph.alpha_p = p->alpha_p;
  ph.dl_max = p->dl_max;
  ph.max_curv = p->max_curvature_scale;
  ph.hyperalpha = p->hyperalpha;
  ph.vs = p->vs;
  ph.vn = p->vn;
  ph.vb = p->vb;
  ctx->flow_time_dep = flow_time_dependent(p->vs.type) ||
                       (p->use_mutual_friction && flow_time_dependent(p->vn.type));
  ph.lia_only = p->LIA_only;
  ph.use_mf = p->use_mutual_friction;
  ph.hyperfriction = p->hyperfriction;
  ctx->lia_only = p->LIA_only != 0;
  ctx->mixed = p->precision == OV_PREC_MIXED;
  ctx->use_tree = p->use_BH != 0;
  ctx->op.box = ctx->box;
  ctx->op.grain = p->BH_grain;
  ctx->op.grain_dl = p->BH_grain * p->dl_max;
  ctx->op.theta = p->BH_resolution;
  ctx->op.quad = p->BH_quadtree != 0;
  ctx->op.dipole = p->multipole == OV_MP_DIPOLE;
  ctx->rnear = p->near_radius > 0 ? p->near_radius : 0.02;
  return OV_OK;
}

// Validate the tangle the way the reference's walks require it and build the
// stream order.  EMPTY points must have no forward link (the reference sums
// every point whose forward link is set, tangle.c:612-615, and every EMPTY
// point it creates has none, tangle.c:94-99, 1063-1067).
int prepare(ov_ctx* ctx, const ov_tangle_state* t, const ov_params* p, bool allow_pins = false) {
  int rc = 0;
  ctx->tree_order_device = false;
  if (!t || !p) return fail(ctx, OV_E_INVALID, "null tangle or params");
  if (t->N < 0) return fail(ctx, OV_E_INVALID, "negative N");
  if (t->N > 0 && (!t->vnodes || !t->connections || !t->status))
    return fail(ctx, OV_E_INVALID, "null tangle arrays");
  for (const ov_flow* f : {&p->vs, &p->vn, &p->vb})
    if (f->type < OV_FLOW_NONE || f->type >= OV_FLOW_COUNT)
      return fail(ctx, OV_E_INVALID, "unknown flow type %d", f->type);
  if (t->bimg.n < 0 || t->bimg.n > MAX_IMAGES - 1 || (t->bimg.n > 0 && !t->bimg.images))
    return fail(ctx, OV_E_INVALID, "bad image list (n=%d)", t->bimg.n);
  const int N = t->N;
  std::vector<int> cid(N, -1);
  ctx->act.clear();
  bool pins = false;
  auto pinned = [&](int k) {
    return t->status[k].status == OV_PINNED || t->status[k].status == OV_PINNED_SLIP;
  };
  for (int k = 0; k < N; ++k) {
    int st = t->status[k].status;
    if (st == OV_EMPTY) {
      if (t->connections[k].forward != -1)
        return fail(ctx, OV_E_INVALID, "EMPTY point %d has a forward link", k);
      continue;
    }
    if (st == OV_PINNED || st == OV_PINNED_SLIP) {
      if (!allow_pins)
        return fail(ctx, OV_E_UNSUPPORTED, "pinned point %d: walls are supported by ov_gpu_velocities, "
                    "ov_gpu_rk4_step/run, ov_gpu_rk4_run_batch, ov_gpu_update_tangents_normals, the "
                    "topology edits, ov_gpu_run and ov_gpu_reconnection_candidates only", k);
      if (t->status[k].pin_wall < 0 || t->status[k].pin_wall > 5)
        return fail(ctx, OV_E_INVALID, "pinned point %d has pin_wall %d", k, t->status[k].pin_wall);
      pins = true;
    } else if (st != OV_FREE) {
      return fail(ctx, OV_E_INVALID, "bad status %d at %d", st, k);
    }
    int f = t->connections[k].forward, r = t->connections[k].reverse;
    // only a pinned point may end an open chain (step_node, tangle.c:176-191)
    const bool pin_k = st != OV_FREE;
    if (f >= N || r >= N || (!pin_k && (f < 0 || r < 0)) || f < -1 || r < -1)
      return fail(ctx, OV_E_INVALID, "%s point %d has links %d/%d", pin_k ? "pinned" : "free", k, f, r);
    if (f < 0 && r < 0)
      return fail(ctx, OV_E_INVALID, "pinned point %d has no links (the reference's walk along it "
                  "does not end)", k);
    for (int nb : {f, r}) {
      if (nb < 0) continue;
      if (t->status[nb].status == OV_EMPTY)
        return fail(ctx, OV_E_INVALID, "point %d links to empty point %d", k, nb);
      if (pinned(nb) && !allow_pins)
        return fail(ctx, OV_E_UNSUPPORTED, "point %d neighbours pinned point %d", k, nb);
    }
    if (f >= 0 && t->connections[f].reverse != k)
      return fail(ctx, OV_E_INVALID, "links of %d and %d disagree", k, f);
    if (r >= 0 && t->connections[r].forward != k)
      return fail(ctx, OV_E_INVALID, "links of %d and %d disagree", r, k);
    cid[k] = (int)ctx->act.size();
    ctx->act.push_back(k);
  }
  if (pins) {
    // a point that is not PINNED but has a PINNED neighbour takes its
    // tangent from both neighbours' tangents (tangle.c:364-430): it needs
    // both links, and neither neighbour may need the same rule (the
    // reference's update_tangent_normal would recurse without end)
    auto hermite = [&](int k) {
      if (t->status[k].status == OV_PINNED) return false;
      int f = t->connections[k].forward, r = t->connections[k].reverse;
      return (f >= 0 && t->status[f].status == OV_PINNED) || (r >= 0 && t->status[r].status == OV_PINNED);
    };
    for (int k : ctx->act) {
      if (!hermite(k)) continue;
      int f = t->connections[k].forward, r = t->connections[k].reverse;
      if (f < 0 || r < 0)
        return fail(ctx, OV_E_UNSUPPORTED, "PINNED_SLIP chain end %d next to a PINNED point (the "
                    "reference reads the tangent of link -1)", k);
      if (hermite(f) || hermite(r))
        return fail(ctx, OV_E_INVALID, "points %d and %d between PINNED points: the reference's "
                    "update_tangent_normal recurses without end on two-point chains", k,
                    hermite(f) ? f : r);
    }
  }
  const int n = (int)ctx->act.size();
  // stream order: loops (and open chains) in order of their lowest caller
  // index; a loop walked forward from that index, a chain from its head
  std::vector<int> order;
  order.reserve(n);
  std::vector<char> seen(n, 0);
  for (int i = 0; i < n; ++i) {
    if (seen[i]) continue;
    int j = i;
    if (pins) {
      // back to the head of an open chain (or once around a loop)
      int steps = 0;
      for (int r = t->connections[ctx->act[j]].reverse; r >= 0 && cid[r] != i && steps <= n; ++steps) {
        j = cid[r];
        r = t->connections[ctx->act[j]].reverse;
      }
      if (t->connections[ctx->act[j]].reverse >= 0) j = i;   // closed loop
    }
    while (j >= 0 && !seen[j]) {
      seen[j] = 1;
      order.push_back(j);
      const int f = t->connections[ctx->act[j]].forward;
      j = f >= 0 ? cid[f] : -1;
    }
  }
  std::vector<int> pos(n);
  for (int q = 0; q < n; ++q) pos[order[q]] = q;
  ctx->pos_by_index = pos;   // compact ids are in ascending tangle index
  std::vector<int> act_s(n);
  ctx->fwd_s.assign(n, 0);
  ctx->rev_s.assign(n, 0);
  ctx->has_pins = pins;
  ctx->pin_s.assign(pins ? n : 0, -1);
  for (int q = 0; q < n; ++q) {
    int k = ctx->act[order[q]];
    act_s[q] = k;
    const int f = t->connections[k].forward, r = t->connections[k].reverse;
    ctx->fwd_s[q] = f >= 0 ? pos[cid[f]] : -1;
    ctx->rev_s[q] = r >= 0 ? pos[cid[r]] : -1;
    if (pins && pinned(k))
      ctx->pin_s[q] = t->status[k].pin_wall | (t->status[k].status == OV_PINNED_SLIP ? 8 : 0) |
                      (f > 0 ? 16 : 0);
  }
  ctx->act.swap(act_s);
  ctx->n = n;
  ctx->n_seg = n;
  for (int q = 0; q < n; ++q) ctx->n_seg -= ctx->fwd_s[q] < 0 ? 1 : 0;

  // box, images and physics
  ImgDev il[MAX_IMAGES];
  int nil = 0;
  if ((rc = setup_box_physics(ctx, &t->box, &t->bimg, p, il, &nil))) return rc;
  ctx->imgs.assign(il, il + nil);
  ctx->rnear = p->near_radius > 0 ? p->near_radius : 0.02;
  ctx->profile = false;

  rc = ensure_capacity(ctx, std::max(n, 1));
  if (rc) return rc;
  ctx->cap = ctx->buf_cap;
  if (ctx->imgs.size() > 0) {
    if (!ctx->img && (rc = dalloc(ctx, (void**)&ctx->img, sizeof(ImgDev) * MAX_IMAGES))) return rc;
  }
  return OV_OK;
}

// caller AoS vec3 array -> stream-order SoA planes on the device
int upload_vec3(ov_ctx* ctx, const ov_vec3d* src, double* dst) {
  const int n = ctx->n;
  const long cap = ctx->cap;
  int rc = ensure_host(ctx, sizeof(double) * 3 * (size_t)cap);
  if (rc) return rc;
  double* h = ctx->hst;
  for (int q = 0; q < n; ++q) {
    const double* v = src[ctx->act[q]].p;
    h[q] = v[0];
    h[cap + q] = v[1];
    h[2 * cap + q] = v[2];
  }
  HIPCHK(hipMemcpyAsync(dst, h, sizeof(double) * 3 * (size_t)cap, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return OV_OK;
}

int download_vec3(ov_ctx* ctx, const double* src, ov_vec3d* dst) {
  const int n = ctx->n;
  const long cap = ctx->cap;
  int rc = ensure_host(ctx, sizeof(double) * 3 * (size_t)cap);
  if (rc) return rc;
  double* h = ctx->hst;
  HIPCHK(hipMemcpyAsync(h, src, sizeof(double) * 3 * (size_t)cap, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int q = 0; q < n; ++q) {
    double* v = dst[ctx->act[q]].p;
    v[0] = h[q];
    v[1] = h[cap + q];
    v[2] = h[2 * cap + q];
  }
  return OV_OK;
}

int download_scalar(ov_ctx* ctx, const double* src, double* dst) {
  const int n = ctx->n;
  int rc = ensure_host(ctx, sizeof(double) * (size_t)ctx->cap);
  if (rc) return rc;
  double* h = ctx->hst;
  HIPCHK(hipMemcpyAsync(h, src, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int q = 0; q < n; ++q) dst[ctx->act[q]] = h[q];
  return OV_OK;
}

// As download_vec3 / download_scalar for a tangle placed at offset o of the planes.
int download_vec3_at(ov_ctx* ctx, const double* src, long o, ov_vec3d* dst) {
  const int n = ctx->n;
  const long cap = ctx->cap;
  double* h = ctx->hst;
  for (int i = 0; i < 3; ++i)
    HIPCHK(hipMemcpyAsync(h + i * (long)n, src + i * cap + o, sizeof(double) * n, hipMemcpyDeviceToHost,
                          ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int q = 0; q < n; ++q) {
    double* v = dst[ctx->act[q]].p;
    v[0] = h[q];
    v[1] = h[n + q];
    v[2] = h[2 * (long)n + q];
  }
  return OV_OK;
}

int download_scalar_at(ov_ctx* ctx, const double* src, long o, double* dst) {
  const int n = ctx->n;
  double* h = ctx->hst;
  HIPCHK(hipMemcpyAsync(h, src + o, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int q = 0; q < n; ++q) dst[ctx->act[q]] = h[q];
  return OV_OK;
}

int upload_topology(ov_ctx* ctx) {
  const int n = ctx->n;
  if (ctx->has_pins) {
    if (ctx->pin_cap < ctx->cap) {
      int rc = dalloc(ctx, (void**)&ctx->pin, sizeof(int) * ctx->cap);
      if (rc) return rc;
      ctx->pin_cap = ctx->cap;
    }
    HIPCHK(hipMemcpyAsync(ctx->pin, ctx->pin_s.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  }
  HIPCHK(hipMemcpyAsync(ctx->fwd, ctx->fwd_s.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->rev, ctx->rev_s.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->img, ctx->imgs.data(), sizeof(ImgDev) * ctx->imgs.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->flag, 0, sizeof(int), ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return OV_OK;
}

inline dim3 grid1(int n) { return dim3((unsigned)((n + 255) / 256)); }

// ---- octree -----------------------------------------------------------------

int tree_point_buffers(ov_ctx* ctx, long need) {
  if (need <= ctx->t_pcap) return OV_OK;
  long cap = std::max<long>(need, 2 * ctx->t_pcap);
  int rc;
  if ((rc = dalloc(ctx, (void**)&ctx->t_order, sizeof(int) * cap))) return rc;
  for (int w = 0; w < 3; ++w) {
    if ((rc = dalloc(ctx, (void**)&ctx->t_k[w], sizeof(uint64_t) * cap))) return rc;
    if ((rc = dalloc(ctx, (void**)&ctx->t_ks[w], sizeof(uint64_t) * cap))) return rc;
  }
  if ((rc = dalloc(ctx, (void**)&ctx->t_ktmp, sizeof(uint64_t) * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->t_kout, sizeof(uint64_t) * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->t_vals, sizeof(int) * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->t_vals2, sizeof(int) * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->t_sorted, sizeof(int) * cap))) return rc;
  // one entry past the points for the traversal's look-ahead load
  if ((rc = dalloc(ctx, (void**)&ctx->t_segs, sizeof(LeafSeg) * (cap + 1)))) return rc;
  if (!ctx->t_maxd && (rc = dalloc(ctx, (void**)&ctx->t_maxd, 256))) return rc;
  if (!ctx->t_stats && (rc = dalloc(ctx, (void**)&ctx->t_stats, 256))) return rc;
  ctx->t_pcap = cap;
  return OV_OK;
}

// Cell buffers for `need` cells; existing cells are kept.
int tree_cell_buffers(ov_ctx* ctx, long need, long keep) {
  if (need <= ctx->t_ccap) return OV_OK;
  long cap = std::max<long>(need, 2 * ctx->t_ccap);
  OctCell* nc = nullptr;
  if (hipMalloc((void**)&nc, sizeof(OctCell) * cap) != hipSuccess)
    return fail(ctx, OV_E_NOMEM, "hipMalloc(%ld cells)", cap);
  if (ctx->t_cells && keep > 0)
    HIPCHK(hipMemcpyAsync(nc, ctx->t_cells, sizeof(OctCell) * keep, hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (ctx->t_cells) hipFree(ctx->t_cells);
  ctx->t_cells = nc;
  int rc;
  if ((rc = dalloc(ctx, (void**)&ctx->t_nchild, sizeof(int) * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->t_offs, sizeof(int) * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->t_size, sizeof(int) * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->t_pre, sizeof(int) * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->t_mom, sizeof(double) * 15 * cap))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->t_nodes, sizeof(OctNode) * (cap + 1)))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->t_nmom, sizeof(OctMom) * cap))) return rc;
  ctx->t_ccap = cap;
  return OV_OK;
}

int grow_bytes_int(ov_ctx* ctx, int** p, long* cap, long need) {
  if (need <= *cap) return OV_OK;
  const long c = std::max<long>(need, 2 * *cap);
  int rc = dalloc(ctx, (void**)p, sizeof(int) * c);
  if (!rc) *cap = c;
  return rc;
}

int cub_temp(ov_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->t_cub_bytes) return OV_OK;
  int rc = dalloc(ctx, &ctx->t_cub, bytes);
  if (rc) return rc;
  ctx->t_cub_bytes = bytes;
  return OV_OK;
}

// Build the reference's octree of the points X (stream order) on the device.
int build_tree(ov_ctx* ctx, const double* X) {
  const int n = ctx->n;
  const long cap = ctx->cap;
  int rc;
  if ((rc = tree_point_buffers(ctx, n))) return rc;
  if (!ctx->tree_order_device)
    HIPCHK(hipMemcpyAsync(ctx->t_order, ctx->pos_by_index.data(), sizeof(int) * n,
                          hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->t_maxd, 0, sizeof(int), ctx->stream));
  const dim3 gp((unsigned)((n + 255) / 256));
  if ((rc = launch(ctx, "k_oct_keys", k_oct_keys, gp, dim3(256), X, cap, (const int*)ctx->t_order, n,
                   ctx->op, ctx->t_k[0], ctx->t_k[1], ctx->t_k[2], ctx->t_vals, ctx->t_maxd)))
    return rc;
  int maxd = 0;
  HIPCHK(hipMemcpyAsync(&maxd, ctx->t_maxd, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const int words = std::max(1, (maxd + 15) / 16);
  // stable LSD passes over the path words, least significant word first
  size_t need = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, ctx->t_ktmp, ctx->t_kout, ctx->t_vals,
                                            ctx->t_vals2, n, 0, 64, ctx->stream));
  if ((rc = cub_temp(ctx, need))) return rc;
  for (int w = words - 1; w >= 0; --w) {
    if (w == words - 1) {
      HIPCHK(hipMemcpyAsync(ctx->t_ktmp, ctx->t_k[w], sizeof(uint64_t) * n, hipMemcpyDeviceToDevice,
                            ctx->stream));
    } else if ((rc = launch(ctx, "k_gather_u64", k_gather_u64, gp, dim3(256),
                            (const uint64_t*)ctx->t_k[w], (const int*)ctx->t_vals, n, ctx->t_ktmp))) {
      return rc;
    }
    size_t tb = ctx->t_cub_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(ctx->t_cub, tb, ctx->t_ktmp, ctx->t_kout, ctx->t_vals,
                                              ctx->t_vals2, n, 0, 64, ctx->stream));
    std::swap(ctx->t_vals, ctx->t_vals2);
  }
  // sorted words and the sorted points' segments
  for (int w = 0; w < 3; ++w) {
    if (w < words) {
      if ((rc = launch(ctx, "k_gather_u64", k_gather_u64, gp, dim3(256), (const uint64_t*)ctx->t_k[w],
                       (const int*)ctx->t_vals, n, ctx->t_ks[w])))
        return rc;
    } else {
      HIPCHK(hipMemsetAsync(ctx->t_ks[w], 0, sizeof(uint64_t) * n, ctx->stream));
    }
  }
  if ((rc = launch(ctx, "k_oct_leafsegs", k_oct_leafsegs, gp, dim3(256), X, cap, (const int*)ctx->fwd,
                   (const int*)ctx->t_order, (const int*)ctx->t_vals, n, ctx->box, ctx->t_sorted,
                   ctx->t_segs)))
    return rc;
  // cells, level by level
  if ((rc = tree_cell_buffers(ctx, 4L * n + 64, 0))) return rc;
  OctCell root{};
  for (int i = 0; i < 3; ++i) {
    root.lo[i] = ctx->box.lo[i];
    root.hi[i] = ctx->box.hi[i];
  }
  root.start = 0;
  root.count = n;
  root.level = 0;
  root.parent = -1;
  root.child_first = -1;
  root.n_child = 0;
  HIPCHK(hipMemcpyAsync(ctx->t_cells, &root, sizeof(OctCell), hipMemcpyHostToDevice, ctx->stream));
  std::vector<std::pair<int, int>> levels;
  int ls = 0, le = 1;
  while (le > ls) {
    levels.push_back({ls, le});
    const int nl = le - ls;
    const dim3 gl((unsigned)((nl + 255) / 256));
    if ((rc = launch(ctx, "k_oct_count", k_oct_count, gl, dim3(256), (const OctCell*)ctx->t_cells, ls, le,
                     (const uint64_t*)ctx->t_ks[0], (const uint64_t*)ctx->t_ks[1],
                     (const uint64_t*)ctx->t_ks[2], ctx->op, ctx->t_nchild)))
      return rc;
    size_t sb = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, ctx->t_nchild, ctx->t_offs, nl, ctx->stream));
    if ((rc = cub_temp(ctx, sb))) return rc;
    sb = ctx->t_cub_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(ctx->t_cub, sb, ctx->t_nchild, ctx->t_offs, nl, ctx->stream));
    int tail[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&tail[0], ctx->t_offs + nl - 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(&tail[1], ctx->t_nchild + nl - 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int total = tail[0] + tail[1];
    if ((rc = tree_cell_buffers(ctx, (long)le + total + 1, le))) return rc;
    if ((rc = launch(ctx, "k_oct_emit", k_oct_emit, gl, dim3(256), ctx->t_cells, ls, le, le,
                     (const uint64_t*)ctx->t_ks[0], (const uint64_t*)ctx->t_ks[1],
                     (const uint64_t*)ctx->t_ks[2], ctx->op, (const int*)ctx->t_offs)))
      return rc;
    ls = le;
    le = le + total;
  }
  const int ncell = le;
  ctx->tree_cells = ncell;
  // cells of more than OCT_BIG points are listed and summed in slices
  if ((rc = grow_bytes_int(ctx, &ctx->t_big, &ctx->t_big_cap, std::max(ncell, 1)))) return rc;
  HIPCHK(hipMemsetAsync(ctx->t_maxd, 0, sizeof(int), ctx->stream));
  if ((rc = launch(ctx, "k_oct_moments", k_oct_moments, dim3((unsigned)((ncell + 3) / 4)), dim3(256),
                   (const OctCell*)ctx->t_cells, ncell, (const LeafSeg*)ctx->t_segs, ctx->t_mom,
                   ctx->t_big, ctx->t_maxd, ctx->op.dipole)))
    return rc;
  int n_big = 0;
  HIPCHK(hipMemcpyAsync(&n_big, ctx->t_maxd, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (n_big > 0) {
    const size_t pb = sizeof(double) * 15 * OCT_SLICES * (size_t)n_big;
    if (pb > ctx->t_big_part_bytes) {
      if ((rc = dalloc(ctx, (void**)&ctx->t_big_part, pb))) return rc;
      ctx->t_big_part_bytes = pb;
    }
    const dim3 gb(OCT_SLICES, (unsigned)n_big);
    if ((rc = launch(ctx, "k_oct_big_sum1", k_oct_big_sum1, gb, dim3(256), (const OctCell*)ctx->t_cells,
                     (const int*)ctx->t_big, (const int*)ctx->t_maxd, (const LeafSeg*)ctx->t_segs,
                     ctx->t_big_part)) ||
        (rc = launch(ctx, "k_oct_big_sum2", k_oct_big_sum2, gb, dim3(256), (const OctCell*)ctx->t_cells,
                     (const int*)ctx->t_big, (const int*)ctx->t_maxd, (const LeafSeg*)ctx->t_segs,
                     ctx->t_big_part, ctx->op.dipole)) ||
        (rc = launch(ctx, "k_oct_big_fin", k_oct_big_fin, dim3((unsigned)((n_big + 63) / 64)), dim3(64),
                     (const OctCell*)ctx->t_cells, (const int*)ctx->t_big, (const int*)ctx->t_maxd,
                     (const double*)ctx->t_big_part, ctx->t_mom)))
      return rc;
  }
  for (int L = (int)levels.size() - 1; L >= 0; --L) {
    const int a = levels[L].first, b = levels[L].second;
    if ((rc = launch(ctx, "k_oct_size", k_oct_size, dim3((unsigned)((b - a + 255) / 256)), dim3(256),
                     (const OctCell*)ctx->t_cells, a, b, ctx->t_size)))
      return rc;
  }
  HIPCHK(hipMemsetAsync(ctx->t_pre, 0, sizeof(int), ctx->stream));
  for (size_t L = 0; L < levels.size(); ++L) {
    const int a = levels[L].first, b = levels[L].second;
    if ((rc = launch(ctx, "k_oct_pre", k_oct_pre, dim3((unsigned)((b - a + 255) / 256)), dim3(256),
                     (const OctCell*)ctx->t_cells, a, b, (const int*)ctx->t_size, ctx->t_pre)))
      return rc;
  }
  if ((rc = launch(ctx, "k_oct_nodes", k_oct_nodes, dim3((unsigned)((ncell + 256) / 256)), dim3(256),
                   (const OctCell*)ctx->t_cells, ncell, (const double*)ctx->t_mom,
                   (const int*)ctx->t_size, (const int*)ctx->t_pre, ctx->t_nodes, ctx->t_nmom,
                   ctx->op.theta, ctx->mixed ? 1 : 0)))
    return rc;
  return OV_OK;
}

// Barnes-Hut sum over the tree of X for every target and image into ctx->bs.
int tree_sum(ov_ctx* ctx, const double* X) {
  const int n = ctx->n;
  const long cap = ctx->cap;
  int rc;
  if ((rc = build_tree(ctx, X))) return rc;
  const int n_img = (int)ctx->imgs.size();
  // images per wavefront: at most oct_images_per_walk, balanced over the
  // groups so that the last group's wavefronts are as full as the others
  const int cap_walk = std::min(std::max(ctx->oct_images_per_walk, 1), n_img);
  const int n_groups = (n_img + cap_walk - 1) / cap_walk;
  const int per_walk = (n_img + n_groups - 1) / n_groups;
  const int tpw = TILE / per_walk;                 // targets per wavefront
  size_t pbytes = sizeof(double) * 3 * (size_t)cap * (size_t)n_img;
  if (pbytes > ctx->partial_bytes) {
    if ((rc = dalloc(ctx, (void**)&ctx->partial, pbytes))) return rc;
    ctx->partial_bytes = pbytes;
  }
  HIPCHK(hipMemsetAsync(ctx->t_stats, 0, 3 * sizeof(long), ctx->stream));
  OctArgs A;
  A.P = X;
  A.cap = cap;
  A.n = n;
  A.nodes = ctx->t_nodes;
  A.nmom = ctx->t_nmom;
  A.n_nodes = (int)ctx->tree_cells;
  A.segs = ctx->t_segs;
  A.img = ctx->img;
  A.n_img = n_img;
  A.img_per_walk = per_walk;
  A.n_groups = n_groups;
  const long n_waves = (n + tpw - 1) / tpw;
  A.n_items = (int)((n_waves + OCT_WAVES - 1) / OCT_WAVES) * n_groups;
  A.per_xcd = (A.n_items + 7) / 8;
  A.op = ctx->op;
  A.kappa4pi = ctx->ph.kappa / 4 / 3.14159265358979323846;
  A.partial = ctx->partial;
  A.stats = ctx->profile ? ctx->t_stats : nullptr;
  const dim3 grid((unsigned)(8 * A.per_xcd));
  if (ctx->mixed)
    rc = launch(ctx, "k_octree_mixed", k_octree_mixed, grid, dim3(256), A);
  else
    rc = launch(ctx, "k_octree_fp64", k_octree_fp64, grid, dim3(256), A);
  if (rc) return rc;
  if ((rc = launch(ctx, "k_reduce", k_reduce, dim3(grid1(n).x, 3), dim3(256), (const double*)ctx->partial, n_img,
                   n, cap, ctx->bs)))
    return rc;
  if (ctx->profile) {
    long st[3] = {0, 0, 0};
    HIPCHK(hipMemcpyAsync(st, ctx->t_stats, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->tree_leaf_pairs += (double)st[0];
    ctx->tree_multipoles += (double)st[1];
  }
  return OV_OK;
}

// One velocity evaluation at positions X: geometry (or the caller's s', s''
// already in Tg/Ng), the direct sum, the epilogue.
// dxi, s', s'' (when geom) and the local induction at positions X: one pass
// for free points, three when points are pinned on walls (dxi, then every
// point but those next to a PINNED one, then those).
int geometry(ov_ctx* ctx, const double* X, bool geom, double* Tg, double* Ng, double* dxi) {
  const int n = ctx->n;
  const long cap = ctx->cap;
  int rc;
  if (!ctx->has_pins)
    return launch(ctx, "k_geometry", k_geometry, grid1(n), dim3(256), X, (const int*)ctx->fwd,
                  (const int*)ctx->rev, n, cap, ctx->box, ctx->ph, geom ? 1 : 0, Tg, Ng, dxi, ctx->lia);
  if (geom && (rc = launch(ctx, "k_wall_dxi", k_wall_dxi, grid1(n), dim3(256), X, (const int*)ctx->fwd,
                           (const int*)ctx->rev, (const int*)ctx->pin, n, cap, ctx->box, dxi, ctx->flag)))
    return rc;
  for (int phase = 1; phase <= (geom ? 2 : 1); ++phase)
    if ((rc = launch(ctx, "k_wall_geometry", k_wall_geometry, grid1(n), dim3(256), X, (const int*)ctx->fwd,
                     (const int*)ctx->rev, (const int*)ctx->pin, n, cap, ctx->box, ctx->ph, geom ? 1 : 0,
                     phase, Tg, Ng, (const double*)dxi, ctx->lia, ctx->flag)))
      return rc;
  return OV_OK;
}

int evaluate(ov_ctx* ctx, const double* X, bool geom, double* Tg, double* Ng, double* dxi,
             double* vs, double* vel, double time) {
  const int n = ctx->n;
  const long cap = ctx->cap;
  int rc;
  if (n == 0) return OV_OK;
  if ((rc = geometry(ctx, X, geom, Tg, Ng, dxi))) return rc;
  int have_bs = 0;
  if (!ctx->lia_only && ctx->use_tree) {
    if ((rc = tree_sum(ctx, X))) return rc;
    have_bs = 1;
    ctx->last_interactions = 0;
  } else if (!ctx->lia_only) {
    const int n_tgt = (n + TILE - 1) / TILE;
    const int nb = (n + 255) / 256;
    const int max_src = (2 * n + TILE - 1) / TILE;
    const int n_img = ctx->n_img;
    // also the target tiles' boxes and, for open chains, adjacency entries of -1
    if ((rc = launch(ctx, "k_cont", k_cont, dim3(nb), dim3(256), X, (const int*)ctx->fwd, n, cap,
                     ctx->box, ctx->brk, ctx->rank, ctx->bsum, ctx->tgt, ctx->has_pins ? 1 : 0, ctx->adj)))
      return rc;
    if ((rc = launch(ctx, "k_scan_blocks", k_scan_blocks, dim3(1), dim3(1024), ctx->bsum, nb, n,
                     ctx->counts)))
      return rc;
    if ((rc = launch(ctx, "k_records", k_records, dim3(nb), dim3(256), X, (const int*)ctx->fwd, n,
                     cap, ctx->box, (const int*)ctx->brk, (const int*)ctx->rank,
                     (const int*)ctx->bsum, ctx->near)))
      return rc;
    // also the adjacency of each target's two segments
    if ((rc = launch(ctx, "k_src_tiles", k_src_tiles, dim3(max_src), dim3(TILE), ctx->near,
                     (const int*)ctx->counts, ctx->src, ctx->far, ctx->adj)))
      return rc;
    // about 2^18 (target tile, chunk) waves, 64 rounds of the 4096 wave slots
    // of 256 CUs at 4 waves per SIMD, so that the last, partly filled round
    // is a small part of the launch (16 rounds: 1.4-1.8 % slower at 0.2M and
    // 1M points); no more chunks than work items per target tile, and at most
    // max_chunks (each chunk's partial plane costs 24 bytes per point to
    // write and to reduce)
    long chunks = (262144 + n_tgt - 1) / n_tgt;
    chunks = std::max<long>(1, std::min<long>(chunks, std::min<long>(ctx->max_chunks, (long)n_tgt * n_img)));
    size_t pbytes = sizeof(double) * 3 * (size_t)cap * (size_t)chunks;
    if (pbytes > ctx->partial_bytes) {
      if ((rc = dalloc(ctx, (void**)&ctx->partial, pbytes))) return rc;
      ctx->partial_bytes = pbytes;
    }
    DirectArgs A;
    A.P = X;
    A.cap = cap;
    A.n = n;
    A.n_tgt_tiles = n_tgt;
    A.tgt = ctx->tgt;
    A.src = ctx->src;
    A.counts = ctx->counts;
    A.far = ctx->far;
    A.near = ctx->near;
    A.img = ctx->img;
    A.adj = ctx->adj;
    A.n_img = n_img;
    A.rnear2 = ctx->rnear * ctx->rnear;
    A.kappa4pi = ctx->ph.kappa / 4 / 3.14159265358979323846;
    A.partial = ctx->partial;
    A.n_chunks = (int)chunks;
    const dim3 grid((unsigned)((n_tgt + 3) / 4), (unsigned)chunks);
    const bool small = n_tgt <= SMALL_TANGLE_TILES;
    if (ctx->mixed && small)
      rc = launch(ctx, "k_direct_mixed_small", k_direct_mixed_small, grid, dim3(256), A);
    else if (ctx->mixed)
      rc = launch(ctx, "k_direct_mixed", k_direct_mixed, grid, dim3(256), A);
    else if (small)
      rc = launch(ctx, "k_direct_fp64_small", k_direct_fp64_small, grid, dim3(256), A);
    else
      rc = launch(ctx, "k_direct_fp64", k_direct_fp64, grid, dim3(256), A);
    if (rc) return rc;
    if ((rc = launch(ctx, "k_reduce", k_reduce, dim3(grid1(n).x, 3), dim3(256), (const double*)ctx->partial,
                     (int)chunks, n, cap, ctx->bs)))
      return rc;
    have_bs = 1;
    // N_tgt (N_seg - 2) + n_img N_tgt N_seg with N_seg the points that start a segment
    const double nd = (double)n, ns = (double)ctx->n_seg;
    ctx->last_interactions = nd * (ns - 2) + (double)(n_img - 1) * nd * ns;
  } else {
    ctx->last_interactions = 0;
  }
  if ((rc = launch(ctx, "k_epilogue", k_epilogue, grid1(n), dim3(256), X, (const double*)Tg,
                   (const double*)Ng, (const double*)ctx->lia, (const double*)ctx->bs, have_bs, n,
                   cap, time, ctx->ph,
                   ctx->has_pins ? (const int*)ctx->pin : (const int*)nullptr, vs, vel, ctx->flag)))
    return rc;
  return OV_OK;
}

// Copy of one vec3 plane set (the n points of 3 planes of stride cap) on the
// context stream.
int copy_planes(ov_ctx* ctx, double* dst, const double* src) {
  if (ctx->n == 0) return OV_OK;
  return launch(ctx, "k_copy3", k_copy3, grid1(ctx->n), dim3(256), src, dst, ctx->n, ctx->cap);
}

// update_tangle(t) + rk4_step(t, dt) on the device-resident state in X0;
// the evaluations run at t, t + dt/2, t + dt/2, t + dt (vortex_dynamics.c:79-92).
// With caller_k1, K[0] already holds k1 and the t-geometry is not produced.
// With `store` (8 plane sets), the step's start positions, k1..k3 and the
// Biot-Savart sums of the four evaluations are kept there, in the order
// X0, K0, K1, K2, BS0, BS1, BS2, BS3 (what the step's VJP needs).
int rk4_device(ov_ctx* ctx, double t, double dt, int rk_variant, bool caller_k1,
               double* store = nullptr) {
  const int n = ctx->n;
  const long cap = ctx->cap;
  const long ps = 3 * cap;
  int rc;
  if (store && (rc = copy_planes(ctx, store, ctx->X0))) return rc;
  if (!caller_k1) {
    if ((rc = evaluate(ctx, ctx->X0, true, ctx->T1, ctx->N1, ctx->dxi1, ctx->vs1, ctx->K[0], t)))
      return rc;
    if (store && (rc = copy_planes(ctx, store + 4 * ps, ctx->bs))) return rc;
  }
  const double h[3] = {dt / 2, dt / 2, rk_variant == OV_RK_CLASSIC ? dt : dt / 2};
  const double ts[3] = {t + dt / 2, t + dt / 2, t + dt};
  for (int st = 0; st < 3; ++st) {
    if ((rc = launch(ctx, "k_stage", k_stage, grid1(n), dim3(256), (const double*)ctx->X0,
                     (const double*)ctx->K[st], h[st], n, cap, ctx->Xs)))
      return rc;
    if ((rc = evaluate(ctx, ctx->Xs, true, ctx->T, ctx->Nn, ctx->dxi, ctx->vs, ctx->K[st + 1], ts[st])))
      return rc;
    if (store && (rc = copy_planes(ctx, store + (5 + st) * ps, ctx->bs))) return rc;
  }
  if (store) {
    for (int st = 0; st < 3; ++st)
      if ((rc = copy_planes(ctx, store + (1 + st) * ps, ctx->K[st]))) return rc;
  }
  return launch(ctx, "k_combine", k_combine, grid1(n), dim3(256), ctx->X0, (const double*)ctx->K[0],
                (const double*)ctx->K[1], (const double*)ctx->K[2], (const double*)ctx->K[3], dt, n,
                cap, ctx->flag);
}

int mark_begin_stream(ov_ctx* ctx, hipEvent_t ev) {
  HIPCHK(hipEventRecord(ev, ctx->stream));
  return OV_OK;
}

int mark_begin(ov_ctx* ctx) {
  HIPCHK(hipEventRecord(ctx->t_begin, ctx->stream));
  return OV_OK;
}

// Device time since mark_begin, into last_device_ms (no flag read).
int mark_end(ov_ctx* ctx) {
  HIPCHK(hipEventRecord(ctx->t_end, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  float ms = 0;
  if (hipEventElapsedTime(&ms, ctx->t_begin, ctx->t_end) == hipSuccess) ctx->last_device_ms = ms;
  return OV_OK;
}

int finish(ov_ctx* ctx, int* flag_out) {
  HIPCHK(hipEventRecord(ctx->t_end, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  float ms = 0;
  if (hipEventElapsedTime(&ms, ctx->t_begin, ctx->t_end) == hipSuccess) ctx->last_device_ms = ms;
  if (ctx->profile) collect_profile(ctx);
  int f = 0;
  HIPCHK(hipMemcpy(&f, ctx->flag, sizeof(int), hipMemcpyDeviceToHost));
  *flag_out = f;
  return OV_OK;
}

// Capture the launches and device copies `body` issues on the context stream
// as a graph and instantiate it (no memsets and no allocations may happen in
// body: callers run one step eagerly first, which sizes every buffer).
template <typename F>
int capture_graph(ov_ctx* ctx, F body, hipGraphExec_t* exec) {
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = body();
  hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
  if (rc) {
    if (graph) hipGraphDestroy(graph);
    return rc;
  }
  if (ec != hipSuccess) return fail(ctx, OV_E_DEVICE, "graph capture: %s", hipGetErrorString(ec));
  ec = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ec != hipSuccess) return fail(ctx, OV_E_DEVICE, "graph instantiate: %s", hipGetErrorString(ec));
  return OV_OK;
}

int launch_graph(ov_ctx* ctx, hipGraphExec_t exec, int times) {
  for (int i = 0; i < times; ++i) {
    hipError_t ec = hipGraphLaunch(exec, ctx->stream);
    if (ec != hipSuccess) return fail(ctx, OV_E_DEVICE, "graph launch: %s", hipGetErrorString(ec));
  }
  return OV_OK;
}

// ---- ensembles ----------------------------------------------------------------

int grow_int(ov_ctx* ctx, int** p, long* cap, long need) {
  if (need <= *cap) return OV_OK;
  long c = std::max<long>(need, 2 * *cap);
  int rc = dalloc(ctx, (void**)p, sizeof(int) * c);
  if (!rc) *cap = c;
  return rc;
}

// Upper bound of the work units of one evaluation at unit size U (a tangle
// of n points has at most ceil(2n / 64) source tiles).
long batch_unit_bound(const ov_ctx* ctx, int U) {
  long units = 0;
  for (const TangleDev& t : ctx->b_tg) {
    const long items = (long)((2 * t.n + TILE - 1) / TILE) * t.n_img;
    units += (long)t.n_tgt * ((items + U - 1) / U);
  }
  return units;
}

// One velocity evaluation of every tangle of the batch at positions X.
int evaluate_batch(ov_ctx* ctx, const double* X, double* Tg, double* Ng, double* dxi, double* vs,
                   double* vel, float* direct_ms, double time) {
  const long cap = ctx->cap;
  const int ntot = ctx->b_ntot, rec_tot = ctx->b_rec_tot, n_tgt = ctx->b_n_tgt;
  const int nb = (ntot + 255) / 256, nbu = (n_tgt + 255) / 256;
  int rc;
  if (!ctx->b_pins) {
    if ((rc = launch(ctx, "kb_geometry", kb_geometry, dim3(nb), dim3(256), X, (const int*)ctx->fwd,
                     (const int*)ctx->rev, cap, (const int*)ctx->b_tid, (const TangleDev*)ctx->b_dtg, 1, Tg,
                     Ng, dxi, ctx->lia, ntot)))
      return rc;
  } else {
    for (int phase = 0; phase <= 2; ++phase)
      if ((rc = launch(ctx, "kb_wall_geometry", kb_wall_geometry, dim3(nb), dim3(256), X,
                       (const int*)ctx->fwd, (const int*)ctx->rev, (const int*)ctx->b_pin, cap,
                       (const int*)ctx->b_tid, (const TangleDev*)ctx->b_dtg, 1, phase, Tg, Ng, dxi,
                       ctx->lia, ctx->flag, ntot)))
        return rc;
    // a chain end lacks one adjacent segment: its entry stays -1
    if ((rc = launch(ctx, "k_fill_int", k_fill_int, grid1(2 * ntot), dim3(256), ctx->adj, 2 * ntot, -1)))
      return rc;
  }
  if ((rc = launch(ctx, "kb_cont", kb_cont, dim3(nb), dim3(256), X, (const int*)ctx->fwd, cap,
                   (const int*)ctx->b_tid, (const TangleDev*)ctx->b_dtg, ctx->brk, ctx->rank, ctx->bsum,
                   ntot)))
    return rc;
  if ((rc = launch(ctx, "k_scan_blocks", k_scan_blocks, dim3(1), dim3(1024), ctx->bsum, nb, 0, ctx->counts)))
    return rc;
  if ((rc = launch(ctx, "kb_records", kb_records, dim3(nb), dim3(256), X, (const int*)ctx->fwd, cap,
                   (const int*)ctx->b_tid, (const TangleDev*)ctx->b_dtg, (const int*)ctx->brk,
                   (const int*)ctx->rank, (const int*)ctx->bsum, ctx->near, ctx->b_tM, ctx->b_tS, ntot)))
    return rc;
  if ((rc = launch(ctx, "kb_src_tiles", kb_src_tiles, dim3(rec_tot / TILE), dim3(TILE), ctx->near,
                   (const int*)ctx->b_tile_tangle, (const TangleDev*)ctx->b_dtg, (const int*)ctx->b_tM,
                   (const int*)ctx->b_tS, ctx->src, ctx->far)))
    return rc;
  if ((rc = launch(ctx, "kb_tgt_tiles", kb_tgt_tiles, dim3(n_tgt), dim3(TILE), X, cap,
                   (const int*)ctx->b_tgt_tangle, (const TangleDev*)ctx->b_dtg, ctx->tgt)))
    return rc;
  if ((rc = launch(ctx, "kb_adjacent", kb_adjacent, dim3((unsigned)((rec_tot + 255) / 256)), dim3(256),
                   (const NearRec*)ctx->near, (const int*)ctx->b_tile_tangle, (const TangleDev*)ctx->b_dtg,
                   (const int*)ctx->b_tM, ctx->adj, rec_tot)))
    return rc;
  if ((rc = launch(ctx, "kb_units", kb_units, dim3(nbu), dim3(256), (const int*)ctx->b_tgt_tangle,
                   (const TangleDev*)ctx->b_dtg, (const int*)ctx->b_tS, ctx->b_U, ctx->b_ucnt, ctx->b_urank,
                   ctx->b_ubsum, n_tgt)))
    return rc;
  if ((rc = launch(ctx, "k_scan_blocks", k_scan_blocks, dim3(1), dim3(1024), ctx->b_ubsum, nbu, 0,
                   ctx->b_ucounts)))
    return rc;
  if ((rc = launch(ctx, "kb_reset", kb_reset, dim3(1), dim3(64), ctx->b_queue))) return rc;
  BatchArgs A;
  A.P = X;
  A.cap = cap;
  A.tg = ctx->b_dtg;
  A.tgt_tangle = ctx->b_tgt_tangle;
  A.n_tgt = n_tgt;
  A.tgt = ctx->tgt;
  A.src = ctx->src;
  A.far = ctx->far;
  A.near = ctx->near;
  A.img = ctx->b_img;
  A.adj = ctx->adj;
  A.tS = ctx->b_tS;
  A.urank = ctx->b_urank;
  A.ubsum = ctx->b_ubsum;
  A.ucounts = ctx->b_ucounts;
  A.queue = ctx->b_queue;
  A.U = ctx->b_U;
  A.rnear2 = ctx->rnear * ctx->rnear;
  A.upart = ctx->b_upart;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (direct_ms) {
    e0 = get_event(ctx);
    e1 = get_event(ctx);
    HIPCHK(hipEventRecord(e0, ctx->stream));
  }
  if (ctx->b_blocks > 0) {   // persistent grid over the unit queue
    if (ctx->b_mixed)
      rc = launch(ctx, "kb_direct_mixed", kb_direct_mixed, dim3(ctx->b_blocks), dim3(256), A);
    else
      rc = launch(ctx, "kb_direct_fp64", kb_direct_fp64, dim3(ctx->b_blocks), dim3(256), A);
  } else {                   // one unit per wavefront
    const unsigned g = (unsigned)((batch_unit_bound(ctx, ctx->b_U) + 3) / 4);
    if (ctx->b_mixed)
      rc = launch(ctx, "kb_grid_mixed", kb_grid_mixed, dim3(std::max(g, 1u)), dim3(256), A);
    else
      rc = launch(ctx, "kb_grid_fp64", kb_grid_fp64, dim3(std::max(g, 1u)), dim3(256), A);
  }
  if (rc) return rc;
  if (direct_ms) {
    HIPCHK(hipEventRecord(e1, ctx->stream));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(direct_ms, e0, e1));
    ctx->ev_pool.push_back(e0);
    ctx->ev_pool.push_back(e1);
  }
  if ((rc = launch(ctx, "kb_reduce", kb_reduce, dim3(nb), dim3(256), (const double*)ctx->b_upart,
                   (const int*)ctx->b_ucnt, (const int*)ctx->b_urank, (const int*)ctx->b_ubsum,
                   (const int*)ctx->b_tid, cap, ctx->bs, ntot)))
    return rc;
  return launch(ctx, "kb_epilogue", kb_epilogue, dim3(nb), dim3(256), X, (const double*)Tg,
                (const double*)Ng, (const double*)ctx->lia, (const double*)ctx->bs, cap, time,
                (const int*)ctx->b_tid, (const TangleDev*)ctx->b_dtg,
                ctx->b_pins ? (const int*)ctx->b_pin : (const int*)nullptr, vs, vel, ctx->flag, ntot);
}

// update_tangle + rk4_step of every tangle of the batch (X0 device-resident).
int rk4_batch(ov_ctx* ctx, double t, double dt, int rk_variant, bool enforce) {
  const long cap = ctx->cap;
  const int ntot = ctx->b_ntot;
  int rc;
  if ((rc = evaluate_batch(ctx, ctx->X0, ctx->T1, ctx->N1, ctx->dxi1, ctx->vs1, ctx->K[0], nullptr, t)))
    return rc;
  const double h[3] = {dt / 2, dt / 2, rk_variant == OV_RK_CLASSIC ? dt : dt / 2};
  const double ts[3] = {t + dt / 2, t + dt / 2, t + dt};
  for (int st = 0; st < 3; ++st) {
    if ((rc = launch(ctx, "k_stage", k_stage, grid1(ntot), dim3(256), (const double*)ctx->X0,
                     (const double*)ctx->K[st], h[st], ntot, cap, ctx->Xs)))
      return rc;
    if ((rc = evaluate_batch(ctx, ctx->Xs, ctx->T, ctx->Nn, ctx->dxi, ctx->vs, ctx->K[st + 1], nullptr,
                             ts[st])))
      return rc;
  }
  if ((rc = launch(ctx, "k_combine", k_combine, grid1(ntot), dim3(256), ctx->X0, (const double*)ctx->K[0],
                   (const double*)ctx->K[1], (const double*)ctx->K[2], (const double*)ctx->K[3], dt, ntot,
                   cap, ctx->flag)))
    return rc;
  if (enforce)
    return launch(ctx, "kb_wrap", kb_wrap, grid1(ntot), dim3(256), ctx->X0, cap, (const int*)ctx->b_tid,
                  (const TangleDev*)ctx->b_dtg, ntot);
  return OV_OK;
}


// ---- adjoint ------------------------------------------------------------------

int grow_bytes(ov_ctx* ctx, double** p, size_t* have, size_t need) {
  if (need <= *have) return OV_OK;
  int rc = dalloc(ctx, (void**)p, need);
  *have = rc ? 0 : need;
  return rc;
}

int adj_buffers(ov_ctx* ctx) {
  if (ctx->a_cap >= ctx->cap) return OV_OK;
  const size_t v3 = sizeof(double) * 3 * (size_t)ctx->cap;
  int rc;
  if ((rc = dalloc(ctx, (void**)&ctx->a_gnb, 5 * v3))) return rc;
  double** b[] = {&ctx->a_gu, &ctx->a_gl,    &ctx->a_SB,    &ctx->a_W,     &ctx->a_G,
                  &ctx->a_Gtmp, &ctx->a_gK[0], &ctx->a_gK[1], &ctx->a_gK[2], &ctx->a_gK[3]};
  for (double** p : b)
    if ((rc = dalloc(ctx, (void**)p, v3))) return rc;
  ctx->a_cap = ctx->cap;
  return OV_OK;
}

// G = (dvels/dX)^T W at positions X whose Biot-Savart sums are `bs`, time t
// (the imposed flows').
int velocity_vjp_dev(ov_ctx* ctx, const double* X, const double* bs, const double* W, double t,
                     double* G) {
  const int n = ctx->n;
  const long cap = ctx->cap;
  const int have_bs = ctx->lia_only ? 0 : 1;
  int rc;
  if ((rc = launch(ctx, "ka_local", ka_local, dim3((unsigned)((n + 127) / 128), 6), dim3(128), X,
                   (const int*)ctx->fwd, (const int*)ctx->rev, bs, have_bs, W, n, cap, ctx->box,
                   ctx->ph, t, ctx->a_gnb, ctx->a_gu)))
    return rc;
  if ((rc = launch(ctx, "ka_gather", ka_gather, grid1(n), dim3(256), (const double*)ctx->a_gnb,
                   (const int*)ctx->fwd, (const int*)ctx->rev, n, cap, ctx->a_gl)))
    return rc;
  if (!have_bs)
    return launch(ctx, "ka_scale", ka_scale, grid1(n), dim3(256), (const double*)ctx->a_gl, 1.0, G,
                  n, cap);
  if ((rc = launch(ctx, "ka_seg_end", ka_seg_end, grid1(n), dim3(256), X, (const int*)ctx->fwd, n,
                   cap, ctx->box, ctx->a_SB)))
    return rc;
  const int n_tiles = (n + TILE - 1) / TILE;
  // about four waves per SIMD of 256 CUs, at least 64 segments (targets) per chunk
  int chunks = (4096 + n_tiles - 1) / n_tiles;
  chunks = std::max(1, std::min(chunks, std::min(256, (n + 63) / 64)));
  // partial planes of n points (rounded up to 64), whatever the buffers' capacity;
  // tangles of few target tiles: one fused pass, segment terms per target tile
  const long ps = (long)n_tiles * TILE;
  const bool fused = (size_t)6 * ps * n_tiles * sizeof(double) <= ctx->a_fused_limit;
  const int s_terms = fused ? n_tiles : chunks;
  if ((rc = grow_bytes(ctx, &ctx->a_tpart, &ctx->a_tpart_bytes, sizeof(double) * 3 * ps * chunks)) ||
      (rc = grow_bytes(ctx, &ctx->a_spart, &ctx->a_spart_bytes, sizeof(double) * 6 * ps * s_terms)))
    return rc;
  AdjArgs A;
  A.P = X;
  A.SB = ctx->a_SB;
  A.fwd = ctx->fwd;
  A.gu = ctx->a_gu;
  A.img = ctx->img;
  A.cap = cap;
  A.ps = ps;
  A.n = n;
  A.n_tiles = n_tiles;
  A.n_img = (int)ctx->imgs.size();
  A.n_chunks = chunks;
  const dim3 grid((unsigned)((n_tiles + 3) / 4), (unsigned)chunks);
  A.part = ctx->a_tpart;
  if (fused) {
    if ((rc = launch(ctx, "ka_bs_fused", ka_bs_fused, grid, dim3(256), A, ctx->a_spart))) return rc;
  } else {
    if ((rc = launch(ctx, "ka_bs_targets", ka_bs_targets, grid, dim3(256), A))) return rc;
    A.part = ctx->a_spart;
    if ((rc = launch(ctx, "ka_bs_sources", ka_bs_sources, grid, dim3(256), A))) return rc;
  }
  if ((rc = launch(ctx, "ka_bs_reduce", ka_bs_reduce, dim3(grid1(n).x, 9), dim3(256),
                   (const double*)ctx->a_tpart, chunks, (const double*)ctx->a_spart, s_terms, n, ps,
                   cap, ctx->a_gnb)))
    return rc;
  return launch(ctx, "ka_bs_finish", ka_bs_finish, grid1(n), dim3(256), (const double*)ctx->a_gl,
                (const double*)ctx->a_gnb, (const int*)ctx->rev, n, cap,
                ctx->ph.kappa / 4 / 3.14159265358979323846, G);
}

// On entry G = cotangent of the step's end positions; on return the cotangent
// of its start positions.  `store` is what rk4_device kept for the step that
// starts at time t (stages at t, t + dt/2, t + dt/2, t + dt, as rk4_device).
int step_vjp_dev(ov_ctx* ctx, const double* store, double t, double dt, int rk_variant, double* G) {
  const int n = ctx->n;
  const long cap = ctx->cap;
  const long ps = 3 * cap;
  const double h[3] = {dt / 2, dt / 2, rk_variant == OV_RK_CLASSIC ? dt : dt / 2};
  const double ts[3] = {t + dt / 2, t + dt / 2, t + dt};
  const double wk[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  int rc;
  for (int i = 0; i < 4; ++i)
    if ((rc = launch(ctx, "ka_scale", ka_scale, grid1(n), dim3(256), (const double*)G, wk[i],
                     ctx->a_gK[i], n, cap)))
      return rc;
  // stage st + 1 was evaluated at X0 + h[st] k_{st+1}
  for (int st = 2; st >= -1; --st) {
    const double* Xs = store;
    if (st >= 0) {
      if ((rc = launch(ctx, "k_stage", k_stage, grid1(n), dim3(256), store, store + (1 + st) * ps,
                       h[st], n, cap, ctx->Xs)))
        return rc;
      Xs = ctx->Xs;
    }
    if ((rc = velocity_vjp_dev(ctx, Xs, store + (5 + st) * ps, ctx->a_gK[st + 1], st >= 0 ? ts[st] : t,
                               ctx->a_Gtmp)))
      return rc;
    if ((rc = launch(ctx, "ka_axpy", ka_axpy, grid1(n), dim3(256), (const double*)G, 1.0,
                     (const double*)ctx->a_Gtmp, G, n, cap)))
      return rc;
    if (st >= 0 && (rc = launch(ctx, "ka_axpy", ka_axpy, grid1(n), dim3(256),
                                (const double*)ctx->a_gK[st], h[st], (const double*)ctx->a_Gtmp,
                                ctx->a_gK[st], n, cap)))
      return rc;
  }
  return OV_OK;
}

// Common checks of the adjoint entries.
int adj_prepare(ov_ctx* ctx, const ov_tangle_state* t, const ov_params* p) {
  int rc = prepare(ctx, t, p);
  if (rc) return rc;
  if (p->use_BH)
    return fail(ctx, OV_E_UNSUPPORTED, "the adjoint differentiates the direct sum only (use_BH=0)");
  if (ctx->n == 0) return OV_OK;
  if ((rc = adj_buffers(ctx))) return rc;
  return upload_topology(ctx);
}

// Fill every caller entry with `pass` (or 0), then scatter the active points'
// gradient over it.
int download_grad(ov_ctx* ctx, const ov_tangle_state* t, const double* G, const ov_vec3d* pass,
                  ov_vec3d* grad) {
  for (int k = 0; k < t->N; ++k) grad[k] = pass ? pass[k] : ov_vec3d{{0, 0, 0}};
  return download_vec3(ctx, G, grad);
}


// ---- reconnection candidates ---------------------------------------------------

int rec_buffers(ov_ctx* ctx, long n) {
  int rc;
  if (ctx->r_cap < ctx->cap) {
    const long c = ctx->cap;
    if ((rc = dalloc(ctx, (void**)&ctx->r_keys, sizeof(uint64_t) * c)) ||
        (rc = dalloc(ctx, (void**)&ctx->r_keys_sorted, sizeof(uint64_t) * c)) ||
        (rc = dalloc(ctx, (void**)&ctx->r_vals, sizeof(int) * c)) ||
        (rc = dalloc(ctx, (void**)&ctx->r_order, sizeof(int) * c)) ||
        (rc = dalloc(ctx, (void**)&ctx->r_act, sizeof(int) * c)))
      return rc;
    ctx->r_cap = c;
  }
  if (!ctx->r_count && (rc = dalloc(ctx, (void**)&ctx->r_count, sizeof(int)))) return rc;
  (void)n;
  return OV_OK;
}

int rec_pair_buffers(ov_ctx* ctx, long need) {
  if (need <= ctx->r_pair_cap) return OV_OK;
  const long c = std::max<long>(need, 2 * ctx->r_pair_cap);
  int rc;
  if ((rc = dalloc(ctx, (void**)&ctx->r_pairs, sizeof(unsigned long long) * c)) ||
      (rc = dalloc(ctx, (void**)&ctx->r_pairs_sorted, sizeof(unsigned long long) * c)))
    return rc;
  ctx->r_pair_cap = c;
  return OV_OK;
}

// Device address of `scratch` per device, looked up once per device with
// that device current (hipGetSymbolAddress); the array's name is only a
// host-side stand-in and is never passed to a kernel or a copy.
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {nullptr};

void* scratch_address(int device) {
  if (device < 0 || device >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

// ---- topology edits in the caller's index space --------------------------------

int ix_buffers(ov_ctx* ctx, int N) {
  if (N <= ctx->ix_alloc) return OV_OK;
  const long c = std::max<long>(N, 64);
  const size_t v3 = sizeof(double) * 3 * (size_t)c, v1 = sizeof(double) * (size_t)c, i1 = sizeof(int) * (size_t)c;
  int rc;
  double** d3[] = {&ctx->ix_X, &ctx->ix_T, &ctx->ix_N3, &ctx->ix_vels, &ctx->ix_vs};
  for (double** b : d3)
    if ((rc = dalloc(ctx, (void**)b, v3))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->ix_dxi, v1))) return rc;
  int** i1b[] = {&ctx->ix_fwd, &ctx->ix_rev, &ctx->ix_status, &ctx->ix_lab[0], &ctx->ix_lab[1],
                 &ctx->ix_jmp[0], &ctx->ix_jmp[1], &ctx->ix_d[0], &ctx->ix_d[1], &ctx->ix_len,
                 &ctx->ix_off, &ctx->ix_pos, &ctx->ix_act, &ctx->ix_bits, &ctx->ix_all,
                 &ctx->ix_pin, &ctx->ix_chain};
  for (int** b : i1b)
    if ((rc = dalloc(ctx, (void**)b, i1))) return rc;
  if ((rc = dalloc(ctx, (void**)&ctx->ix_state, sizeof(int) * 8))) return rc;
  ctx->ix_alloc = c;
  return OV_OK;
}

// Validate a tangle for the index-space edits and set box/physics: free
// points on closed loops with consistent links, or, with points pinned on
// walls, the open chains and pins that the velocity path takes (validated
// by prepare()).
int ix_prepare(ov_ctx* ctx, const ov_tangle_state* t, const ov_params* p) {
  if (!t) return fail(ctx, OV_E_INVALID, "null tangle");
  ctx->tree_order_device = false;
  const int N = t->N;
  if (N < 0) return fail(ctx, OV_E_INVALID, "negative N");
  if (N > 0 && (!t->vnodes || !t->connections || !t->status))
    return fail(ctx, OV_E_INVALID, "null tangle arrays");
  ov_params q;
  if (p)
    q = *p;
  else
    ov_params_default(&q);
  bool pins = false;
  for (int k = 0; k < N && !pins; ++k)
    pins = t->status[k].status == OV_PINNED || t->status[k].status == OV_PINNED_SLIP;
  ctx->ix_has_pins = pins;
  if (pins) {
    int rc = prepare(ctx, t, &q, true);
    if (rc) return rc;
  }
  for (int k = 0; k < N && !pins; ++k) {
    const int st = t->status[k].status;
    if (st == OV_EMPTY) continue;
    if (st != OV_FREE) return fail(ctx, OV_E_INVALID, "bad status %d at %d", st, k);
    const int f = t->connections[k].forward, r = t->connections[k].reverse;
    if (f < 0 || f >= N || r < 0 || r >= N || t->status[f].status != OV_FREE ||
        t->status[r].status != OV_FREE || t->connections[f].reverse != k)
      return fail(ctx, OV_E_INVALID, "point %d: links %d/%d inconsistent", k, f, r);
  }
  ImgDev il[MAX_IMAGES];
  int nil = 0;
  int rc = setup_box_physics(ctx, &t->box, &t->bimg, &q, il, &nil);
  if (rc) return rc;
  ctx->imgs.assign(il, il + nil);
  ctx->ix_N = N;
  return ix_buffers(ctx, N);
}

// caller AoS -> index-space planes (vec3) / scalars, through the staging buffer
int ix_up3(ov_ctx* ctx, const ov_vec3d* src, double* dst) {
  const int N = ctx->ix_N;
  const long c = ctx->ix_alloc;
  int rc = ensure_host(ctx, sizeof(double) * 3 * (size_t)c);
  if (rc) return rc;
  double* h = ctx->hst;
  for (int k = 0; k < N; ++k)
    for (int i = 0; i < 3; ++i) h[(size_t)i * c + k] = src ? src[k].p[i] : 0.0;
  HIPCHK(hipMemcpyAsync(dst, h, sizeof(double) * 3 * (size_t)c, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return OV_OK;
}
int ix_down3(ov_ctx* ctx, const double* src, ov_vec3d* dst, const int* only) {
  const int N = ctx->ix_N;
  const long c = ctx->ix_alloc;
  int rc = ensure_host(ctx, sizeof(double) * 3 * (size_t)c);
  if (rc) return rc;
  double* h = ctx->hst;
  HIPCHK(hipMemcpyAsync(h, src, sizeof(double) * 3 * (size_t)c, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < N; ++k)
    if (!only || only[k] != OV_EMPTY)
      for (int i = 0; i < 3; ++i) dst[k].p[i] = h[(size_t)i * c + k];
  return OV_OK;
}

// Upload positions, links, status and (with geom) dxi / s' / s'' and
// vels / vs of the caller's tangle into the index-space arrays.
int ix_upload(ov_ctx* ctx, const ov_tangle_state* t, bool geom, bool vel) {
  const int N = ctx->ix_N;
  int rc;
  if ((rc = ix_up3(ctx, t->vnodes, ctx->ix_X))) return rc;
  if (geom) {
    if (!t->tangents || !t->normals || !t->dxi) return fail(ctx, OV_E_INVALID, "geometry arrays required");
    if ((rc = ix_up3(ctx, t->tangents, ctx->ix_T)) || (rc = ix_up3(ctx, t->normals, ctx->ix_N3))) return rc;
  }
  if (vel) {
    if ((rc = ix_up3(ctx, t->vels, ctx->ix_vels)) || (rc = ix_up3(ctx, t->vs, ctx->ix_vs))) return rc;
  }
  std::vector<int> f(N > 0 ? N : 1), r(N > 0 ? N : 1), st(N > 0 ? N : 1), pc(N > 0 ? N : 1, -1);
  std::vector<double> dx(N > 0 ? N : 1, 0.0);
  ctx->ix_status_in.assign(N, OV_EMPTY);
  for (int k = 0; k < N; ++k) {
    f[k] = t->connections[k].forward;
    r[k] = t->connections[k].reverse;
    const int s0 = t->status[k].status;
    st[k] = s0 == OV_EMPTY ? OV_EMPTY : ctx->ix_has_pins ? s0 : OV_FREE;
    if (s0 == OV_PINNED || s0 == OV_PINNED_SLIP)
      pc[k] = t->status[k].pin_wall | (s0 == OV_PINNED_SLIP ? 8 : 0) | (f[k] > 0 ? 16 : 0);
    ctx->ix_status_in[k] = st[k];
    if (geom) dx[k] = t->dxi[k];
  }
  if (N > 0) {
    HIPCHK(hipMemcpyAsync(ctx->ix_fwd, f.data(), sizeof(int) * N, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->ix_rev, r.data(), sizeof(int) * N, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->ix_status, st.data(), sizeof(int) * N, hipMemcpyHostToDevice, ctx->stream));
    if (ctx->ix_has_pins)
      HIPCHK(hipMemcpyAsync(ctx->ix_pin, pc.data(), sizeof(int) * N, hipMemcpyHostToDevice, ctx->stream));
    if (geom)
      HIPCHK(hipMemcpyAsync(ctx->ix_dxi, dx.data(), sizeof(double) * N, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  return OV_OK;
}

// Write the index-space state back into the caller's tangle: links and
// status of every slot (new points get pin_wall NOT_A_FACE, as add_point
// sets it), positions and, with geom, dxi / s' / s'' of the active points,
// with vel also vels / vs.
int ix_download(ov_ctx* ctx, ov_tangle_state* t, bool geom, bool vel) {
  const int N = ctx->ix_N;
  if (N == 0) return OV_OK;
  std::vector<int> f(N), r(N), st(N);
  HIPCHK(hipMemcpyAsync(f.data(), ctx->ix_fwd, sizeof(int) * N, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(r.data(), ctx->ix_rev, sizeof(int) * N, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(st.data(), ctx->ix_status, sizeof(int) * N, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  int rc;
  if ((rc = ix_down3(ctx, ctx->ix_X, t->vnodes, st.data()))) return rc;
  if (geom) {
    if ((rc = ix_down3(ctx, ctx->ix_T, t->tangents, st.data())) ||
        (rc = ix_down3(ctx, ctx->ix_N3, t->normals, st.data())))
      return rc;
    std::vector<double> dx(N);
    HIPCHK(hipMemcpyAsync(dx.data(), ctx->ix_dxi, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < N; ++k)
      if (st[k] != OV_EMPTY) t->dxi[k] = dx[k];
  }
  if (vel) {
    if ((rc = ix_down3(ctx, ctx->ix_vels, t->vels, st.data())) ||
        (rc = ix_down3(ctx, ctx->ix_vs, t->vs, st.data())))
      return rc;
  }
  for (int k = 0; k < N; ++k) {
    t->connections[k].forward = f[k];
    t->connections[k].reverse = r[k];
    // a slot that holds a point add_point made (FREE now, not FREE on entry)
    if (st[k] == OV_FREE && ctx->ix_status_in[k] != OV_FREE) t->status[k].pin_wall = -1;
    t->status[k].status = st[k];
  }
  return OV_OK;
}

TopoArrays ix_arrays(ov_ctx* ctx) {
  TopoArrays A;
  A.X = ctx->ix_X;
  A.T = ctx->ix_T;
  A.Nn = ctx->ix_N3;
  A.dxi = ctx->ix_dxi;
  A.fwd = ctx->ix_fwd;
  A.rev = ctx->ix_rev;
  A.status = ctx->ix_status;
  A.pin = ctx->ix_has_pins ? ctx->ix_pin : nullptr;
  A.cap = ctx->ix_alloc;
  A.N = ctx->ix_N;
  return A;
}

// remesh on the index-space tangle; next_free in/out.  OV_E_NOMEM when the
// reference would have to expand the tangle (no EMPTY slot left).
int ix_remesh(ov_ctx* ctx, double min_dist, double max_dist, double max_curv, double dl_max,
              int* next_free, int* added) {
  int st[8] = {*next_free, 0, 0, 0, 0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(ctx->ix_state, st, sizeof st, hipMemcpyHostToDevice, ctx->stream));
  int rc;
  if ((rc = launch(ctx, "k_remesh", k_remesh, dim3(1), dim3(64), ix_arrays(ctx), ctx->box, min_dist,
                   max_dist, max_curv, dl_max, ctx->ix_state)))
    return rc;
  // "if (added) enforce_boundaries" (tangle.c:870-872), decided on the device
  if ((rc = launch(ctx, "k_wrap_index", k_wrap_index, grid1(ctx->ix_N), dim3(256), ctx->ix_X,
                   (const int*)ctx->ix_status, ctx->ix_N, (long)ctx->ix_alloc, ctx->box,
                   ctx->wall_periodic_mask, (const int*)(ctx->ix_state + 1))))
    return rc;
  HIPCHK(hipMemcpyAsync(st, ctx->ix_state, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *next_free = st[0];
  if (added) *added = st[1];
  if (st[2] == 1)
    return fail(ctx, OV_E_NOMEM,
                "remesh: no EMPTY slot left in a tangle of %d points (the reference would "
                "expand_tangle to %d): pass a tangle with more capacity", ctx->ix_N, 2 * ctx->ix_N);
  if (st[2] == 2)
    return fail(ctx, OV_E_INVALID,
                "remesh left two neighbouring points between PINNED points (the reference's "
                "update_tangent_normal recurses without end there)");
  return OV_OK;
}

// Loop labels (lowest index of each loop) and ranks along the loop; leaves
// the label in ix_lab[0] and the rank in ix_d[0].
int ix_label_loops(ov_ctx* ctx) {
  const int N = ctx->ix_N;
  int rc;
  int rounds = 1;
  while ((1L << rounds) < (long)N + 1) ++rounds;
  ++rounds;
  if ((rc = launch(ctx, "k_loop_init", k_loop_init, grid1(N), dim3(256), (const int*)ctx->ix_fwd,
                   (const int*)ctx->ix_status, N, ctx->ix_lab[0], ctx->ix_jmp[0])))
    return rc;
  for (int q = 0; q < rounds; ++q) {
    const int a = q & 1, b = a ^ 1;
    if ((rc = launch(ctx, "k_loop_min", k_loop_min, grid1(N), dim3(256), (const int*)ctx->ix_lab[a],
                     (const int*)ctx->ix_jmp[a], N, ctx->ix_lab[b], ctx->ix_jmp[b])))
      return rc;
  }
  if (rounds & 1)
    HIPCHK(hipMemcpyAsync(ctx->ix_lab[0], ctx->ix_lab[1], sizeof(int) * N, hipMemcpyDeviceToDevice,
                          ctx->stream));
  const bool ch = ctx->ix_has_pins;
  if (ch) {
    // open chains: the lowest index over the reverse links too, the chain
    // flags (count in ix_state[4]) and ranks from the chain heads
    if ((rc = launch(ctx, "k_jmp_rev", k_jmp_rev, grid1(N), dim3(256), (const int*)ctx->ix_rev,
                     (const int*)ctx->ix_lab[0], N, ctx->ix_jmp[0])))
      return rc;
    for (int q = 0; q < rounds; ++q) {
      const int a = q & 1, b = a ^ 1;
      if ((rc = launch(ctx, "k_loop_min", k_loop_min, grid1(N), dim3(256), (const int*)ctx->ix_lab[a],
                       (const int*)ctx->ix_jmp[a], N, ctx->ix_lab[b], ctx->ix_jmp[b])))
        return rc;
    }
    if (rounds & 1)
      HIPCHK(hipMemcpyAsync(ctx->ix_lab[0], ctx->ix_lab[1], sizeof(int) * N, hipMemcpyDeviceToDevice,
                            ctx->stream));
    if ((rc = launch(ctx, "k_fill_int", k_fill_int, grid1(N), dim3(256), ctx->ix_chain, N, 0)) ||
        (rc = launch(ctx, "k_fill_int", k_fill_int, dim3(1), dim3(256), ctx->ix_state + 4, 1, 0)) ||
        (rc = launch(ctx, "k_chain_flag", k_chain_flag, grid1(N), dim3(256), (const int*)ctx->ix_rev,
                     (const int*)ctx->ix_lab[0], N, ctx->ix_chain, ctx->ix_state + 4)) ||
        (rc = launch(ctx, "k_rank_init_ch", k_rank_init_ch, grid1(N), dim3(256), (const int*)ctx->ix_rev,
                     (const int*)ctx->ix_lab[0], (const int*)ctx->ix_chain, N, ctx->ix_d[0],
                     ctx->ix_jmp[0])))
      return rc;
  } else if ((rc = launch(ctx, "k_rank_init", k_rank_init, grid1(N), dim3(256), (const int*)ctx->ix_rev,
                          (const int*)ctx->ix_lab[0], N, ctx->ix_d[0], ctx->ix_jmp[0]))) {
    return rc;
  }
  for (int q = 0; q < rounds; ++q) {
    const int a = q & 1, b = a ^ 1;
    if ((rc = launch(ctx, "k_rank_step", k_rank_step, grid1(N), dim3(256), (const int*)ctx->ix_d[a],
                     (const int*)ctx->ix_jmp[a], N, ctx->ix_d[b], ctx->ix_jmp[b])))
      return rc;
  }
  if (rounds & 1)
    HIPCHK(hipMemcpyAsync(ctx->ix_d[0], ctx->ix_d[1], sizeof(int) * N, hipMemcpyDeviceToDevice,
                          ctx->stream));
  if (ch)
    return (rc = launch(ctx, "k_fill_int", k_fill_int, grid1(N), dim3(256), ctx->ix_len, N, 0))
               ? rc
               : launch(ctx, "k_loop_len_ch", k_loop_len_ch, grid1(N), dim3(256), (const int*)ctx->ix_fwd,
                        (const int*)ctx->ix_rev, (const int*)ctx->ix_lab[0], (const int*)ctx->ix_d[0],
                        (const int*)ctx->ix_chain, N, ctx->ix_len);
  return launch(ctx, "k_loop_len", k_loop_len, grid1(N), dim3(256), (const int*)ctx->ix_rev,
                (const int*)ctx->ix_lab[0], (const int*)ctx->ix_d[0], N, ctx->ix_len);
}

// The reference's loop-region eliminations (tangle.c:947-1015) walk closed
// loops only (a chain's -1 link would be followed).
int ix_check_regions(ov_ctx* ctx, const ov_topology* tp) {
  if (ctx->ix_has_pins &&
      (tp->eliminate_origin_loops || tp->eliminate_zaxis_loops || tp->eliminate_outer_loops))
    return fail(ctx, OV_E_UNSUPPORTED, "loop eliminations near the origin / the z axis walk closed "
                "loops only (the reference follows a chain's missing link)");
  return OV_OK;
}

int ix_eliminate(ov_ctx* ctx, const ov_topology* tp) {
  const int N = ctx->ix_N;
  if (N == 0) return OV_OK;
  int rc;
  if ((rc = ix_check_regions(ctx, tp))) return rc;
  if ((rc = ix_label_loops(ctx))) return rc;
  const int mask = (tp->eliminate_origin_loops ? 1 : 0) | (tp->eliminate_zaxis_loops ? 2 : 0) |
                   (tp->eliminate_outer_loops ? 4 : 0);
  if ((rc = launch(ctx, "k_loop_region", k_loop_region, grid1(N), dim3(256), (const double*)ctx->ix_X,
                   (const int*)ctx->ix_lab[0], N, (long)ctx->ix_alloc,
                   tp->eliminate_loops_origin_cutoff, tp->eliminate_loops_zaxis_cutoff,
                   tp->eliminate_outer_loops_cutoff, ctx->ix_bits)) ||
      (rc = launch(ctx, "k_fill_int", k_fill_int, grid1(N), dim3(256), ctx->ix_all, N, 7)) ||
      (rc = launch(ctx, "k_loop_and", k_loop_and, grid1(N), dim3(256), (const int*)ctx->ix_lab[0],
                   (const int*)ctx->ix_bits, N, ctx->ix_all)))
    return rc;
  return launch(ctx, "k_loop_delete", k_loop_delete, grid1(N), dim3(256), (const int*)ctx->ix_lab[0],
                (const int*)ctx->ix_len, (const int*)ctx->ix_all, N, mask, tp->small_loop_cutoff,
                ctx->ix_fwd, ctx->ix_rev, ctx->ix_status);
}

// Stream order of the index-space tangle (loops in order of their lowest
// index, each walked forward from it, as prepare() orders them): fills
// ctx->n, the stream buffers' positions X0 and links fwd/rev, and
// ix_act/ix_pos.
int ix_to_stream(ov_ctx* ctx) {
  const int N = ctx->ix_N;
  int rc;
  if ((rc = ix_label_loops(ctx))) return rc;
  size_t sb = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, ctx->ix_len, ctx->ix_off, N, ctx->stream));
  if ((rc = cub_temp(ctx, sb))) return rc;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(ctx->t_cub, sb, ctx->ix_len, ctx->ix_off, N, ctx->stream));
  int last[3] = {0, 0, 0};
  HIPCHK(hipMemcpyAsync(&last[0], ctx->ix_off + (N - 1), sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(&last[1], ctx->ix_len + (N - 1), sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  if (ctx->ix_has_pins)
    HIPCHK(hipMemcpyAsync(&last[2], ctx->ix_state + 4, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const int n = last[0] + last[1];
  ctx->n = n;
  ctx->n_seg = n - last[2];   // a chain's last point starts no segment
  ctx->has_pins = ctx->ix_has_pins;
  if ((rc = ensure_capacity(ctx, std::max(n, 1)))) return rc;
  ctx->cap = ctx->buf_cap;
  if (ctx->has_pins && ctx->pin_cap < ctx->cap) {
    if ((rc = dalloc(ctx, (void**)&ctx->pin, sizeof(int) * ctx->cap))) return rc;
    ctx->pin_cap = ctx->cap;
  }
  if ((rc = launch(ctx, "k_stream_pos", k_stream_pos, grid1(N), dim3(256), (const int*)ctx->ix_lab[0],
                   (const int*)ctx->ix_d[0], (const int*)ctx->ix_off, N, ctx->ix_pos, ctx->ix_act)))
    return rc;
  if (n == 0) return OV_OK;
  if (ctx->has_pins &&
      (rc = launch(ctx, "k_gather_pin", k_gather_pin, grid1(n), dim3(256), (const int*)ctx->ix_pin,
                   (const int*)ctx->ix_fwd, (const int*)ctx->ix_act, n, ctx->pin)))
    return rc;
  return launch(ctx, "k_gather_stream", k_gather_stream, grid1(n), dim3(256), (const double*)ctx->ix_X,
                (long)ctx->ix_alloc, (const int*)ctx->ix_fwd, (const int*)ctx->ix_rev,
                (const int*)ctx->ix_act, (const int*)ctx->ix_pos, n, ctx->cap, ctx->X0, ctx->fwd,
                ctx->rev);
}

// The octree's insertion order of the device-built stream (what prepare()'s
// pos_by_index holds for a host-built one), into t_order.
int ix_tree_order(ov_ctx* ctx) {
  const int N = ctx->ix_N;
  int rc;
  if ((rc = tree_point_buffers(ctx, ctx->n))) return rc;
  if ((rc = launch(ctx, "k_active_flags", k_active_flags, grid1(N), dim3(256), (const int*)ctx->ix_pos, N,
                   ctx->ix_bits)))
    return rc;
  size_t sb = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, ctx->ix_bits, ctx->ix_all, N, ctx->stream));
  if ((rc = cub_temp(ctx, sb))) return rc;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(ctx->t_cub, sb, ctx->ix_bits, ctx->ix_all, N, ctx->stream));
  if ((rc = launch(ctx, "k_tree_order", k_tree_order, grid1(N), dim3(256), (const int*)ctx->ix_pos,
                   (const int*)ctx->ix_all, N, ctx->t_order)))
    return rc;
  ctx->tree_order_device = true;
  return OV_OK;
}

int ix_scatter3(ov_ctx* ctx, const double* src, double* dst) {
  if (ctx->n == 0) return OV_OK;
  return launch(ctx, "k_scatter3", k_scatter3, grid1(ctx->n), dim3(256), src, ctx->cap,
                (const int*)ctx->ix_act, ctx->n, (long)ctx->ix_alloc, dst);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

void ov_params_default(ov_params* p) {
  memset(p, 0, sizeof *p);
  p->kappa = 9.97e-4;
  p->vortex_width = 1e-8;
  p->alpha = 0.1;
  p->alpha_p = 0.01;
  p->dl_max = 5e-3;
  p->LIA_only = 0;
  p->use_mutual_friction = 1;
  p->hyperfriction = 1;
  p->use_BH = 0;
  p->max_curvature_scale = 0.2;
  p->hyperalpha = 0.5;
  p->BH_resolution = 0.1;
  p->BH_grain = 1e-8;
  p->BH_quadtree = 0;
  p->precision = OV_PREC_FP64;
  p->rk_variant = OV_RK_REFERENCE;
  p->multipole = OV_MP_REFERENCE;
  p->near_radius = 0;
  p->vn.type = p->vs.type = p->vb.type = OV_FLOW_NONE;
}

ov_ctx* ov_ctx_create(int device) {
  ov_ctx* ctx = new (std::nothrow) ov_ctx();
  if (!ctx) return nullptr;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&ctx->flag, 256) != hipSuccess || hipEventCreate(&ctx->t_begin) != hipSuccess ||
      hipEventCreate(&ctx->t_end) != hipSuccess) {
    delete ctx;
    return nullptr;
  }
  // set-up of the allocation-free entry point: this device's scratch address
  ctx->scratch_ptr = (unsigned char*)scratch_address(device);
  return ctx;
}

void ov_ctx_destroy(ov_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  double** v[] = {&ctx->X0, &ctx->Xs, &ctx->K[0], &ctx->K[1], &ctx->K[2], &ctx->K[3], &ctx->T,
                  &ctx->Nn, &ctx->lia, &ctx->bs,  &ctx->vs,   &ctx->vel_tmp, &ctx->T1, &ctx->N1,
                  &ctx->vs1, &ctx->dxi, &ctx->dxi1, &ctx->partial};
  for (double** b : v)
    if (*b) hipFree(*b);
  void* o[] = {ctx->fwd,  ctx->rev, ctx->flag, ctx->brk, ctx->rank, ctx->bsum, ctx->counts, ctx->adj,
               ctx->tgt,  ctx->src, ctx->far,  ctx->near, ctx->img, ctx->pin};
  for (void* b : o)
    if (b) hipFree(b);
  void* tb[] = {ctx->t_order, ctx->t_k[0], ctx->t_k[1], ctx->t_k[2], ctx->t_ks[0], ctx->t_ks[1],
                ctx->t_ks[2], ctx->t_ktmp, ctx->t_kout, ctx->t_vals, ctx->t_vals2, ctx->t_sorted,
                ctx->t_maxd, ctx->t_segs, ctx->t_cells, ctx->t_nchild, ctx->t_offs, ctx->t_size,
                ctx->t_pre, ctx->t_mom, ctx->t_nodes, ctx->t_nmom, ctx->t_cub, ctx->t_stats, ctx->t_big,
                ctx->t_big_part};
  for (void* b : tb)
    if (b) hipFree(b);
  void* bb[] = {ctx->b_dtg,  ctx->b_tid,   ctx->b_tile_tangle, ctx->b_tgt_tangle, ctx->b_tM,
                ctx->b_tS,   ctx->b_ucnt,  ctx->b_urank,       ctx->b_ubsum,      ctx->b_ucounts,
                ctx->b_queue, ctx->b_img,  ctx->b_upart, ctx->b_pin};
  for (void* b : bb)
    if (b) hipFree(b);
  void* ab[] = {ctx->a_gnb,   ctx->a_gu,    ctx->a_gl,    ctx->a_SB,    ctx->a_W,
                ctx->a_G,     ctx->a_Gtmp,  ctx->a_gK[0], ctx->a_gK[1], ctx->a_gK[2],
                ctx->a_gK[3], ctx->a_tpart, ctx->a_spart, ctx->a_ck,    ctx->a_store};
  for (void* b : ab)
    if (b) hipFree(b);
  void* rb[] = {ctx->r_keys, ctx->r_keys_sorted, ctx->r_vals,  ctx->r_order,
                ctx->r_act,  ctx->r_count,       ctx->r_pairs, ctx->r_pairs_sorted};
  for (void* b : rb)
    if (b) hipFree(b);
  void* ib[] = {ctx->ix_X, ctx->ix_T, ctx->ix_N3, ctx->ix_dxi, ctx->ix_vels, ctx->ix_vs, ctx->ix_fwd,
                ctx->ix_rev, ctx->ix_status, ctx->ix_lab[0], ctx->ix_lab[1], ctx->ix_jmp[0],
                ctx->ix_jmp[1], ctx->ix_d[0], ctx->ix_d[1], ctx->ix_len, ctx->ix_off, ctx->ix_pos,
                ctx->ix_act, ctx->ix_bits, ctx->ix_all, ctx->ix_state, ctx->ix_pin, ctx->ix_chain};
  for (void* b : ib)
    if (b) hipFree(b);
  if (ctx->hst) hipHostFree(ctx->hst);
  for (auto& p : ctx->pending) {
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  for (auto e : ctx->ev_pool) hipEventDestroy(e);
  if (ctx->t_begin) hipEventDestroy(ctx->t_begin);
  if (ctx->t_end) hipEventDestroy(ctx->t_end);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* ov_last_error(const ov_ctx* ctx) { return ctx ? ctx->err : "null context"; }

void* ov_ctx_stream(ov_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

long ov_ctx_kernel_stats(ov_ctx* ctx, const char* kernel, double* total_ms) {
  if (!ctx || !kernel) return 0;
  auto it = ctx->stats.find(kernel);
  if (it == ctx->stats.end()) {
    if (total_ms) *total_ms = 0;
    return 0;
  }
  if (total_ms) *total_ms = it->second.ms;
  return it->second.launches;
}

void ov_ctx_reset_stats(ov_ctx* ctx) {
  if (ctx) ctx->stats.clear();
}

double ov_ctx_last_interactions(const ov_ctx* ctx) { return ctx ? ctx->last_interactions : 0; }

double ov_ctx_last_device_ms(const ov_ctx* ctx) { return ctx ? ctx->last_device_ms : 0; }

void ov_ctx_tree_stats(ov_ctx* ctx, double* cells, double* leaf_pairs, double* multipoles) {
  if (!ctx) return;
  if (cells) *cells = (double)ctx->tree_cells;
  if (leaf_pairs) *leaf_pairs = ctx->tree_leaf_pairs;
  if (multipoles) *multipoles = ctx->tree_multipoles;
  ctx->tree_leaf_pairs = 0;
  ctx->tree_multipoles = 0;
}

int ov_gpu_update_tangents_normals(ov_ctx* ctx, ov_tangle_state* t) {
  if (!ctx) return OV_E_INVALID;
  ov_params p;
  ov_params_default(&p);
  p.LIA_only = 1;
  HIPCHK(hipSetDevice(ctx->device));
  int rc = prepare(ctx, t, &p, true);
  if (rc) return rc;
  if (ctx->n == 0) return OV_OK;
  if ((rc = upload_topology(ctx)) || (rc = upload_vec3(ctx, t->vnodes, ctx->X0))) return rc;
  if ((rc = mark_begin(ctx))) return rc;
  if ((rc = geometry(ctx, ctx->X0, true, ctx->T, ctx->Nn, ctx->dxi))) return rc;
  int flag = 0;
  if ((rc = finish(ctx, &flag))) return rc;
  if ((rc = download_vec3(ctx, ctx->T, t->tangents)) || (rc = download_vec3(ctx, ctx->Nn, t->normals)) ||
      (rc = download_scalar(ctx, ctx->dxi, t->dxi)))
    return rc;
  return OV_OK;
}

int ov_gpu_velocities(ov_ctx* ctx, ov_tangle_state* t, double time, const ov_params* p, int flags) {
  if (!ctx) return OV_E_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  int rc = prepare(ctx, t, p, true);
  if (rc) return rc;
  if (ctx->n == 0) return OV_OK;
  ctx->profile = (flags & OV_PROFILE) != 0;
  const bool geom = (flags & OV_UPDATE_GEOMETRY) != 0;
  if ((rc = upload_topology(ctx)) || (rc = upload_vec3(ctx, t->vnodes, ctx->X0))) return rc;
  if (!geom) {
    if (!t->tangents || !t->normals) return fail(ctx, OV_E_INVALID, "tangents/normals required");
    if ((rc = upload_vec3(ctx, t->tangents, ctx->T)) || (rc = upload_vec3(ctx, t->normals, ctx->Nn)))
      return rc;
  }
  if ((rc = mark_begin(ctx))) return rc;
  if ((rc = evaluate(ctx, ctx->X0, geom, ctx->T, ctx->Nn, ctx->dxi, ctx->vs, ctx->K[0], time))) return rc;
  int flag = 0;
  if ((rc = finish(ctx, &flag))) return rc;
  if ((rc = download_vec3(ctx, ctx->vs, t->vs)) || (rc = download_vec3(ctx, ctx->K[0], t->vels)))
    return rc;
  if (geom) {
    if ((rc = download_vec3(ctx, ctx->T, t->tangents)) || (rc = download_vec3(ctx, ctx->Nn, t->normals)) ||
        (rc = download_scalar(ctx, ctx->dxi, t->dxi)))
      return rc;
  }
  if (flag) return fail(ctx, OV_E_NONFINITE, "non-finite velocity");
  return OV_OK;
}

static int rk4_common(ov_ctx* ctx, ov_tangle_state* t, double t0, double dt, int nsteps,
                      const ov_params* p, int flags) {
  if (!ctx) return OV_E_INVALID;
  if (nsteps < 0) return fail(ctx, OV_E_INVALID, "negative step count");
  HIPCHK(hipSetDevice(ctx->device));
  int rc = prepare(ctx, t, p, true);
  if (rc) return rc;
  if (ctx->n == 0 || nsteps == 0) return OV_OK;
  ctx->profile = (flags & OV_PROFILE) != 0;
  const bool caller_k1 = (flags & OV_USE_CALLER_K1) != 0;
  if ((rc = upload_topology(ctx)) || (rc = upload_vec3(ctx, t->vnodes, ctx->X0))) return rc;
  if (caller_k1) {
    if (!t->vels) return fail(ctx, OV_E_INVALID, "vels required as k1");
    if ((rc = upload_vec3(ctx, t->vels, ctx->K[0]))) return rc;
  }
  if ((rc = mark_begin(ctx))) return rc;
  // step i starts at t0 + dt + ... + dt, accumulated as the reference's
  // driver does (main.c:127)
  std::vector<double> times(nsteps);
  double tt = t0;
  for (int i = 0; i < nsteps; ++i) {
    times[i] = tt;
    tt += dt;
  }
  auto step = [&](int i) {
    const bool use_k1 = caller_k1 && i == 0;
    int r = rk4_device(ctx, times[i], dt, p->rk_variant, use_k1);
    if (!r && (flags & OV_ENFORCE_BOUNDARIES))
      r = launch(ctx, "k_wrap", k_wrap, grid1(ctx->n), dim3(256), ctx->X0, ctx->n, ctx->cap, ctx->box,
                 ctx->wall_periodic_mask);
    return r;
  };
  // The first step runs eagerly (it sizes every buffer); the direct-sum steps
  // after it are one captured step replayed as a graph, which removes the
  // per-launch overhead that dominates small tangles.  Profiling, the octree
  // (host-read cell counts) and flows that depend on t (a graph replays its
  // launch arguments) keep eager launches.
  const bool graphs = !ctx->profile && !ctx->use_tree && !ctx->flow_time_dep;
  int i = 0;
  for (; i < nsteps && (i == 0 || !graphs); ++i)
    if ((rc = step(i))) return rc;
  if (i < nsteps) {
    hipGraphExec_t exec = nullptr;
    if ((rc = capture_graph(ctx, [&]() { return step(i); }, &exec))) return rc;
    rc = launch_graph(ctx, exec, nsteps - i);
    hipGraphExecDestroy(exec);
    if (rc) return rc;
  }
  const bool produced_t_state = !(caller_k1 && nsteps == 1);
  int flag = 0;
  if ((rc = finish(ctx, &flag))) return rc;
  if ((rc = download_vec3(ctx, ctx->X0, t->vnodes))) return rc;
  if (produced_t_state) {
    if ((rc = download_vec3(ctx, ctx->K[0], t->vels)) || (rc = download_vec3(ctx, ctx->vs1, t->vs)) ||
        (rc = download_vec3(ctx, ctx->T1, t->tangents)) || (rc = download_vec3(ctx, ctx->N1, t->normals)) ||
        (rc = download_scalar(ctx, ctx->dxi1, t->dxi)))
      return rc;
  }
  if (flag) return fail(ctx, OV_E_NONFINITE, "non-finite %s", (flag & 2) ? "position" : "velocity");
  return OV_OK;
}

int ov_gpu_rk4_run_batch(ov_ctx* ctx, ov_tangle_state* const* tangles, int ntangles, double t0,
                         double dt, int nsteps, const ov_params* params, int flags) {
  if (!ctx) return OV_E_INVALID;
  if (!tangles || !params || ntangles <= 0 || nsteps < 0)
    return fail(ctx, OV_E_INVALID, "batch: bad arguments");
  HIPCHK(hipSetDevice(ctx->device));
  const int K = ntangles;
  const int prec = params[0].precision, rkv = params[0].rk_variant;
  ctx->b_tg.assign(K, TangleDev{});
  ctx->b_act.assign(K, {});
  std::vector<std::vector<int>> fwd_l(K), rev_l(K);
  std::vector<ImgDev> imgs;
  std::vector<int> npts(K);
  std::vector<std::vector<int>> pin_l(K);
  double rnear = 0;
  bool time_dep = false;
  ctx->b_pins = false;
  for (int k = 0; k < K; ++k) {
    if (params[k].use_BH) return fail(ctx, OV_E_UNSUPPORTED, "batch: octree not supported (tangle %d)", k);
    if (params[k].precision != prec || params[k].rk_variant != rkv)
      return fail(ctx, OV_E_INVALID, "batch: tangles must share precision and RK variant");
    int rc = prepare(ctx, tangles[k], &params[k], true);
    if (rc) {
      char msg[512];
      snprintf(msg, sizeof msg, "tangle %d: %s", k, ctx->err);
      return fail(ctx, rc, "%s", msg);
    }
    if (ctx->n == 0) return fail(ctx, OV_E_INVALID, "batch: tangle %d is empty", k);
    time_dep = time_dep || ctx->flow_time_dep;
    TangleDev& t = ctx->b_tg[k];
    t.box = ctx->box;
    t.ph = ctx->ph;
    t.kappa4pi = ctx->ph.kappa / 4 / 3.14159265358979323846;
    t.n = ctx->n;
    t.lia_only = ctx->lia_only;
    t.wall_mask = ctx->wall_periodic_mask;
    t.pins = ctx->has_pins ? 1 : 0;
    if (ctx->has_pins) {
      pin_l[k] = ctx->pin_s;
      ctx->b_pins = true;
    }
    t.img_base = (int)imgs.size();
    t.n_img = (int)ctx->imgs.size();
    imgs.insert(imgs.end(), ctx->imgs.begin(), ctx->imgs.end());
    ctx->b_act[k] = ctx->act;
    fwd_l[k] = ctx->fwd_s;
    rev_l[k] = ctx->rev_s;
    npts[k] = ctx->n;
    rnear = ctx->rnear;
  }
  ctx->rnear = rnear;
  ctx->b_mixed = prec == OV_PREC_MIXED;
  // layout
  long pt = 0, rec = 0;
  for (int k = 0; k < K; ++k) {
    TangleDev& t = ctx->b_tg[k];
    t.pt_base = (int)pt;
    t.tgt_base = (int)(pt / TILE);
    t.n_tgt = (t.n + TILE - 1) / TILE;
    pt += (long)t.n_tgt * TILE;
    t.rec_base = (int)rec;
    rec += (long)((2 * t.n + TILE - 1) / TILE) * TILE;
  }
  const int ntot = (int)pt, rec_tot = (int)rec, n_tgt = ntot / TILE;
  ctx->b_K = K;
  ctx->b_ntot = ntot;
  ctx->b_rec_tot = rec_tot;
  ctx->b_n_tgt = n_tgt;
  ctx->n = ntot;
  int rc = ensure_capacity(ctx, std::max<long>(ntot, (rec_tot + 1) / 2));
  if (rc) return rc;
  ctx->cap = ctx->buf_cap;
  const long cap = ctx->cap;
  std::vector<int> tid(ntot, -1), fwd(ntot), rev(ntot), tile_tangle(rec_tot / TILE), tgt_tangle(n_tgt);
  for (int s = 0; s < ntot; ++s) fwd[s] = rev[s] = s;
  for (int k = 0; k < K; ++k) {
    const TangleDev& t = ctx->b_tg[k];
    for (int q = 0; q < t.n; ++q) {
      tid[t.pt_base + q] = k;
      fwd[t.pt_base + q] = fwd_l[k][q] >= 0 ? t.pt_base + fwd_l[k][q] : -1;   // -1: a chain end
      rev[t.pt_base + q] = rev_l[k][q] >= 0 ? t.pt_base + rev_l[k][q] : -1;
    }
    for (int j = t.rec_base / TILE; j < (t.rec_base + ((2 * t.n + TILE - 1) / TILE) * TILE) / TILE; ++j)
      tile_tangle[j] = k;
    for (int T = t.tgt_base; T < t.tgt_base + t.n_tgt; ++T) tgt_tangle[T] = k;
  }
  if ((rc = grow_int(ctx, &ctx->b_tid, &ctx->b_cap_tid, ntot)) ||
      (rc = grow_int(ctx, &ctx->b_tile_tangle, &ctx->b_cap_rec, rec_tot / TILE + 1)))
    return rc;
  if (ctx->b_pins) {
    std::vector<int> pin(ntot, -1);
    for (int k = 0; k < K; ++k)
      for (int q = 0; q < (int)pin_l[k].size(); ++q) pin[ctx->b_tg[k].pt_base + q] = pin_l[k][q];
    if ((rc = grow_int(ctx, &ctx->b_pin, &ctx->b_cap_pin, ntot))) return rc;
    HIPCHK(hipMemcpyAsync(ctx->b_pin, pin.data(), sizeof(int) * ntot, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  if (n_tgt > ctx->b_cap_tgt) {
    long c = std::max<long>(n_tgt, 2 * ctx->b_cap_tgt);
    if ((rc = dalloc(ctx, (void**)&ctx->b_tgt_tangle, sizeof(int) * c)) ||
        (rc = dalloc(ctx, (void**)&ctx->b_ucnt, sizeof(int) * c)) ||
        (rc = dalloc(ctx, (void**)&ctx->b_urank, sizeof(int) * c)) ||
        (rc = dalloc(ctx, (void**)&ctx->b_ubsum, sizeof(int) * (c / 256 + 2))))
      return rc;
    ctx->b_cap_tgt = c;
  }
  if (K > ctx->b_cap_K) {
    long c = std::max<long>(K, 2 * ctx->b_cap_K);
    if ((rc = dalloc(ctx, (void**)&ctx->b_dtg, sizeof(TangleDev) * c)) ||
        (rc = dalloc(ctx, (void**)&ctx->b_tM, sizeof(int) * c)) ||
        (rc = dalloc(ctx, (void**)&ctx->b_tS, sizeof(int) * c)))
      return rc;
    ctx->b_cap_K = c;
  }
  if ((long)imgs.size() > ctx->b_cap_img) {
    long c = std::max<long>((long)imgs.size(), 2 * ctx->b_cap_img);
    if ((rc = dalloc(ctx, (void**)&ctx->b_img, sizeof(ImgDev) * c))) return rc;
    ctx->b_cap_img = c;
  }
  if (!ctx->b_ucounts && (rc = dalloc(ctx, (void**)&ctx->b_ucounts, 256))) return rc;
  if (!ctx->b_queue && (rc = dalloc(ctx, (void**)&ctx->b_queue, 256))) return rc;
  // unit partials for the smallest unit size the tuner tries
  const int Us[4] = {128, 256, 512, 1024};
  size_t units = 0;
  for (int k = 0; k < K; ++k) {
    const TangleDev& t = ctx->b_tg[k];
    const long items = (long)((2 * t.n + TILE - 1) / TILE) * t.n_img;
    units += (size_t)t.n_tgt * (size_t)((items + Us[0] - 1) / Us[0]);
  }
  const size_t ub = sizeof(double) * 3 * TILE * std::max<size_t>(units, 1);
  if (ub > ctx->b_upart_bytes) {
    if ((rc = dalloc(ctx, (void**)&ctx->b_upart, ub))) return rc;
    ctx->b_upart_bytes = ub;
  }
  // upload
  HIPCHK(hipMemcpyAsync(ctx->b_tid, tid.data(), sizeof(int) * ntot, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->b_tile_tangle, tile_tangle.data(), sizeof(int) * tile_tangle.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->b_tgt_tangle, tgt_tangle.data(), sizeof(int) * n_tgt, hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->fwd, fwd.data(), sizeof(int) * ntot, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->rev, rev.data(), sizeof(int) * ntot, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->b_dtg, ctx->b_tg.data(), sizeof(TangleDev) * K, hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->b_img, imgs.data(), sizeof(ImgDev) * imgs.size(), hipMemcpyHostToDevice,
                        ctx->stream));
  for (int q = 0; q < 4; ++q) HIPCHK(hipMemsetAsync(ctx->K[q], 0, sizeof(double) * 3 * cap, ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->flag, 0, sizeof(int), ctx->stream));
  if ((rc = ensure_host(ctx, sizeof(double) * 3 * (size_t)cap))) return rc;
  {
    double* h = ctx->hst;
    for (int k = 0; k < K; ++k) {
      const TangleDev& t = ctx->b_tg[k];
      const ov_vec3d* src = tangles[k]->vnodes;
      for (int q = 0; q < t.n_tgt * TILE; ++q) {
        const int s = t.pt_base + q;
        const double* v = src[ctx->b_act[k][q < t.n ? q : 0]].p;   // padding: a copy of a point
        h[s] = v[0];
        h[cap + s] = v[1];
        h[2 * cap + s] = v[2];
      }
    }
    HIPCHK(hipMemcpyAsync(ctx->X0, h, sizeof(double) * 3 * cap, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  ctx->profile = false;
  // tuner: unit size x persistent grid, timed on one evaluation of these tangles
  const auto key = std::make_pair((long)ntot, K);
  auto it = ctx->b_tuned.find(key);
  if (it == ctx->b_tuned.end()) {
    int dev_cus = 256;
    hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    float best = 1e30f;
    std::pair<int, int> pick(256, dev_cus * 4);
    double tune_ms = 0;
    for (int U : Us) {
      for (int m : {0, 2, 4, 8}) {   // 0: one unit per wavefront, else a persistent grid
        ctx->b_U = U;
        ctx->b_blocks = dev_cus * m;
        float ms = 0;
        if ((rc = evaluate_batch(ctx, ctx->X0, ctx->T1, ctx->N1, ctx->dxi1, ctx->vs1, ctx->K[0], &ms, t0)))
          return rc;
        tune_ms += ms;
        if (ms < best) {
          best = ms;
          pick = {U, dev_cus * m};
        }
      }
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->b_tune_ms = tune_ms;
    it = ctx->b_tuned.emplace(key, pick).first;
    for (int q = 0; q < 4; ++q) HIPCHK(hipMemsetAsync(ctx->K[q], 0, sizeof(double) * 3 * cap, ctx->stream));
  }
  ctx->b_U = it->second.first;
  ctx->b_blocks = it->second.second;
  // one RK4 step captured as a graph, replayed nsteps times; with flows that
  // depend on t, eager steps at the accumulated times of main.c:127
  const bool enforce = (flags & OV_ENFORCE_BOUNDARIES) != 0;
  if ((rc = mark_begin(ctx))) return rc;
  if (nsteps > 0 && time_dep) {
    double tt = t0;
    for (int i = 0; i < nsteps; ++i) {
      if ((rc = rk4_batch(ctx, tt, dt, rkv, enforce))) return rc;
      tt += dt;
    }
  } else if (nsteps > 0) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = rk4_batch(ctx, t0, dt, rkv, enforce);
    hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
    if (rc) {
      if (graph) hipGraphDestroy(graph);
      return rc;
    }
    if (ec != hipSuccess) return fail(ctx, OV_E_DEVICE, "graph capture: %s", hipGetErrorString(ec));
    ec = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ec != hipSuccess) {
      hipGraphDestroy(graph);
      return fail(ctx, OV_E_DEVICE, "graph instantiate: %s", hipGetErrorString(ec));
    }
    for (int i = 0; i < nsteps && ec == hipSuccess; ++i) ec = hipGraphLaunch(exec, ctx->stream);
    hipGraphExecDestroy(exec);
    hipGraphDestroy(graph);
    if (ec != hipSuccess) return fail(ctx, OV_E_DEVICE, "graph launch: %s", hipGetErrorString(ec));
  }
  int flag = 0;
  if ((rc = finish(ctx, &flag))) return rc;
  // results back into each tangle
  for (int k = 0; k < K; ++k) {
    const TangleDev& t = ctx->b_tg[k];
    ctx->act = ctx->b_act[k];
    ctx->n = t.n;
    const long o = t.pt_base;
    if ((rc = download_vec3_at(ctx, ctx->X0, o, tangles[k]->vnodes))) return rc;
    if (nsteps > 0) {
      if ((rc = download_vec3_at(ctx, ctx->K[0], o, tangles[k]->vels)) ||
          (rc = download_vec3_at(ctx, ctx->vs1, o, tangles[k]->vs)) ||
          (rc = download_vec3_at(ctx, ctx->T1, o, tangles[k]->tangents)) ||
          (rc = download_vec3_at(ctx, ctx->N1, o, tangles[k]->normals)) ||
          (rc = download_scalar_at(ctx, ctx->dxi1, o, tangles[k]->dxi)))
        return rc;
    }
  }
  ctx->n = ntot;
  if (flag) return fail(ctx, OV_E_NONFINITE, "non-finite %s", (flag & 2) ? "position" : "velocity");
  return OV_OK;
}

void ov_ctx_set_adjoint_partial_limit(ov_ctx* ctx, size_t bytes) {
  if (ctx) ctx->a_fused_limit = bytes;
}

void ov_ctx_batch_tuning(const ov_ctx* ctx, int* unit, int* blocks, double* tune_ms) {
  if (!ctx) return;
  if (unit) *unit = ctx->b_U;
  if (blocks) *blocks = ctx->b_blocks;
  if (tune_ms) *tune_ms = ctx->b_tune_ms;
}

int ov_gpu_rk4_step(ov_ctx* ctx, ov_tangle_state* t, double time, double dt, const ov_params* p,
                    int flags) {
  return rk4_common(ctx, t, time, dt, 1, p, flags & ~OV_ENFORCE_BOUNDARIES);
}

int ov_gpu_rk4_run(ov_ctx* ctx, ov_tangle_state* t, double t0, double dt, int nsteps,
                   const ov_params* p, int flags) {
  return rk4_common(ctx, t, t0, dt, nsteps, p, flags);
}


int ov_gpu_velocity_vjp(ov_ctx* ctx, const ov_tangle_state* t, double time, const ov_params* p,
                        const ov_vec3d* cot_vels, ov_vec3d* grad_vnodes, int flags) {
  if (!ctx) return OV_E_INVALID;
  if (!cot_vels || !grad_vnodes) return fail(ctx, OV_E_INVALID, "null cotangent or gradient");
  HIPCHK(hipSetDevice(ctx->device));
  int rc = adj_prepare(ctx, t, p);
  if (rc) return rc;
  if (ctx->n == 0) {
    for (int k = 0; k < t->N; ++k) grad_vnodes[k] = ov_vec3d{{0, 0, 0}};
    return OV_OK;
  }
  ctx->profile = (flags & OV_PROFILE) != 0;
  if ((rc = upload_vec3(ctx, t->vnodes, ctx->X0)) || (rc = upload_vec3(ctx, cot_vels, ctx->a_W)))
    return rc;
  if ((rc = mark_begin(ctx))) return rc;
  if ((rc = evaluate(ctx, ctx->X0, true, ctx->T, ctx->Nn, ctx->dxi, ctx->vs, ctx->K[0], time)) ||
      (rc = velocity_vjp_dev(ctx, ctx->X0, ctx->bs, ctx->a_W, time, ctx->a_G)) ||
      (rc = launch(ctx, "ka_finite", ka_finite, grid1(ctx->n), dim3(256), (const double*)ctx->a_G,
                   ctx->n, ctx->cap, ctx->flag)))
    return rc;
  int flag = 0;
  if ((rc = finish(ctx, &flag))) return rc;
  if ((rc = download_grad(ctx, t, ctx->a_G, nullptr, grad_vnodes))) return rc;
  if (flag) return fail(ctx, OV_E_NONFINITE, "non-finite %s", (flag & 4) ? "gradient" : "velocity");
  return OV_OK;
}

int ov_gpu_rk4_run_vjp(ov_ctx* ctx, const ov_tangle_state* t, double t0, double dt, int nsteps,
                       const ov_params* p, int flags, int checkpoint_every,
                       const ov_vec3d* cot_final, ov_vec3d* grad_initial) {
  if (!ctx) return OV_E_INVALID;
  if (!cot_final || !grad_initial) return fail(ctx, OV_E_INVALID, "null cotangent or gradient");
  if (nsteps < 0) return fail(ctx, OV_E_INVALID, "negative step count");
  if (flags & OV_USE_CALLER_K1)
    return fail(ctx, OV_E_UNSUPPORTED, "the adjoint recomputes k1 from the positions");
  HIPCHK(hipSetDevice(ctx->device));
  int rc = adj_prepare(ctx, t, p);
  if (rc) return rc;
  if (ctx->n == 0 || nsteps == 0) {   // nothing moves: the map is the identity
    for (int k = 0; k < t->N; ++k) grad_initial[k] = cot_final[k];
    return OV_OK;
  }
  ctx->profile = (flags & OV_PROFILE) != 0;
  const bool enforce = (flags & OV_ENFORCE_BOUNDARIES) != 0;
  const int n = ctx->n;
  const long ps = 3 * ctx->cap;
  const int k = checkpoint_every > 0 ? std::min(checkpoint_every, nsteps) : nsteps;
  const int n_seg = (nsteps + k - 1) / k;
  if ((rc = grow_bytes(ctx, &ctx->a_ck, &ctx->a_ck_bytes, sizeof(double) * ps * n_seg)) ||
      (rc = grow_bytes(ctx, &ctx->a_store, &ctx->a_store_bytes, sizeof(double) * ps * 8 * k)))
    return rc;
  if ((rc = upload_vec3(ctx, t->vnodes, ctx->X0)) || (rc = upload_vec3(ctx, cot_final, ctx->a_G)))
    return rc;
  if ((rc = mark_begin(ctx))) return rc;
  auto wrap = [&]() {
    return enforce ? launch(ctx, "k_wrap", k_wrap, grid1(n), dim3(256), ctx->X0, n, ctx->cap,
                            ctx->box, ctx->wall_periodic_mask)
                   : OV_OK;
  };
  // segment bodies: `count` steps from X0 (keeping each step's stage data when
  // differentiating), and the VJPs of the kept steps in reverse order
  auto forward = [&](int s0, int count, bool keep) {
    int r = OV_OK;
    for (int j = 0; j < count && !r; ++j) {
      r = rk4_device(ctx, t0 + (s0 + j) * dt, dt, p->rk_variant, false,
                     keep ? ctx->a_store + (long)j * 8 * ps : nullptr);
      if (!r) r = wrap();
    }
    return r;
  };
  auto backward = [&](int s0, int count) {
    int r = OV_OK;
    for (int j = count - 1; j >= 0 && !r; --j)
      r = step_vjp_dev(ctx, ctx->a_store + (long)j * 8 * ps, t0 + (s0 + j) * dt, dt, p->rk_variant,
                       ctx->a_G);
    return r;
  };
  // full segments after the first eager one are graphs (no profiling; none
  // when an imposed flow depends on t, which a captured launch would fix)
  const bool graphs = !ctx->profile && !ctx->flow_time_dep;
  hipGraphExec_t g_fwd = nullptr, g_seg = nullptr;
  auto cleanup = [&](int r) {
    if (g_fwd) hipGraphExecDestroy(g_fwd);
    if (g_seg) hipGraphExecDestroy(g_seg);
    return r;
  };
  // forward sweep keeping the start of every segment (one segment: the input)
  if ((rc = copy_planes(ctx, ctx->a_ck, ctx->X0))) return rc;
  for (int seg = 0; seg + 1 < n_seg; ++seg) {
    if (graphs && seg > 0) {
      if (!g_fwd && (rc = capture_graph(ctx, [&]() { return forward(seg * k, k, false); }, &g_fwd)))
        return cleanup(rc);
      if ((rc = launch_graph(ctx, g_fwd, 1))) return cleanup(rc);
    } else if ((rc = forward(seg * k, k, false))) {
      return cleanup(rc);
    }
    if ((rc = copy_planes(ctx, ctx->a_ck + (seg + 1) * ps, ctx->X0))) return cleanup(rc);
  }
  // reverse sweep: recompute a segment keeping every step, then its VJPs
  bool sized = false;
  for (int seg = n_seg - 1; seg >= 0; --seg) {
    const int s0 = seg * k, count = std::min(nsteps, s0 + k) - s0;
    if ((rc = copy_planes(ctx, ctx->X0, ctx->a_ck + seg * ps))) return cleanup(rc);
    if (graphs && sized && count == k) {
      if (!g_seg && (rc = capture_graph(ctx, [&]() {
                       int r = forward(s0, k, true);
                       return r ? r : backward(s0, k);
                     }, &g_seg)))
        return cleanup(rc);
      if ((rc = launch_graph(ctx, g_seg, 1))) return cleanup(rc);
    } else {
      if ((rc = forward(s0, count, true)) || (rc = backward(s0, count))) return cleanup(rc);
      sized = true;
    }
  }
  cleanup(OV_OK);
  g_fwd = g_seg = nullptr;
  if ((rc = launch(ctx, "ka_finite", ka_finite, grid1(n), dim3(256), (const double*)ctx->a_G, n,
                   ctx->cap, ctx->flag)))
    return rc;
  int flag = 0;
  if ((rc = finish(ctx, &flag))) return rc;
  if ((rc = download_grad(ctx, t, ctx->a_G, cot_final, grad_initial))) return rc;
  if (flag)
    return fail(ctx, OV_E_NONFINITE, "non-finite %s",
                (flag & 4) ? "gradient" : (flag & 2) ? "position" : "velocity");
  return OV_OK;
}


int ov_gpu_reconnection_candidates(ov_ctx* ctx, const ov_tangle_state* t, double rec_dist,
                                   double rec_angle, int* pairs, int max_pairs, int* n_pairs) {
  if (!ctx) return OV_E_INVALID;
  if (!n_pairs || (max_pairs > 0 && !pairs)) return fail(ctx, OV_E_INVALID, "null output");
  if (!(rec_dist > 0) || !std::isfinite(rec_dist) || !std::isfinite(rec_angle))
    return fail(ctx, OV_E_INVALID, "bad rec_dist %g or rec_angle %g", rec_dist, rec_angle);
  *n_pairs = 0;
  HIPCHK(hipSetDevice(ctx->device));
  ov_params p;
  ov_params_default(&p);
  int rc = prepare(ctx, t, &p, true);
  if (rc) return rc;
  const int n = ctx->n;
  if (n < 2) return OV_OK;
  if (!t->tangents) return fail(ctx, OV_E_INVALID, "tangents required");
  if (ctx->has_pins) {
    if (ctx->pin_cap < ctx->cap) {
      if ((rc = dalloc(ctx, (void**)&ctx->pin, sizeof(int) * ctx->cap))) return rc;
      ctx->pin_cap = ctx->cap;
    }
    HIPCHK(hipMemcpyAsync(ctx->pin, ctx->pin_s.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  }
  ctx->profile = false;
  const long cap = ctx->cap;
  if ((rc = rec_buffers(ctx, n)) || (rc = upload_vec3(ctx, t->vnodes, ctx->X0)) ||
      (rc = upload_vec3(ctx, t->tangents, ctx->T)))
    return rc;
  HIPCHK(hipMemcpyAsync(ctx->r_act, ctx->act.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  // cells of edge >= rec_dist; a periodic axis is cut into whole cells
  CellGrid g{};
  for (int ax = 0; ax < 3; ++ax) {
    g.lo[ax] = ctx->box.lo[ax];
    if (ctx->box.periodic[ax] && ctx->box.L[ax] > 0) {
      const double cells = std::floor(ctx->box.L[ax] / rec_dist);
      g.dims[ax] = (int)std::max(1.0, std::min(cells, 1048576.0));
      g.h[ax] = ctx->box.L[ax] / g.dims[ax];
    } else {
      g.dims[ax] = 0;
      g.h[ax] = rec_dist;
    }
  }
  if ((rc = mark_begin(ctx))) return rc;
  if ((rc = launch(ctx, "kr_keys", kr_keys, grid1(n), dim3(256), (const double*)ctx->X0, n, cap, g,
                   ctx->r_keys, ctx->r_vals)))
    return rc;
  size_t need = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, ctx->r_keys, ctx->r_keys_sorted,
                                            ctx->r_vals, ctx->r_order, n, 0, 63, ctx->stream));
  if ((rc = cub_temp(ctx, need))) return rc;
  size_t tb = ctx->t_cub_bytes;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(ctx->t_cub, tb, ctx->r_keys, ctx->r_keys_sorted,
                                            ctx->r_vals, ctx->r_order, n, 0, 63, ctx->stream));
  RecArgs A;
  A.P = ctx->X0;
  A.T = ctx->T;
  A.act = ctx->r_act;
  A.keys = ctx->r_keys_sorted;
  A.order = ctx->r_order;
  A.pin = ctx->has_pins ? ctx->pin : nullptr;
  A.cap = cap;
  A.n = n;
  A.g = g;
  A.box = ctx->box;
  A.rec_dist = rec_dist;
  A.cos_angle = std::cos(rec_angle);
  A.count = ctx->r_count;
  if ((rc = rec_pair_buffers(ctx, std::max<long>(4L * n, 1024)))) return rc;
  int count = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    A.pairs = ctx->r_pairs;
    A.max_pairs = (int)std::min<long>(ctx->r_pair_cap, 0x7fffffffL);
    HIPCHK(hipMemsetAsync(ctx->r_count, 0, sizeof(int), ctx->stream));
    if ((rc = launch(ctx, "kr_pairs", kr_pairs, grid1(n), dim3(256), A))) return rc;
    HIPCHK(hipMemcpyAsync(&count, ctx->r_count, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (count < 0) return fail(ctx, OV_E_NOMEM, "more than 2^31 candidate pairs");
    if (count <= A.max_pairs) break;
    if ((rc = rec_pair_buffers(ctx, count))) return rc;   // grow and search again
  }
  if (count > 1) {
    need = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, need, ctx->r_pairs, ctx->r_pairs_sorted, count,
                                             0, 64, ctx->stream));
    if ((rc = cub_temp(ctx, need))) return rc;
    tb = ctx->t_cub_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortKeys(ctx->t_cub, tb, ctx->r_pairs, ctx->r_pairs_sorted, count,
                                             0, 64, ctx->stream));
  }
  int flag = 0;
  if ((rc = finish(ctx, &flag))) return rc;
  *n_pairs = count;
  const long out = std::min<long>(count, std::max(max_pairs, 0));
  if (out > 0) {
    std::vector<unsigned long long> h(out);
    HIPCHK(hipMemcpy(h.data(), count > 1 ? ctx->r_pairs_sorted : ctx->r_pairs,
                     sizeof(unsigned long long) * out, hipMemcpyDeviceToHost));
    for (long i = 0; i < out; ++i) {
      pairs[2 * i] = (int)(h[i] >> 32);
      pairs[2 * i + 1] = (int)(h[i] & 0xffffffffu);
    }
  }
  return OV_OK;
}

// update_tangle + update_velocities (the direct sum) on device-resident
// planes, with every intermediate buffer in the file-scope scratch array: no
// device or host allocation.  See include/openvort_gpu.h.
int ov_gpu_velocities_noalloc(ov_ctx* ctx, const double* pos, const int* fwd, const int* rev, int n,
                              long stride, const ov_domain_box* box, const ov_boundary_images* bimg,
                              double time, const ov_params* p, double* tangents, double* normals,
                              double* dxi, double* vs, double* vels) {
  if (!ctx) return OV_E_INVALID;
  if (!box || !bimg || !p) return fail(ctx, OV_E_INVALID, "noalloc: null box, images or params");
  if (n < 0 || stride < n) return fail(ctx, OV_E_INVALID, "noalloc: n = %d, stride = %ld", n, stride);
  if (stride > NOALLOC_MAX_POINTS)
    return fail(ctx, OV_E_INVALID, "noalloc: stride %ld above the maximum of %ld points", stride,
                NOALLOC_MAX_POINTS);
  if (p->use_BH) return fail(ctx, OV_E_UNSUPPORTED, "noalloc: direct sum only (use_BH = 0)");
  if (n == 0) return OV_OK;
  if (!pos || !fwd || !rev || !tangents || !normals || !dxi || !vs || !vels)
    return fail(ctx, OV_E_INVALID, "noalloc: null device array");
  if (!ctx->scratch_ptr) return fail(ctx, OV_E_DEVICE, "noalloc: no scratch address on device %d", ctx->device);
  HIPCHK(hipSetDevice(ctx->device));

  // a view of the context whose buffers all lie in this device's scratch
  ov_ctx v;
  v.device = ctx->device;
  v.stream = ctx->stream;
  ImgDev il[MAX_IMAGES];
  int nil = 0;
  int rc = setup_box_physics(&v, box, bimg, p, il, &nil);
  if (rc) {
    memcpy(ctx->err, v.err, sizeof ctx->err);
    return rc;
  }
  if (v.use_tree) return fail(ctx, OV_E_UNSUPPORTED, "noalloc: direct sum only (use_BH = 0)");
  v.n = n;
  v.n_seg = n;
  v.cap = stride;
  v.max_chunks = NOALLOC_CHUNKS;
  v.profile = false;
  const NoAllocLayout L = noalloc_layout(NOALLOC_MAX_POINTS);
  unsigned char* b = ctx->scratch_ptr;
  v.fwd = const_cast<int*>(fwd);
  v.rev = const_cast<int*>(rev);
  v.lia = (double*)(b + L.lia);
  v.bs = (double*)(b + L.bs);
  v.partial = (double*)(b + L.partial);
  v.partial_bytes = L.near - L.partial;
  v.near = (NearRec*)(b + L.near);
  v.far = (FarRec*)(b + L.far);
  v.src = (SrcTile*)(b + L.src);
  v.tgt = (TgtTile*)(b + L.tgt);
  v.brk = (int*)(b + L.brk);
  v.rank = (int*)(b + L.rank);
  v.bsum = (int*)(b + L.bsum);
  v.counts = (int*)(b + L.counts);
  v.adj = (int*)(b + L.adj);
  v.img = (ImgDev*)(b + L.img);
  v.flag = (int*)(b + L.flag);
  ImgList lst;
  for (int i = 0; i < nil; ++i) lst.im[i] = il[i];
  if (!rc) rc = launch(&v, "k_set_images", k_set_images, dim3(1), dim3(64), lst, nil, v.img);
  if (!rc) rc = launch(&v, "k_zero_int", k_zero_int, dim3(1), dim3(64), v.flag);
  if (!rc) rc = mark_begin_stream(&v, ctx->t_begin);
  if (!rc) rc = evaluate(&v, pos, true, tangents, normals, dxi, vs, vels, time);
  int flag = 0;
  if (!rc) {
    hipError_t e = hipEventRecord(ctx->t_end, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, v.flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = fail(&v, OV_E_DEVICE, "noalloc: %s", hipGetErrorString(e));
  }
  if (rc) {
    memcpy(ctx->err, v.err, sizeof ctx->err);
    return rc;
  }
  float ms = 0;
  if (hipEventElapsedTime(&ms, ctx->t_begin, ctx->t_end) == hipSuccess) ctx->last_device_ms = ms;
  ctx->last_interactions = v.last_interactions;
  if (flag) return fail(ctx, OV_E_NONFINITE, "non-finite velocity");
  return OV_OK;
}

void ov_topology_default(ov_topology* tp) {
  if (!tp) return;
  tp->dl_min = 1e-3;
  tp->dl_max = 5e-3;
  tp->small_loop_cutoff = 5;
  tp->eliminate_origin_loops = 0;
  tp->eliminate_loops_origin_cutoff = 3e-2;
  tp->eliminate_zaxis_loops = 0;
  tp->eliminate_loops_zaxis_cutoff = 2e-2;
  tp->eliminate_outer_loops = 0;
  tp->eliminate_outer_loops_cutoff = 1;
}

int ov_gpu_remesh(ov_ctx* ctx, ov_tangle_state* t, const ov_params* p, double min_dist,
                  double max_dist) {
  if (!ctx) return OV_E_INVALID;
  if (!p) return fail(ctx, OV_E_INVALID, "null params");
  HIPCHK(hipSetDevice(ctx->device));
  ctx->profile = false;
  int rc = ix_prepare(ctx, t, p);
  if (rc || ctx->ix_N == 0) return rc;
  if ((rc = ix_upload(ctx, t, true, false)) || (rc = mark_begin(ctx))) return rc;
  int nf = t->next_free;
  if ((rc = ix_remesh(ctx, min_dist, max_dist, p->max_curvature_scale, p->dl_max, &nf, nullptr)) ||
      (rc = mark_end(ctx)))
    return rc;
  if ((rc = ix_download(ctx, t, true, false))) return rc;
  t->next_free = nf;
  return OV_OK;
}

int ov_gpu_eliminate_small_loops(ov_ctx* ctx, ov_tangle_state* t, const ov_topology* tp) {
  if (!ctx) return OV_E_INVALID;
  if (!tp) return fail(ctx, OV_E_INVALID, "null topology settings");
  HIPCHK(hipSetDevice(ctx->device));
  ctx->profile = false;
  int rc = ix_prepare(ctx, t, nullptr);
  if (rc || ctx->ix_N == 0) return rc;
  if ((rc = ix_upload(ctx, t, false, false)) || (rc = mark_begin(ctx)) || (rc = ix_eliminate(ctx, tp)) ||
      (rc = mark_end(ctx)))
    return rc;
  return ix_download(ctx, t, false, false);
}

int ov_gpu_enforce_boundaries(ov_ctx* ctx, ov_tangle_state* t) {
  if (!ctx) return OV_E_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  ctx->profile = false;
  int rc = ix_prepare(ctx, t, nullptr);
  if (rc || ctx->ix_N == 0) return rc;
  if ((rc = ix_upload(ctx, t, false, false)) || (rc = mark_begin(ctx))) return rc;
  if ((rc = launch(ctx, "k_wrap_index", k_wrap_index, grid1(ctx->ix_N), dim3(256), ctx->ix_X,
                   (const int*)ctx->ix_status, ctx->ix_N, (long)ctx->ix_alloc, ctx->box,
                   ctx->wall_periodic_mask, (const int*)nullptr)) ||
      (rc = mark_end(ctx)))
    return rc;
  return ix_download(ctx, t, false, false);
}

int ov_gpu_run(ov_ctx* ctx, ov_tangle_state* t, double t0, double dt, int nsteps, const ov_params* p,
               const ov_topology* tp, int flags) {
  if (!ctx) return OV_E_INVALID;
  if (!p || !tp) return fail(ctx, OV_E_INVALID, "null params or topology settings");
  if (nsteps < 0) return fail(ctx, OV_E_INVALID, "negative step count");
  if (flags & OV_USE_CALLER_K1) return fail(ctx, OV_E_INVALID, "ov_gpu_run evaluates k1 itself");
  HIPCHK(hipSetDevice(ctx->device));
  int rc = ix_prepare(ctx, t, p);
  if (rc || (rc = ix_check_regions(ctx, tp))) return rc;
  if (ctx->ix_N == 0 || nsteps == 0) return OV_OK;
  if (!t->vels || !t->vs || !t->tangents || !t->normals || !t->dxi)
    return fail(ctx, OV_E_INVALID, "null tangle arrays");
  ctx->profile = false;
  if (!ctx->img && (rc = dalloc(ctx, (void**)&ctx->img, sizeof(ImgDev) * MAX_IMAGES))) return rc;
  HIPCHK(hipMemcpyAsync(ctx->img, ctx->imgs.data(), sizeof(ImgDev) * ctx->imgs.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  if ((rc = ix_upload(ctx, t, false, true))) return rc;
  ctx->profile = (flags & OV_PROFILE) != 0;
  if ((rc = launch(ctx, "k_zero_int", k_zero_int, dim3(1), dim3(64), ctx->flag))) return rc;
  if ((rc = mark_begin(ctx))) return rc;
  double time = t0;
  int nf = t->next_free;
  for (int step = 0; step < nsteps; ++step) {
    // main.c:112-117: update_tangle; rk4_step; update_tangents_normals;
    // remesh; eliminate_small_loops; enforce_boundaries
    if ((rc = ix_to_stream(ctx))) return rc;
    if (ctx->n == 0) break;
    if (ctx->use_tree && (rc = ix_tree_order(ctx))) return rc;
    if ((rc = rk4_device(ctx, time, dt, p->rk_variant, false))) return rc;
    if ((rc = ix_scatter3(ctx, ctx->K[0], ctx->ix_vels)) || (rc = ix_scatter3(ctx, ctx->vs1, ctx->ix_vs)))
      return rc;
    if ((rc = geometry(ctx, ctx->X0, true, ctx->T, ctx->Nn, ctx->dxi))) return rc;
    if ((rc = ix_scatter3(ctx, ctx->X0, ctx->ix_X)) || (rc = ix_scatter3(ctx, ctx->T, ctx->ix_T)) ||
        (rc = ix_scatter3(ctx, ctx->Nn, ctx->ix_N3)) ||
        (rc = launch(ctx, "k_scatter1", k_scatter1, grid1(ctx->n), dim3(256), (const double*)ctx->dxi,
                     (const int*)ctx->ix_act, ctx->n, ctx->ix_dxi)))
      return rc;
    if ((rc = ix_remesh(ctx, tp->dl_min, tp->dl_max, p->max_curvature_scale, p->dl_max, &nf, nullptr)))
      return rc;
    if ((rc = ix_eliminate(ctx, tp))) return rc;
    if ((rc = launch(ctx, "k_wrap_index", k_wrap_index, grid1(ctx->ix_N), dim3(256), ctx->ix_X,
                     (const int*)ctx->ix_status, ctx->ix_N, (long)ctx->ix_alloc, ctx->box,
                     ctx->wall_periodic_mask, (const int*)nullptr)))
      return rc;
    time += dt;
  }
  int flag = 0;
  if ((rc = finish(ctx, &flag))) return rc;
  if ((rc = ix_download(ctx, t, true, true))) return rc;
  t->next_free = nf;
  if (flag) return fail(ctx, OV_E_NONFINITE, "non-finite %s", (flag & 2) ? "position" : "velocity");
  return OV_OK;
}

}  // extern "C"
