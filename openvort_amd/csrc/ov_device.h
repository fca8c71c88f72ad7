// ov_device.h — device-side data layout shared by the kernels and the host
// orchestration of libopenvort_gpu.
//
// Points live in HBM as SoA FP64 planes in "stream order": the active
// (non-EMPTY) points of the caller's tangle renumbered loop by loop along
// their forward links.  Target tiles are 64 consecutive stream points (one
// wavefront, one target per lane).
//
// Sources are a "record stream": every stream point s is one record, and a
// segment s -> fwd[s] whose end is not exactly the next point (loop closure,
// periodic wrap) adds one record holding that end.  A record with link >= 0
// closes the segment that starts at the previous record, so a pass over the
// records evaluates each segment once while computing each end point's
// distance once.  The stream is padded to whole 64-record source tiles with
// unlinked records.
#pragma once
#include <stdint.h>

#include "../../include/openvort_gpu.h"

namespace ov {

constexpr int TILE = 64;          // entries per tile = wavefront width
constexpr int MAX_IMAGES = 32;    // primary + up to 31 boundary images

// Periodic box (reference vec3_maths.h:61-68): low corner, lengths, the
// periodic flag per axis (the low wall's type, vec3_maths.c:117-136).
struct BoxDev {
  double lo[3], hi[3], L[3];
  int periodic[3];
  int pad;
};

// One term of the image sum (reference tangle.c:39-79, 627-636):
// rs = r - shift; for a mirror image rs[axis] = wall2 - r[axis] and the
// image velocity's axis component changes sign.  primary = 1 only for the
// unshifted sum, the one that skips the two segments adjacent to the target.
struct ImgDev {
  double shift[3];
  double wall2;
  int axis;      // -1: no reflection
  int primary;
};

// Target tile: bounding box of its points.
struct TgtTile {
  double lo[3];
  double hi[3];
};

// Source tile: centre c of the bounding box of its records (and of the
// record before it), the FP32 frame origin; the box; the record before the
// tile (start of a segment closed by the tile's first record), in FP64 and
// in the tile frame; the largest squared segment length of its FP32 records.
struct SrcTile {
  double c[3];
  double lo[3];
  double hi[3];
  double prev[3];
  float prev_loc[3];
  float dmax2;
};

// FP32 far-field record: point relative to the tile centre and half the
// segment vector from the previous record (FP64 difference rounded once,
// then halved exactly; zero when the record closes no segment).  The half
// vector makes the far pass's cross product R x d/2, so |R x d/2|^2 u^2 u'^2
// is sin^2/4, the term the small-angle form needs, without an extra scaling.
struct FarRec {
  float p[3];
  float d[3];
};

// FP64 record: point, the segment it closes (stream index of the segment's
// start, -1 for none) and that segment's forward point (adjacent-segment
// skip of the primary sum, reference tangle.c:612-615).
struct NearRec {
  double p[3];
  int seg;
  int seg_fwd;
};

// Physics block of the per-point epilogue (reference globals,
// vortex_constants.c:24-95; imposed flows vs_conf, vn_conf,
// external_velocity.c:76-78).
struct PhysDev {
  double kappa, a0, alpha, alpha_p, dl_max;
  double max_curv, hyperalpha;
  ov_flow vs, vn, vb;   // vb: velocity of PINNED points (get_vb)
  int lia_only, use_mf, hyperfriction, pad;
};

}  // namespace ov
