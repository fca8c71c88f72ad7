// ov_flow.h — imposed flows on the device: the functions of the reference's
// flow table (external_velocity.c:85-174, 176-512), evaluated per point at
// its position r and the evaluation's time t.
//
// The arithmetic follows each reference function operation for operation,
// without fused multiply-add, so everything but cos/sin/exp rounds as the
// reference does; those three come from the device math library (within an
// ulp or two of glibc's).  Parameter slots are those of ov_flow
// (include/openvort_gpu.h).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/openvort_gpu.h"

namespace ov {

constexpr double FLOW_PI = 3.14159265358979323846;

// vec3_d (vec3_maths.c:338-345): sum of squares from zero, then sqrt
__device__ __host__ inline double flow_norm(double x, double y, double z) {
#pragma clang fp contract(off)
  double o = 0;
  o += x * x;
  o += y * y;
  o += z * z;
  return sqrt(o);
}

// "spherical" (external_velocity.c:197-223): strength / (4 pi r^2)
// exp(-(cutoff / r)^2) along r / |r|, zero where r < 1e-8 or the exponent
// exceeds 100.
__device__ inline void flow_spherical(const ov_flow& f, const double r[3], double out[3]) {
#pragma clang fp contract(off)
  const double strength = f.a[0], cutoff = f.a[1];
  const double d = flow_norm(r[0], r[1], r[2]);
  const double attn = (cutoff / d) * (cutoff / d);
  if (d < 1e-8 || attn > 100) {
    out[0] = out[1] = out[2] = 0;
    return;
  }
  const double factor = exp(-attn);
  const double inv = 1 / flow_norm(r[0], r[1], r[2]);   // vec3_normalize
  const double m = strength / (4 * FLOW_PI * d * d) * factor;
  for (int i = 0; i < 3; ++i) out[i] = (r[i] * inv) * m;
}

// "cylindrical" (external_velocity.c:243-270): the same about the z axis,
// strength / (2 pi r).
__device__ inline void flow_cylindrical(const ov_flow& f, const double r[3], double out[3]) {
#pragma clang fp contract(off)
  const double strength = f.a[0], cutoff = f.a[1];
  const double d = sqrt(r[0] * r[0] + r[1] * r[1]);
  const double attn = (cutoff / d) * (cutoff / d);
  if (d < 1e-8 || attn > 100) {
    out[0] = out[1] = out[2] = 0;
    return;
  }
  const double factor = exp(-attn);
  const double inv = 1 / flow_norm(r[0], r[1], 0.0);
  const double m = strength / (2 * FLOW_PI * d) * factor;
  out[0] = (r[0] * inv) * m;
  out[1] = (r[1] * inv) * m;
  out[2] = (0.0 * inv) * m;
}

// z x r (vec3_cross with u = (0, 0, 1), vec3_maths.c:304-312)
__device__ inline void flow_zcross(const double r[3], double v[3]) {
#pragma clang fp contract(off)
  v[0] = 0.0 * r[2] - 1.0 * r[1];
  v[1] = 1.0 * r[0] - 0.0 * r[2];
  v[2] = 0.0 * r[1] - 0.0 * r[0];
}

// The flow of descriptor f at position r, time t.
__device__ inline void flow_eval(const ov_flow& f, const double r[3], double t, double out[3]) {
#pragma clang fp contract(off)
  const double x = r[0], y = r[1], z = r[2];
  switch (f.type) {
    case OV_FLOW_SIMPLE:
      out[0] = f.v[0];
      out[1] = f.v[1];
      out[2] = f.v[2];
      return;
    case OV_FLOW_SPHERICAL:
      flow_spherical(f, r, out);
      return;
    case OV_FLOW_SPHERICAL_AND_SIMPLE: {
      double s[3];
      flow_spherical(f, r, s);
      for (int i = 0; i < 3; ++i) out[i] = s[i] + f.v[i];
      return;
    }
    case OV_FLOW_CYLINDRICAL:
      flow_cylindrical(f, r, out);
      return;
    case OV_FLOW_OSCILLATING: {   // :272-294
      const double strength = f.a[0], freq = f.a[1];
      double kr = 0;
      kr += f.w[0] * x;
      kr += f.w[1] * y;
      kr += f.w[2] * z;
      const double X = cos(2 * FLOW_PI * kr);
      const double T = sin(2 * FLOW_PI * freq * t);
      const double m = strength * X * T;
      for (int i = 0; i < 3; ++i) out[i] = f.q[i] * m;
      return;
    }
    case OV_FLOW_OSCILLATING_SLIT: {   // :296-322 (strength unused there)
      const double freq = f.a[1], cutoff = f.a[2];
      const double rr = sqrt(x * x + z * z);
      const double dir[3] = {x / rr, 0.0, z / rr};
      const double attn = exp(-cutoff * cutoff / rr / rr);
      const double osc = cos(2 * FLOW_PI * freq * t);
      const double m = osc * attn;
      for (int i = 0; i < 3; ++i) out[i] = dir[i] * m;
      return;
    }
    case OV_FLOW_COSCOS: {   // :324-344
      const double k = f.a[0], v0 = f.a[1], l = f.a[2];
      out[0] = out[1] = 0;
      out[2] = v0 * (1 + cos(k * x) * cos(k * y) * exp(-z / l));
      return;
    }
    case OV_FLOW_COSCOS_DIVFREE: {   // :347-370
      const double k = f.a[0], v0 = f.a[1], l = f.a[2];
      out[0] = v0 / 2 / l / k * sin(k * x) * cos(k * y) * exp(-z / l);
      out[1] = v0 / 2 / l / k * cos(k * x) * sin(k * y) * exp(-z / l);
      out[2] = v0 * (1 + cos(k * x) * cos(k * y) * exp(-z / l));
      return;
    }
    case OV_FLOW_COS_DIVFREE: {   // :372-394
      const double k = f.a[0], v0 = f.a[1], l = f.a[2];
      out[0] = v0 / l / k * sin(k * x) * exp(-z / l);
      out[1] = 0;
      out[2] = v0 * (1 + cos(k * x) * exp(-z / l));
      return;
    }
    case OV_FLOW_SIMPLE_SHEAR: {   // :396-413
      const double rate = f.a[0], v0 = f.a[1];
      out[0] = out[1] = 0;
      out[2] = v0 + rate * x;
      return;
    }
    case OV_FLOW_ROTOSC:
    case OV_FLOW_ROTOSC_BOUNDARY: {   // :415-433 (both table entries)
      const double freq = f.a[0], amp = f.a[1];
      double v[3];
      flow_zcross(r, v);
      const double m = amp * cos(2 * FLOW_PI * freq * t);
      for (int i = 0; i < 3; ++i) out[i] = v[i] * m;
      return;
    }
    case OV_FLOW_DG: {   // :438-457
      const double rot = f.a[0], vz = f.a[1];
      double v[3];
      flow_zcross(r, v);
      for (int i = 0; i < 3; ++i) out[i] = v[i] * rot;
      out[2] += vz;
      return;
    }
    case OV_FLOW_CONSTANT_ACCELERATION:   // :483-495
      out[0] = f.a[0] * t;
      out[1] = out[2] = 0;
      return;
    case OV_FLOW_PARABOLIC: {   // :497-512
      const double H = f.a[0], amp = f.a[1];
      out[0] = amp * z * (H - z);
      out[1] = out[2] = 0;
      return;
    }
    default:   // OV_FLOW_NONE
      out[0] = out[1] = out[2] = 0;
      return;
  }
}

// Flows whose value depends on t (a captured launch sequence bakes t in).
inline bool flow_time_dependent(int type) {
  return type == OV_FLOW_OSCILLATING || type == OV_FLOW_OSCILLATING_SLIT || type == OV_FLOW_ROTOSC ||
         type == OV_FLOW_ROTOSC_BOUNDARY || type == OV_FLOW_CONSTANT_ACCELERATION;
}
// Flows whose value depends on the position.
inline bool flow_position_dependent(int type) {
  return !(type == OV_FLOW_NONE || type == OV_FLOW_SIMPLE || type == OV_FLOW_CONSTANT_ACCELERATION);
}

}  // namespace ov
