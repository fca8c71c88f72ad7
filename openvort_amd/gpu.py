"""ctypes binding of libopenvort_gpu.so (include/openvort_gpu.h).

The product path: every call here runs the HIP kernels.  Loading fails
loudly when the library is missing — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .tangle import Params, Tangle, TangleState, Topology

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libopenvort_gpu.so")

OV_OK, OV_E_INVALID, OV_E_UNSUPPORTED, OV_E_NONFINITE, OV_E_DEVICE, OV_E_NOMEM = 0, -1, -2, -3, -4, -5
OV_UPDATE_GEOMETRY, OV_USE_CALLER_K1, OV_ENFORCE_BOUNDARIES, OV_PROFILE = 1, 2, 4, 8

# every symbol the header declares
SYMBOLS = (
    "ov_params_default",
    "ov_ctx_create",
    "ov_ctx_destroy",
    "ov_last_error",
    "ov_ctx_stream",
    "ov_gpu_update_tangents_normals",
    "ov_gpu_velocities",
    "ov_gpu_rk4_step",
    "ov_gpu_rk4_run",
    "ov_ctx_kernel_stats",
    "ov_ctx_reset_stats",
    "ov_ctx_last_interactions",
    "ov_ctx_last_device_ms",
    "ov_ctx_tree_stats",
    "ov_gpu_rk4_run_batch",
    "ov_ctx_batch_tuning",
    "ov_ctx_set_adjoint_partial_limit",
    "ov_gpu_velocity_vjp",
    "ov_gpu_rk4_run_vjp",
    "ov_gpu_reconnection_candidates",
    "ov_gpu_velocities_noalloc",
    "ov_topology_default",
    "ov_gpu_remesh",
    "ov_gpu_eliminate_small_loops",
    "ov_gpu_enforce_boundaries",
    "ov_gpu_run",
)

_lib = None


class OVError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"openvort_gpu error {code}: {msg}")
        self.code = code


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load and type the library (no device calls)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with python -m openvort_amd.build")
    lib = C.CDLL(path)
    P = C.POINTER
    lib.ov_params_default.argtypes = [P(Params)]
    lib.ov_params_default.restype = None
    lib.ov_ctx_create.argtypes = [C.c_int]
    lib.ov_ctx_create.restype = C.c_void_p
    lib.ov_ctx_destroy.argtypes = [C.c_void_p]
    lib.ov_ctx_destroy.restype = None
    lib.ov_last_error.argtypes = [C.c_void_p]
    lib.ov_last_error.restype = C.c_char_p
    lib.ov_ctx_stream.argtypes = [C.c_void_p]
    lib.ov_ctx_stream.restype = C.c_void_p
    lib.ov_gpu_update_tangents_normals.argtypes = [C.c_void_p, P(TangleState)]
    lib.ov_gpu_update_tangents_normals.restype = C.c_int
    lib.ov_gpu_velocities.argtypes = [C.c_void_p, P(TangleState), C.c_double, P(Params), C.c_int]
    lib.ov_gpu_velocities.restype = C.c_int
    lib.ov_gpu_rk4_step.argtypes = [C.c_void_p, P(TangleState), C.c_double, C.c_double, P(Params), C.c_int]
    lib.ov_gpu_rk4_step.restype = C.c_int
    lib.ov_gpu_rk4_run.argtypes = [C.c_void_p, P(TangleState), C.c_double, C.c_double, C.c_int,
                                   P(Params), C.c_int]
    lib.ov_gpu_rk4_run.restype = C.c_int
    lib.ov_ctx_kernel_stats.argtypes = [C.c_void_p, C.c_char_p, P(C.c_double)]
    lib.ov_ctx_kernel_stats.restype = C.c_long
    lib.ov_ctx_reset_stats.argtypes = [C.c_void_p]
    lib.ov_ctx_reset_stats.restype = None
    lib.ov_ctx_last_interactions.argtypes = [C.c_void_p]
    lib.ov_ctx_last_interactions.restype = C.c_double
    lib.ov_ctx_last_device_ms.argtypes = [C.c_void_p]
    lib.ov_ctx_last_device_ms.restype = C.c_double
    lib.ov_ctx_tree_stats.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_double)]
    lib.ov_ctx_tree_stats.restype = None
    lib.ov_gpu_rk4_run_batch.argtypes = [C.c_void_p, P(P(TangleState)), C.c_int, C.c_double, C.c_double,
                                         C.c_int, P(Params), C.c_int]
    lib.ov_gpu_rk4_run_batch.restype = C.c_int
    lib.ov_ctx_batch_tuning.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int), P(C.c_double)]
    lib.ov_ctx_batch_tuning.restype = None
    lib.ov_ctx_set_adjoint_partial_limit.argtypes = [C.c_void_p, C.c_size_t]
    lib.ov_ctx_set_adjoint_partial_limit.restype = None
    V3 = C.c_void_p   # ov_vec3d* as a numpy (N, 3) float64 buffer
    lib.ov_gpu_velocity_vjp.argtypes = [C.c_void_p, P(TangleState), C.c_double, P(Params), V3, V3, C.c_int]
    lib.ov_gpu_velocity_vjp.restype = C.c_int
    lib.ov_gpu_rk4_run_vjp.argtypes = [C.c_void_p, P(TangleState), C.c_double, C.c_double, C.c_int,
                                       P(Params), C.c_int, C.c_int, V3, V3]
    lib.ov_gpu_rk4_run_vjp.restype = C.c_int
    lib.ov_gpu_reconnection_candidates.argtypes = [C.c_void_p, P(TangleState), C.c_double, C.c_double,
                                                   C.c_void_p, C.c_int, P(C.c_int)]
    lib.ov_gpu_reconnection_candidates.restype = C.c_int
    lib.ov_gpu_velocities_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                              C.c_long, C.c_void_p, C.c_void_p, C.c_double, P(Params),
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p]
    lib.ov_gpu_velocities_noalloc.restype = C.c_int
    lib.ov_topology_default.argtypes = [P(Topology)]
    lib.ov_topology_default.restype = None
    lib.ov_gpu_remesh.argtypes = [C.c_void_p, P(TangleState), P(Params), C.c_double, C.c_double]
    lib.ov_gpu_remesh.restype = C.c_int
    lib.ov_gpu_eliminate_small_loops.argtypes = [C.c_void_p, P(TangleState), P(Topology)]
    lib.ov_gpu_eliminate_small_loops.restype = C.c_int
    lib.ov_gpu_enforce_boundaries.argtypes = [C.c_void_p, P(TangleState)]
    lib.ov_gpu_enforce_boundaries.restype = C.c_int
    lib.ov_gpu_run.argtypes = [C.c_void_p, P(TangleState), C.c_double, C.c_double, C.c_int, P(Params),
                               P(Topology), C.c_int]
    lib.ov_gpu_run.restype = C.c_int
    _lib = lib
    return lib


class Context:
    """One device context (stream + buffers) of libopenvort_gpu."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        self.ptr = self.lib.ov_ctx_create(device)
        if not self.ptr:
            raise OVError(OV_E_DEVICE, f"ov_ctx_create({device}) failed (no HIP device?)")

    def close(self):
        if self.ptr:
            self.lib.ov_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, allow_nonfinite=False) -> int:
        if rc == OV_OK or (allow_nonfinite and rc == OV_E_NONFINITE):
            return rc
        raise OVError(rc, self.lib.ov_last_error(self.ptr).decode(errors="replace"))

    # -- entry points ------------------------------------------------------
    def update_tangents_normals(self, t: Tangle) -> int:
        ts = t.cstruct()
        return self._check(self.lib.ov_gpu_update_tangents_normals(self.ptr, C.byref(ts)))

    def velocities(self, t: Tangle, time: float, params: Params, geometry=True, profile=False,
                   allow_nonfinite=False) -> int:
        ts = t.cstruct()
        flags = (OV_UPDATE_GEOMETRY if geometry else 0) | (OV_PROFILE if profile else 0)
        rc = self.lib.ov_gpu_velocities(self.ptr, C.byref(ts), time, C.byref(params), flags)
        return self._check(rc, allow_nonfinite)

    def rk4_step(self, t: Tangle, time: float, dt: float, params: Params, caller_k1=False,
                 profile=False, allow_nonfinite=False) -> int:
        ts = t.cstruct()
        flags = (OV_USE_CALLER_K1 if caller_k1 else 0) | (OV_PROFILE if profile else 0)
        rc = self.lib.ov_gpu_rk4_step(self.ptr, C.byref(ts), time, dt, C.byref(params), flags)
        return self._check(rc, allow_nonfinite)

    def rk4_run(self, t: Tangle, t0: float, dt: float, nsteps: int, params: Params,
                enforce_boundaries=False, profile=False, allow_nonfinite=False) -> int:
        ts = t.cstruct()
        flags = (OV_ENFORCE_BOUNDARIES if enforce_boundaries else 0) | (OV_PROFILE if profile else 0)
        rc = self.lib.ov_gpu_rk4_run(self.ptr, C.byref(ts), t0, dt, nsteps, C.byref(params), flags)
        return self._check(rc, allow_nonfinite)

    def rk4_run_batch(self, tangles, t0: float, dt: float, nsteps: int, params,
                      enforce_boundaries=False, allow_nonfinite=False) -> int:
        """nsteps RK4 steps of independent tangles in one launch per stage;
        params is one Params or a list (one per tangle)."""
        if isinstance(params, Params):
            params = [params] * len(tangles)
        structs = [t.cstruct() for t in tangles]
        arr = (C.POINTER(TangleState) * len(tangles))(*[C.pointer(x) for x in structs])
        parr = (Params * len(tangles))(*params)
        flags = OV_ENFORCE_BOUNDARIES if enforce_boundaries else 0
        rc = self.lib.ov_gpu_rk4_run_batch(self.ptr, arr, len(tangles), t0, dt, nsteps, parr, flags)
        return self._check(rc, allow_nonfinite)

    @staticmethod
    def _vec3(a, n):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (n, 3):
            raise ValueError(f"expected an ({n}, 3) array, got {a.shape}")
        return a

    def velocity_vjp(self, t: Tangle, time: float, params: Params, cot_vels, profile=False):
        """(d vels / d vnodes)^T cot_vels at t's positions; (N, 3) arrays
        indexed like the tangle (EMPTY rows of the result are 0)."""
        w = self._vec3(cot_vels, t.N)
        g = np.zeros((t.N, 3))
        ts = t.cstruct()
        rc = self.lib.ov_gpu_velocity_vjp(self.ptr, C.byref(ts), time, C.byref(params),
                                          w.ctypes.data, g.ctypes.data, OV_PROFILE if profile else 0)
        self._check(rc)
        return g

    def rk4_run_vjp(self, t: Tangle, t0: float, dt: float, nsteps: int, params: Params, cot_final,
                    checkpoint_every=0, enforce_boundaries=False, profile=False):
        """(d final vnodes / d initial vnodes)^T cot_final for nsteps RK4
        steps from t's positions (t is not modified)."""
        w = self._vec3(cot_final, t.N)
        g = np.zeros((t.N, 3))
        ts = t.cstruct()
        flags = (OV_ENFORCE_BOUNDARIES if enforce_boundaries else 0) | (OV_PROFILE if profile else 0)
        rc = self.lib.ov_gpu_rk4_run_vjp(self.ptr, C.byref(ts), t0, dt, nsteps, C.byref(params), flags,
                                         checkpoint_every, w.ctypes.data, g.ctypes.data)
        self._check(rc)
        return g

    def reconnection_candidates(self, t: Tangle, rec_dist: float, rec_angle: float):
        """(n, 2) int32 array of candidate pairs (k < l, caller indices) in
        (k, l) order: the pass-invariant tests of the reference's reconnect."""
        n = C.c_int(0)
        cap = max(4 * t.N, 1024)
        while True:
            buf = np.zeros((cap, 2), dtype=np.int32)
            ts = t.cstruct()
            rc = self.lib.ov_gpu_reconnection_candidates(self.ptr, C.byref(ts), rec_dist, rec_angle,
                                                         buf.ctypes.data, cap, C.byref(n))
            self._check(rc)
            if n.value <= cap:
                return buf[:n.value].copy()
            cap = n.value

    def velocities_noalloc(self, pos, fwd, rev, n: int, stride: int, t: Tangle, time: float,
                           params: Params, tangents, normals, dxi, vs, vels) -> int:
        """ov_gpu_velocities_noalloc on device arrays given as integer
        addresses (planes of `stride` doubles; int32 links); t supplies the
        box and the image list only."""
        ts = t.cstruct()
        rc = self.lib.ov_gpu_velocities_noalloc(self.ptr, pos, fwd, rev, n, stride,
                                                C.addressof(ts.box), C.addressof(ts.bimg), time,
                                                C.byref(params), tangents, normals, dxi, vs, vels)
        return self._check(rc)

    def remesh(self, t: Tangle, params: Params, min_dist: float, max_dist: float) -> int:
        """remesh (tangle.c:824-873) on the GPU; t's geometry must be current."""
        ts = t.cstruct()
        rc = self._check(self.lib.ov_gpu_remesh(self.ptr, C.byref(ts), C.byref(params), min_dist, max_dist))
        t.next_free = ts.next_free
        return rc

    def eliminate_small_loops(self, t: Tangle, tp: Topology) -> int:
        ts = t.cstruct()
        return self._check(self.lib.ov_gpu_eliminate_small_loops(self.ptr, C.byref(ts), C.byref(tp)))

    def enforce_boundaries(self, t: Tangle) -> int:
        ts = t.cstruct()
        return self._check(self.lib.ov_gpu_enforce_boundaries(self.ptr, C.byref(ts)))

    def run(self, t: Tangle, t0: float, dt: float, nsteps: int, params: Params, tp: Topology,
            profile=False, allow_nonfinite=False) -> int:
        """nsteps of the reference main loop body (main.c:112-117) device-resident."""
        ts = t.cstruct()
        rc = self.lib.ov_gpu_run(self.ptr, C.byref(ts), t0, dt, nsteps, C.byref(params), C.byref(tp),
                                 OV_PROFILE if profile else 0)
        if rc in (OV_OK, OV_E_NONFINITE):
            t.next_free = ts.next_free
        return self._check(rc, allow_nonfinite)

    def set_adjoint_partial_limit(self, nbytes: int) -> None:
        """Memory for the adjoint's fused segment pass (0: two passes)."""
        self.lib.ov_ctx_set_adjoint_partial_limit(self.ptr, int(nbytes))

    def batch_tuning(self):
        u, b, ms = C.c_int(), C.c_int(), C.c_double()
        self.lib.ov_ctx_batch_tuning(self.ptr, C.byref(u), C.byref(b), C.byref(ms))
        return u.value, b.value, ms.value

    # -- profiling -----------------------------------------------------------
    def kernel_stats(self, name: str):
        ms = C.c_double(0)
        n = self.lib.ov_ctx_kernel_stats(self.ptr, name.encode(), C.byref(ms))
        return int(n), float(ms.value)

    def reset_stats(self):
        self.lib.ov_ctx_reset_stats(self.ptr)

    def last_interactions(self) -> float:
        return float(self.lib.ov_ctx_last_interactions(self.ptr))

    def last_device_ms(self) -> float:
        return float(self.lib.ov_ctx_last_device_ms(self.ptr))

    def tree_stats(self):
        """(cells, leaf pairs, multipoles) of the octree evaluations since the last call."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self.lib.ov_ctx_tree_stats(self.ptr, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def stream(self) -> int:
        return int(self.lib.ov_ctx_stream(self.ptr) or 0)
