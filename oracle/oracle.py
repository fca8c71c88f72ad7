"""ctypes access to the test-only checkers.

TEST INFRASTRUCTURE ONLY — imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of bench.py, never by the product package.

* ``Oracle``: the FP64 C restatement (liboracle.so, ov_oracle.c).
* ``Reference``: the reference's own hot-path C compiled from its sources
  (oracle/_ref/libovref.so, built by oracle/Makefile where the reference
  tree exists); ``Reference.available()`` is False elsewhere.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(_HERE))

from openvort_amd.tangle import Flow, Params, Tangle, TangleState  # noqa: E402

ORACLE_LIB = os.path.join(_HERE, "liboracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libovref.so")

P = C.POINTER


class Oracle:
    _lib = None

    def __init__(self, threads: int = 0):
        if Oracle._lib is None:
            if not os.path.exists(ORACLE_LIB):
                raise RuntimeError(f"{ORACLE_LIB} missing: run make -C oracle")
            lib = C.CDLL(ORACLE_LIB)
            lib.ovo_update_tangents_normals.argtypes = [P(TangleState)]
            lib.ovo_update_velocities.argtypes = [P(TangleState), C.c_double, P(Params)]
            lib.ovo_velocities_sample.argtypes = [P(TangleState), C.c_double, P(Params), P(C.c_int),
                                                  C.c_int, P(C.c_double), P(C.c_double)]
            lib.ovo_rk4_step.argtypes = [P(TangleState), C.c_double, C.c_double, P(Params)]
            lib.ovo_enforce_boundaries.argtypes = [P(TangleState)]
            lib.ovo_enforce_boundaries.restype = None
            lib.ovo_segment_field.argtypes = [P(C.c_double), P(C.c_double), P(C.c_double), C.c_double,
                                              P(C.c_double)]
            lib.ovo_segment_field.restype = None
            lib.ovo_octree_stats.argtypes = [P(C.c_long), P(C.c_long), P(C.c_long)]
            lib.ovo_octree_stats.restype = None
            lib.ovo_set_threads.argtypes = [C.c_int]
            lib.ovo_set_threads.restype = None
            lib.ovo_reconnection_candidates.argtypes = [P(TangleState), C.c_double, C.c_double,
                                                        P(C.c_int), C.c_long]
            lib.ovo_reconnection_candidates.restype = C.c_long
            lib.ovo_flow.argtypes = [P(Flow), P(C.c_double), C.c_double, P(C.c_double)]
            lib.ovo_cell_far_field.argtypes = [P(TangleState), P(C.c_int), C.c_int, P(C.c_double),
                                               P(Params), P(C.c_double)]
            Oracle._lib = lib
        self.lib = Oracle._lib
        self.lib.ovo_set_threads(threads)

    def _rc(self, rc):
        if rc != 0:
            raise RuntimeError(f"oracle returned {rc}")

    def update_tangents_normals(self, t: Tangle):
        ts = t.cstruct()
        self._rc(self.lib.ovo_update_tangents_normals(C.byref(ts)))

    def update_velocities(self, t: Tangle, time: float, p: Params):
        ts = t.cstruct()
        self._rc(self.lib.ovo_update_velocities(C.byref(ts), time, C.byref(p)))

    def update_tangle(self, t: Tangle, time: float, p: Params):
        self.update_tangents_normals(t)
        self.update_velocities(t, time, p)

    def velocities_sample(self, t: Tangle, time: float, p: Params, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        vs = np.zeros((len(idx), 3))
        vel = np.zeros((len(idx), 3))
        ts = t.cstruct()
        self._rc(self.lib.ovo_velocities_sample(C.byref(ts), time, C.byref(p),
                                                idx.ctypes.data_as(P(C.c_int)), len(idx),
                                                vs.ctypes.data_as(P(C.c_double)),
                                                vel.ctypes.data_as(P(C.c_double))))
        return vs, vel

    def rk4_step(self, t: Tangle, time: float, dt: float, p: Params):
        """rk4_step alone: k1 = t.vels as given."""
        ts = t.cstruct()
        self._rc(self.lib.ovo_rk4_step(C.byref(ts), time, dt, C.byref(p)))

    def step(self, t: Tangle, time: float, dt: float, p: Params):
        """update_tangle(t) then rk4_step(t, dt) (reference main.c:112-113)."""
        self.update_tangle(t, time, p)
        self.rk4_step(t, time, dt, p)

    def enforce_boundaries(self, t: Tangle):
        ts = t.cstruct()
        self.lib.ovo_enforce_boundaries(C.byref(ts))

    def segment_field(self, a, b, r, kappa):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros(3)
        self.lib.ovo_segment_field(a.ctypes.data_as(P(C.c_double)), b.ctypes.data_as(P(C.c_double)),
                                   r.ctypes.data_as(P(C.c_double)), kappa,
                                   out.ctypes.data_as(P(C.c_double)))
        return out

    def reconnection_candidates(self, t: Tangle, rec_dist: float, rec_angle: float):
        """(n, 2) int32 pairs k < l passing the pass-invariant reconnection tests."""
        ts = t.cstruct()
        n = self.lib.ovo_reconnection_candidates(C.byref(ts), rec_dist, rec_angle, None, 0)
        out = np.zeros((max(n, 1), 2), dtype=np.int32)
        self.lib.ovo_reconnection_candidates(C.byref(ts), rec_dist, rec_angle,
                                             out.ctypes.data_as(P(C.c_int)), n)
        return out[:n]

    def cell_far_field(self, t: Tangle, ids, r, p: Params):
        """Multipole value at r of one octree cell holding the points ids."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros(3)
        ts = t.cstruct()
        self._rc(self.lib.ovo_cell_far_field(C.byref(ts), ids.ctypes.data_as(P(C.c_int)), len(ids),
                                             r.ctypes.data_as(P(C.c_double)), C.byref(p),
                                             out.ctypes.data_as(P(C.c_double))))
        return out

    def octree_stats(self):
        a, b, c = C.c_long(), C.c_long(), C.c_long()
        self.lib.ovo_octree_stats(C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def flow(self, f: Flow, r, t: float):
        """The imposed flow f at position r, time t."""
        return _flow_call(self.lib.ovo_flow, f, r, t)


def _flow_call(fn, f, r, t):
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(3)
    if fn(C.byref(f), r.ctypes.data_as(P(C.c_double)), t, out.ctypes.data_as(P(C.c_double))) != 0:
        raise ValueError(f"flow type {f.type}")
    return out


class Reference:
    """The reference's own C (update_tangents_normals, update_velocities,
    update_velocity, rk4_step, enforce_boundaries) through ref_harness.c."""

    _lib = None

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_LIB)

    def __init__(self, threads: int = 1):
        if Reference._lib is None:
            lib = C.CDLL(REF_LIB)
            lib.ovref_set_params.argtypes = [P(Params), C.c_int]
            lib.ovref_update_tangents_normals.argtypes = [P(TangleState)]
            lib.ovref_update_tangents_normals.restype = None
            lib.ovref_update_velocities.argtypes = [P(TangleState), C.c_double]
            lib.ovref_update_velocities.restype = None
            lib.ovref_velocities_sample.argtypes = [P(TangleState), C.c_double, P(C.c_int), C.c_int,
                                                    P(C.c_double), P(C.c_double)]
            lib.ovref_velocities_sample.restype = None
            lib.ovref_rk4_step.argtypes = [P(TangleState), C.c_double, C.c_double]
            lib.ovref_rk4_step.restype = None
            lib.ovref_enforce_boundaries.argtypes = [P(TangleState)]
            lib.ovref_enforce_boundaries.restype = None
            lib.ovref_reconnect.argtypes = [P(TangleState), C.c_double, C.c_double, C.c_double]
            lib.ovref_reconnect.restype = C.c_int
            lib.ovref_reconnect_candidates.argtypes = [P(TangleState), C.c_double, P(C.c_int), C.c_long]
            lib.ovref_reconnect_candidates.restype = C.c_int
            lib.ovref_flow.argtypes = [P(Flow), P(C.c_double), C.c_double, P(C.c_double)]
            lib.ovref_save_tangle.argtypes = [P(TangleState), C.c_char_p]
            lib.ovref_save_tangle.restype = None
            lib.ovref_load_tangle.argtypes = [C.c_char_p, C.c_int, P(C.c_double), P(C.c_int)]
            lib.ovref_load_tangle.restype = C.c_int
            lib.ovref_vconf_size.argtypes = []
            lib.ovref_vconf_size.restype = C.c_int
            lib.ovref_vconf.argtypes = [P(Flow), C.c_void_p]
            lib.ovref_vconf.restype = C.c_int
            lib.ovref_remesh.argtypes = [P(TangleState), C.c_double, C.c_double, C.c_int]
            lib.ovref_remesh.restype = C.c_int
            lib.ovref_eliminate_small_loops.argtypes = [P(TangleState), C.c_int, C.c_int, C.c_double,
                                                        C.c_int, C.c_double, C.c_int, C.c_double]
            lib.ovref_eliminate_small_loops.restype = None
            Reference._lib = lib
        self.lib = Reference._lib
        self.threads = threads

    def set_params(self, p: Params):
        rc = self.lib.ovref_set_params(C.byref(p), self.threads)
        if rc:
            raise RuntimeError("reference: unsupported flow")

    def update_tangents_normals(self, t: Tangle):
        ts = t.cstruct()
        self.lib.ovref_update_tangents_normals(C.byref(ts))

    def update_velocities(self, t: Tangle, time: float, p: Params):
        self.set_params(p)
        ts = t.cstruct()
        self.lib.ovref_update_velocities(C.byref(ts), time)

    def update_tangle(self, t: Tangle, time: float, p: Params):
        self.update_tangents_normals(t)
        self.update_velocities(t, time, p)

    def velocities_sample(self, t: Tangle, time: float, p: Params, idx):
        self.set_params(p)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        vs = np.zeros((len(idx), 3))
        vel = np.zeros((len(idx), 3))
        ts = t.cstruct()
        self.lib.ovref_velocities_sample(C.byref(ts), time, idx.ctypes.data_as(P(C.c_int)), len(idx),
                                         vs.ctypes.data_as(P(C.c_double)),
                                         vel.ctypes.data_as(P(C.c_double)))
        return vs, vel

    def rk4_step(self, t: Tangle, time: float, dt: float, p: Params):
        self.set_params(p)
        ts = t.cstruct()
        self.lib.ovref_rk4_step(C.byref(ts), time, dt)

    def step(self, t: Tangle, time: float, dt: float, p: Params):
        self.update_tangle(t, time, p)
        self.rk4_step(t, time, dt, p)

    def enforce_boundaries(self, t: Tangle):
        ts = t.cstruct()
        self.lib.ovref_enforce_boundaries(C.byref(ts))

    def reconnect(self, t: Tangle, time: float, p: Params, rec_dist: float, rec_angle: float) -> int:
        """The reference's reconnect pass (vortex_dynamics.c:140-315)."""
        self.set_params(p)
        ts = t.cstruct()
        return self.lib.ovref_reconnect(C.byref(ts), time, rec_dist, rec_angle)

    def reconnect_candidates(self, t: Tangle, time: float, p: Params, pairs) -> int:
        """The same pass driven by a sorted candidate list (ref_harness.c)."""
        self.set_params(p)
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        ts = t.cstruct()
        return self.lib.ovref_reconnect_candidates(C.byref(ts), time, pairs.ctypes.data_as(P(C.c_int)),
                                                   len(pairs))

    def flow(self, f: Flow, r, t: float):
        """The reference's own flow function (its table entry for f's type,
        parameters set by name) at position r, time t."""
        return _flow_call(self.lib.ovref_flow, f, r, t)

    def save_tangle(self, t: Tangle, path: str):
        """The reference's save_tangle."""
        ts = t.cstruct()
        self.lib.ovref_save_tangle(C.byref(ts), path.encode())

    def load_tangle(self, path: str, n: int):
        """The reference's load_tangle into a tangle of capacity n: (rc,
        positions (n, 3), links (n, 2) forward/reverse)."""
        pos = np.zeros((n, 3))
        links = np.zeros((n, 2), dtype=np.int32)
        rc = self.lib.ovref_load_tangle(path.encode(), n, pos.ctypes.data_as(P(C.c_double)),
                                        links.ctypes.data_as(P(C.c_int)))
        return rc, pos, links

    def vconf(self, f: Flow):
        """Bytes of the reference's v_conf_t for flow f (the buffer must stay
        alive while its address is used)."""
        buf = C.create_string_buffer(self.lib.ovref_vconf_size())
        if self.lib.ovref_vconf(C.byref(f), buf) != 0:
            raise RuntimeError("unknown flow")
        return buf

    def remesh(self, t: Tangle, min_dist: float, max_dist: float, need_free: int = 1):
        """The reference's remesh (globals from set_params: dl_max and
        max_curvature_scale of add_point); t.next_free is updated."""
        ts = t.cstruct()
        if self.lib.ovref_remesh(C.byref(ts), min_dist, max_dist, need_free) != 0:
            raise RuntimeError("not enough EMPTY slots for the reference's remesh")
        t.next_free = ts.next_free

    def eliminate_small_loops(self, t: Tangle, loop_length: int, origin=(0, 3e-2), zaxis=(0, 2e-2),
                              outer=(0, 1.0)):
        ts = t.cstruct()
        self.lib.ovref_eliminate_small_loops(C.byref(ts), loop_length, origin[0], origin[1], zaxis[0],
                                             zaxis[1], outer[0], outer[1])

