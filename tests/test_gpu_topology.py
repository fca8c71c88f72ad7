"""Device-resident topology edits against the reference's own functions
(compiled from src/tangle.c into oracle/_ref): remesh, eliminate_small_loops,
enforce_boundaries, and the main loop body (main.c:112-117) run on the
device for several steps.

Links, status and next_free must be identical.  Positions: the points
remesh moves or inserts are computed from s' and s'' (tangle.c:1033-1052,
1096-1140); wherever the reference recomputes those with its own stencil
expressions the device's agree to ~1e-12 relative, which moves such a point
by ~1e-17 cm, so positions are compared to 1e-15 cm (the box is 0.1 cm).
"""
import math

import numpy as np
import pytest

from openvort_amd import tangle as T
from openvort_amd.gpu import OV_E_NOMEM, OVError
from oracle.oracle import Reference

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(), reason="reference library not built")]

POS_TOL = 1e-15


def uneven_tangle(seed=5, n_rings=4, capacity=3.0, box=0.1):
    """Rings whose points sit at jittered angles, so that segments run from
    well below dl_min = 1e-3 to above dl_max = 2e-3; one ring crosses the
    periodic box; spare EMPTY capacity for the points remesh adds."""
    rng = np.random.default_rng(seed)
    b = T.TangleBuilder()
    for q in range(n_rings):
        r = rng.uniform(0.006, 0.012)
        npts = int(2 * math.pi * r / 1.2e-3)
        phi = np.sort(2 * math.pi * (np.arange(npts) + rng.uniform(-0.45, 0.45, npts)) / npts)
        c = rng.uniform(0.02, 0.08, 3) if q else np.array([0.005, 0.05, 0.05])
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        u = np.cross(d, [1.0, 0.0, 0.0] if abs(d[0]) < 0.9 else [0.0, 1.0, 0.0])
        u /= np.linalg.norm(u)
        v = np.cross(d, u)
        b.add_loop(c + r * (np.cos(phi)[:, None] * u + np.sin(phi)[:, None] * v))
    n = b.n
    t = b.build((0, 0, 0), (box, box, box), (T.WALL_PERIODIC,) * 6, T.periodic_images(26),
                holes=int(capacity * n) - n, seed=seed)
    t.wrap_into_box()
    t.next_free = 0
    return t


def params(**kw):
    base = dict(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0, dl_max=2e-3)
    base.update(kw)
    return T.default_params(**base)


def _same_topology(a, b):
    np.testing.assert_array_equal(b.status[:, 0], a.status[:, 0])
    np.testing.assert_array_equal(b.connections, a.connections)
    assert a.next_free == b.next_free
    act = a.active
    d = np.abs(a.vnodes[act] - b.vnodes[act]).max()
    return d


def test_remesh_matches_reference(ctx):
    R = Reference()
    p = params()
    R.set_params(p)
    a = uneven_tangle()
    R.update_tangents_normals(a)
    b = a.copy()
    n0 = len(a.active)
    R.remesh(a, 1e-3, 2e-3, need_free=n0)
    ctx.remesh(b, p, 1e-3, 2e-3)
    d = _same_topology(a, b)
    act = a.active
    print(f"remesh: {n0} -> {len(act)} points, next_free {a.next_free}, position diff {d:.3g}", flush=True)
    assert len(act) != n0
    assert d < POS_TOL
    np.testing.assert_allclose(b.dxi[act], a.dxi[act], rtol=1e-13, atol=0)
    # the link structure is a set of closed loops
    for k in act:
        assert b.connections[b.connections[k, 0], 1] == k


def test_remesh_nomem_leaves_tangle(ctx):
    """With no EMPTY slot the reference would expand_tangle: the GPU call
    reports it and writes nothing."""
    p = params()
    b0 = T.TangleBuilder()
    b0.add_ring((0.05, 0.05, 0.05), (0.0, 0.0, 1.0), 0.01, 25)     # 2.5e-3 cm segments, none short
    a = b0.build((0, 0, 0), (0.1, 0.1, 0.1), (T.WALL_PERIODIC,) * 6, T.periodic_images(26))
    a.next_free = 0
    Reference().update_tangents_normals(a)
    b = a.copy()
    with pytest.raises(OVError) as e:
        ctx.remesh(b, p, 1e-3, 2e-3)
    assert e.value.code == OV_E_NOMEM
    np.testing.assert_array_equal(b.vnodes, a.vnodes)
    np.testing.assert_array_equal(b.connections, a.connections)


@pytest.mark.parametrize("case", ["small", "origin", "zaxis", "outer"])
def test_eliminate_small_loops_matches_reference(ctx, case):
    b0 = T.TangleBuilder()
    rng = np.random.default_rng(11)
    for npts in (3, 4, 5, 6, 7, 12, 40):
        c = rng.uniform(-0.02, 0.02, 3)
        b0.add_ring(tuple(c), (0.0, 0.0, 1.0), 0.002 + 1e-4 * npts, npts)
    b0.add_ring((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.004, 30)          # near the origin
    b0.add_ring((0.0, 0.0, 0.03), (0.0, 0.0, 1.0), 0.003, 30)         # near the z axis
    b0.add_ring((0.09, 0.0, 0.0), (1.0, 0.0, 0.0), 0.004, 30)         # far from the z axis
    t = b0.build((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1), holes=50, seed=2)
    kw = {"small": {}, "origin": dict(eliminate_origin_loops=1, eliminate_loops_origin_cutoff=0.01),
          "zaxis": dict(eliminate_zaxis_loops=1, eliminate_loops_zaxis_cutoff=0.005),
          "outer": dict(eliminate_outer_loops=1, eliminate_outer_loops_cutoff=0.06)}[case]
    tp = T.default_topology(**kw)
    a, b = t.copy(), t.copy()
    Reference().eliminate_small_loops(a, tp.small_loop_cutoff,
                                      (tp.eliminate_origin_loops, tp.eliminate_loops_origin_cutoff),
                                      (tp.eliminate_zaxis_loops, tp.eliminate_loops_zaxis_cutoff),
                                      (tp.eliminate_outer_loops, tp.eliminate_outer_loops_cutoff))
    ctx.eliminate_small_loops(b, tp)
    np.testing.assert_array_equal(b.status[:, 0], a.status[:, 0])
    np.testing.assert_array_equal(b.connections, a.connections)
    print(case, "removed", len(t.active) - len(a.active), "points", flush=True)
    assert len(a.active) < len(t.active)


def test_enforce_boundaries_matches_reference(ctx):
    t = uneven_tangle(seed=8)
    act = t.active
    rng = np.random.default_rng(3)
    t.vnodes[act] += rng.uniform(-0.25, 0.25, (len(act), 3))
    a, b = t.copy(), t.copy()
    Reference().enforce_boundaries(a)
    ctx.enforce_boundaries(b)
    np.testing.assert_array_equal(b.vnodes[act], a.vnodes[act])


@pytest.mark.parametrize("prec,bh", [(T.PREC_FP64, 0), (T.PREC_MIXED, 0), (T.PREC_FP64, 1), (T.PREC_MIXED, 1)])
def test_run_matches_host_sequence(ctx, prec, bh):
    """ov_gpu_run over 4 steps equals the same steps driven from the host
    with the single-call entry points (ov_gpu_rk4_step,
    ov_gpu_update_tangents_normals, ov_gpu_remesh,
    ov_gpu_eliminate_small_loops, ov_gpu_enforce_boundaries) bit for bit:
    the device rebuilds the point stream exactly as the host does, and with
    the octree (bh) the tree's insertion order as well."""
    p = params(precision=prec, near_radius=0.01, use_BH=bh, BH_resolution=0.5)
    tp = T.default_topology(dl_min=1e-3, dl_max=2e-3)
    t = uneven_tangle(seed=21)
    a, b = t.copy(), t.copy()
    dt = 1e-5
    for k in range(4):
        ctx.rk4_step(a, k * dt, dt, p)
        ctx.update_tangents_normals(a)
        ctx.remesh(a, p, tp.dl_min, tp.dl_max)
        ctx.eliminate_small_loops(a, tp)
        ctx.enforce_boundaries(a)
    ctx.run(b, 0.0, dt, 4, p, tp)
    d = _same_topology(a, b)
    act = a.active
    print(f"run vs host sequence: {len(t.active)} -> {len(act)} points, position diff {d}", flush=True)
    assert d == 0
    np.testing.assert_array_equal(b.dxi[act], a.dxi[act])


def test_run_matches_reference_main_loop(ctx):
    """Three steps of the reference's loop body with the reference's own
    functions on the CPU (update_tangle, rk4_step, update_tangents_normals,
    remesh, eliminate_small_loops, enforce_boundaries) against ov_gpu_run
    in FP64: identical links and status, positions to 1e-14 cm."""
    R = Reference()
    p = params()
    tp = T.default_topology(dl_min=1e-3, dl_max=2e-3)
    R.set_params(p)
    t = uneven_tangle(seed=34, n_rings=3)
    a, b = t.copy(), t.copy()
    dt = 1e-5
    n0 = len(t.active)
    for k in range(3):
        R.step(a, k * dt, dt, p)
        R.update_tangents_normals(a)
        R.remesh(a, tp.dl_min, tp.dl_max, need_free=n0)
        R.eliminate_small_loops(a, tp.small_loop_cutoff)
        R.enforce_boundaries(a)
    ctx.run(b, 0.0, dt, 3, p, tp)
    d = _same_topology(a, b)
    print(f"run vs reference loop: {n0} -> {len(a.active)} points, position diff {d}", flush=True)
    assert d < 1e-14
