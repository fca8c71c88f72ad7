"""Edge cases of the velocity and RK4 entry points through the C ABI:
empty tangles, tangles of EMPTY slots only, the smallest loops, tile
boundaries (63 / 64 / 65 points), and a context reused across tangles whose
sizes grow and shrink (device buffers grown x2 and kept).  FP64 results
against the oracle to 1e-11 relative, as in test_gpu_parity.py.
"""
import numpy as np
import pytest

from openvort_amd import tangle as T
from tests.helpers import rel_err, ring_params, tangle_params

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11


def _vs_oracle(ctx, oracle, t, p):
    a, b = t.copy(), t.copy()
    oracle.update_tangle(a, 0.0, p)
    ctx.velocities(b, 0.0, p)
    act = t.active
    e = rel_err(b.vels[act], a.vels[act])
    assert e[0] < FP64_TOL, e
    return b


def test_empty_and_all_empty_tangles(ctx):
    p = ring_params()
    empty = T.Tangle(np.zeros((0, 3)), np.zeros((0, 2), dtype=np.int32), np.zeros((0, 2), dtype=np.int32))
    assert ctx.velocities(empty, 0.0, p) == 0
    assert ctx.rk4_step(empty, 0.0, 1e-5, p) == 0
    st = np.tile(np.array([[T.EMPTY, -1]], dtype=np.int32), (10, 1))
    holes = T.Tangle(np.zeros((10, 3)), np.full((10, 2), -1, dtype=np.int32), st)
    before = holes.vnodes.copy()
    assert ctx.velocities(holes, 0.0, p) == 0
    assert ctx.rk4_run(holes, 0.0, 1e-5, 3, p) == 0
    np.testing.assert_array_equal(holes.vnodes, before)


@pytest.mark.parametrize("npts", [3, 4, 5])
def test_smallest_loops(ctx, oracle, npts):
    """Loops of 3-5 points: the five-point stencil wraps onto the loop
    itself (s_{+2} = s_{-1} on a 3-point loop), as the reference's
    step_node walks it."""
    b = T.TangleBuilder()
    b.add_ring((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.01, npts)
    b.add_ring((0.02, 0.01, 0.0), (0.3, 0.4, 0.866), 0.008, 40)
    t = b.build((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1))
    _vs_oracle(ctx, oracle, t, ring_params())


@pytest.mark.parametrize("npts", [63, 64, 65, 127, 129])
def test_tile_boundaries(ctx, oracle, npts):
    """Point counts around the 64-point tile size: partial last tiles,
    a loop closing exactly at a tile boundary, periodic images."""
    t = T.ring(npts)
    _vs_oracle(ctx, oracle, t, ring_params())
    b = T.TangleBuilder()
    b.add_ring((0.05, 0.05, 0.05), (0.0, 1.0, 0.0), 0.01, npts)
    b.add_ring((0.02, 0.08, 0.03), (0.0, 0.0, 1.0), 0.006, 23)
    tp = b.build((0, 0, 0), (0.1, 0.1, 0.1), (T.WALL_PERIODIC,) * 6, T.periodic_images(26))
    _vs_oracle(ctx, oracle, tp, tangle_params())


def test_context_reuse_across_sizes(ctx, oracle):
    """One context for tangles of ~100, ~22,000, 14 and ~3,700 points (device
    buffers grow and are kept): each result matches the oracle on a seeded
    sample, and a repeated small call gives the same bits as the first."""
    p = tangle_params()
    first = None
    rng = np.random.default_rng(7)
    for n_rings, spacing in ((3, 2e-3), (64, 2e-4), (1, 4e-3), (24, 5e-4)):
        t = T.random_rings(3000 + n_rings, 0.1, n_rings=n_rings, spacing=spacing)
        b = t.copy()
        ctx.velocities(b, 0.0, p)
        act = t.active
        idx = np.sort(rng.choice(act, size=min(256, len(act)), replace=False))
        o = t.copy()
        oracle.update_tangents_normals(o)
        _, vel = oracle.velocities_sample(o, 0.0, p, idx)
        assert rel_err(b.vels[idx], vel)[0] < FP64_TOL
        if first is None:
            first = (t, b.vels.copy())
    t0, v0 = first
    again = t0.copy()
    ctx.velocities(again, 0.0, p)
    np.testing.assert_array_equal(again.vels, v0)
