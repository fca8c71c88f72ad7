"""Small tangles for the adjoint checks (CPU-sized: the torch restatement
sums every pair densely)."""
import numpy as np

from openvort_amd import tangle as T
from tests.helpers import ring_params, tangle_params


def _wavy_ring(npts=64, r=0.01, amp=1.5e-3, seed=7, centre=(0.0, 0.0, 0.0)):
    t = T.kelvin_ring(seed=seed, npts=npts, r=r, amp_hi=amp)
    t.vnodes += np.asarray(centre)
    return t


def case(name):
    """(tangle, params) of a named case."""
    if name == "ring_mf_flows":
        # imposed flows, mutual friction, hyperfriction on some points only
        t = _wavy_ring()
        p = ring_params(hyperfriction=1, max_curvature_scale=0.5, vs=(0.01, 0.02, 0.0),
                        vn=(0.1, 0.0, 0.05))
        return t, p
    if name == "lia_only":
        return _wavy_ring(48, seed=9), ring_params(LIA_only=1, hyperfriction=1)
    if name == "periodic_holes":
        t = T.random_rings(31, 0.05, n_rings=3, spacing=1.2e-3, holes=5)
        return t, tangle_params()
    if name == "mirror":
        b = T.TangleBuilder()
        b.add_ring((0.5, 0.5, 0.02), (0.3, 0.0, 0.954), 0.01, 48)
        t = b.build((0, 0, 0), (1, 1, 1), (T.WALL_OPEN,) * 4 + (T.WALL_MIRROR, T.WALL_OPEN),
                    [(0, 0, -1, T.Z_L)])
        return t, ring_params()
    if name == "small_loops":
        # loops of 4 and 5 points: stencil slots repeat points
        b = T.TangleBuilder()
        b.add_ring((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.01, 40)
        b.add_ring((0.013, 0.002, 0.001), (0.2, 0.9, 0.4), 0.002, 5)
        b.add_ring((-0.004, 0.011, -0.003), (0.7, -0.1, 0.7), 0.0015, 4)
        return b.build(), ring_params()
    raise KeyError(name)


NAMES = ("ring_mf_flows", "lia_only", "periodic_holes", "mirror", "small_loops")


def seeded(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)
