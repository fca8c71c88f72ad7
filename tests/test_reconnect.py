"""Reconnection candidates: the oracle's restatement of the pass-invariant
tests of the reference's pair loop, pinned by an independent numpy brute
force and by the reference's own reconnect pass (oracle/_ref): driving the
reference's dynamic tests and rewiring with the candidate list must leave the
tangle exactly as reconnect() does.  CPU only."""
import math

import numpy as np
import pytest

from openvort_amd import tangle as T
from oracle.oracle import Reference
from tests.helpers import ring_params, tangle_params

needs_ref = pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built")


def touching_rings():
    """Two rings touching at (0.01, 0, 0) with perpendicular tangents, plus a
    third ring crossing both."""
    b = T.TangleBuilder()
    b.add_ring((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.01, 160)
    b.add_ring((0.0199, 0.0, 0.0), (0.0, 1.0, 0.0), 0.01, 160)
    b.add_ring((0.01, 0.0, 0.004), (1.0, 0.2, 0.3), 0.006, 90)
    return b.build(), ring_params()


def dense_periodic(seed=71, holes=0):
    t = T.random_rings(seed, 0.02, n_rings=6, spacing=5e-4, r_lo=0.002, r_hi=0.006, holes=holes)
    return t, tangle_params()


CASES = {
    "touching_rings": touching_rings,
    "dense_periodic": dense_periodic,
    "dense_periodic_holes": lambda: dense_periodic(73, holes=7),
}


def numpy_candidates(t, rec_dist, rec_angle):
    """Independent restatement: all pairs at once, the reference's operation
    order (sums from zero, x then y then z; no fused multiply-add in numpy)."""
    act = np.nonzero(t.status[:, 0] == T.FREE)[0]
    X, Tg = t.vnodes[act], t.tangents[act]
    lo, hi = np.asarray(t.box_lo, float), np.asarray(t.box_hi, float)
    L = hi - lo
    d = X[None, :, :] - X[:, None, :]           # b - a for a = row, b = column
    w = X[None, :, :].repeat(len(act), 0)
    for i in range(3):
        if t.walls[2 * i] == T.WALL_PERIODIC:
            w[..., i] = np.where(d[..., i] > L[i] / 2, w[..., i] - L[i],
                                 np.where(d[..., i] < -L[i] / 2, w[..., i] + L[i], w[..., i]))
    v = X[:, None, :] - w
    dist = np.sqrt(v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1] + v[..., 2] * v[..., 2])
    nt = np.sqrt(Tg[:, 0] * Tg[:, 0] + Tg[:, 1] * Tg[:, 1] + Tg[:, 2] * Tg[:, 2])
    dot = (Tg[:, None, 0] * Tg[None, :, 0] + Tg[:, None, 1] * Tg[None, :, 1]) + Tg[:, None, 2] * Tg[None, :, 2]
    cosa = dot / nt[:, None] / nt[None, :]
    ok = (dist <= rec_dist) & (nt[:, None] >= 0.5) & (nt[None, :] >= 0.5) & ~(cosa > math.cos(rec_angle))
    iu = np.triu(np.ones_like(ok), 1)
    k, l = np.nonzero(ok & iu.astype(bool))
    return np.stack([act[k], act[l]], 1).astype(np.int32)


def prepared(name, oracle):
    t, p = CASES[name]()
    oracle.update_tangents_normals(t)
    return t, p


@pytest.mark.parametrize("rec_dist,rec_angle", [(1e-3, 0.087), (6e-4, 0.3), (2e-3, 0.087)])
@pytest.mark.parametrize("name", list(CASES))
def test_oracle_candidates_match_numpy(oracle, name, rec_dist, rec_angle):
    t, _ = prepared(name, oracle)
    a = oracle.reconnection_candidates(t, rec_dist, rec_angle)
    b = numpy_candidates(t, rec_dist, rec_angle)
    print(name, rec_dist, rec_angle, "candidates", len(a))
    if rec_dist >= 1e-3:
        assert len(a) > 0
    assert np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("rec_dist", [1e-3, 6e-4])
@pytest.mark.parametrize("name", list(CASES))
def test_candidate_pass_equals_reference_reconnect(oracle, name, rec_dist):
    """reconnect() of the reference vs the candidate-driven pass: same
    reconnection count and the same links afterwards."""
    ref = Reference()
    t, p = prepared(name, oracle)
    a, b = t.copy(), t.copy()
    na = ref.reconnect(a, 0.0, p, rec_dist, 0.087)
    cand = oracle.reconnection_candidates(b, rec_dist, 0.087)
    nb = ref.reconnect_candidates(b, 0.0, p, cand)
    print(name, rec_dist, "reconnections", na, nb, "candidates", len(cand))
    assert na == nb
    assert np.array_equal(a.connections, b.connections)


@needs_ref
def test_reconnections_happen_in_the_cases(oracle):
    """The equivalence above is not vacuous: these tangles reconnect."""
    ref = Reference()
    total = 0
    for name in CASES:
        t, p = prepared(name, oracle)
        total += ref.reconnect(t, 0.0, p, 1e-3, 0.087)
    assert total > 0
