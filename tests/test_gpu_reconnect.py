"""GPU reconnection candidates (ov_gpu_reconnection_candidates) against the
oracle's pair loop (pinned in tests/test_reconnect.py) — identical (k, l)
lists, bit for bit — and, driving the reference's own dynamic tests and
rewiring, against the reference's reconnect pass."""
import numpy as np
import pytest

from openvort_amd import tangle as T
from openvort_amd.gpu import OV_E_INVALID, OVError
from oracle.oracle import Reference
from tests.helpers import tangle_params
from tests.test_reconnect import CASES, prepared

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rec_dist,rec_angle", [(1e-3, 0.087), (6e-4, 0.3), (2e-3, 0.087), (5e-3, 0.087)])
@pytest.mark.parametrize("name", list(CASES))
def test_candidates_match_oracle(ctx, oracle, name, rec_dist, rec_angle):
    t, _ = prepared(name, oracle)
    g = ctx.reconnection_candidates(t, rec_dist, rec_angle)
    o = oracle.reconnection_candidates(t, rec_dist, rec_angle)
    print(name, rec_dist, rec_angle, "gpu", len(g), "oracle", len(o), flush=True)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("rec_dist", [0.02, 0.03, 0.06])
def test_candidates_few_periodic_cells(ctx, oracle, rec_dist):
    """Periodic axes cut into 2 cells, 1 cell, and rec_dist beyond the box."""
    t = T.random_rings(79, 0.05, n_rings=3, spacing=2e-3)
    oracle.update_tangents_normals(t)
    g = ctx.reconnection_candidates(t, rec_dist, 0.087)
    o = oracle.reconnection_candidates(t, rec_dist, 0.087)
    print("cells", rec_dist, len(g), len(o), flush=True)
    assert len(o) > 0 and np.array_equal(g, o)


def test_candidates_large_periodic(ctx, oracle):
    t = T.random_rings(83, 0.1, n_points_target=20000, spacing=2e-4, r_lo=0.002, r_hi=0.02)
    oracle.update_tangents_normals(t)
    g = ctx.reconnection_candidates(t, 1e-3, 0.087)
    o = oracle.reconnection_candidates(t, 1e-3, 0.087)
    print("20k points", t.N, "candidates", len(g), flush=True)
    assert len(o) > 0 and np.array_equal(g, o)


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_candidate_pass_equals_reference_reconnect(ctx, oracle, name):
    ref = Reference()
    t, p = prepared(name, oracle)
    a, b = t.copy(), t.copy()
    na = ref.reconnect(a, 0.0, p, 1e-3, 0.087)
    nb = ref.reconnect_candidates(b, 0.0, p, ctx.reconnection_candidates(b, 1e-3, 0.087))
    print(name, "reconnections", na, nb, flush=True)
    assert na == nb and np.array_equal(a.connections, b.connections)


def test_candidates_reject_bad_distance(ctx, oracle):
    t, _ = prepared("touching_rings", oracle)
    with pytest.raises(OVError) as e:
        ctx.reconnection_candidates(t, 0.0, 0.087)
    assert e.value.code == OV_E_INVALID
