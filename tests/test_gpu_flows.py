"""Imposed flows on the GPU (every entry of the reference's flow table,
src/external_velocity.c:85-512) against the FP64 oracle, whose flows are
pinned against the reference's own functions in tests/test_flows.py.

Tolerance: FP64 path vs oracle 1e-11 (max |dv|/|v|), as the other FP64
parity tests; the device's cos/sin/exp may differ from glibc's by an ulp or
two, far below it.  Positions after RK4 steps: 1e-15 absolute."""
import numpy as np
import pytest

from openvort_amd import tangle as T
from openvort_amd.gpu import OV_E_NONFINITE, OVError
from tests import adjoint_cases as AC
from tests.flow_cases import TIME_DEPENDENT, flow_params
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-11


@pytest.mark.parametrize("name", list(T.FLOWS))
def test_velocities_with_flow(ctx, oracle, name):
    """The flow as vs and as vn (mutual friction on), at t = 1.3e-4 s."""
    t = T.random_rings(1001, 0.1, n_rings=4)
    p = flow_params(vs=name, vn=name)
    a, b = t.copy(), t.copy()
    oracle.update_tangle(a, 1.3e-4, p)
    ctx.velocities(b, 1.3e-4, p, geometry=True)
    act = a.active
    e = rel_err(b.vels[act], a.vels[act])
    print(name, "vels err", e, flush=True)
    assert e[0] < TOL and rel_err(b.vs[act], a.vs[act])[0] < TOL


@pytest.mark.parametrize("vs,vn", [("oscillating", "rotosc"), ("simple", "constant acceleration"),
                                   ("coscos-divfree", "DG"), ("cylindrical", "cos-divfree")])
def test_rk4_run_with_flows(ctx, oracle, vs, vn):
    """Five device-resident steps against the oracle stepped at accumulated
    times (t += dt, main.c:127).  Flows that depend on t run eager steps;
    the others replay a captured step."""
    t = T.random_rings(1003, 0.1, n_rings=3)
    p = flow_params(vs=vs, vn=vn)
    a, b = t.copy(), t.copy()
    time = 2e-5
    for _ in range(5):
        oracle.step(a, time, 5e-6, p)
        time += 5e-6
    ctx.rk4_run(b, 2e-5, 5e-6, 5, p)
    d = np.abs(a.vnodes - b.vnodes).max()
    mode = "eager" if {vs, vn} & TIME_DEPENDENT else "graph"
    print(vs, vn, mode, "5-step position diff", d, flush=True)
    assert d < 1e-15
    moved = np.abs(b.vnodes - t.vnodes).max()
    assert moved > 1e-7


def test_batch_with_flows(ctx):
    """Three tangles with different flows (one depends on t) in one batch
    against the same tangles run one by one."""
    tangles = [T.random_rings(1004 + k, 0.1, n_rings=2) for k in range(3)]
    params = [flow_params(vn="coscos"), flow_params(vn="rotosc"), flow_params(vs="spherical", vn="simple")]
    batch = [x.copy() for x in tangles]
    ctx.rk4_run_batch(batch, 1e-5, 5e-6, 3, params)
    for k, (x, p) in enumerate(zip(tangles, params)):
        y = x.copy()
        ctx.rk4_run(y, 1e-5, 5e-6, 3, p)
        d = np.abs(y.vnodes - batch[k].vnodes).max()
        print("tangle", k, "batch vs single", d, flush=True)
        assert d < 1e-15


def _fd_pair(run, t, w, h, seed):
    """Central difference of sum w . run(X) along a seeded direction."""
    act = t.active
    d = np.zeros((t.N, 3))
    d[act] = np.random.default_rng(seed).standard_normal((len(act), 3))
    a, b = t.copy(), t.copy()
    a.vnodes += h * d
    b.vnodes -= h * d
    return float(((run(a) - run(b))[act] * w[act]).sum() / (2 * h)), d


@pytest.mark.parametrize("name", list(T.FLOWS))
def test_velocity_vjp_with_flow(ctx, oracle, name):
    """The adjoint carries each flow's spatial Jacobian: the flow as vs and
    as vn (mutual friction on) at t = 1.3e-4 s, against central differences
    of the oracle (h = 1e-9 cm, 1e-6 relative) on a 64-point wavy ring (on a
    dense tangle the differences' rounding noise exceeds 1e-6)."""
    t = AC._wavy_ring(centre=(0.05, 0.04, 0.06))
    p = flow_params(vs=name, vn=name)
    w = np.random.default_rng(61).standard_normal((t.N, 3))
    g = ctx.velocity_vjp(t, 1.3e-4, p, w)

    def vels(x):
        oracle.update_tangle(x, 1.3e-4, p)
        return x.vels

    for s in range(2):
        fd, d = _fd_pair(vels, t, w, 1e-9, 400 + s)
        ad = float((g * d).sum())
        print(name, s, "fd", fd, "ad", ad, flush=True)
        assert abs(fd - ad) <= 1e-6 * abs(ad)


@pytest.mark.parametrize("vs,vn", [("oscillating", "rotosc"), ("coscos-divfree", "constant acceleration")])
def test_rk4_run_vjp_with_flows(ctx, vs, vn):
    """Six steps from t0 = 2e-5 s with flows that depend on t (eager
    segments): the adjoint against central differences of the GPU forward."""
    t = AC._wavy_ring(centre=(0.05, 0.04, 0.06))
    p = flow_params(vs=vs, vn=vn)
    w = np.random.default_rng(62).standard_normal((t.N, 3))
    g = ctx.rk4_run_vjp(t, 2e-5, 5e-6, 6, p, w, checkpoint_every=4)

    def run(x):
        ctx.rk4_run(x, 2e-5, 5e-6, 6, p)
        return x.vnodes

    for s in range(2):
        fd, d = _fd_pair(run, t, w, 1e-9, 500 + s)
        ad = float((g * d).sum())
        print(vs, vn, s, "fd", fd, "ad", ad, flush=True)
        assert abs(fd - ad) <= 1e-6 * abs(ad)


def test_nonfinite_flow_is_reported(ctx, oracle):
    """The oscillating slit divides by sqrt(x^2 + z^2) (external_velocity.c:
    313-314): a point on the y axis gives NaN in the reference; here the
    velocity is written and the call returns OV_E_NONFINITE."""
    t = T.ring(64)   # point 0 is (0, -0.01, 0)
    p = flow_params(vn="oscillating slit")
    b = t.copy()
    with pytest.raises(OVError) as e:
        ctx.velocities(b, 1e-4, p)
    assert e.value.code == OV_E_NONFINITE
    assert not np.isfinite(b.vels[0]).all() and np.isfinite(b.vels[1:]).all()
