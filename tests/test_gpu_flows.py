"""Imposed flows on the GPU (every entry of the reference's flow table,
src/external_velocity.c:85-512) against the FP64 oracle, whose flows are
pinned against the reference's own functions in tests/test_flows.py.

Tolerance: FP64 path vs oracle 1e-11 (max |dv|/|v|), as the other FP64
parity tests; the device's cos/sin/exp may differ from glibc's by an ulp or
two, far below it.  Positions after RK4 steps: 1e-15 absolute."""
import numpy as np
import pytest

from openvort_amd import tangle as T
from openvort_amd.gpu import OV_E_NONFINITE, OV_E_UNSUPPORTED, OVError
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


def test_adjoint_rejects_varying_flows(ctx):
    t = T.ring(64)
    p = flow_params(vn="coscos")
    with pytest.raises(OVError) as e:
        ctx.velocity_vjp(t, 0.0, p, np.ones((t.N, 3)))
    assert e.value.code == OV_E_UNSUPPORTED


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
