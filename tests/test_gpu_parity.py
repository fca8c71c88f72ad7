"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Tolerances (max over points of |dv|/|v|):
  FP64 path vs oracle: 1e-11 (summation order only; pair terms agree to
  a few ulps, the reference's zero rules are decided in the same FP64 form);
  mixed path vs oracle: 1e-6 (the stated accuracy target of the mixed form).
"""
import numpy as np
import pytest

from openvort_amd import tangle as T
from openvort_amd.gpu import OV_E_INVALID, OV_E_UNSUPPORTED, OVError
from tests.helpers import rel_err, ring_params, tangle_params

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11
MIXED_TOL = 1e-6


def _both(t, fn_o, fn_g):
    a, b = t.copy(), t.copy()
    fn_o(a)
    fn_g(b)
    return a, b


def _check_vel(a, b, tol, geom=True):
    act = a.active
    if geom:
        assert rel_err(b.tangents[act], a.tangents[act])[0] < 1e-12
        assert rel_err(b.normals[act], a.normals[act])[0] < 1e-10
        np.testing.assert_allclose(b.dxi[act], a.dxi[act], rtol=1e-15, atol=0)
    e_vs = rel_err(b.vs[act], a.vs[act])
    e_v = rel_err(b.vels[act], a.vels[act])
    print("vs err", e_vs, "vels err", e_v, flush=True)
    assert e_vs[0] < tol and e_v[0] < tol


def test_ring_velocities_fp64(ctx, oracle):
    t = T.ring(256)
    p = ring_params()
    a, b = _both(t, lambda x: oracle.update_tangle(x, 0.0, p),
                 lambda x: ctx.velocities(x, 0.0, p, geometry=True))
    _check_vel(a, b, FP64_TOL)


def test_ring_velocities_mixed(ctx, oracle):
    t = T.ring(256)
    p = ring_params(precision=T.PREC_MIXED, near_radius=1e-3)
    a, b = _both(t, lambda x: oracle.update_tangle(x, 0.0, p),
                 lambda x: ctx.velocities(x, 0.0, p, geometry=True))
    _check_vel(a, b, MIXED_TOL)


@pytest.mark.parametrize("n_images", [26, 6, 0])
def test_periodic_tangle_fp64(ctx, oracle, n_images):
    t = T.random_rings(1001, 0.1, n_rings=6, n_images=n_images)
    p = tangle_params()
    a, b = _both(t, lambda x: oracle.update_tangle(x, 0.0, p),
                 lambda x: ctx.velocities(x, 0.0, p, geometry=True))
    _check_vel(a, b, FP64_TOL)


@pytest.mark.parametrize("rnear", [0.002, 0.005, 0.02])
def test_periodic_tangle_mixed(ctx, oracle, rnear):
    t = T.random_rings(1001, 0.1, n_rings=6)
    p = tangle_params(precision=T.PREC_MIXED, near_radius=rnear)
    a, b = _both(t, lambda x: oracle.update_tangle(x, 0.0, p),
                 lambda x: ctx.velocities(x, 0.0, p, geometry=True))
    _check_vel(a, b, MIXED_TOL)


def test_caller_geometry_and_holes(ctx, oracle):
    """EMPTY slots scattered through the arrays; s', s'' supplied by the caller."""
    b0 = T.TangleBuilder()
    b0.add_ring((0.05, 0.05, 0.05), (0.0, 0.6, 0.8), 0.01, 300)
    b0.add_ring((0.06, 0.04, 0.05), (1.0, 0.0, 0.0), 0.007, 220)
    t = b0.build((0, 0, 0), (0.1, 0.1, 0.1), (T.WALL_PERIODIC,) * 6, T.periodic_images(26),
                 holes=333, seed=5)
    p = tangle_params(hyperfriction=1, dl_max=1e-3)
    oracle.update_tangents_normals(t)
    a, b = _both(t, lambda x: oracle.update_velocities(x, 0.0, p),
                 lambda x: ctx.velocities(x, 0.0, p, geometry=False))
    _check_vel(a, b, FP64_TOL, geom=False)
    # EMPTY slots are untouched
    emp = np.nonzero(t.status[:, 0] == T.EMPTY)[0]
    assert np.all(b.vels[emp] == 0) and np.all(b.vs[emp] == 0)


def test_lia_only(ctx, oracle):
    t = T.random_rings(7, 0.1, n_rings=3)
    p = tangle_params(LIA_only=1)
    a, b = _both(t, lambda x: oracle.update_tangle(x, 0.0, p),
                 lambda x: ctx.velocities(x, 0.0, p, geometry=True))
    _check_vel(a, b, 1e-13)


def test_rk4_step_ring(ctx, oracle):
    t = T.ring(256)
    p = ring_params()
    a, b = _both(t, lambda x: oracle.step(x, 0.0, 1e-5, p),
                 lambda x: ctx.rk4_step(x, 0.0, 1e-5, p))
    d = np.abs(a.vnodes - b.vnodes).max()
    print("rk4 ring position diff", d, flush=True)
    assert d < 1e-16
    _check_vel(a, b, FP64_TOL)


@pytest.mark.parametrize("variant", [T.RK_REFERENCE, T.RK_CLASSIC])
def test_rk4_step_periodic(ctx, oracle, variant):
    t = T.random_rings(1001, 0.1, n_rings=4)
    p = tangle_params(rk_variant=variant)
    a, b = _both(t, lambda x: oracle.step(x, 0.0, 5e-6, p),
                 lambda x: ctx.rk4_step(x, 0.0, 5e-6, p))
    d = np.abs(a.vnodes - b.vnodes).max()
    print("rk4 periodic position diff", d, flush=True)
    assert d < 1e-16


def test_rk4_caller_k1(ctx, oracle):
    """rk4_step alone: k1 is whatever the caller left in vels (reference
    vortex_dynamics.c:100)."""
    t = T.random_rings(3, 0.1, n_rings=3)
    p = tangle_params()
    oracle.update_tangle(t, 0.0, p)
    t.vels *= 0.5
    a, b = _both(t, lambda x: oracle.rk4_step(x, 0.0, 5e-6, p),
                 lambda x: ctx.rk4_step(x, 0.0, 5e-6, p, caller_k1=True))
    assert np.abs(a.vnodes - b.vnodes).max() < 1e-16


def test_rk4_run_ring_100_steps(ctx, oracle):
    """The 256-point ring, 100 steps of dt = 1e-5 s, device-resident."""
    t = T.ring(256)
    p = ring_params()
    a, b = t.copy(), t.copy()
    for k in range(100):
        oracle.step(a, k * 1e-5, 1e-5, p)
    ctx.rk4_run(b, 0.0, 1e-5, 100, p)
    d = np.abs(a.vnodes - b.vnodes).max()
    print("100-step ring position diff", d, flush=True)
    assert d < 1e-14
    # the ring translates along +z at close to the Kelvin speed
    dz = b.vnodes[:, 2].mean() - t.vnodes[:, 2].mean()
    assert 0.1 < dz / 1e-3 < 0.13


def test_rk4_run_enforce_boundaries(ctx, oracle):
    t = T.random_rings(11, 0.1, n_rings=3)
    p = tangle_params(vs=(300.0, 0, 0))
    a, b = t.copy(), t.copy()
    for k in range(5):
        oracle.step(a, k * 1e-6, 1e-6, p)
        oracle.enforce_boundaries(a)
    ctx.rk4_run(b, 0.0, 1e-6, 5, p, enforce_boundaries=True)
    assert np.any(np.abs(b.vnodes - t.vnodes) > 0.05)  # some points wrapped
    assert np.abs(a.vnodes - b.vnodes).max() < 1e-14


def test_errors(ctx):
    t = T.ring(32)
    p = ring_params()
    bad = t.copy()
    bad.status[3, 0] = T.PINNED          # pinned, but on no wall (pin_wall -1)
    with pytest.raises(OVError) as e:
        ctx.velocities(bad, 0.0, p)
    assert e.value.code == OV_E_INVALID
    bad = t.copy()
    bad.status[3] = (T.PINNED, T.Z_L)    # walls: not in the adjoint
    with pytest.raises(OVError) as e:
        ctx.velocity_vjp(bad, 0.0, p, np.ones((t.N, 3)))
    assert e.value.code == OV_E_UNSUPPORTED
    bad = t.copy()
    bad.connections[5, 0] = -1           # a free point must have both links
    with pytest.raises(OVError) as e:
        ctx.velocities(bad, 0.0, p)
    assert e.value.code == OV_E_INVALID
    bad = t.copy()
    bad.status[3, 0] = T.EMPTY
    with pytest.raises(OVError) as e:
        ctx.velocities(bad, 0.0, p)
    assert e.value.code == OV_E_INVALID
    # the context keeps working after an error
    ok = t.copy()
    ctx.velocities(ok, 0.0, p)
    assert np.isfinite(ok.vels).all()


def test_mixed_vs_fp64_larger(ctx, oracle):
    """~20k points, periodic-26: mixed vs FP64 on the device everywhere, and
    FP64 device vs the oracle on a seeded 512-point sample."""
    t = T.random_rings(1001, 0.1, n_rings=64)
    p64 = tangle_params()
    pmx = tangle_params(precision=T.PREC_MIXED, near_radius=0.005)
    a, b = t.copy(), t.copy()
    ctx.velocities(a, 0.0, p64)
    ctx.velocities(b, 0.0, pmx)
    act = t.active
    e = rel_err(b.vels[act], a.vels[act])
    print("N", len(act), "mixed vs fp64", e, flush=True)
    assert e[0] < MIXED_TOL
    rng = np.random.default_rng(2002)
    idx = np.sort(rng.choice(act, size=512, replace=False))
    o = t.copy()
    oracle.update_tangents_normals(o)
    vs, vel = oracle.velocities_sample(o, 0.0, p64, idx)
    e2 = rel_err(a.vels[idx], vel)
    print("fp64 vs oracle sample", e2, flush=True)
    assert e2[0] < FP64_TOL


from tests import fixtures  # noqa: E402


@pytest.mark.parametrize("name", [n for n in fixtures.names() if not n.startswith("octree")])
def test_gpu_matches_golden(ctx, name):
    """The HIP path against the reference's own outputs (tests/golden)."""
    t, p, z = fixtures.load(name)
    b = t.copy()
    ctx.velocities(b, 0.0, p, geometry=True)
    act = t.active
    assert rel_err(b.tangents[act], z["t0_tangents"][act])[0] < 1e-12
    assert rel_err(b.normals[act], z["t0_normals"][act])[0] < 1e-10
    e = rel_err(b.vels[act], z["t0_vels"][act])
    print(name, "vels vs reference", e, flush=True)
    assert e[0] < FP64_TOL
    assert rel_err(b.vs[act], z["t0_vs"][act])[0] < FP64_TOL
    if "out_vnodes_after" in z:
        dt = float(z["dt"])
        steps = int(z["steps"]) if "steps" in z else 1
        x = t.copy()
        ctx.rk4_run(x, 0.0, dt, steps, p)
        d = np.abs(x.vnodes[act] - z["out_vnodes_after"][act]).max()
        print(name, "positions after", steps, "steps:", d, flush=True)
        assert d < 1e-16 * max(1, steps)


@pytest.mark.parametrize("rnear", [0.005, 0.02])
def test_mixed_skips_wrapped_adjacent_segment(ctx, oracle, rnear):
    """A ring crossing the x faces of a periodic box: at the crossing the
    segment into a target wraps, so in the primary frame it lies one box
    length away, in an FP32 tile; the primary sum must still skip it."""
    b0 = T.TangleBuilder()
    b0.add_ring((0.001, 0.05, 0.05), (0.0, 0.0, 1.0), 0.01, 300)
    t = b0.build((0, 0, 0), (0.1, 0.1, 0.1), (T.WALL_PERIODIC,) * 6, T.periodic_images(26))
    t.wrap_into_box()
    p = tangle_params(precision=T.PREC_MIXED, near_radius=rnear)
    a, b = _both(t, lambda x: oracle.update_tangle(x, 0.0, p),
                 lambda x: ctx.velocities(x, 0.0, p, geometry=True))
    _check_vel(a, b, MIXED_TOL)


@pytest.mark.parametrize("precision", [T.PREC_FP64, T.PREC_MIXED])
def test_rk4_run_graph_replay_matches_eager_steps(ctx, precision):
    """ov_gpu_rk4_run replays steps 2..n as a captured HIP graph; it must give
    the bits of n separate ov_gpu_rk4_step calls (eager launches, host round
    trip between steps) and of a profiled (eager) run, on every call."""
    t = T.random_rings(61, 0.05, n_rings=3, spacing=1e-3)
    p = tangle_params(precision=precision)
    a = t.copy()
    for i in range(6):
        ctx.rk4_step(a, i * 2e-6, 2e-6, p)
    b = t.copy()
    ctx.rk4_run(b, 0.0, 2e-6, 6, p, profile=True)
    assert np.array_equal(a.vnodes, b.vnodes)
    for _ in range(2):
        c = t.copy()
        ctx.rk4_run(c, 0.0, 2e-6, 6, p)
        assert np.array_equal(c.vnodes, a.vnodes)
        assert np.array_equal(c.vels, b.vels)


@pytest.mark.parametrize("variant", [T.RK_REFERENCE, T.RK_CLASSIC])
def test_rk4_linear_flow_bitwise(ctx, oracle, variant):
    """LIA only on a straight periodic line in the linear "DG" flow (local
    induction exactly zero): the stage positions and the combine are each
    rounded as the reference rounds them (no fused multiply-add), so the GPU
    positions equal the oracle's bit for bit, and x + iy is multiplied by the
    variant's amplification polynomial in z = i rot dt."""
    from tests.test_oracle import _dg_line
    rot, vz, dt = 5e4, 0.3, 1e-5
    t = _dg_line()
    p = T.default_params(use_mutual_friction=0, hyperfriction=0, LIA_only=1, rk_variant=variant,
                         vs=("DG", {"rot": rot, "vz": vz}))
    a, b = _both(t, lambda x: oracle.step(x, 0.0, dt, p), lambda x: ctx.rk4_step(x, 0.0, dt, p))
    np.testing.assert_array_equal(b.vnodes, a.vnodes)
    z = 1j * rot * dt
    amp = (1 + z + 5 * z**2 / 12 + z**3 / 8 + z**4 / 48) if variant == T.RK_REFERENCE else \
        (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24)
    w0 = t.vnodes[:, 0] + 1j * t.vnodes[:, 1]
    np.testing.assert_allclose(b.vnodes[:, 0] + 1j * b.vnodes[:, 1], amp * w0, rtol=1e-14, atol=0)
