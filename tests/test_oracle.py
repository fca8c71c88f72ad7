"""Pins of the FP64 oracle (CPU only).

1. Golden fixtures written from the reference's own C (tests/golden/).
2. The compiled reference itself (oracle/_ref), where it is built.
3. What the mathematics fixes: the straight-segment law in closed form and
   by quadrature, guard cases, the FD stencil on lines and regular polygons,
   the Kelvin ring speed, exact translation under RK4, periodic images as an
   explicitly replicated open tangle, sign reversal of the links, the octree
   at theta = 0.
"""
import math

import numpy as np
import pytest

from openvort_amd import tangle as T
from oracle.oracle import Oracle, Reference
from tests import fixtures
from tests.helpers import rel_err, ring_params, tangle_params

KAPPA = 9.97e-4
needs_ref = pytest.mark.skipif(not Reference.available(), reason="reference library not built here")


# ---- 1. golden fixtures -------------------------------------------------------

@pytest.mark.parametrize("name", fixtures.names())
def test_oracle_matches_golden(name, oracle):
    t, p, z = fixtures.load(name)
    oracle.update_tangle(t, 0.0, p)
    act = t.active
    assert rel_err(t.tangents[act], z["t0_tangents"][act])[0] < 1e-13
    assert rel_err(t.normals[act], z["t0_normals"][act])[0] < 1e-11
    np.testing.assert_array_equal(t.dxi[act], z["t0_dxi"][act])
    e_vs = rel_err(t.vs[act], z["t0_vs"][act])
    e_v = rel_err(t.vels[act], z["t0_vels"][act])
    assert e_vs[0] < 1e-11 and e_v[0] < 1e-11, (e_vs, e_v)
    if "out_vnodes_after" in z:
        dt = float(z["dt"])
        steps = int(z["steps"]) if "steps" in z else 1
        x = t.copy()
        x.vnodes[:] = z["in_vnodes"]
        for k in range(steps):
            oracle.step(x, k * dt, dt, p)
        assert np.abs(x.vnodes[act] - z["out_vnodes_after"][act]).max() < 1e-16


# ---- 2. the compiled reference --------------------------------------------------

@needs_ref
@pytest.mark.parametrize("case", ["ring", "periodic", "hyper", "images6", "images18"])
def test_oracle_matches_reference(case, oracle):
    R = Reference(threads=4)
    if case == "ring":
        t, p = T.ring(200, r=0.007), ring_params()
    elif case == "periodic":
        t, p = T.random_rings(31, 0.1, n_rings=5, holes=77), tangle_params()
    elif case == "hyper":
        t, p = T.random_rings(5, 0.1, n_rings=4, spacing=5e-4), tangle_params(hyperfriction=1, dl_max=0.0004)
    else:
        k = 6 if case == "images6" else 18
        t, p = T.random_rings(9, 0.1, n_rings=4, n_images=k), tangle_params()
    a, b = t.copy(), t.copy()
    oracle.update_tangle(a, 0.0, p)
    R.update_tangle(b, 0.0, p)
    act = t.active
    assert rel_err(a.vels[act], b.vels[act])[0] < 1e-12
    assert rel_err(a.vs[act], b.vs[act])[0] < 1e-12
    oracle.rk4_step(a, 0.0, 2e-6, p)
    R.rk4_step(b, 0.0, 2e-6, p)
    assert np.abs(a.vnodes[act] - b.vnodes[act]).max() < 1e-16


@needs_ref
@pytest.mark.parametrize("theta", [0.2, 0.4, 0.6, 0.9])
@pytest.mark.parametrize("quad", [0, 1])
def test_oracle_octree_matches_reference(theta, quad, oracle):
    R = Reference(threads=4)
    t = T.random_rings(41, 0.1, n_rings=6, spacing=4e-4)
    p = tangle_params(use_BH=1, BH_resolution=theta, BH_quadtree=quad)
    a, b = t.copy(), t.copy()
    oracle.update_tangle(a, 0.0, p)
    R.update_tangle(b, 0.0, p)
    act = t.active
    assert rel_err(a.vels[act], b.vels[act])[0] < 1e-12


@needs_ref
def test_oracle_enforce_boundaries_matches_reference(oracle):
    t = T.random_rings(3, 0.1, n_rings=4, n_images=26)
    t.vnodes[t.active] += np.array([0.13, -0.02, 0.051])
    a, b = t.copy(), t.copy()
    oracle.enforce_boundaries(a)
    Reference().enforce_boundaries(b)
    np.testing.assert_array_equal(a.vnodes, b.vnodes)


# ---- 3. closed forms and invariants -----------------------------------------------

def _textbook_segment(a, b, r, kappa):
    """(kappa / 4 pi h) (cos t1 - cos t2) e_phi for a straight segment a -> b:
    h = distance of r from the segment's line, t1, t2 the angles between the
    segment direction and r - a, r - b."""
    e = (b - a) / np.linalg.norm(b - a)
    ra, rb = r - a, r - b
    c1 = np.dot(e, ra) / np.linalg.norm(ra)
    c2 = np.dot(e, rb) / np.linalg.norm(rb)
    perp = ra - np.dot(ra, e) * e
    h = np.linalg.norm(perp)
    ephi = np.cross(e, perp / h)
    return kappa / (4 * math.pi * h) * (c1 - c2) * ephi


def test_segment_law_closed_form(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b, r = rng.normal(size=3), rng.normal(size=3), rng.normal(size=3) * 2
        v = oracle.segment_field(a, b, r, KAPPA)
        ref = _textbook_segment(a, b, r, KAPPA)
        assert np.allclose(v, ref, rtol=1e-11, atol=1e-18)


def test_segment_law_quadrature(oracle):
    """Gauss-Legendre quadrature of kappa/4pi int ds x (r - s)/|r - s|^3."""
    x, w = np.polynomial.legendre.leggauss(200)
    rng = np.random.default_rng(2)
    for _ in range(20):
        a, b = rng.normal(size=3), rng.normal(size=3)
        r = rng.normal(size=3) * 3
        s = a[None] + (b - a)[None] * (0.5 * (x[:, None] + 1))
        ds = (b - a) * 0.5
        d = r[None] - s
        integrand = np.cross(ds[None], d) / np.linalg.norm(d, axis=1)[:, None] ** 3
        q = KAPPA / (4 * math.pi) * (w[:, None] * integrand).sum(0)
        v = oracle.segment_field(a, b, r, KAPPA)
        assert np.allclose(v, q, rtol=1e-9, atol=1e-16)


def test_segment_law_guards(oracle):
    a, b = np.array([0.0, 0, 0]), np.array([1e-4, 0, 0])
    # target on the segment's line beyond b: R, R' parallel -> exactly 0
    assert np.all(oracle.segment_field(a, b, np.array([0.5, 0, 0]), KAPPA) == 0)
    # target within 1e-8 of an endpoint -> exactly 0
    assert np.all(oracle.segment_field(a, b, np.array([1e-4 + 5e-9, 0, 0]), KAPPA) == 0)
    assert np.all(oracle.segment_field(a, b, np.array([0, 3e-9, 0]), KAPPA) == 0)
    # just inside / just outside the 1e-8 band of the normalised dot product:
    # the angle between R and R' is about y / 2e-4 for a target at (2e-4, y, 0)
    for y, zero in ((2.6e-8, True), (3.0e-8, False), (2.7e-8, True), (2.9e-8, False)):
        r = np.array([2e-4, y, 0.0])
        R, Rp = a - r, b - r
        nd = np.dot(R, Rp) / np.linalg.norm(R) / np.linalg.norm(Rp)
        assert (abs(nd - 1) < 1e-8) == zero
        v = oracle.segment_field(a, b, r, KAPPA)
        assert np.all(v == 0) == zero


def test_fd_stencil_straight_irregular_line(oracle):
    """Collinear points at irregular spacing: chord length is the parameter,
    so s' is the unit direction and s'' = 0 exactly up to rounding."""
    rng = np.random.default_rng(3)
    sp = rng.uniform(0.5e-4, 2e-4, size=64)
    d = np.array([0.3, -0.4, 0.866])
    d /= np.linalg.norm(d)
    x = np.cumsum(sp)
    b = T.TangleBuilder()
    # a long straight run closed by a far detour (the stencil of interior
    # points only sees the straight part)
    pts = 0.5 + x[:, None] * d[None]
    detour = pts[-1] + np.array([0.0, 0.05, 0.0])
    back = pts[0] + np.array([0.0, 0.05, 0.0])
    b.add_loop(np.vstack([pts, detour, back]))
    t = b.build()
    oracle.update_tangents_normals(t)
    for k in range(3, 60):
        np.testing.assert_allclose(t.tangents[k], d, rtol=0, atol=1e-9)
        assert np.abs(t.normals[k]).max() < 1e-3 * (1 / sp.min())


def test_fd_stencil_regular_polygon(oracle):
    """Regular N-gon: equal chords c; the stencil reduces to the symmetric
    five-point formulas s' = (8(s1 - s-1) - (s2 - s-2)) / (12 c),
    s'' = (16(s1 + s-1) - (s2 + s-2) - 30 s0) / (12 c^2)."""
    N, R = 40, 0.01
    t = T.ring(N, r=R)
    oracle.update_tangents_normals(t)
    c = 2 * R * math.sin(math.pi / N)
    s = t.vnodes
    for k in range(N):
        sm2, sm1, s0, s1, s2 = (s[(k + j) % N] for j in (-2, -1, 0, 1, 2))
        tp = (8 * (s1 - sm1) - (s2 - sm2)) / (12 * c)
        npp = (16 * (s1 + sm1) - (s2 + sm2) - 30 * s0) / (12 * c * c)
        np.testing.assert_allclose(t.tangents[k], tp, rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(t.normals[k], npp, rtol=1e-9, atol=1e-8)
        np.testing.assert_allclose(t.dxi[k], c, rtol=1e-13)


def test_ring_speed_kelvin(oracle):
    """256-point ring, no friction: self-induced speed against Kelvin's
    kappa/(4 pi R) (ln(8R/a0) - 1/4); the discretisation limits agreement to
    the 1e-3 level."""
    R, a0 = 0.01, 1.3e-8
    t = T.ring(256, r=R)
    p = ring_params(use_mutual_friction=0)
    oracle.update_tangle(t, 0.0, p)
    vk = KAPPA / (4 * math.pi * R) * (math.log(8 * R / a0) - 0.25)
    vz = t.vels[:, 2].copy()
    assert np.allclose(vz, vz.mean(), rtol=1e-10)
    # the local cutoff at the chord lengths plus the non-adjacent segments
    # converge to Kelvin's speed with a constant offset of ~0.5 % (0.067 in
    # the logarithm); it does not shrink with N
    assert abs(vz.mean() / vk - 1) < 6e-3
    # the a0 dependence is the LIA term alone: dv = kappa/4pi ln(a1/a2) |s' x s''|
    t2 = T.ring(256, r=R)
    oracle.update_tangle(t2, 0.0, ring_params(use_mutual_friction=0, vortex_width=a0 * 10))
    dv = vz.mean() - t2.vels[:, 2].mean()
    assert abs(dv / (KAPPA / (4 * math.pi) * math.log(10) / R) - 1) < 1e-3
    # mutual friction: alpha' slows it by (1 - alpha'), alpha shrinks it
    p2 = ring_params()
    oracle.update_tangle(t, 0.0, p2)
    # (vn = 0, s' in the ring plane, |s'| = 1 up to the stencil's error)
    sp = np.linalg.norm(t.tangents, axis=1)
    assert np.abs(sp - 1).max() < 1e-4
    np.testing.assert_allclose(t.vels[:, 2], vz * (1 - 0.016 * sp * sp), rtol=1e-10)
    rad = np.einsum("ij,ij->i", t.vels[:, :2], t.vnodes[:, :2]) / R
    np.testing.assert_allclose(rad, -0.098 * sp * vz, rtol=1e-8)


def test_rk4_uniform_translation(oracle):
    """Straight periodic lines in a uniform superfluid flow, local
    induction only (zero for a straight line), no friction: every stage
    velocity equals the imposed flow, so one step moves every point by
    exactly dt v (weights 1/6 + 1/3 + 1/3 + 1/6 = 1)."""
    b = T.TangleBuilder()
    n = 50
    L = 0.1
    x = (np.arange(n) + 0.3) * (L / n)
    b.add_loop(np.stack([x, np.full(n, 0.03), np.full(n, 0.04)], 1))
    b.add_loop(np.stack([np.full(n, 0.07), x[::-1], np.full(n, 0.02)], 1))
    t = b.build((0, 0, 0), (L, L, L), (T.WALL_PERIODIC,) * 6, T.periodic_images(26))
    v = np.array([0.31, -0.7, 0.13])
    p = T.default_params(use_mutual_friction=0, hyperfriction=0, LIA_only=1, vs=tuple(v))
    x0 = t.vnodes.copy()
    oracle.step(t, 0.0, 1e-3, p)
    np.testing.assert_allclose(t.vnodes - x0, np.broadcast_to(v * 1e-3, x0.shape), rtol=0, atol=1e-14)


def test_periodic_images_equal_replicated_open_tangle(oracle):
    """Periodic-26 sum == open sum over 27 explicit copies of the tangle
    (copies at +L S, the primary in the middle), targets the primary's points."""
    L = 0.1
    b = T.TangleBuilder()
    b.add_ring((0.05, 0.05, 0.05), (0.0, 0.6, 0.8), 0.02, 150)
    b.add_ring((0.03, 0.07, 0.05), (1.0, 0.0, 0.0), 0.015, 110)
    per = b.build((0, 0, 0), (L, L, L), (T.WALL_PERIODIC,) * 6, T.periodic_images(26))
    rep = T.TangleBuilder()
    shifts = [(0, 0, 0)] + [s[:3] for s in T.periodic_images(26)]
    for s in shifts:
        off = L * np.array(s, dtype=float)
        for (start, cnt) in b.loops:
            rep.add_loop(per.vnodes[start:start + cnt] + off)
    big = rep.build((-5, -5, -5), (5, 5, 5))
    p = tangle_params()
    oracle.update_tangents_normals(per)
    oracle.update_velocities(per, 0.0, p)
    oracle.update_tangents_normals(big)
    idx = np.arange(per.N)
    vs, vel = oracle.velocities_sample(big, 0.0, p, idx)
    assert rel_err(per.vs, vs)[0] < 1e-12
    assert rel_err(per.vels, vel)[0] < 1e-12


def test_reversed_links_flip_vs(oracle):
    """Open boundaries: reversing every link reverses every segment, so vs
    changes sign (under periodic wrapping a reversed crossing segment would
    sit one box length away, outside the symmetry)."""
    t = T.random_rings(13, 0.1, n_rings=3, n_images=0)
    p = T.default_params(use_mutual_friction=0, hyperfriction=0, vortex_width=1.3e-8)
    a = t.copy()
    bt = t.copy()
    bt.connections[:] = bt.connections[:, ::-1]
    oracle.update_tangle(a, 0.0, p)
    oracle.update_tangle(bt, 0.0, p)
    act = t.active
    np.testing.assert_allclose(bt.vs[act], -a.vs[act], rtol=1e-10, atol=1e-14)


def test_octree_theta_zero_is_direct(oracle):
    t = T.random_rings(19, 0.1, n_rings=4, spacing=4e-4)
    p0 = tangle_params()
    pt = tangle_params(use_BH=1, BH_resolution=0.0)
    a, b = t.copy(), t.copy()
    oracle.update_tangle(a, 0.0, p0)
    oracle.update_tangle(b, 0.0, pt)
    act = t.active
    assert rel_err(b.vels[act], a.vels[act])[0] < 1e-11
    cells, pairs, mps = oracle.octree_stats()
    assert mps == 0 and cells > len(act)


# ---- RK variants against their amplification polynomials -----------------------

def _dg_line(x0=0.01, y0=0.02, n=40, L=0.1):
    """A straight periodic line along z at (x0, y0): its local induction is
    exactly zero (s' and s'' have no x or y component), so with LIA only, no
    friction and the "DG" flow v = rot (-y, x, 0) + vz z as the superfluid flow
    every point moves by v(r) alone, and the line stays straight."""
    b = T.TangleBuilder()
    z = (np.arange(n) + 0.37) * (L / n) - L / 2
    b.add_loop(np.stack([np.full(n, x0), np.full(n, y0), z], 1))
    return b.build((-L / 2,) * 3, (L / 2,) * 3, (T.WALL_PERIODIC,) * 6, ())


@pytest.mark.parametrize("variant", [T.RK_REFERENCE, T.RK_CLASSIC])
def test_rk4_linear_flow_amplification(oracle, variant):
    """In the linear flow d(x + iy)/dt = i rot (x + iy) one step multiplies
    x + iy by the scheme's amplification polynomial in z = i rot dt:
    1 + z + 5z^2/12 + z^3/8 + z^4/48 for the reference's k4 at x + dt/2 k3
    (vortex_dynamics.c:89), 1 + z + z^2/2 + z^3/6 + z^4/24 for the classic
    scheme.  At rot dt = 0.5 the two differ by 2e-2 relative, so a wrong
    offset, weight or stage fails."""
    rot, vz, dt = 5e4, 0.3, 1e-5
    t = _dg_line()
    p = T.default_params(use_mutual_friction=0, hyperfriction=0, LIA_only=1, rk_variant=variant,
                         vs=("DG", {"rot": rot, "vz": vz}))
    w0 = t.vnodes[:, 0] + 1j * t.vnodes[:, 1]
    z0 = t.vnodes[:, 2].copy()
    oracle.step(t, 0.0, dt, p)
    z = 1j * rot * dt
    if variant == T.RK_REFERENCE:
        amp = 1 + z + 5 * z**2 / 12 + z**3 / 8 + z**4 / 48
    else:
        amp = 1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24
    w1 = t.vnodes[:, 0] + 1j * t.vnodes[:, 1]
    np.testing.assert_allclose(w1, amp * w0, rtol=1e-14, atol=0)
    np.testing.assert_allclose(t.vnodes[:, 2] - z0, vz * dt, rtol=0, atol=2e-17)
    other = (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24) if variant == T.RK_REFERENCE else \
        (1 + z + 5 * z**2 / 12 + z**3 / 8 + z**4 / 48)
    assert abs(amp - other) / abs(amp) > 1e-2


def test_rk4_stage_times(oracle):
    """A time-dependent linear flow ("rotosc": rot(t) = amp cos(2 pi f t))
    fixes the stage times t, t + dt/2, t + dt/2, t + dt (vortex_dynamics.c:
    79-92): the step equals the scheme run in complex scalars with those
    times, and differs from it at every other placement of the k4 time."""
    amp_, f, dt, t0 = 4e4, 2e4, 1e-5, 3e-6
    lam = lambda tt: 1j * amp_ * math.cos(2 * math.pi * f * tt)

    def scheme(w, h4, t4):
        k1 = lam(t0) * w
        k2 = lam(t0 + dt / 2) * (w + dt / 2 * k1)
        k3 = lam(t0 + dt / 2) * (w + dt / 2 * k2)
        k4 = lam(t4) * (w + h4 * k3)
        return w + dt / 6 * k1 + dt / 3 * k2 + dt / 3 * k3 + dt / 6 * k4

    for variant, h4 in ((T.RK_REFERENCE, dt / 2), (T.RK_CLASSIC, dt)):
        t = _dg_line()
        p = T.default_params(use_mutual_friction=0, hyperfriction=0, LIA_only=1, rk_variant=variant,
                             vs=("rotosc", {"freq": f, "amp": amp_}))
        w0 = t.vnodes[:, 0] + 1j * t.vnodes[:, 1]
        oracle.step(t, t0, dt, p)
        w1 = t.vnodes[:, 0] + 1j * t.vnodes[:, 1]
        np.testing.assert_allclose(w1, scheme(w0, h4, t0 + dt), rtol=1e-13, atol=0)
        wrong = scheme(w0, h4, t0 + dt / 2)
        assert np.abs(w1 - wrong).max() / np.abs(w1).max() > 1e-4


# ---- the dipole multipole against the direct sum it expands -------------------------

def _loop_tangle(points):
    b = T.TangleBuilder()
    b.add_loop(points)
    return b.build((-1, -1, -1), (1, 1, 1))


def _direct(oracle, pts, closed, r):
    n = len(pts)
    v = np.zeros(3)
    for k in range(n if closed else n - 1):
        v += oracle.segment_field(pts[k], pts[(k + 1) % n], r, KAPPA)
    return v


def _far_errors(oracle, t, ids, seg_pts, closed, centre, direction, dists, mode):
    p = T.default_params(multipole=mode)
    out = []
    for D in dists:
        r = centre + D * direction
        ff = oracle.cell_far_field(t, ids, r, p)
        ex = _direct(oracle, seg_pts, closed, r)
        out.append(np.linalg.norm(ff - ex) / np.linalg.norm(ex))
    return np.array(out)


def test_dipole_multipole_closed_loop_convergence(oracle):
    """A cell holding one small closed loop (zero circulation): the dipole
    expansion's error against the direct sum over its segments falls like
    a/D (the omitted quadrupole relative to the dipole field), halving per
    doubling of D.  A flipped or transposed eps:M term, a wrong factor on
    3 (M R) x R or moments about segment starts instead of midpoints leave an
    error that does not fall with D (the first two of order 1, the last of
    order pi/N).  The reference's form (octree.c:303-328, quirks kept) is
    shown not to converge, which the test relies on to tell the forms apart."""
    a, n = 1e-3, 96
    ph = 2 * np.pi * np.arange(n) / n
    e1 = np.array([0.8, 0.6, 0.0])
    e2 = np.array([-0.36, 0.48, 0.8])
    e3 = np.cross(e1, e2)
    c0 = np.array([0.1, -0.05, 0.02])
    pts = c0 + a * (np.cos(ph)[:, None] * e1 + 0.6 * np.sin(ph)[:, None] * e2
                    + 0.3 * np.sin(2 * ph)[:, None] * e3 + 0.2 * np.cos(3 * ph)[:, None] * e1)
    t = _loop_tangle(pts)
    ids = np.arange(n)
    centre = pts.mean(axis=0)
    direction = np.array([0.3, -0.5, 0.81])
    direction /= np.linalg.norm(direction)
    dists = a * np.array([8, 16, 32, 64, 128])
    e = _far_errors(oracle, t, ids, pts, True, centre, direction, dists, T.MP_DIPOLE)
    ratios = e[1:] / e[:-1]
    assert np.all((ratios > 0.42) & (ratios < 0.58)), (e, ratios)
    assert e[-1] < 3e-2
    e_ref = _far_errors(oracle, t, ids, pts, True, centre, direction, dists, T.MP_REFERENCE)
    assert e_ref.min() > 0.1, e_ref


def test_dipole_multipole_open_chain_convergence(oracle):
    """A cell holding an open piece of filament (non-zero circulation Gamma):
    monopole plus exact first-order term, so the error falls like (L/D)^2,
    a quarter per doubling of D; a wrong sign of Gamma x R or of either
    first-order term, or start-point moments, leave errors of order L/D."""
    n = 24
    s = np.linspace(0, 1, n + 1)
    L = 2e-3
    c0 = np.array([-0.2, 0.1, 0.05])
    pts = c0 + L * np.stack([s, 0.3 * np.sin(3 * s), 0.2 * s * s], 1)
    # the cell holds points 0..n-1; the segment of point n-1 ends at point n
    t = _loop_tangle(np.vstack([pts, pts[::-1][1:-1] + np.array([0, 0, 0.5])]))
    ids = np.arange(n)
    centre = pts[:n].mean(axis=0)
    direction = np.array([-0.2, 0.9, 0.39])
    direction /= np.linalg.norm(direction)
    dists = L * np.array([8, 16, 32, 64, 128])
    e = _far_errors(oracle, t, ids, pts, False, centre, direction, dists, T.MP_DIPOLE)
    ratios = e[1:] / e[:-1]
    assert np.all((ratios > 0.18) & (ratios < 0.32)), (e, ratios)
