"""Host-side configuration and frame I/O (include/openvort_host.h), CPU only.

* The reference's own example configuration (tests/golden/example_config.cfg,
  a data file of the reference) read key by key, with libconfig's typing: the
  integer it gives for the float key reconnection_angle_cutoff is not applied
  (configuration.c:258-259 reads it with config_lookup_float).
* libconfig's syntax cases and the errors the reference would abort on.
* Every flow of the reference's table: the reference's own v_conf_t (built by
  the compiled reference) mapped back by name to the ov_flow it came from.
* Frames: ov_frame_save writes the file the reference's save_tangle writes,
  byte for byte; ov_frame_load restores it (topology exactly, values to the
  format's 15 digits, exactly with 17); the reference's own load_tangle does
  not (it reads 13 of the 19 columns).
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from openvort_amd import host as H
from openvort_amd import tangle as T
from openvort_amd.gpu import OV_E_INVALID, OVError
from oracle.oracle import Reference

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(not Reference.available(), reason="reference library not built here")

BASE = '''domain = ([0.0, 0.0, 0.0], [0.1, 0.1, 0.1]);
boundaries = "periodic-26";
init_mode = "random loops";
init_loops_n = 3;
'''


def test_example_config_values():
    c = H.load_config(os.path.join(GOLDEN, "example_config.cfg"))
    assert tuple(c.box.bottom_left_back.p) == (-0.05, -0.05, 0.0)
    assert tuple(c.box.top_right_front.p) == (0.05, 0.05, 0.3)
    assert c.boundaries == b"wall-1-6"
    # wall-1-6: periodic in x and y, a mirror at z_L, open at z_H (configuration.c:360-364)
    assert tuple(c.box.wall) == (T.WALL_PERIODIC,) * 4 + (T.WALL_MIRROR, T.WALL_OPEN)
    assert c.n_images == 5
    imgs = [(tuple(c.images[j].shift), c.images[j].reflect) for j in range(5)]
    assert imgs == [((-1, 0, 0), -1), ((0, -1, 0), -1), ((0, 0, -1), 4), ((0, 1, 0), -1), ((1, 0, 0), -1)]
    assert c.use_mutual_friction == 1
    assert c.alpha == 0.481 and c.alpha_p == -2.412e-2
    assert c.init_mode == b"random loops" and c.init_loops_n == 50
    assert c.frame_shots == 5 and c.dt == 5e-05
    assert c.dl_min == 0.001 and c.dl_max == 0.002
    assert c.small_loop_cutoff == 5
    # "reconnection_angle_cutoff = 5;" is an integer: config_lookup_float
    # ignores it, the default 0.087 rad stays
    assert c.reconnection_angle_cutoff == 0.087
    assert c.rec_dist == 0.002
    assert c.num_threads == 2                      # no terminating ';' in the file
    assert c.loop_injection == 1 and c.loop_injection_frequency == 50.0
    for f, v0 in ((c.vn, 0.1305), (c.vs, -0.3695)):
        assert f.type == T.FLOWS["cos-divfree"][0]
        assert (f.a[0], f.a[1], f.a[2]) == (188.5, v0, 0.1)      # k, v0, l in table order
    assert c.vn_set and c.vs_set and not c.vb_set and c.vb.type == T.FLOW_NONE
    # the rest keeps the reference defaults (vortex_constants.c:24-95)
    d = H.default_config()
    for name in ("max_steps", "kappa", "use_BH", "BH_resolution", "hyperfriction", "hyperalpha",
                 "max_curvature_scale", "LIA_only", "pin_mode"):
        assert getattr(c, name) == getattr(d, name), name
    p = H.params_from_config(c, T.default_params(vortex_width=1.3e-8))
    assert p.alpha == 0.481 and p.dl_max == 0.002 and p.vortex_width == 1.3e-8
    assert p.vn.type == c.vn.type and p.vs.a[1] == -0.3695


@pytest.mark.parametrize("text,key,value", [
    ("alpha: 0.25;", "alpha", 0.25),                        # ':' as well as '='
    ("alpha = 0.25,", "alpha", 0.25),                       # ',' terminator
    ("alpha = 2.5e-1  // comment\n", "alpha", 0.25),
    ("/* a\nb */ alpha = .25;", "alpha", 0.25),
    ("alpha = 1;", "alpha", 0.1),                           # an int for a float key: ignored
    ("small_loop_cutoff = 7.0;", "small_loop_cutoff", 5),   # a float for an int key: ignored
    ("small_loop_cutoff = 0x10;", "small_loop_cutoff", 16),
    ("small_loop_cutoff = 7L;", "small_loop_cutoff", 5),    # a 64-bit int is not an int
    ("use_BH = TRUE;", "use_BH", 1),
    ("use_BH = 1;", "use_BH", 0),                           # an int for a bool key: ignored
    ("reconnection_angle_cutoff = 10.0;", "reconnection_angle_cutoff", math.pi / 180 * 10.0),
    ('wall_type = "sl" "ip";', "pin_mode", T.PINNED_SLIP),   # adjacent strings concatenate
])
def test_config_syntax_and_typing(text, key, value):
    c = H.parse_config(BASE + text)
    assert getattr(c, key) == value


def test_config_groups_and_flows():
    c = H.parse_config(BASE + '''
        vn_conf = { type = "oscillating"; strength = 2.0; frequency = 30.0;
                    wave = [1.0, 0.0, 0.0]; polarization = [0.0, 0.5, 0.0]; };
        vs_conf = { type = "simple"; external_v = [0.0, 0.0, -0.25]; };
        vb_conf = { type = "rotosc-boundary"; freq = 3.0; amp = 0.1; delta = 9.0; };
    ''')
    assert c.vn.type == T.FLOWS["oscillating"][0]
    assert (c.vn.a[0], c.vn.a[1]) == (2.0, 30.0)
    assert tuple(c.vn.w) == (1.0, 0.0, 0.0) and tuple(c.vn.q) == (0.0, 0.5, 0.0)
    assert c.vs.type == T.FLOW_SIMPLE and tuple(c.vs.v) == (0.0, 0.0, -0.25)
    # rotosc-boundary counts two parameters (external_velocity.c:156-163)
    assert c.vb.type == T.FLOWS["rotosc-boundary"][0] and (c.vb.a[0], c.vb.a[1], c.vb.a[2]) == (3.0, 0.1, 0.0)


@pytest.mark.parametrize("boundaries,n,walls", [
    ("periodic-26", 26, (1,) * 6), ("periodic-18", 18, (1,) * 6), ("periodic-6", 6, (1,) * 6),
    ("open", 0, (0,) * 6), ("wall-2-26", 6, (1, 1, 1, 1, 2, 2)), ("channel-z", 6, (2, 2, 2, 2, 1, 1)),
    ("periodic-z-open-xy", 2, (0, 0, 0, 0, 1, 1)), ("wall-1-open", 1, (0, 0, 0, 0, 2, 0)),
])
def test_boundary_presets(boundaries, n, walls):
    c = H.parse_config(BASE.replace("periodic-26", boundaries))
    assert c.n_images == n and tuple(c.box.wall) == walls
    if boundaries.startswith("periodic-") and n > 2:
        got = [tuple(c.images[j].shift) + (c.images[j].reflect,) for j in range(n)]
        assert got == [tuple(s) for s in T.periodic_images(n)]


@pytest.mark.parametrize("text,what", [
    ('domain = ([0.0, 0.0, 0.0], [1.0, 1.0, 1.0]);\ninit_mode = "random loops"; init_loops_n = 1;', "boundaries"),
    ('boundaries = "open"; init_mode = "random loops"; init_loops_n = 1;', "domain"),
    (BASE.replace("periodic-26", "torus"), "unknown boundary"),
    (BASE.replace('init_mode = "random loops";', ""), "initialization mode"),
    (BASE.replace("init_loops_n = 3;", ""), "initial condition"),
    (BASE + 'vn_conf = { type = "vortex street"; };', "unknown flow type"),
    (BASE + 'vn_conf = { type = "DG"; rot = 1.0; };', "could not find parameter vz"),
    (BASE + 'wall_type = "glue";', "wall type"),
    (BASE + "alpha = 0.1;\nalpha = 0.2;", "duplicate"),
    (BASE + 'x = [1.0, 2];', "mismatched"),
    (BASE + 'x = "open', "line 5"),
    (BASE + "x = ;", "line 5"),
])
def test_config_errors(text, what):
    with pytest.raises(OVError) as e:
        H.parse_config(text)
    assert e.value.code == OV_E_INVALID
    assert what.lower() in str(e.value).lower(), str(e.value)


@needs_ref
@pytest.mark.parametrize("name", sorted(T.FLOWS))
def test_flow_from_reference_vconf(name):
    """The reference's own v_conf_t of each table entry (built by the
    compiled reference from an ov_flow with distinct parameter values) maps
    back to that ov_flow."""
    R = Reference()
    code, pnames = T.FLOWS[name]
    params = {}
    for k, pn in enumerate(pnames):
        params[pn] = (0.5 + k, -1.0 - k, 2.0 * k) if pn in T.FLOW_VECTORS else 1.25 + k
    f = T.set_flow(T.Flow(), name, **params)
    buf = R.vconf(f)
    assert C.sizeof(H.VConf) == len(buf.raw)
    g = H.flow_from_vconf(C.addressof(buf))
    assert g.type == f.type
    n_used = 2 if name == "rotosc-boundary" else len(pnames)   # the table's n_params
    used = pnames[:n_used]
    scalars = [pn for pn in used if pn not in T.FLOW_VECTORS]
    for j, pn in enumerate(scalars):
        assert g.a[j] == f.a[j]
    for pn in used:
        if pn in T.FLOW_VECTORS:
            slot = T.FLOW_VECTORS[pn]
            assert tuple(getattr(g, slot)) == tuple(getattr(f, slot))


def _frame_tangle():
    b = T.TangleBuilder()
    b.add_ring((0.02, 0.03, 0.04), (0.0, 0.6, 0.8), 0.01, 40)
    b.add_ring((0.06, 0.05, 0.03), (1.0, 0.0, 0.0), 0.007, 25)
    t = b.build((0, 0, 0), (0.1, 0.1, 0.1), (T.WALL_PERIODIC,) * 6, T.periodic_images(26), holes=17, seed=3)
    rng = np.random.default_rng(9)
    act = t.active
    for a in ("vels", "tangents", "normals"):
        getattr(t, a)[act] = rng.standard_normal((len(act), 3)) * 1e-3
    t.dxi[act] = rng.uniform(1e-4, 2e-3, len(act))
    return t


@needs_ref
def test_frame_save_matches_reference_bytes(tmp_path):
    t = _frame_tangle()
    ours, ref = str(tmp_path / "ours.dat"), str(tmp_path / "ref.dat")
    H.save_frame(ours, t)
    Reference().save_tangle(t, ref)
    assert open(ours, "rb").read() == open(ref, "rb").read()


def _loops(t):
    """Each loop as the tuple of its positions' rows, started at its lowest row
    value, following forward links."""
    seen, out = set(), []
    for k in t.active:
        if k in seen:
            continue
        cyc, j = [], k
        while j not in seen:
            seen.add(j)
            cyc.append(j)
            j = t.connections[j, 0]
        out.append(cyc)
    return out


@pytest.mark.parametrize("digits", [15, 17])
def test_frame_round_trip(tmp_path, digits):
    t = _frame_tangle()
    path = str(tmp_path / "f.dat")
    H.save_frame(path, t, digits=digits)
    u = H.load_frame(path)
    assert u.N == len(t.active)
    # topology: the same loops, in the same point order
    lt, lu = _loops(t), _loops(u)
    assert [len(c) for c in lt] == [len(c) for c in lu]
    for ct, cu in zip(lt, lu):
        pt, pu = t.vnodes[ct], u.vnodes[cu]
        tol = 0 if digits == 17 else 1e-15
        for a in ("vnodes", "vels", "tangents", "normals"):
            x, y = getattr(t, a)[ct], getattr(u, a)[cu]
            if digits == 17:
                np.testing.assert_array_equal(x, y)
            else:
                np.testing.assert_allclose(y, x, rtol=1e-14, atol=0)
        np.testing.assert_allclose(u.dxi[cu], t.dxi[ct], rtol=tol * 10, atol=0)
    assert np.all(u.status[:, 0] == T.FREE)
    for k in range(u.N):        # links consistent
        assert u.connections[u.connections[k, 0], 1] == k


@needs_ref
def test_reference_load_tangle_misreads_its_own_frames(tmp_path):
    """load_tangle reads 13 columns of save_tangle's 19, so the second
    point's position comes from the first line's trailing columns."""
    t = _frame_tangle()
    path = str(tmp_path / "r.dat")
    Reference().save_tangle(t, path)
    n = len(t.active)
    rc, pos, links = Reference().load_tangle(path, 2 * n)
    ours = H.load_frame(path)
    assert not np.allclose(pos[1], ours.vnodes[1])
    np.testing.assert_allclose(pos[0], ours.vnodes[0], rtol=1e-15)


def test_frame_13_columns_closes_loops(tmp_path):
    """13-column rows (load_tangle's format): each vortex becomes a loop in
    file order."""
    path = tmp_path / "old.dat"
    lines = []
    for v, cnt in ((0, 4), (1, 3)):
        for k in range(cnt):
            row = [v] + [0.1 * v + 0.01 * k, 0.2, 0.3] + [0.0] * 9
            lines.append("\t".join(str(x) for x in row))
    path.write_text("\n".join(lines) + "\n")
    u = H.load_frame(str(path))
    assert u.N == 7
    assert list(u.connections[:, 0]) == [1, 2, 3, 0, 5, 6, 4]
    assert list(u.connections[:, 1]) == [3, 0, 1, 2, 6, 4, 5]


def test_frame_errors(tmp_path):
    bad = tmp_path / "bad.dat"
    bad.write_text("0\t1\t2\n")
    with pytest.raises(OVError):
        H.load_frame(str(bad))
