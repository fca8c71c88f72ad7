"""Tangle container and seeded workload generators.

``Tangle`` holds numpy arrays with exactly the memory layout of the
reference's ``struct tangle_state`` (reference include/tangle.h:82-106):
AoS ``double[3]`` per point for positions, velocities and the FD geometry,
``{forward, reverse}`` int pairs, ``{status, pin_wall}`` int pairs, the
domain box and the image list.  ``Tangle.cstruct()`` yields the
``ov_tangle_state`` view (include/openvort_gpu.h) of those arrays.

The generators build the workloads the path is measured on; all randomness
comes from a SplitMix64 stream so that a seed gives the same tangle on every
machine (up to libm rounding of sin/cos).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

# ---- ctypes mirror of include/openvort_gpu.h -------------------------------

EMPTY, FREE, PINNED, PINNED_SLIP = -1, 0, 1, 2
WALL_OPEN, WALL_PERIODIC, WALL_MIRROR = 0, 1, 2
X_L, X_H, Y_L, Y_H, Z_L, Z_H = 0, 1, 2, 3, 4, 5


class Vec3(C.Structure):
    _fields_ = [("p", C.c_double * 3)]


class DomainBox(C.Structure):
    _fields_ = [("bottom_left_back", Vec3), ("top_right_front", Vec3), ("wall", C.c_int * 6)]


class ImageTangle(C.Structure):
    _fields_ = [("shift", C.c_int * 3), ("reflect", C.c_int)]


class BoundaryImages(C.Structure):
    _fields_ = [("images", C.POINTER(ImageTangle)), ("n", C.c_int)]


class TangleState(C.Structure):
    _fields_ = [
        ("vnodes", C.c_void_p),
        ("vnodes_new", C.c_void_p),
        ("vels", C.c_void_p),
        ("vs", C.c_void_p),
        ("tangents", C.c_void_p),
        ("normals", C.c_void_p),
        ("dxi", C.c_void_p),
        ("recalculate", C.c_void_p),
        ("connections", C.c_void_p),
        ("status", C.c_void_p),
        ("box", DomainBox),
        ("bimg", BoundaryImages),
        ("N", C.c_int),
        ("next_free", C.c_int),
        ("total_free", C.c_int),
    ]


class Flow(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("v", C.c_double * 3),
        ("a", C.c_double * 4),
        ("w", C.c_double * 3),
        ("q", C.c_double * 3),
    ]


# The reference's flow table (src/external_velocity.c:85-174), in order: the
# name of each entry, its type code and its parameters by name -> slot of
# ov_flow ("external_v" -> v, scalars -> a[i] in table order, "wave" -> w,
# "polarization" -> q).
FLOWS = {
    "no flow": (0, []),
    "simple": (1, ["external_v"]),
    "spherical": (2, ["strength", "cutoff"]),
    "spherical and simple": (3, ["external_v", "strength", "cutoff"]),
    "cylindrical": (4, ["strength", "cutoff"]),
    "oscillating": (5, ["strength", "frequency", "wave", "polarization"]),
    "oscillating slit": (6, ["strength", "frequency", "cutoff"]),
    "coscos": (7, ["k", "v0", "l"]),
    "coscos-divfree": (8, ["k", "v0", "l"]),
    "cos-divfree": (9, ["k", "v0", "l"]),
    "simple-shear": (10, ["rate", "v0"]),
    "rotosc": (11, ["freq", "amp"]),
    "DG": (12, ["rot", "vz"]),
    "rotosc-boundary": (13, ["freq", "amp", "delta"]),
    "constant acceleration": (14, ["accel"]),
    "parabolic": (15, ["H", "amp"]),
}
FLOW_NONE, FLOW_SIMPLE = 0, 1
FLOW_VECTORS = {"external_v": "v", "wave": "w", "polarization": "q"}


def set_flow(f: Flow, name: str = "no flow", **params) -> Flow:
    """Fill an ov_flow with the reference flow `name` and its parameters
    (missing ones are 0, as the reference's table initialises them)."""
    if name not in FLOWS:
        raise ValueError(f"unknown flow {name!r}")
    code, names = FLOWS[name]
    unknown = set(params) - set(names)
    if unknown:
        raise ValueError(f"flow {name!r} has no parameters {sorted(unknown)}")
    f.type = code
    for i in range(3):
        f.v[i] = f.w[i] = f.q[i] = 0.0
    for i in range(4):
        f.a[i] = 0.0
    scalar = 0
    for pn in names:
        if pn in FLOW_VECTORS:
            dst = getattr(f, FLOW_VECTORS[pn])
            val = params.get(pn, (0.0, 0.0, 0.0))
            for i in range(3):
                dst[i] = float(val[i])
        else:
            f.a[scalar] = float(params.get(pn, 0.0))
            scalar += 1
    return f
class Topology(C.Structure):
    """ov_topology: settings of remesh and eliminate_small_loops (reference
    globals, vortex_constants.c:39-66)."""
    _fields_ = [
        ("dl_min", C.c_double), ("dl_max", C.c_double), ("small_loop_cutoff", C.c_int),
        ("eliminate_origin_loops", C.c_int), ("eliminate_loops_origin_cutoff", C.c_double),
        ("eliminate_zaxis_loops", C.c_int), ("eliminate_loops_zaxis_cutoff", C.c_double),
        ("eliminate_outer_loops", C.c_int), ("eliminate_outer_loops_cutoff", C.c_double),
    ]


def default_topology(**kw) -> Topology:
    """Reference defaults (vortex_constants.c:39-66)."""
    tp = Topology(1e-3, 5e-3, 5, 0, 3e-2, 0, 2e-2, 0, 1.0)
    for k, v in kw.items():
        setattr(tp, k, v)
    return tp


PREC_FP64, PREC_MIXED = 0, 1
RK_REFERENCE, RK_CLASSIC = 0, 1
MP_REFERENCE, MP_DIPOLE = 0, 1


class Params(C.Structure):
    _fields_ = [
        ("kappa", C.c_double),
        ("vortex_width", C.c_double),
        ("alpha", C.c_double),
        ("alpha_p", C.c_double),
        ("dl_max", C.c_double),
        ("LIA_only", C.c_int),
        ("use_mutual_friction", C.c_int),
        ("hyperfriction", C.c_int),
        ("use_BH", C.c_int),
        ("max_curvature_scale", C.c_double),
        ("hyperalpha", C.c_double),
        ("BH_resolution", C.c_double),
        ("BH_grain", C.c_double),
        ("BH_quadtree", C.c_int),
        ("precision", C.c_int),
        ("rk_variant", C.c_int),
        ("multipole", C.c_int),
        ("near_radius", C.c_double),
        ("vn", Flow),
        ("vs", Flow),
        ("vb", Flow),
    ]


def default_params(**kw) -> Params:
    """Reference defaults (src/vortex_constants.c:24-95), no imposed flows."""
    p = Params()
    p.kappa = 9.97e-4
    p.vortex_width = 1e-8
    p.alpha = 0.1
    p.alpha_p = 0.01
    p.dl_max = 5e-3
    p.LIA_only = 0
    p.use_mutual_friction = 1
    p.hyperfriction = 1
    p.use_BH = 0
    p.max_curvature_scale = 0.2
    p.hyperalpha = 0.5
    p.BH_resolution = 0.1
    p.BH_grain = 1e-8
    p.BH_quadtree = 0
    p.precision = PREC_FP64
    p.rk_variant = RK_REFERENCE
    p.multipole = MP_REFERENCE
    p.near_radius = 0.0
    for name, val in kw.items():
        if name in ("vn", "vs", "vb"):
            # None: "no flow"; a 3-vector: "simple"; (flow name, {params}) or
            # a flow name alone: that entry of the flow table
            f = getattr(p, name)
            if val is None:
                set_flow(f)
            elif isinstance(val, str):
                set_flow(f, val)
            elif isinstance(val, tuple) and len(val) == 2 and isinstance(val[0], str):
                set_flow(f, val[0], **val[1])
            else:
                set_flow(f, "simple", external_v=val)
        else:
            setattr(p, name, val)
    return p


# ---- images (reference src/boundary_images.c:26-163) -----------------------

def periodic_images(kind: int):
    """Shift lists of the periodic presets: 6 face, 18 face+edge, 26 all
    neighbour cells, in the reference's order (boundary_images.c:29-47)."""
    out = []
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            for sz in (-1, 0, 1):
                nz = abs(sx) + abs(sy) + abs(sz)
                if nz == 0:
                    continue
                if kind == 6 and nz != 1:
                    continue
                if kind == 18 and nz == 3:
                    continue
                out.append((sx, sy, sz, -1))
    if kind == 6:  # the 6-image table is listed -x, -y, -z, +z, +y, +x
        order = [(-1, 0, 0), (0, -1, 0), (0, 0, -1), (0, 0, 1), (0, 1, 0), (1, 0, 0)]
        out = [(a, b, c, -1) for a, b, c in order]
    return out


# ---- SplitMix64 -------------------------------------------------------------

class SplitMix64:
    """SplitMix64 generator (Steele, Lea, Flood 2014); uniform doubles from
    the top 53 bits."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.s = seed & self.MASK

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0) -> float:
        return a + (b - a) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def unit_vector(self):
        z = self.uniform(-1.0, 1.0)
        ph = self.uniform(0.0, 2 * math.pi)
        s = math.sqrt(max(0.0, 1 - z * z))
        return (s * math.cos(ph), s * math.sin(ph), z)


# ---- the tangle -------------------------------------------------------------

@dataclass
class Tangle:
    vnodes: np.ndarray
    connections: np.ndarray
    status: np.ndarray
    box_lo: tuple = (0.0, 0.0, 0.0)
    box_hi: tuple = (1.0, 1.0, 1.0)
    walls: tuple = (WALL_OPEN,) * 6
    images: list = field(default_factory=list)
    vels: np.ndarray = None
    vs: np.ndarray = None
    tangents: np.ndarray = None
    normals: np.ndarray = None
    dxi: np.ndarray = None
    next_free: int = None      # the reference's next_free (None: N)

    def __post_init__(self):
        N = self.N
        self.vnodes = np.ascontiguousarray(self.vnodes, dtype=np.float64).reshape(N, 3)
        self.connections = np.ascontiguousarray(self.connections, dtype=np.int32).reshape(N, 2)
        self.status = np.ascontiguousarray(self.status, dtype=np.int32).reshape(N, 2)
        for name in ("vels", "vs", "tangents", "normals"):
            if getattr(self, name) is None:
                setattr(self, name, np.zeros((N, 3)))
        if self.dxi is None:
            self.dxi = np.zeros(N)
        self.recalculate = np.zeros(N, dtype=np.int32)
        self.vnodes_new = np.zeros((N, 3))

    @property
    def N(self) -> int:
        return int(np.asarray(self.vnodes).size // 3)

    @property
    def active(self) -> np.ndarray:
        return np.nonzero(self.status[:, 0] != EMPTY)[0]

    def copy(self) -> "Tangle":
        t = Tangle(self.vnodes.copy(), self.connections.copy(), self.status.copy(), self.box_lo,
                   self.box_hi, self.walls, list(self.images), self.vels.copy(), self.vs.copy(),
                   self.tangents.copy(), self.normals.copy(), self.dxi.copy(), self.next_free)
        return t

    def expand(self, new_N: int) -> "Tangle":
        """A copy with capacity new_N: the points keep their indices and the
        slots above N are EMPTY (the reference's expand_tangle, tangle.c:
        114-150); next_free stays, or becomes N when it was unset."""
        N = self.N
        if new_N < N:
            raise ValueError(f"expand to {new_N} below the capacity {N}")
        k = new_N - N
        st = np.vstack([self.status, np.tile(np.array([[EMPTY, -1]], dtype=np.int32), (k, 1))])
        t = Tangle(np.vstack([self.vnodes, np.zeros((k, 3))]),
                   np.vstack([self.connections, np.full((k, 2), -1, dtype=np.int32)]), st,
                   self.box_lo, self.box_hi, self.walls, list(self.images),
                   np.vstack([self.vels, np.zeros((k, 3))]), np.vstack([self.vs, np.zeros((k, 3))]),
                   np.vstack([self.tangents, np.zeros((k, 3))]),
                   np.vstack([self.normals, np.zeros((k, 3))]),
                   np.concatenate([self.dxi, np.zeros(k)]),
                   N if self.next_free is None else self.next_free)
        return t

    def cstruct(self) -> TangleState:
        """ov_tangle_state view of the arrays (valid while self lives)."""
        ts = TangleState()
        ts.vnodes = self.vnodes.ctypes.data
        ts.vnodes_new = self.vnodes_new.ctypes.data
        ts.vels = self.vels.ctypes.data
        ts.vs = self.vs.ctypes.data
        ts.tangents = self.tangents.ctypes.data
        ts.normals = self.normals.ctypes.data
        ts.dxi = self.dxi.ctypes.data
        ts.recalculate = self.recalculate.ctypes.data
        ts.connections = self.connections.ctypes.data
        ts.status = self.status.ctypes.data
        for i in range(3):
            ts.box.bottom_left_back.p[i] = self.box_lo[i]
            ts.box.top_right_front.p[i] = self.box_hi[i]
        for i in range(6):
            ts.box.wall[i] = self.walls[i]
        n = len(self.images)
        self._img = (ImageTangle * max(n, 1))()
        for j, (sx, sy, sz, rf) in enumerate(self.images):
            self._img[j].shift[0], self._img[j].shift[1], self._img[j].shift[2] = sx, sy, sz
            self._img[j].reflect = rf
        ts.bimg.images = C.cast(self._img, C.POINTER(ImageTangle)) if n else None
        ts.bimg.n = n
        ts.N = self.N
        ts.next_free = self.N if self.next_free is None else self.next_free
        ts.total_free = 0
        return ts

    def wrap_into_box(self) -> None:
        """Periodic wrap of every point (reference enforce_boundaries,
        src/tangle.c:766-819)."""
        lo = np.asarray(self.box_lo)
        hi = np.asarray(self.box_hi)
        L = hi - lo
        act = self.active
        for ax in range(3):
            if self.walls[2 * ax] != WALL_PERIODIC:
                continue
            x = self.vnodes[act, ax]
            x = np.where(x > hi[ax], x - L[ax] * np.ceil((x - hi[ax]) / L[ax]), x)
            x = np.where(x < lo[ax], x + L[ax] * np.ceil((lo[ax] - x) / L[ax]), x)
            self.vnodes[act, ax] = x


class TangleBuilder:
    """Accumulates closed filaments and open chains pinned on walls, and
    produces a Tangle."""

    def __init__(self):
        self.pts = []
        self.n = 0
        self.loops = []
        self.chains = {}      # start -> ((head status, wall), (tail status, wall))

    def add_loop(self, pts: np.ndarray) -> None:
        pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
        self.pts.append(pts)
        self.loops.append((self.n, len(pts)))
        self.n += len(pts)

    def add_chain(self, pts: np.ndarray, head=(PINNED, Z_L), tail=(PINNED, Z_H)) -> None:
        """An open filament from pts[0] to pts[-1]: the head has no reverse
        link, the tail no forward link, both are pinned, (status, pin_wall)
        each (a vortex line ending on walls, reference tangle.h:36-44)."""
        start = self.n
        self.add_loop(pts)
        self.chains[start] = (head, tail)

    def add_ring(self, center, direction, r: float, npts: int) -> None:
        """Ring geometry of the reference add_circle (src/vortex_utils.c:29-99):
        in-plane axes u = perp x dir, v = dir x u with perp the unit part of
        x (or y) orthogonal to dir; p_k = c + r cos(phi_k) u + r sin(phi_k) v,
        phi_k = 2 pi k / npts; links k -> k+1 around the ring."""
        d = np.asarray(direction, dtype=np.float64)
        a = np.array([1.0, 0.0, 0.0])
        zd = a - d * np.dot(a, d)
        if np.linalg.norm(zd) > 1e-8:
            zd = zd / np.linalg.norm(zd)
        else:
            b = np.array([0.0, 1.0, 0.0])
            zd = b - d * np.dot(b, d)
            zd = zd / np.linalg.norm(zd)
        u = np.cross(zd, d)
        v = np.cross(d, u)
        u = u / np.linalg.norm(u)
        v = v / np.linalg.norm(v)
        k = np.arange(npts)
        phi = 2 * np.pi / npts * k
        pts = (u[None, :] * (r * np.cos(phi))[:, None] + v[None, :] * (r * np.sin(phi))[:, None]
               + np.asarray(center, dtype=np.float64)[None, :])
        self.add_loop(pts)

    def build(self, box_lo=(0, 0, 0), box_hi=(1, 1, 1), walls=(WALL_OPEN,) * 6, images=(),
              capacity=None, holes=0, seed=0) -> Tangle:
        n = self.n
        N = max(n + holes, capacity or 0)
        vn = np.zeros((N, 3))
        con = np.full((N, 2), -1, dtype=np.int32)
        st = np.zeros((N, 2), dtype=np.int32)
        st[:, 0] = EMPTY
        st[:, 1] = -1
        # slots: active points at a seeded permutation of the capacity when
        # holes are requested, else in order
        if N > n:
            rng = np.random.default_rng(seed)
            slots = np.sort(rng.choice(N, size=n, replace=False))
        else:
            slots = np.arange(n)
        allp = np.concatenate(self.pts) if self.pts else np.zeros((0, 3))
        for (start, cnt) in self.loops:
            idx = slots[start:start + cnt]
            vn[idx] = allp[start:start + cnt]
            con[idx, 0] = np.roll(idx, -1)
            con[idx, 1] = np.roll(idx, 1)
            st[idx, 0] = FREE
            if start in self.chains:
                (hs, hw), (ts_, tw) = self.chains[start]
                con[idx[0], 1] = -1
                con[idx[-1], 0] = -1
                st[idx[0]] = (hs, hw)
                st[idx[-1]] = (ts_, tw)
        return Tangle(vn, con, st, tuple(box_lo), tuple(box_hi), tuple(walls), list(images))


# ---- workloads --------------------------------------------------------------

def ring(npts=256, r=0.01, center=(0.0, 0.0, 0.0), direction=(0.0, 0.0, 1.0)) -> Tangle:
    """Single ring in open boundaries (the 256-point CPU-runnable case)."""
    b = TangleBuilder()
    b.add_ring(center, direction, r, npts)
    return b.build()


def random_rings(seed: int, box: float, n_rings: int | None = None, r_lo=0.005, r_hi=0.02,
                 spacing=2e-4, log_radius=False, n_points_target: int | None = None,
                 n_images=26, holes=0) -> Tangle:
    """Rings with seeded centres (uniform in the box), isotropic normals and
    radii uniform (or log-uniform) in [r_lo, r_hi], npts = round(2 pi r /
    spacing), wrapped into the periodic box.  Either n_rings rings or rings
    until n_points_target points."""
    rng = SplitMix64(seed)
    b = TangleBuilder()
    k = 0
    while True:
        if n_rings is not None and k >= n_rings:
            break
        if n_points_target is not None and b.n >= n_points_target:
            break
        if log_radius:
            r = math.exp(rng.uniform(math.log(r_lo), math.log(r_hi)))
        else:
            r = rng.uniform(r_lo, r_hi)
        c = (rng.uniform(0, box), rng.uniform(0, box), rng.uniform(0, box))
        d = rng.unit_vector()
        npts = max(5, int(round(2 * math.pi * r / spacing)))
        b.add_ring(c, d, r, npts)
        k += 1
    imgs = periodic_images(n_images) if n_images else []
    walls = (WALL_PERIODIC,) * 6 if n_images else (WALL_OPEN,) * 6
    t = b.build((0, 0, 0), (box, box, box), walls, imgs, holes=holes, seed=seed)
    if n_images:
        t.wrap_into_box()
    return t


def bh_scale(seed=1002, box=0.2, n_points=1_000_000, n_lines=64, spacing=1e-4, amp=1e-3,
             r_lo=0.002, r_hi=0.03, n_images=26) -> Tangle:
    """Barnes-Hut scale tangle: box-spanning periodic lines along a seeded
    axis with a sinusoidal transverse displacement (amplitude amp, seeded
    mode 1..4, seeded phase and axial offset so that no point sits on a
    dyadic cell face), then rings with log-uniform radii until n_points."""
    rng = SplitMix64(seed)
    b = TangleBuilder()
    n_line = int(round(box / spacing))
    for _ in range(n_lines):
        ax = int(rng.uniform(0, 3)) % 3
        t1, t2 = [i for i in range(3) if i != ax]
        pos = (rng.uniform(0, box), rng.uniform(0, box))
        mode = 1 + int(rng.uniform(0, 4)) % 4
        phase = rng.uniform(0, 2 * math.pi)
        shift = rng.uniform(0.1, 0.9)
        sgn = 1.0 if rng.uniform() < 0.5 else -1.0
        s = (np.arange(n_line) + shift) * (box / n_line)
        if sgn < 0:
            s = s[::-1]
        pts = np.zeros((n_line, 3))
        pts[:, ax] = s
        pts[:, t1] = pos[0] + amp * np.sin(2 * np.pi * mode * s / box + phase)
        pts[:, t2] = pos[1] + amp * np.cos(2 * np.pi * mode * s / box + phase)
        b.add_loop(pts)
    while b.n < n_points:
        r = math.exp(rng.uniform(math.log(r_lo), math.log(r_hi)))
        c = (rng.uniform(0, box), rng.uniform(0, box), rng.uniform(0, box))
        d = rng.unit_vector()
        npts = max(5, int(round(2 * math.pi * r / spacing)))
        b.add_ring(c, d, r, npts)
    imgs = periodic_images(n_images) if n_images else []
    t = b.build((0, 0, 0), (box, box, box), (WALL_PERIODIC,) * 6, imgs)
    t.wrap_into_box()
    return t


def kelvin_ring(seed=1004, npts=4096, r=0.01, modes=range(2, 9), amp_hi=2e-4) -> Tangle:
    """Ring carrying circularly polarised Kelvin waves: point k at angle
    phi_k is displaced by sum_m A_m [cos(m phi + ph_m) e_r + sin(m phi + ph_m) e_z],
    A_m ~ U[0, amp_hi], ph_m ~ U[0, 2 pi) (seeded)."""
    rng = SplitMix64(seed)
    amps = [rng.uniform(0, amp_hi) for _ in modes]
    phases = [rng.uniform(0, 2 * math.pi) for _ in modes]
    phi = 2 * np.pi / npts * np.arange(npts)
    er = np.stack([np.sin(phi), -np.cos(phi), np.zeros_like(phi)], 1)
    ez = np.array([0.0, 0.0, 1.0])
    pts = r * er
    for m, A, ph in zip(modes, amps, phases):
        pts = pts + A * (np.cos(m * phi + ph)[:, None] * er + np.sin(m * phi + ph)[:, None] * ez)
    b = TangleBuilder()
    b.add_loop(pts)
    return b.build()


# He II mutual-friction coefficients (alpha, alpha') at 1.3, 1.6, 1.9 K, the
# rows of the reference's python/tcal/mutual_friction.csv at those temperatures.
MUTUAL_FRICTION = {1.3: (0.034, 1.383e-2), 1.6: (0.097, 1.608e-2), 1.9: (0.206, 8.340e-3)}


def ensemble(n_tangles=256, seed=1003, box=0.05, n_lo=500, n_hi=20000, spacing=2e-4,
             r_lo=0.005, r_hi=0.02):
    """Independent periodic tangles: tangle i uses seed seed * 2**32 + i; its
    point target is log-uniform in [n_lo, n_hi], filled with seeded rings as
    random_rings; its counterflow vn = (U[0.1, 1.0], 0, 0) cm/s and its
    temperature one of 1.3/1.6/1.9 K (alpha, alpha' from MUTUAL_FRICTION).
    Returns [(tangle, kwargs for default_params)]."""
    out = []
    for i in range(n_tangles):
        rng = SplitMix64(seed * 2 ** 32 + i)
        target = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        vn = rng.uniform(0.1, 1.0)
        temp = (1.3, 1.6, 1.9)[int(rng.uniform(0, 3)) % 3]
        t = random_rings(seed * 2 ** 32 + i + 7, box, n_points_target=target, r_lo=r_lo, r_hi=r_hi,
                         spacing=spacing)
        a, ap = MUTUAL_FRICTION[temp]
        out.append((t, dict(alpha=a, alpha_p=ap, vn=(vn, 0.0, 0.0), temperature=temp)))
    return out
