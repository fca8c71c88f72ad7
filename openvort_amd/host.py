"""ctypes binding of the host-side configuration and frame I/O
(include/openvort_host.h): the reference's configuration files read without
libconfig, the reference's flow configurations mapped by name, and frames in
save_tangle's format, read back completely.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import tangle as T
from .gpu import OVError, load_library

HOST_SYMBOLS = (
    "ov_config_default",
    "ov_config_parse",
    "ov_config_load",
    "ov_params_from_config",
    "ov_flow_from_vconf",
    "ov_frame_save",
    "ov_frame_count",
    "ov_frame_load",
)


class VParam(C.Structure):
    """struct v_param_t (external_velocity.h:54-62)"""
    class _V(C.Union):
        _fields_ = [("vector", T.Vec3), ("scalar", C.c_double)]
    _fields_ = [("name", C.c_char_p), ("type", C.c_int), ("value", _V)]


class VConf(C.Structure):
    """struct v_conf_t (external_velocity.h:78-83)"""
    _fields_ = [("name", C.c_char_p), ("fun", C.c_void_p), ("n_params", C.c_int),
                ("v_params", VParam * 32)]


class Config(C.Structure):
    _fields_ = [
        ("max_steps", C.c_int), ("frame_shots", C.c_int), ("kappa", C.c_double),
        ("num_threads", C.c_int), ("pin_mode", C.c_int),
        ("use_BH", C.c_int), ("BH_resolution", C.c_double), ("BH_quadtree", C.c_int),
        ("BH_grain", C.c_double),
        ("use_mutual_friction", C.c_int), ("alpha", C.c_double), ("alpha_p", C.c_double),
        ("LIA_only", C.c_int), ("hyperfriction", C.c_int), ("hyperalpha", C.c_double),
        ("max_curvature_scale", C.c_double),
        ("dt", C.c_double), ("dl_min", C.c_double), ("dl_max", C.c_double),
        ("small_loop_cutoff", C.c_int), ("reconnection_angle_cutoff", C.c_double),
        ("rec_dist", C.c_double),
        ("eliminate_origin_loops", C.c_int), ("eliminate_loops_origin_cutoff", C.c_double),
        ("eliminate_zaxis_loops", C.c_int), ("eliminate_loops_zaxis_cutoff", C.c_double),
        ("eliminate_outer_loops", C.c_int), ("eliminate_outer_loops_cutoff", C.c_double),
        ("loop_injection", C.c_int), ("loop_injection_frequency", C.c_double),
        ("line_injection", C.c_int), ("line_injection_frequency", C.c_double),
        ("line_injection_n", C.c_int), ("line_injection_polarized", C.c_int),
        ("vn", T.Flow), ("vs", T.Flow), ("vb", T.Flow),
        ("vn_set", C.c_int), ("vs_set", C.c_int), ("vb_set", C.c_int),
        ("box", T.DomainBox), ("boundaries", C.c_char * 32),
        ("images", T.ImageTangle * 26), ("n_images", C.c_int),
        ("init_mode", C.c_char * 64), ("init_file", C.c_char * 256),
        ("init_loops_n", C.c_int), ("loop_center", C.c_double * 3), ("loop_dir", C.c_double * 3),
        ("loop_r", C.c_double), ("loop_N", C.c_int), ("ring_r", C.c_double), ("ring_N", C.c_int),
        ("init_npairs", C.c_int), ("init_line_points", C.c_int), ("init_nvorts", C.c_int),
        ("init_direction", C.c_int), ("init_lattice_shells", C.c_int),
        ("init_lattice_constant", C.c_double), ("init_separation", C.c_double),
        ("dipole_direction", C.c_int), ("init_line_direction", C.c_int), ("init_k_KW", C.c_int),
        ("init_r_KW", C.c_double),
    ]


_typed = False


def _lib():
    global _typed
    lib = load_library()
    if not _typed:
        P = C.POINTER
        lib.ov_config_default.argtypes = [P(Config)]
        lib.ov_config_default.restype = None
        lib.ov_config_parse.argtypes = [C.c_char_p, P(Config), C.c_char_p, C.c_int]
        lib.ov_config_parse.restype = C.c_int
        lib.ov_config_load.argtypes = [C.c_char_p, P(Config), C.c_char_p, C.c_int]
        lib.ov_config_load.restype = C.c_int
        lib.ov_params_from_config.argtypes = [P(Config), P(T.Params)]
        lib.ov_params_from_config.restype = None
        lib.ov_flow_from_vconf.argtypes = [C.c_void_p, P(T.Flow)]
        lib.ov_flow_from_vconf.restype = C.c_int
        lib.ov_frame_save.argtypes = [C.c_char_p, P(T.TangleState), C.c_int]
        lib.ov_frame_save.restype = C.c_int
        lib.ov_frame_count.argtypes = [C.c_char_p, P(C.c_int)]
        lib.ov_frame_count.restype = C.c_int
        lib.ov_frame_load.argtypes = [C.c_char_p, P(T.TangleState), C.c_char_p, C.c_int]
        lib.ov_frame_load.restype = C.c_int
        _typed = True
    return lib


def _check(rc: int, err: C.Array) -> None:
    if rc != 0:
        raise OVError(rc, err.value.decode(errors="replace"))


def default_config() -> Config:
    c = Config()
    _lib().ov_config_default(C.byref(c))
    return c


def parse_config(text: str, cfg: Config | None = None) -> Config:
    """Settings of a reference configuration text (load_conf + setup_init)."""
    c = cfg if cfg is not None else default_config()
    err = C.create_string_buffer(512)
    _check(_lib().ov_config_parse(text.encode(), C.byref(c), err, 512), err)
    return c


def load_config(path: str, cfg: Config | None = None) -> Config:
    c = cfg if cfg is not None else default_config()
    err = C.create_string_buffer(512)
    _check(_lib().ov_config_load(path.encode(), C.byref(c), err, 512), err)
    return c


def params_from_config(c: Config, p: T.Params | None = None) -> T.Params:
    p = p if p is not None else T.default_params()
    _lib().ov_params_from_config(C.byref(c), C.byref(p))
    return p


def flow_from_vconf(addr: int) -> T.Flow:
    """ov_flow of a reference v_conf_t at address addr."""
    f = T.Flow()
    rc = _lib().ov_flow_from_vconf(addr, C.byref(f))
    if rc:
        raise OVError(rc, "unknown flow configuration")
    return f


def save_frame(path: str, t: T.Tangle, digits: int = 15) -> None:
    ts = t.cstruct()
    rc = _lib().ov_frame_save(path.encode(), C.byref(ts), digits)
    if rc:
        raise OVError(rc, f"cannot write {path}")


def load_frame(path: str) -> T.Tangle:
    """A tangle holding exactly the frame's points (box: the default open
    unit box; the frame format carries no box)."""
    lib = _lib()
    n = C.c_int(0)
    rc = lib.ov_frame_count(path.encode(), C.byref(n))
    if rc:
        raise OVError(rc, f"cannot read {path}")
    N = n.value
    st = np.zeros((N, 2), dtype=np.int32)
    st[:, 0] = T.EMPTY
    st[:, 1] = -1
    t = T.Tangle(np.zeros((N, 3)), np.full((N, 2), -1, dtype=np.int32), st)
    ts = t.cstruct()
    err = C.create_string_buffer(512)
    _check(lib.ov_frame_load(path.encode(), C.byref(ts), err, 512), err)
    return t


def tangle_from_config(c: Config, t: T.Tangle) -> T.Tangle:
    """t with the configuration's box, walls and image list."""
    t.box_lo = tuple(c.box.bottom_left_back.p)
    t.box_hi = tuple(c.box.top_right_front.p)
    t.walls = tuple(c.box.wall)
    t.images = [tuple(c.images[j].shift) + (c.images[j].reflect,) for j in range(c.n_images)]
    return t
