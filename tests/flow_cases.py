"""Imposed-flow cases shared by the CPU and GPU flow tests: for every entry
of the reference's flow table (src/external_velocity.c:85-174) a parameter
set that gives velocities comparable to the self-induced ones (~0.1-1 cm/s)
on tangles in the box [0, 0.1]^3 over times of ~1e-4 s."""
from openvort_amd import tangle as T

FLOW_CASES = {
    "no flow": {},
    "simple": dict(external_v=(0.3, -0.1, 0.2)),
    "spherical": dict(strength=2e-3, cutoff=2e-3),
    "spherical and simple": dict(external_v=(0.1, 0.0, -0.05), strength=2e-3, cutoff=2e-3),
    "cylindrical": dict(strength=3e-2, cutoff=2e-3),
    "oscillating": dict(strength=0.2, frequency=3e3, wave=(20.0, 5.0, -8.0),
                        polarization=(0.3, 1.0, -0.4)),
    "oscillating slit": dict(strength=5.0, frequency=2e3, cutoff=3e-3),
    "coscos": dict(k=300.0, v0=0.2, l=0.05),
    "coscos-divfree": dict(k=300.0, v0=0.2, l=0.05),
    "cos-divfree": dict(k=300.0, v0=0.2, l=0.05),
    "simple-shear": dict(rate=20.0, v0=0.1),
    "rotosc": dict(freq=500.0, amp=30.0),
    "DG": dict(rot=15.0, vz=0.1),
    "rotosc-boundary": dict(freq=500.0, amp=30.0, delta=0.01),
    "constant acceleration": dict(accel=1e3),
    "parabolic": dict(H=0.1, amp=50.0),
}

TIME_DEPENDENT = {"oscillating", "oscillating slit", "rotosc", "rotosc-boundary",
                  "constant acceleration"}


def flow(name, **override):
    f = T.Flow()
    kw = dict(FLOW_CASES[name])
    kw.update(override)
    return T.set_flow(f, name, **kw)


def flow_params(vs=None, vn=None, **kw):
    """tangle parameters (a0 = 1.3e-8, T = 1.6 K, hyperfriction off) with
    imposed flows given by table name (the case parameters above)."""
    p = T.default_params(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0, **kw)
    if vs is not None:
        T.set_flow(p.vs, vs, **FLOW_CASES[vs])
    if vn is not None:
        T.set_flow(p.vn, vn, **FLOW_CASES[vn])
    return p
