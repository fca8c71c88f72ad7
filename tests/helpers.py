"""Shared error measures and workload parameters for the tests."""
import numpy as np

from openvort_amd import tangle as T


def rel_err(x, y):
    """(max_k |x_k - y_k| / |y_k|, rms of the same, max |x - y| / rms |y|)"""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    y = np.asarray(y, dtype=np.float64).reshape(-1, 3)
    d = np.linalg.norm(x - y, axis=1)
    n = np.linalg.norm(y, axis=1)
    r = d / np.maximum(n, 1e-300)
    return float(r.max()), float(np.sqrt(np.mean(r * r))), float(d.max() / np.sqrt(np.mean(n * n)))


def ring_params(**kw):
    """Single-ring case: a0 = 1.3e-8 cm, T = 1.6 K (alpha 0.098, alpha' 0.016),
    no imposed flows, hyperfriction off."""
    base = dict(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0)
    base.update(kw)
    return T.default_params(**base)


def tangle_params(**kw):
    """Periodic random-ring case: as ring_params with counterflow vn = 0.5 x."""
    base = dict(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0, vn=(0.5, 0, 0))
    base.update(kw)
    return T.default_params(**base)
