"""Loading of the golden fixtures (tests/golden/*.npz, written from the
reference's own C by tests/golden/make_golden.py)."""
import glob
import os

import numpy as np

from openvort_amd import tangle as T
from tests.golden.make_golden import PARAM_FIELDS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    return sorted(os.path.basename(f)[:-4] for f in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    t = T.Tangle(z["in_vnodes"], z["in_connections"], z["in_status"], tuple(z["in_box_lo"]),
                 tuple(z["in_box_hi"]), tuple(int(w) for w in z["in_walls"]),
                 [tuple(int(v) for v in row) for row in z["in_images"]])
    p = T.default_params()
    for k in PARAM_FIELDS:
        v = z["p_" + k]
        setattr(p, k, type(getattr(p, k))(v))
    for fl in ("vn", "vs"):
        a = z["p_" + fl]
        f = getattr(p, fl)
        f.type = int(a[0])
        for i in range(3):
            f.v[i] = float(a[1 + i])
    return t, p, z
