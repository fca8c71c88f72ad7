#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference's own C.

Run in a container that has the reference tree (the library is built from
its sources by oracle/Makefile into oracle/_ref/libovref.so):

    make -C oracle ref && python tests/golden/make_golden.py

Each fixture is a small .npz of inputs (positions, links, box, images,
parameters) and the reference's outputs for them.  Nothing of the reference
is stored except these numbers.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from openvort_amd import tangle as T  # noqa: E402
from oracle.oracle import Reference  # noqa: E402

PARAM_FIELDS = ("kappa", "vortex_width", "alpha", "alpha_p", "dl_max", "LIA_only",
                "use_mutual_friction", "hyperfriction", "use_BH", "max_curvature_scale",
                "hyperalpha", "BH_resolution", "BH_grain", "BH_quadtree", "multipole")


def pack_params(p):
    d = {f"p_{k}": getattr(p, k) for k in PARAM_FIELDS}
    d["p_vn"] = np.array([p.vn.type] + list(p.vn.v))
    d["p_vs"] = np.array([p.vs.type] + list(p.vs.v))
    return d


def pack_tangle(t, prefix="in_"):
    return {prefix + "vnodes": t.vnodes.copy(), prefix + "connections": t.connections.copy(),
            prefix + "status": t.status.copy(), prefix + "box_lo": np.array(t.box_lo),
            prefix + "box_hi": np.array(t.box_hi), prefix + "walls": np.array(t.walls),
            prefix + "images": np.array(t.images, dtype=np.int32).reshape(-1, 4)}


def out_fields(t, prefix="out_"):
    return {prefix + k: getattr(t, k).copy() for k in ("vnodes", "tangents", "normals", "dxi", "vs", "vels")}


def main():
    if not Reference.available():
        sys.exit("oracle/_ref/libovref.so is missing: make -C oracle ref")
    R = Reference(threads=4)

    ring_p = T.default_params(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0)
    per_p = T.default_params(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0,
                             vn=(0.5, 0, 0))

    # 1. the 256-point ring: update_tangle, then 10 steps of update_tangle + rk4_step
    t = T.ring(256)
    fx = {**pack_tangle(t), **pack_params(ring_p)}
    R.update_tangle(t, 0.0, ring_p)
    fx.update(out_fields(t, "t0_"))
    for k in range(10):
        R.step(t, k * 1e-5, 1e-5, ring_p)
    fx["steps"] = 10
    fx["dt"] = 1e-5
    fx["out_vnodes_after"] = t.vnodes.copy()
    np.savez_compressed(os.path.join(HERE, "ring256.npz"), **fx)

    # 2. small periodic-26 tangle with rings crossing the box faces, holes,
    #    counterflow; update_tangle then one rk4_step
    b = T.TangleBuilder()
    b.add_ring((0.098, 0.05, 0.05), (0.0, 0.6, 0.8), 0.012, 160)
    b.add_ring((0.05, 0.003, 0.05), (1.0, 0.0, 0.0), 0.008, 110)
    b.add_ring((0.04, 0.06, 0.097), (0.3, -0.4, 0.866), 0.01, 130)
    t = b.build((0, 0, 0), (0.1, 0.1, 0.1), (T.WALL_PERIODIC,) * 6, T.periodic_images(26),
                holes=41, seed=3)
    t.wrap_into_box()
    fx = {**pack_tangle(t), **pack_params(per_p)}
    R.update_tangle(t, 0.0, per_p)
    fx.update(out_fields(t, "t0_"))
    R.rk4_step(t, 0.0, 5e-6, per_p)
    fx["dt"] = 5e-6
    fx["out_vnodes_after"] = t.vnodes.copy()
    np.savez_compressed(os.path.join(HERE, "periodic26.npz"), **fx)

    # 3. hyperfriction on (threshold inside the range of |s''|)
    t = T.random_rings(17, 0.1, n_rings=3, spacing=4e-4)
    hp = T.default_params(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=1,
                          dl_max=0.002, vn=(0.0, 0.3, 0))
    R.update_tangents_normals(t)
    spp = np.linalg.norm(t.normals[t.active], axis=1)
    thr = np.median(spp) / 1.0
    hp.max_curvature_scale = thr * hp.dl_max * 1.0000001
    fx = {**pack_tangle(t), **pack_params(hp)}
    R.update_tangle(t, 0.0, hp)
    fx.update(out_fields(t, "t0_"))
    np.savez_compressed(os.path.join(HERE, "hyperfriction.npz"), **fx)

    # 4. mirror images: a ring near the z = 0 wall, wall_1_open (one mirror
    #    image reflected in Z_L, reference boundary_images.c:106)
    b = T.TangleBuilder()
    b.add_ring((0.5, 0.5, 0.02), (0.3, 0.0, 0.954), 0.01, 120)
    t = b.build((0, 0, 0), (1, 1, 1), (T.WALL_OPEN,) * 4 + (T.WALL_MIRROR, T.WALL_OPEN),
                [(0, 0, -1, T.Z_L)])
    fx = {**pack_tangle(t), **pack_params(ring_p)}
    R.update_tangle(t, 0.0, ring_p)
    fx.update(out_fields(t, "t0_"))
    np.savez_compressed(os.path.join(HERE, "mirror.npz"), **fx)

    # 5. octree velocities, theta in {0.3, 0.5, 0.7}, octree and quadtree
    t = T.random_rings(23, 0.1, n_rings=5, spacing=5e-4)
    for quad in (0, 1):
        for th in (0.3, 0.5, 0.7):
            p = T.default_params(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0,
                                 vn=(0.5, 0, 0), use_BH=1, BH_resolution=th, BH_quadtree=quad)
            x = t.copy()
            fx = {**pack_tangle(x), **pack_params(p)}
            R.update_tangle(x, 0.0, p)
            fx.update(out_fields(x, "t0_"))
            np.savez_compressed(os.path.join(HERE, f"octree_q{quad}_th{int(th * 10)}.npz"), **fx)

    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
