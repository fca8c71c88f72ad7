"""The benchmark configurations at their own sizes on the GPU.

* Barnes-Hut scale (1.0M points, box 0.2 cm, 26 images): the GPU octree at
  theta 0.3 / 0.5 / 0.7 against the oracle's octree on seeded targets; the
  mixed direct sum against the FP64 direct sum over every point, and the FP64
  sum against the oracle on seeded targets.
* The mixed form at other box and spacing scales than the benchmark's.
* The 256-tangle ensemble: one batched step against the per-tangle loop, and
  the smallest, median and largest members against the oracle.
* The differentiation workload (Kelvin ring, 4096 points, 200 steps at
  dt = 1e-7 s, checkpoints every 20): the adjoint gradient against torch
  autograd through the FP64 restatement, and that restatement's gradient
  against its central differences with the collinearity cut-off off.
"""
import math
import time

import numpy as np
import pytest
import torch  # imported before any test initialises the device through the library

from openvort_amd import tangle as T
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

COMMON = dict(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0)
MIXED_TOL = 1e-6


@pytest.fixture(scope="module")
def bh1m():
    return T.bh_scale(seed=1002, n_points=1_000_000)


def _sample(t, n, seed=2002):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(t.active, size=n, replace=False)).astype(np.int32)


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_bh1m_octree_matches_oracle(ctx, oracle, bh1m, theta):
    """Same cells (count) and the same velocities (1e-10) as the oracle's
    octree on 256 seeded targets of the 1.0M-point tangle."""
    p = T.default_params(precision=T.PREC_FP64, use_BH=1, BH_resolution=theta, **COMMON)
    g = bh1m.copy()
    ctx.tree_stats()
    t0 = time.time()
    ctx.velocities(g, 0.0, p, geometry=True, profile=True)
    cells, pairs, mps = ctx.tree_stats()
    idx = _sample(bh1m, 256)
    c = bh1m.copy()
    oracle.update_tangents_normals(c)
    _, vel = oracle.velocities_sample(c, 0.0, p, idx)
    ocells, _, _ = oracle.octree_stats()
    e = rel_err(g.vels[idx], vel)
    print(f"theta {theta}: cells {cells:.0f}/{ocells}, leaf pairs {pairs:.3g}, multipoles "
          f"{mps:.3g}, err {e}, {time.time() - t0:.1f}s", flush=True)
    assert int(cells) == ocells
    assert e[0] < 1e-10


def test_bh1m_mixed_direct_vs_fp64(ctx, oracle, bh1m):
    """Mixed against FP64 over all 1.0M points (max |dv|/|v| <= 1e-6), and
    the FP64 GPU sum against the oracle on 64 seeded targets (1e-11)."""
    pm = T.default_params(precision=T.PREC_MIXED, near_radius=0.02, **COMMON)
    p64 = T.default_params(precision=T.PREC_FP64, **COMMON)
    a, b = bh1m.copy(), bh1m.copy()
    t0 = time.time()
    ctx.velocities(a, 0.0, pm)
    ctx.velocities(b, 0.0, p64)
    act = bh1m.active
    e = rel_err(a.vels[act], b.vels[act])
    e_vs = rel_err(a.vs[act], b.vs[act])
    idx = _sample(bh1m, 64)
    c = bh1m.copy()
    oracle.update_tangents_normals(c)
    _, vel = oracle.velocities_sample(c, 0.0, p64, idx)
    e64 = rel_err(b.vels[idx], vel)
    print(f"1M mixed vs fp64 vels {e}, vs {e_vs}; fp64 vs oracle {e64}; {time.time() - t0:.1f}s",
          flush=True)
    assert e[0] < MIXED_TOL and e_vs[0] < MIXED_TOL
    assert e64[0] < 1e-11


@pytest.mark.parametrize("box,spacing,n_points,rnear", [
    (0.05, 2.5e-5, 150_000, 0.005),     # four times finer than the benchmark
    (0.8, 4e-4, 150_000, 0.08),          # four times coarser
    (0.2, 1e-4, 150_000, 0.01),          # the benchmark's scale, half the near radius
])
def test_mixed_precision_across_scales(ctx, box, spacing, n_points, rnear):
    """The mixed form's FP32 far field (rsq bias compensation, small-angle
    series, zero-rule bands) at box and spacing scales other than the
    benchmark's: max |dv|/|v| against the FP64 form stays below 1e-6."""
    t = T.bh_scale(seed=1302, box=box, n_points=n_points, spacing=spacing, amp=5 * spacing,
                   r_lo=10 * spacing, r_hi=0.15 * box)
    pm = T.default_params(precision=T.PREC_MIXED, near_radius=rnear, **COMMON)
    p64 = T.default_params(precision=T.PREC_FP64, **COMMON)
    a, b = t.copy(), t.copy()
    ctx.velocities(a, 0.0, pm)
    ctx.velocities(b, 0.0, p64)
    act = t.active
    e = rel_err(a.vels[act], b.vels[act])
    print(f"box {box} spacing {spacing} rnear {rnear}: {len(act)} points, mixed vs fp64 {e}", flush=True)
    assert e[0] < MIXED_TOL


def _ens_params(kw, prec):
    kw = dict(kw)
    kw.pop("temperature")
    return T.default_params(vortex_width=1.3e-8, hyperfriction=0, precision=prec, near_radius=0.02, **kw)


@pytest.mark.parametrize("prec", [T.PREC_FP64, T.PREC_MIXED])
def test_ensemble256_batch_vs_loop_and_oracle(ctx, oracle, prec):
    """256 tangles of 500-20000 points: one batched RK4 step equals the
    per-tangle loop (positions to 1e-16); the smallest, median and largest
    members' k1 velocities against the oracle on 128 seeded targets each
    (FP64 1e-11, mixed 1e-6); the smallest member's step against the
    oracle's step (FP64: positions to 1e-16)."""
    ens = T.ensemble(n_tangles=256)
    ps = [_ens_params(kw, prec) for _, kw in ens]
    bt = [t.copy() for t, _ in ens]
    t0 = time.time()
    ctx.rk4_run_batch(bt, 0.0, 5e-6, 1, ps)
    t_batch = time.time() - t0
    sizes = np.array([len(t.active) for t, _ in ens])
    order = np.argsort(sizes, kind="stable")
    picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    worst = 0.0
    for i, (t, _) in enumerate(ens):
        a = t.copy()
        ctx.rk4_run(a, 0.0, 5e-6, 1, ps[i])
        act = t.active
        worst = max(worst, float(np.abs(a.vnodes[act] - bt[i].vnodes[act]).max()))
    print(f"prec {prec}: {sizes.sum()} points, batch {t_batch:.1f}s, batch vs loop {worst:.3g}", flush=True)
    assert worst < 1e-16
    tol = 1e-11 if prec == T.PREC_FP64 else MIXED_TOL
    for i in picks:
        t = ens[i][0]
        c = t.copy()
        oracle.update_tangents_normals(c)
        idx = _sample(t, min(128, len(t.active)), seed=3003 + i)
        _, vel = oracle.velocities_sample(c, 0.0, _ens_params(ens[i][1], T.PREC_FP64), idx)
        e = rel_err(bt[i].vels[idx], vel)
        print(f"member {i} ({sizes[i]} points) k1 vs oracle {e}", flush=True)
        assert e[0] < tol
    if prec == T.PREC_FP64:
        i = picks[0]
        a = ens[i][0].copy()
        oracle.step(a, 0.0, 5e-6, ps[i])
        act = a.active
        assert np.abs(a.vnodes[act] - bt[i].vnodes[act]).max() < 1e-16


def _ref_run_checkpointed(X, topo, dt, nsteps):
    """nsteps of the torch restatement, each step a checkpoint (its stage
    graph is rebuilt during the backward pass)."""
    from torch.utils.checkpoint import checkpoint
    from tests import torch_ref as TR
    for _ in range(nsteps):
        X = checkpoint(lambda Y: TR.rk4_run(Y, topo, dt, 1), X, use_reentrant=False)
    return X


def test_kelvin_ring_200_steps_gradient(ctx):
    """The differentiation workload at full length: 4096 points, 200
    reference-variant steps at dt = 1e-7 s, checkpoints every 20.
    (1) The adjoint's gradient of the Kelvin loss with respect to the
    initial positions equals torch autograd through the FP64 restatement
    (the reference gradient, collinearity cut-off 1e-8 as in tangle.c:
    505-506) to 1e-10.  (2) With the cut-off off (a smooth function), the
    restatement's gradient matches its central differences (h = 1e-9 cm) to
    1e-5 along a random direction and along the m = 2 mode shape.  The number
    of zeroed pairs summed over the 800 evaluations of the two perturbed runs
    is printed: where it differs, pairs crossed the cut-off."""
    from openvort_amd import adjoint as A
    from tests import torch_ref as TR
    dev = torch.device("cuda:0")
    k = A.kelvin_setup()
    base = T.kelvin_ring()
    p = T.default_params(**COMMON)
    dt, nsteps, ck = 1e-7, 200, 20
    t0 = time.time()
    run = A.Run(ctx, base, p, 0.0, dt, nsteps, checkpoint_every=ck)
    X0 = torch.from_numpy(base.vnodes.copy()).requires_grad_(True)
    A.kelvin_loss(A.rk4_run(X0, run), k.target).backward()
    g = X0.grad.numpy().copy()
    t1 = time.time()
    grads = {}
    for tol in (1e-8, 0.0):
        topo = TR.Topo(base, p)
        topo.ndot_tol = tol
        Xr = torch.from_numpy(base.vnodes.copy()).to(dev).requires_grad_(True)
        A.kelvin_loss(_ref_run_checkpointed(Xr, topo, dt, nsteps).cpu(), k.target).backward()
        grads[tol] = Xr.grad.cpu().numpy()
        del Xr
        torch.cuda.empty_cache()
    e = float(np.linalg.norm(g - grads[1e-8], axis=1).max() /
              np.sqrt(np.mean(np.sum(grads[1e-8] ** 2, 1))))
    t2 = time.time()
    print(f"adjoint vs autograd (200 steps): {e:.3g}; adjoint {t1 - t0:.1f}s, autograd {t2 - t1:.1f}s",
          flush=True)
    assert e < 1e-10
    # central differences of the smooth restatement
    topo = TR.Topo(base, p)
    topo.ndot_tol = 0.0
    X0d = torch.from_numpy(base.vnodes.copy()).to(dev)
    rng = np.random.default_rng(4004)
    phi = 2 * np.pi / k.npts * np.arange(k.npts)
    er = np.stack([np.sin(phi), -np.cos(phi), np.zeros_like(phi)], 1)
    arg = 2 * phi + k.phases[0]
    mode2 = np.cos(arg)[:, None] * er + np.sin(arg)[:, None] * np.array([0.0, 0.0, 1.0])
    h = 1e-9
    with torch.no_grad():
        L = lambda Y: float(A.kelvin_loss(TR.rk4_run(Y, topo, dt, nsteps).cpu(), k.target))
        for name, d in (("random", rng.standard_normal((k.npts, 3))), ("mode 2", mode2)):
            dd = torch.from_numpy(d).to(dev)
            fd = (L(X0d + h * dd) - L(X0d - h * dd)) / (2 * h)
            ad = float((grads[0.0] * d).sum())
            print(f"cut-off off, {name}: fd {fd:.12g} ad {ad:.12g} rel {abs(fd - ad) / abs(ad):.3g}",
                  flush=True)
            assert abs(fd - ad) <= 1e-5 * abs(ad)
    # pairs at the reference's cut-off along the same perturbation
    topo_c = TR.Topo(base, p)
    counts = {}
    with torch.no_grad():
        d = torch.from_numpy(rng.standard_normal((k.npts, 3))).to(dev)
        for sgn in (1, -1):
            topo_c.cut_counts = []
            TR.rk4_run(X0d + sgn * h * d, topo_c, dt, nsteps)
            counts[sgn] = np.array(topo_c.cut_counts)
    flips = int(np.abs(counts[1] - counts[-1]).sum())
    print(f"zeroed pairs per evaluation at t=0: {counts[1][0]}; evaluations whose count differs "
          f"between X0 +- h d: {int((counts[1] != counts[-1]).sum())} of {len(counts[1])}, "
          f"summed |difference| {flips}; {time.time() - t0:.1f}s", flush=True)
    torch.cuda.empty_cache()
