#!/usr/bin/env python3
"""Benchmark of the vortex-filament velocity / RK4 path on one MI355X.

Workload: the Barnes-Hut-scale periodic tangle (box 0.2 cm, periodic-26
images, 64 perturbed box-spanning lines + log-uniform rings at 1e-4 cm
spacing, ~1.0M points, seed 1002) with velocities from the direct
Biot-Savart sum + LIA + mutual friction.

A step is one RK4 step (ov_gpu_rk4_run, device-resident): 4 velocity
evaluations + geometry + stage updates, in the mixed form (FP32 far field,
FP64 near field / LIA / RK).  The timed region is device time (HIP events on
the library's stream); inputs are already in HBM.

value = segment-point interactions per second of the timed RK4 steps.
Also reported: RK4 steps/s, the plain FP64 form's rate on one evaluation of
the same positions, the mixed form's error against it, the roofline of the
dominant kernel (k_direct) and a CPU baseline: the reference's own OpenMP
update_velocity (oracle/_ref, or the FP64 restatement where that is absent)
on a seeded sample of targets of the same tangle, on this host's cores.
The Barnes-Hut octree (theta 0.3/0.5/0.7, FP64 and mixed) is evaluated on the
same positions and compared with the FP64 direct sum on 4096 seeded targets.
An ensemble of 256 independent periodic tangles (box 0.05 cm, 500-20000
points each) is stepped in one launch per stage and, for comparison, one
tangle at a time.  The adjoint is timed on the Kelvin-ring loss (4096
points, 200 RK4 steps) and checked against central differences.  The
reconnection candidate search (the reference reconnect's pair loop) is timed
on the workload's positions.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_INTERACTION = 43          # 40 add/mul (FMA = 2) + 2 sqrt + 1 div
PEAK_FP32 = 157.3e12               # MI355X vector FP32, FLOP/s
PEAK_FP64 = 78.6e12                # MI355X vector FP64, FLOP/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def err_stats(x, y):
    d = np.linalg.norm(x - y, axis=1)
    n = np.linalg.norm(y, axis=1)
    r = d / np.maximum(n, 1e-300)
    return {"max_rel": float(r.max()), "rms_rel": float(np.sqrt(np.mean(r * r))),
            "max_abs_over_rms": float(d.max() / np.sqrt(np.mean(n * n)))}


# octree build passes timed per evaluation (the radix sort runs inside hipcub)
OCT_BUILD = ("k_oct_keys", "k_gather_u64", "k_oct_leafsegs", "k_oct_count", "k_oct_emit",
             "k_oct_moments", "k_oct_big_sum1", "k_oct_big_sum2", "k_oct_big_fin", "k_oct_size",
             "k_oct_pre", "k_oct_nodes")


def tree_hbm_bytes():
    """HBM bytes of each octree build pass per evaluation of the 1M-point
    workload, from the committed PMC summary (copy-calibrated FETCH_SIZE +
    WRITE_SIZE); {} when absent."""
    path = os.path.join(ROOT, "profiles", "octree_build_bh1m_pmc.md")
    try:
        text = open(path).read()
        return json.loads(text.split("```json", 1)[1].split("```", 1)[0])["hbm_bytes_per_evaluation"]
    except (OSError, IndexError, ValueError, KeyError):
        return {}


def measured_traffic(points, seed, near_radius):
    """HBM bytes per launch of k_direct_mixed from the committed PMC summary
    (rocprofv3 FETCH_SIZE / WRITE_SIZE, copy-calibrated), when it covers
    this workload; None otherwise."""
    path = os.path.join(ROOT, "profiles", "k_direct_mixed_bh1m_pmc.md")
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        d = json.loads(block)
    except (OSError, IndexError, ValueError):
        return None
    w = d.get("workload", {})
    if w.get("points") != points or w.get("seed") != seed or w.get("near_radius_cm") != near_radius:
        return None
    return d["hbm_bytes_per_launch"]


def adjoint_section(ctx, T):
    """Gradient of the Kelvin-ring loss (centroid target + Kelvin-wave energy
    of the final state) with respect to the initial positions and the mode
    amplitudes: 4096 points, 200 reference-variant RK4 steps at dt = 1e-7 s
    (1e-5 s destroys the ring at this resolution), checkpoints every 20
    steps.  Checked against central differences (h = 1e-9 cm) of the FP64
    forward along single coordinates, a random direction and the amplitudes;
    the mixed-precision forward's gradient against the FP64 one."""
    import torch
    from openvort_amd import adjoint as A
    k = A.kelvin_setup()
    base = T.kelvin_ring()
    dt, nsteps, ck = 1e-7, 200, 20
    res = {"points": k.npts, "steps": nsteps, "dt_s": dt, "checkpoint_every": ck}
    grads = {}
    for name, prec in (("fp64", T.PREC_FP64), ("mixed", T.PREC_MIXED)):
        # hyperfriction off: its curvature switch (tangle.c:641-650) makes the
        # trajectory only piecewise smooth in the initial positions
        p = T.default_params(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0,
                             precision=prec)
        run = A.Run(ctx, base, p, 0.0, dt, nsteps, checkpoint_every=ck)
        amps = torch.tensor(k.amps, requires_grad=True)
        X0 = A.kelvin_positions(k, amps)
        X0.retain_grad()
        w0 = time.time()
        XT = A.rk4_run(X0, run)
        fwd_ms = ctx.last_device_ms()
        loss = A.kelvin_loss(XT, k.target)
        w1 = time.time()
        loss.backward()
        grad_ms = ctx.last_device_ms()
        w2 = time.time()
        # the same gradient once more with per-kernel timing (eager launches)
        ctx.reset_stats()
        ctx.rk4_run_vjp(base, 0.0, dt, nsteps, p, np.ones((k.npts, 3)), checkpoint_every=ck, profile=True)
        prof_ms = ctx.last_device_ms()
        kstats = {}
        for kn in ("ka_bs_fused", "ka_bs_targets", "ka_bs_sources", "ka_bs_reduce", "ka_local",
                   "ka_gather", "ka_bs_finish", "k_direct_fp64_small", "k_direct_mixed_small",
                   "k_geometry", "k_epilogue"):
            nl, ms = ctx.kernel_stats(kn)
            if nl:
                kstats[kn] = {"launches": nl, "total_ms": round(ms, 2)}
        lv = loss.item()
        grads[name] = (X0.grad.numpy().copy(), amps.grad.numpy().copy(), run, lv)
        res[name] = {"loss": lv, "forward_device_ms": round(fwd_ms, 2),
                     "gradient_device_ms": round(grad_ms, 2),
                     "forward_wall_s": round(w1 - w0, 3), "gradient_wall_s": round(w2 - w1, 3),
                     "gradient_profiled_device_ms": round(prof_ms, 2), "gradient_kernels": kstats}
    gX, gA, run, _ = grads["fp64"]
    X0 = A.kelvin_positions(k, torch.tensor(k.amps)).detach()
    L = lambda X: A.kelvin_loss(A.rk4_run(X, run), k.target).item()
    h = 1e-9
    # single coordinates of seeded points: few pairs near the collinearity
    # cut-off move, so central differences see the smooth part
    rng = np.random.default_rng(4004)
    fd_pt = []
    for i in rng.choice(k.npts, size=6, replace=False):
        ax = int(rng.integers(3))
        e = torch.zeros_like(X0)
        e[i, ax] = 1.0
        fd = (L(X0 + h * e) - L(X0 - h * e)) / (2 * h)
        fd_pt.append(abs(fd - gX[i, ax]) / abs(gX[i, ax]))
    # a random direction and the mode amplitudes move every point: ~1e5
    # pairs sit near the reference's |R.R'/(l l') - 1| < 1e-8 cut-off
    # (tangle.c:505-506) and the flipped ones bias central differences
    d = torch.from_numpy(rng.standard_normal((k.npts, 3)))
    fd = (L(X0 + h * d) - L(X0 - h * d)) / (2 * h)
    ad = float((torch.from_numpy(gX) * d).sum())
    fd_amp = []
    for j in range(len(k.modes)):
        e = torch.zeros(len(k.modes), dtype=torch.float64)
        e[j] = h
        amps = torch.tensor(k.amps)
        fda = (L(A.kelvin_positions(k, amps + e)) - L(A.kelvin_positions(k, amps - e))) / (2 * h)
        fd_amp.append(float(abs(fda - gA[j]) / max(abs(gA[j]), 1e-300)))
    res["fd_h_cm"] = h
    res["fd_rel_err_single_points_max"] = float(max(fd_pt))
    res["fd_rel_err_random_direction"] = float(abs(fd - ad) / abs(ad))
    res["fd_rel_err_amplitudes_max"] = max(fd_amp)
    res["fd_note"] = ("over 200 steps a perturbation along any direction reaches pairs near the "
                      "reference's collinearity cut-off (|R.R'/(l l')-1|<1e-8 zeroes a pair), and "
                      "the flipped pairs bias central differences independently of h; the gradient "
                      "equals autograd through the pinned FP64 restatement over all 200 steps to "
                      "~2e-12 (GPU test), and with the cut-off off that restatement's central "
                      "differences agree with its autograd to ~3e-8")
    res["grad_amplitudes"] = [float(x) for x in gA]
    gm = grads["mixed"][0]
    res["mixed_vs_fp64_grad_rel"] = float(np.linalg.norm(gm - gX) / np.linalg.norm(gX))
    return res


def host_threads():
    """Threads the CPU baseline runs on: OMP_NUM_THREADS when set (the box
    gives each GPU a share of its cores and sets it), else every core this
    process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def tree_section(ctx, T, tg, common, b):
    """The octree at theta 0.3 / 0.5 / 0.7: the reference's multipole form
    (octree.c:303-328) in FP64 and with FP32 multipoles, and the true dipole
    expansion; per evaluation the device time, the traversal kernel's time,
    cells, leaf pairs and multipoles, and (when b holds the FP64 direct sum)
    the error on a 4096-target sample."""
    rng_t = np.random.default_rng(2002)
    samp = np.sort(rng_t.choice(tg.active, size=min(4096, len(tg.active)), replace=False))
    tree = []
    tree_samples = {}   # theta -> (octree entry, GPU FP64 reference-multipole vels on the sample)
    for prec, pname, mp in ((T.PREC_FP64, "fp64", T.MP_REFERENCE), (T.PREC_MIXED, "mixed", T.MP_REFERENCE),
                            (T.PREC_FP64, "fp64", T.MP_DIPOLE)):
        for th in (0.3, 0.5, 0.7):
            xt = tg.copy()
            ctx.reset_stats()
            ctx.tree_stats()
            ctx.velocities(xt, 0.0, T.default_params(precision=prec, use_BH=1, BH_resolution=th,
                                                     multipole=mp, **common), profile=True)
            ms = ctx.last_device_ms()
            cells, pairs, mps = ctx.tree_stats()
            kname = "k_octree_fp64" if prec == T.PREC_FP64 else "k_octree_mixed"
            tree.append({"theta": th, "precision": pname,
                         "multipole": "reference" if mp == T.MP_REFERENCE else "dipole",
                         "eval_ms": round(ms, 2),
                         "traversal_ms": round(ctx.kernel_stats(kname)[1], 2),
                         "cells": int(cells), "leaf_pairs": pairs, "multipoles": mps,
                         "build_kernels_ms": {kn: round(ctx.kernel_stats(kn)[1], 3) for kn in OCT_BUILD}})
            if b is not None:
                tree[-1]["err_vels_vs_fp64_direct_sample"] = err_stats(xt.vels[samp], b.vels[samp])
            if mp == T.MP_REFERENCE and prec == T.PREC_FP64:
                tree_samples[th] = (len(tree) - 1, xt.vels[samp].copy())
            log(f"tree {pname} {tree[-1]['multipole']} theta {th}: {ms:.1f} ms, "
                f"err {tree[-1].get('err_vels_vs_fp64_direct_sample')}")
    hbm = tree_hbm_bytes()
    build_hbm = None
    if hbm:
        build = tree[0]["build_kernels_ms"]
        build_hbm = {kn: {"hbm_bytes": hbm[kn], "ms": build[kn],
                          "GB_per_s": round(hbm[kn] / (build[kn] / 1e3) / 1e9, 1) if build.get(kn) else None}
                     for kn in OCT_BUILD if kn in hbm}
    return samp, tree, tree_samples, build_hbm


def configs_section(ctx, T, ref, threads, full):
    """The per-configuration CPU plan: each configuration's GPU figure beside
    the reference's own C (oracle/_ref, `threads` OpenMP threads).
    256-point ring, open, FP64, dt 1e-5 s, 100 steps, hyperfriction
    off; GPU and CPU positions compared.  configs[1]: 64 rings (seed 1001,
    periodic-26, v_n = 0.5 cm/s along x): one FP64 and one mixed evaluation on
    the GPU, mixed vs FP64 over every point, 200 mixed RK4 steps at dt 5e-6 s;
    the CPU evaluation of every point (full) or of 1024 seeded targets,
    scaled to the whole evaluation.  configs[4]: the 4096-point Kelvin ring,
    FP64 reference RK4 at dt 1e-7 s: 200 GPU steps, 3 CPU steps scaled to 200."""
    res = {}
    p0 = T.default_params(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0)
    g = T.ring(256)
    ctx.rk4_run(g, 0.0, 1e-5, 100, p0)
    c0 = {"points": 256, "steps": 100, "gpu_device_ms": round(ctx.last_device_ms(), 2)}
    if ref is not None:
        a = T.ring(256)
        ref.set_params(p0)
        w0 = time.time()
        for k in range(100):
            ref.step(a, k * 1e-5, 1e-5, p0)
        c0["cpu_reference_ms"] = round((time.time() - w0) * 1e3, 1)
        c0["max_position_diff_cm"] = float(np.abs(a.vnodes - g.vnodes).max())
    res["cfg0_ring256"] = c0

    t1 = T.random_rings(1001, 0.1, n_rings=64)
    n1 = len(t1.active)
    common = dict(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, vn=(0.5, 0.0, 0.0))
    p64 = T.default_params(**common)
    pmx = T.default_params(precision=T.PREC_MIXED, near_radius=0.01, **common)
    a, b = t1.copy(), t1.copy()
    ctx.velocities(a, 0.0, p64)
    f64 = ctx.last_device_ms()
    inter = ctx.last_interactions()
    ctx.velocities(b, 0.0, pmx)
    mx = ctx.last_device_ms()
    act = t1.active
    x = t1.copy()
    ctx.rk4_run(x, 0.0, 5e-6, 200, pmx, enforce_boundaries=True)
    c1 = {"points": n1, "interactions_per_eval": inter, "gpu_fp64_eval_ms": round(f64, 3),
          "gpu_mixed_eval_ms": round(mx, 3), "mixed_vs_fp64": round(f64 / mx, 3),
          "err_vels_mixed_vs_fp64": err_stats(b.vels[act], a.vels[act]),
          "gpu_mixed_rk4_steps": 200, "gpu_mixed_ms_per_step": round(ctx.last_device_ms() / 200, 3)}
    if ref is not None:
        c = t1.copy()
        ref.update_tangents_normals(c)
        idx = act if full else np.sort(np.random.default_rng(2002).choice(act, 1024, replace=False))
        idx = idx.astype(np.int32)
        w0 = time.time()
        _, vel = ref.velocities_sample(c, 0.0, p64, idx)
        cpu_s = time.time() - w0
        c1["cpu_targets"] = int(len(idx))
        c1["cpu_threads"] = threads
        c1["cpu_reference_eval_s"] = round(cpu_s * n1 / len(idx), 2)
        c1["cpu_interactions_per_s"] = inter * len(idx) / n1 / cpu_s
        c1["err_vels_gpu_fp64_vs_reference"] = err_stats(a.vels[idx], vel)
    res["cfg1_64_rings"] = c1

    k4 = T.kelvin_ring()
    p4 = T.default_params(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0)
    g4 = k4.copy()
    ctx.rk4_run(g4, 0.0, 1e-7, 200, p4)
    c4 = {"points": len(k4.active), "steps": 200, "dt_s": 1e-7, "gpu_fp64_ms": round(ctx.last_device_ms(), 2)}
    if ref is not None:
        a4 = k4.copy()
        ref.set_params(p4)
        w0 = time.time()
        for k in range(3):
            ref.step(a4, k * 1e-7, 1e-7, p4)
        c4["cpu_reference_ms_for_200_steps"] = round((time.time() - w0) * 1e3 * 200 / 3, 1)
        c4["cpu_steps_timed"] = 3
    res["cfg4_kelvin4096"] = c4
    return res


def only_section(args, ctx, T, tg, common):
    """One optional section on its own (--only), printed as one JSON line."""
    t0 = time.time()
    res = {"section": args.only, "points": int(len(tg.active)), "seed": args.seed}
    if args.only == "octree":
        res["octree"] = tree_section(ctx, T, tg, common, None)[1]
    elif args.only == "topology":
        res["topology"] = topology_section(ctx, T, tg, None, 0)
    elif args.only == "reconnection":
        res["reconnection_candidates"] = reconnection_section(ctx, tg)
    elif args.only == "ensemble":
        res["ensemble"] = ensemble_section(ctx, T, args)
    elif args.only == "configs":
        from oracle.oracle import Reference
        thr = args.cpu_threads or host_threads()
        ref = Reference(thr) if Reference.available() else None
        res["configs"] = configs_section(ctx, T, ref, thr, True)
    else:
        res["adjoint"] = adjoint_section(ctx, T)
    res["wall_s"] = round(time.time() - t0, 1)
    print(json.dumps(res), flush=True)


def reconnection_section(ctx, tg):
    """Reconnection candidates (the pass-invariant tests of the reference's
    pair loop) on the workload at rec_dist 1e-3 and 2e-4 cm."""
    n = len(tg.active)
    xr = tg.copy()
    ctx.update_tangents_normals(xr)
    rec = []
    for rd in (1e-3, 2e-4):
        ctx.reconnection_candidates(xr, rd, 0.087)          # sizes the buffers
        w0 = time.time()
        cand = ctx.reconnection_candidates(xr, rd, 0.087)
        wall_ms = (time.time() - w0) * 1e3
        rec.append({"rec_dist_cm": rd, "rec_angle": 0.087, "candidates": int(len(cand)),
                    "device_ms": round(ctx.last_device_ms(), 2), "wall_ms": round(wall_ms, 2),
                    "reference_pair_tests": n * (n - 1) // 2})
        log(f"reconnection candidates rec_dist {rd}: {len(cand)} pairs, "
            f"{rec[-1]['device_ms']} ms device, {wall_ms:.1f} ms wall")
    return rec


def main():
    t_start = time.time()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="RK4 steps to time (clamped by --budget)")
    ap.add_argument("--warmup", type=int, default=1,
                    help="untimed RK4 steps of the workload before the timed ones (at most 1 is run)")
    ap.add_argument("--budget", type=float, default=450.0,
                    help="wall seconds for the whole run: the timed steps stop when the next one "
                         "would not fit, the optional sections run only while time is left")
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=1002)
    ap.add_argument("--dt", type=float, default=1e-6)
    ap.add_argument("--near-radius", type=float, default=0.02)
    ap.add_argument("--fp64", type=int, default=1, help="FP64 evaluation for the comparison (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=256, help="targets of the CPU baseline (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--ensemble", type=int, default=256,
                    help="tangles of the batched-ensemble measurement (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=2)
    ap.add_argument("--adjoint", type=int, default=1,
                    help="Kelvin-ring gradient (4096 points, 200 steps, checkpoints every 20; 0: skip)")
    ap.add_argument("--reconnect", type=int, default=1,
                    help="reconnection candidate search on the workload (0: skip)")
    ap.add_argument("--only", choices=("octree", "topology", "reconnection", "ensemble", "adjoint", "configs"),
                    help="run this section alone on the same workload (no headline steps), for profiles")
    ap.add_argument("--configs", type=int, default=1,
                    help="the other configurations beside the reference's own C (0: skip)")
    ap.add_argument("--topology", type=int, default=1,
                    help="device-resident topology edits and the 200-step main loop (0: skip)")
    ap.add_argument("--tree", type=int, default=1,
                    help="octree evaluations at theta 0.3/0.5/0.7 against the FP64 direct sum (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device (HBM-traffic counter calibration) and exit")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = 256 * 1024 * 1024          # 1 GiB of float32
        x = torch.ones(n, device="cuda", dtype=torch.float32)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): this build has no "
              "multi-GPU path (targets would be sharded by stream tiles with one all-gather "
              "of positions per RK stage)", flush=True)
        sys.exit(2)

    def left():
        return args.budget - (time.time() - t_start)

    from openvort_amd import tangle as T
    from openvort_amd.gpu import Context

    tg = T.bh_scale(seed=args.seed, n_points=args.points)
    n = len(tg.active)
    log(f"workload: {n} points ({time.time() - t_start:.1f}s)")
    common = dict(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0)
    p_mx = T.default_params(precision=T.PREC_MIXED, near_radius=args.near_radius, **common)
    p_64 = T.default_params(precision=T.PREC_FP64, **common)

    ctx = Context(0)
    # load every code object of the step on a tiny tangle (not a step of the
    # workload, not profiled)
    for pp in (p_mx, p_64):
        ctx.rk4_run(T.ring(128), 0.0, 1e-6, 1, pp)
    if args.only:
        only_section(args, ctx, T, tg, common)
        ctx.close()
        return

    out = {}
    # ---- one mixed and one FP64 evaluation of the workload's positions ------
    a = tg.copy()
    ctx.reset_stats()
    ctx.velocities(a, 0.0, p_mx, profile=True)
    mx_ms = ctx.last_device_ms()
    I_eval = ctx.last_interactions()
    out["mixed_eval_ms"] = round(mx_ms, 2)
    out["mixed_interactions_per_s_eval"] = I_eval / (mx_ms / 1e3)
    log(f"mixed evaluation: {mx_ms:.1f} ms ({time.time() - t_start:.0f}s)")
    b = None
    if args.fp64 > 0:
        b = tg.copy()
        ctx.reset_stats()
        ctx.velocities(b, 0.0, p_64, profile=True)
        f64_ms = ctx.last_device_ms()
        nl64, kd64 = ctx.kernel_stats("k_direct_fp64")
        act = tg.active
        out["fp64_eval_ms"] = round(f64_ms, 2)
        out["fp64_interactions_per_s"] = I_eval / (f64_ms / 1e3)
        out["mixed_vs_fp64"] = round(f64_ms / mx_ms, 3)
        out["err_vels_mixed_vs_fp64"] = err_stats(a.vels[act], b.vels[act])
        out["err_vs_mixed_vs_fp64"] = err_stats(a.vs[act], b.vs[act])
        fp64_ach = FLOP_PER_INTERACTION * I_eval / (kd64 / nl64 / 1e3)
        out["fp64_roofline"] = {"kernel": "k_direct_fp64", "bound": "valu",
                                "achieved": round(fp64_ach / 1e12, 3), "peak": PEAK_FP64 / 1e12,
                                "unit": "TFLOP/s", "frac": round(fp64_ach / PEAK_FP64, 4)}
        log(f"fp64 evaluation {f64_ms:.1f} ms, err {out['err_vels_mixed_vs_fp64']} "
            f"({time.time() - t_start:.0f}s)")

    # ---- RK4 steps: at most one warm-up step, then timed steps while the
    # budget holds (a step is about four evaluations) -------------------------
    est_step_s = 4.2 * mx_ms / 1e3 + 0.5
    # CPU leg + print, and the optional sections' measured ~60 s (octree, the
    # reference's octree, reconnection, ensemble, adjoint, topology, configs)
    optional = (args.tree, args.reconnect, args.ensemble, args.adjoint, args.topology, args.configs)
    reserve_s = 40.0 + (30.0 if args.cpu_sample > 0 else 0.0) + (70.0 if any(optional) else 0.0)
    warm = 0
    x = tg.copy()
    t_sim = 0.0
    if args.warmup > 0 and left() - reserve_s > 2 * est_step_s:
        ctx.rk4_run(x, t_sim, args.dt, 1, p_mx)
        t_sim += args.dt
        warm = 1
        est_step_s = ctx.last_device_ms() / 1e3 + 0.5
        log(f"warm-up step: {ctx.last_device_ms():.1f} ms ({time.time() - t_start:.0f}s)")
    ctx.reset_stats()
    step_ms = []
    while len(step_ms) < max(args.steps, 1):
        if step_ms and left() - reserve_s < est_step_s:
            break
        ctx.rk4_run(x, t_sim, args.dt, 1, p_mx, profile=True)
        t_sim += args.dt
        step_ms.append(ctx.last_device_ms())
        est_step_s = max(step_ms) / 1e3 + 0.5
        log(f"timed step {len(step_ms)}: {step_ms[-1]:.1f} ms ({time.time() - t_start:.0f}s)")
    steps = len(step_ms)
    dev_ms = float(sum(step_ms))
    ms_per_step = dev_ms / steps
    value = 4 * I_eval * steps / (dev_ms / 1e3)
    n_launch, kd_ms = ctx.kernel_stats("k_direct_mixed")
    kd_avg_s = kd_ms / max(n_launch, 1) / 1e3
    achieved = FLOP_PER_INTERACTION * I_eval / kd_avg_s
    kernels = {}
    for k in ("k_geometry", "k_cont", "k_scan_blocks", "k_records", "k_src_tiles",
              "k_direct_mixed", "k_reduce", "k_epilogue", "k_stage", "k_combine"):
        nl, ms = ctx.kernel_stats(k)
        if nl:
            kernels[k] = {"launches": nl, "avg_ms": round(ms / nl, 4)}

    # ---- CPU baseline: the reference's update_velocity on sampled targets ---
    cpu = None
    c = None
    eng = None
    kind = None
    threads = args.cpu_threads or host_threads()
    if args.cpu_sample > 0:
        from oracle.oracle import Oracle, Reference
        rng = np.random.default_rng(2002)
        idx = np.sort(rng.choice(tg.active, size=min(args.cpu_sample, n), replace=False)).astype(np.int32)
        c = tg.copy()
        Oracle(threads).update_tangents_normals(c)
        if Reference.available():
            kind, eng = "reference", Reference(threads)
        else:
            kind, eng = "port", Oracle(threads)
        c0 = time.time()
        vs_c, vel_c = eng.velocities_sample(c, 0.0, p_64, idx)
        cpu_s = time.time() - c0
        inter = len(idx) * (27 * n - 2)
        cpu = {"value": inter / cpu_s, "unit": "seg-interactions/s", "cores": threads, "kind": kind,
               "nproc": os.cpu_count(),
               "sample": f"{len(idx)} seeded targets of the same tangle (seed 2002), FP64 direct sum "
                         f"+ 26 images, reference update_velocity per target ({inter:.3g} "
                         f"interactions, {cpu_s:.1f} s)"}
        log(f"cpu baseline ({kind}, {threads} threads): {cpu['value']:.4g} int/s in {cpu_s:.1f}s")
        if b is not None:
            out["err_vels_fp64_vs_cpu_sample"] = err_stats(b.vels[idx], vel_c)
        out["err_vels_mixed_vs_cpu_sample"] = err_stats(a.vels[idx], vel_c)

    res = {
        "metric": "Biot-Savart seg-interactions/s & RK4 steps/s, mixed vs FP64, 1M pts @1e-6 rel",
        "value": value,
        "unit": "seg-interactions/s",
        "n_gpus": 1,
        "steps": steps,
        "steps_requested": args.steps,
        "warmup": warm,
        "warmup_requested": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "ms_per_step_median": round(float(np.median(step_ms)), 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "Barnes-Hut-scale periodic tangle: box 0.2 cm, 26 images, 64 perturbed "
                               "box-spanning lines + log-uniform rings (r 0.002-0.03 cm), spacing "
                               "1e-4 cm, seed %d" % args.seed,
                   "points": n, "images": 26, "method": "direct", "precision": "mixed",
                   "near_radius_cm": args.near_radius, "dt_s": args.dt,
                   "interactions_per_eval": I_eval},
        "rk4_steps_per_s": 1e3 / ms_per_step,
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3), "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32, 4),
                     "traffic": measured_traffic(n, args.seed, args.near_radius),
                     "kernel": "k_direct_mixed", "avg_launch_ms": round(kd_avg_s * 1e3, 2),
                     "launches": n_launch, "flop_per_interaction": FLOP_PER_INTERACTION},
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    res.update(out)
    print(json.dumps(res), flush=True)
    log(f"headline printed ({time.time() - t_start:.0f}s)")

    # ---- optional sections, each only while the budget allows it ----------
    skipped = []
    tree_samples = {}   # theta -> (octree entry, GPU FP64 reference-multipole vels on the sample)
    samp = None
    if args.tree > 0 and left() > 40:
        samp, tree, tree_samples, build_hbm = tree_section(ctx, T, tg, common, b)
        if build_hbm:
            out["octree_build_hbm"] = build_hbm
        out["octree"] = tree
    elif args.tree > 0:
        skipped.append("octree")

    # the reference's own octree on the octree section's 4096-target sample
    # (build + evaluation per theta), against the GPU tree's velocities
    if tree_samples and kind == "reference" and left() > 30:
        for th, (entry, vel_g) in tree_samples.items():
            pr = T.default_params(precision=T.PREC_FP64, use_BH=1, BH_resolution=th, **common)
            c0 = time.time()
            _, vel_r = eng.velocities_sample(c, 0.0, pr, samp.astype(np.int32))
            out["octree"][entry]["cpu_reference"] = {
                "threads": threads, "targets": int(len(samp)),
                "build_plus_eval_s": round(time.time() - c0, 2),
                "err_vels_gpu_vs_reference_tree": err_stats(vel_g, vel_r)}
            log(f"cpu reference octree theta {th}: {out['octree'][entry]['cpu_reference']}")
    elif tree_samples and kind == "reference":
        skipped.append("octree_cpu_reference")

    if args.reconnect > 0 and left() > 30:
        out["reconnection_candidates"] = reconnection_section(ctx, tg)
    elif args.reconnect > 0:
        skipped.append("reconnection_candidates")

    if args.ensemble > 0 and left() > 40:
        out["ensemble"] = ensemble_section(ctx, T, args)
        log(f"ensemble: {out['ensemble']}")
    elif args.ensemble > 0:
        skipped.append("ensemble")

    if args.adjoint > 0 and left() > 45:
        out["adjoint"] = adjoint_section(ctx, T)
        log(f"adjoint: {out['adjoint']}")
    elif args.adjoint > 0:
        skipped.append("adjoint")

    if args.topology > 0 and left() > 30:
        out["topology"] = topology_section(ctx, T, tg, eng if kind == "reference" else None, threads)
        log(f"topology: {out['topology']}")
    elif args.topology > 0:
        skipped.append("topology")

    if args.configs > 0 and left() > 30:
        out["configs"] = configs_section(ctx, T, eng if kind == "reference" else None, threads, False)
        log(f"configs: {out['configs']}")
    elif args.configs > 0:
        skipped.append("configs")

    ctx.close()
    out["sections_skipped_for_budget"] = skipped
    out["wall_s"] = round(time.time() - t_start, 1)
    res.update(out)
    print(json.dumps(res), flush=True)


TOPO_KERNELS = ("k_remesh", "k_wrap_index", "k_loop_init", "k_loop_min", "k_rank_init", "k_rank_step",
                "k_loop_len", "k_loop_region", "k_fill_int", "k_loop_and", "k_loop_delete",
                "k_stream_pos", "k_gather_stream", "k_scatter3", "k_scatter1")


def jittered(t, spacing, amp, seed, capacity):
    """t's points each moved by a seeded uniform offset in [-amp, amp] x
    spacing per axis (segment lengths spread around the spacing), in a copy
    of the given capacity."""
    x = t.expand(capacity)
    act = x.active
    x.vnodes[act] += np.random.default_rng(seed).uniform(-amp, amp, (len(act), 3)) * spacing
    return x


def topology_section(ctx, T, tg, ref, threads):
    """Device-resident topology edits.  (a) The 1M-point tangle, each point
    moved by a seeded offset of up to 0.15 x spacing per axis, 15 % spare
    capacity: remesh with dl_min / dl_max = 0.6 / 1.2 x spacing (the example
    configuration's 1:2 ratio; ~6 % of the segments are too long),
    eliminate_small_loops, enforce_boundaries.  (b) The 64-ring periodic
    tangle (0.1 cm box, 2e-4 cm spacing) moved the same way: remesh on the
    GPU and with the reference's own remesh on the host (its OpenMP loops on
    `threads` threads), links compared.  (c) The reference's main-loop body
    (update_tangle, rk4_step, update_tangents_normals, remesh,
    eliminate_small_loops, enforce_boundaries) for 200 steps of the 64-ring
    tangle with v_n = 0.5 cm/s along x and dt = 5e-6 s, dl_min / dl_max =
    1.2e-4 / 2.4e-4 cm, mixed direct sum, resident on the device
    (ov_gpu_run).  (d) The same loop on the 1M-point tangle with the octree
    (theta 0.5, FP32 multipoles), 5 steps at dt = 1e-6 s."""
    res = {}
    x = jittered(tg, 1e-4, 0.15, 5, tg.N + (tg.N * 3) // 20)
    ctx.update_tangents_normals(x)
    lo, hi = 0.6e-4, 1.2e-4
    p = T.default_params(dl_max=hi)
    n0 = len(x.active)
    w0 = time.time()
    ctx.remesh(x, p, lo, hi)
    wall = (time.time() - w0) * 1e3
    rm = ctx.last_device_ms()
    n1 = len(x.active)
    inserted = int(np.count_nonzero(x.status[tg.N:, 0] != T.EMPTY))
    w0 = time.time()
    ctx.eliminate_small_loops(x, T.default_topology())
    el_wall, el = (time.time() - w0) * 1e3, ctx.last_device_ms()
    w0 = time.time()
    ctx.enforce_boundaries(x)
    en_wall, en = (time.time() - w0) * 1e3, ctx.last_device_ms()
    res["remesh_1m"] = {"points": n0, "capacity": x.N, "dl_min_cm": lo, "dl_max_cm": hi,
                        "points_after": n1, "inserted_above_N": inserted,
                        "device_ms": round(rm, 2), "wall_ms": round(wall, 1),
                        "eliminate_small_loops_device_ms": round(el, 3), "eliminate_wall_ms": round(el_wall, 1),
                        "enforce_boundaries_device_ms": round(en, 3), "enforce_wall_ms": round(en_wall, 1)}

    base = T.random_rings(1001, 0.1, n_rings=64)
    c1 = jittered(base, 2e-4, 0.15, 7, 2 * base.N)
    ctx.update_tangents_normals(c1)
    lo1, hi1 = 1.2e-4, 2.4e-4
    p1 = T.default_params(dl_max=hi1)
    g = c1.copy()
    ctx.remesh(g, p1, lo1, hi1)
    r1 = {"points": int(len(c1.active)), "dl_min_cm": lo1, "dl_max_cm": hi1,
          "points_after": int(len(g.active)), "gpu_device_ms": round(ctx.last_device_ms(), 3)}
    if ref is not None:
        # the reference's own remesh (oracle/_ref) on the same tangle
        a = c1.copy()
        ref.set_params(p1)
        ref.update_tangents_normals(a)
        c0 = time.time()
        ref.remesh(a, lo1, hi1, need_free=len(c1.active))
        r1["cpu_reference_ms"] = round((time.time() - c0) * 1e3, 2)
        r1["cpu_threads"] = threads
        r1["links_identical"] = bool(np.array_equal(a.connections, g.connections)
                                     and np.array_equal(a.status[:, 0], g.status[:, 0]))
        r1["max_position_diff_cm"] = float(np.abs(a.vnodes[a.active] - g.vnodes[a.active]).max())
    res["remesh_64_rings"] = r1

    c2 = T.random_rings(1001, 0.1, n_rings=64)
    c2 = c2.expand(2 * c2.N)
    p2 = T.default_params(vortex_width=1.3e-8, precision=T.PREC_MIXED, near_radius=0.01, dl_max=2.4e-4,
                          vn=(0.5, 0.0, 0.0))
    tp = T.default_topology(dl_min=1.2e-4, dl_max=2.4e-4)
    ctx.update_tangents_normals(c2)
    n_start = len(c2.active)
    w0 = time.time()
    ctx.run(c2, 0.0, 5e-6, 200, p2, tp)
    run_wall = time.time() - w0
    run_ms = ctx.last_device_ms()
    n_end = int(len(c2.active))
    # 20 more steps with per-kernel timing, for the breakdown
    ctx.reset_stats()
    ctx.run(c2, 200 * 5e-6, 5e-6, 20, p2, tp, profile=True)
    topo_ms = sum(ctx.kernel_stats(k)[1] for k in TOPO_KERNELS)
    res["main_loop_64_rings"] = {
        "steps": 200, "dt_s": 5e-6, "points_start": n_start, "points_end": n_end,
        "device_ms_per_step": round(run_ms / 200, 3), "wall_s": round(run_wall, 2),
        "profiled_steps": 20,
        "topology_kernels_ms_per_step": round(topo_ms / 20, 3),
        "k_remesh_ms_per_step": round(ctx.kernel_stats("k_remesh")[1] / 20, 3),
        "k_direct_mixed_ms_per_step": round(ctx.kernel_stats("k_direct_mixed")[1] / 20, 3)}

    # (d) the main loop on the 1M-point tangle with the octree (theta 0.5,
    # FP32 multipoles), dt 1e-6 s, dl_min / dl_max = 0.6 / 1.2 x spacing
    x = tg.expand(tg.N + tg.N // 20)
    pb = T.default_params(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, precision=T.PREC_MIXED,
                          use_BH=1, BH_resolution=0.5, dl_max=1.2e-4)
    tpb = T.default_topology(dl_min=0.6e-4, dl_max=1.2e-4)
    ctx.update_tangents_normals(x)
    n_start = int(len(x.active))
    ctx.reset_stats()
    ctx.run(x, 0.0, 1e-6, 5, pb, tpb, profile=True)
    res["main_loop_1m_octree"] = {
        "steps": 5, "dt_s": 1e-6, "theta": 0.5, "points_start": n_start, "points_end": int(len(x.active)),
        "device_ms_per_step": round(ctx.last_device_ms() / 5, 2),
        "k_octree_mixed_ms_per_step": round(ctx.kernel_stats("k_octree_mixed")[1] / 5, 2),
        "topology_kernels_ms_per_step": round(sum(ctx.kernel_stats(k)[1] for k in TOPO_KERNELS) / 5, 2)}
    return res


def ensemble_section(ctx, T, args):
    """The 256-tangle ensemble (box 0.05 cm, 500-20000 points each, own
    counterflow and temperature per tangle): batched RK4 steps, one launch per
    evaluation for the whole ensemble, against a per-tangle loop of
    ov_gpu_rk4_run on the same tangles."""
    ens = T.ensemble(n_tangles=args.ensemble)

    def ep(kw):
        kw = dict(kw)
        kw.pop("temperature")
        return T.default_params(vortex_width=1.3e-8, hyperfriction=0, precision=T.PREC_MIXED,
                                near_radius=args.near_radius, **kw)
    ps = [ep(kw) for _, kw in ens]
    ts = [t.copy() for t, _ in ens]
    ctx.rk4_run_batch(ts, 0.0, 5e-6, 0, ps)          # layout + tuner, no step
    ts = [t.copy() for t, _ in ens]
    ctx.rk4_run_batch(ts, 0.0, 5e-6, args.ensemble_steps, ps)
    b_ms = ctx.last_device_ms()
    U, blocks, tune_ms = ctx.batch_tuning()
    loop_ms = 0.0
    ts2 = [t.copy() for t, _ in ens]
    for x, pp in zip(ts2, ps):
        ctx.rk4_run(x, 0.0, 5e-6, args.ensemble_steps, pp)
        loop_ms += ctx.last_device_ms()
    npts = sum(len(t.active) for t, _ in ens)
    inter = sum(27.0 * len(t.active) ** 2 for t, _ in ens)
    return {
        "tangles": args.ensemble, "points": npts, "steps": args.ensemble_steps,
        "batched_ms_per_step": round(b_ms / args.ensemble_steps, 2),
        "loop_ms_per_step": round(loop_ms / args.ensemble_steps, 2),
        "batched_vs_loop": round(loop_ms / b_ms, 3),
        "aggregate_steps_per_s_batched": args.ensemble * args.ensemble_steps / (b_ms / 1e3),
        "aggregate_steps_per_s_loop": args.ensemble * args.ensemble_steps / (loop_ms / 1e3),
        "interactions_per_s_batched": 4 * inter * args.ensemble_steps / (b_ms / 1e3),
        "tuned_unit_items": U, "tuned_blocks": blocks, "tuning_ms": round(tune_ms, 1),
        "max_position_diff_batched_vs_loop": float(max(np.abs(a.vnodes - c.vnodes).max()
                                                     for a, c in zip(ts, ts2)))}


if __name__ == "__main__":
    main()
