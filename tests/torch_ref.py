"""Torch float64 restatement of the point velocity, for autograd.

TEST INFRASTRUCTURE ONLY: the gradients of the hand-written adjoint are
checked against torch.autograd through this restatement; the restatement's
values are checked against the oracle (tests/test_adjoint_ref.py), and its
gradients against central differences of the oracle, so it is pinned to the
reference on both counts.

Follows reference update_velocity (src/tangle.c:583-684): geometry by the
five-node stencil (src/tangle.c:432-479), LIA (src/tangle.c:530-551), the
straight-segment law (src/tangle.c:486-514) over segments and images
(src/tangle.c:39-79, 553-581), hyperfriction and mutual friction
(src/tangle.c:641-675).  Minimum-image shifts and the zero rules are decided
on values (no gradient flows through a decision).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from openvort_amd import tangle as T


class Topo:
    """Index maps and constants of a tangle (active points only)."""

    def __init__(self, t: T.Tangle, p: T.Params):
        self.act = t.active
        pos = -np.ones(t.N, dtype=np.int64)
        pos[self.act] = np.arange(len(self.act))
        self.fwd = torch.as_tensor(pos[t.connections[self.act, 0]])
        self.rev = torch.as_tensor(pos[t.connections[self.act, 1]])
        lo, hi = np.asarray(t.box_lo, float), np.asarray(t.box_hi, float)
        self.lo, self.hi, self.L = lo, hi, hi - lo
        self.periodic = [t.walls[2 * i] == T.WALL_PERIODIC for i in range(3)]
        self.images = [((0.0, 0.0, 0.0), -1, True)]
        for sx, sy, sz, rf in t.images:
            self.images.append(((self.L[0] * sx, self.L[1] * sy, self.L[2] * sz), rf, False))
        self.p = p
        # collinearity cut-off of the segment law (src/tangle.c:505-506); the
        # reference's 1e-8 unless a check sets it otherwise
        self.ndot_tol = 1e-8
        # when a list: the number of pairs the cut-off zeroes, per image sum
        self.cut_counts = None


def zero_adj(adjacent, primary):
    """the pairs the primary sum skips anyway"""
    return adjacent if primary else torch.zeros_like(adjacent)


def pwrap(r1, r2, topo):
    """minimum image of r2 seen from r1 (rows)"""
    out = []
    for i in range(3):
        x = r2[:, i]
        if topo.periodic[i]:
            d = (r2[:, i] - r1[:, i]).detach()
            L = float(topo.L[i])
            zero = torch.zeros_like(d)
            sh = torch.where(d > L / 2, zero - L, torch.where(d < -L / 2, zero + L, zero))
            x = x + sh
        out.append(x)
    return torch.stack(out, 1)


def geometry(X, topo):
    """(s', s'', l+, l-) of every point."""
    f, r = topo.fwd.to(X.device), topo.rev.to(X.device)
    p0, pf, pff, pr, prr = X, X[f], X[f[f]], X[r], X[r[r]]
    chord = lambda a, b: torch.linalg.norm(a - pwrap(a, b, topo), dim=1)
    dk, df = chord(p0, pf), chord(pf, pff)
    dr, drr = chord(pr, p0), chord(prr, pr)
    x = [dk + df, dk, -dr, -(dr + drr)]
    pj = [pff, pf, pr, prr]
    t = torch.zeros_like(X)
    nm = torch.zeros_like(X)
    for j in range(4):
        o = [x[m] for m in range(4) if m != j]
        den = x[j] * (x[j] - o[0]) * (x[j] - o[1]) * (x[j] - o[2])
        w1 = -(o[0] * o[1] * o[2]) / den
        w2 = 2 * (o[0] * o[1] + o[0] * o[2] + o[1] * o[2]) / den
        d = pj[j] - pwrap(pj[j], p0, topo)
        t = t + w1[:, None] * d
        nm = nm + w2[:, None] * d
    return t, nm, dk, dr


def biot_savart(X, topo):
    """kappa/4pi sum over segments and images of the straight-segment law."""
    p = topo.p
    n = X.shape[0]
    a = X
    fwd = topo.fwd.to(X.device)
    b = pwrap(a, X[fwd], topo)
    idx = torch.arange(n, device=X.device)
    adjacent = (idx[:, None] == idx[None, :]) | (fwd[None, :] == idx[:, None])
    out = torch.zeros_like(X)
    one = torch.tensor([1.0, 0.0, 0.0], dtype=X.dtype, device=X.device)
    two = torch.tensor([0.0, 1.0, 0.0], dtype=X.dtype, device=X.device)
    for shift, rf, primary in topo.images:
        rs = X - torch.tensor(shift, dtype=X.dtype, device=X.device)
        if rf >= 0:
            ax = rf // 2
            wall = topo.hi[ax] if rf & 1 else topo.lo[ax]
            rs = rs.clone()
            rs[:, ax] = 2 * wall - X[:, ax]
        R = a[None, :, :] - rs[:, None, :]      # [target, segment, 3]
        Rp = b[None, :, :] - rs[:, None, :]
        with torch.no_grad():
            l = torch.linalg.norm(R, dim=2)
            lp = torch.linalg.norm(Rp, dim=2)
            nd = (R * Rp).sum(2) / l / lp
            cut = (nd - 1).abs() < topo.ndot_tol
            zero = (l < 1e-8) | (lp < 1e-8) | cut
            if primary:
                zero = zero | adjacent
            if topo.cut_counts is not None:
                topo.cut_counts.append(int((cut & ~zero_adj(adjacent, primary)).sum()))
        keep = ~zero
        Rs = torch.where(keep[..., None], R, one)
        Rps = torch.where(keep[..., None], Rp, two)
        l = torch.linalg.norm(Rs, dim=2)
        lp = torch.linalg.norm(Rps, dim=2)
        d = (Rs * Rps).sum(2)
        f = (l + lp) / (l * lp * (l * lp + d))
        v = f[..., None] * torch.cross(Rs, Rps, dim=2)
        v = torch.where(keep[..., None], v, torch.zeros_like(v)).sum(1)
        if rf >= 0:
            v = v.clone()
            v[:, rf // 2] = -v[:, rf // 2]
        out = out + v
    return out * (p.kappa / 4 / math.pi)


def velocities(X, topo):
    """(vs, vels) of every active point at positions X (active rows)."""
    p = topo.p
    t, nm, dk, dr = geometry(X, topo)
    beta = p.kappa / 4 / math.pi * torch.log(2 * torch.sqrt(dk * dr) / math.exp(0.25) / p.vortex_width)
    lia = beta[:, None] * torch.cross(t, nm, dim=1)
    vs_ext = torch.tensor(list(p.vs.v) if p.vs.type == T.FLOW_SIMPLE else [0.0] * 3, dtype=X.dtype,
                          device=X.device)
    vn_ext = torch.tensor(list(p.vn.v) if p.vn.type == T.FLOW_SIMPLE else [0.0] * 3, dtype=X.dtype,
                          device=X.device)
    vs = lia + vs_ext
    if not p.LIA_only:
        vs = vs + biot_savart(X, topo)
    vel = vs
    if p.hyperfriction:
        spp = torch.linalg.norm(nm, dim=1).detach()
        on = (spp > p.max_curvature_scale / p.dl_max).to(X.dtype)
        vel = vel - p.hyperalpha * on[:, None] * torch.cross(t, lia, dim=1)
    if p.use_mutual_friction:
        dv = vn_ext - vs
        a1 = torch.cross(t, dv.expand_as(t), dim=1)
        a2 = torch.cross(t, a1, dim=1)
        vel = vel + p.alpha * a1 - p.alpha_p * a2
    return vs, vel


def rk4_run(X, topo, dt, nsteps, variant=T.RK_REFERENCE):
    """nsteps of the reference RK4 (src/vortex_dynamics.c:58-121; classic
    with variant RK_CLASSIC), no boundary wrap."""
    h3 = dt if variant == T.RK_CLASSIC else dt / 2
    for _ in range(nsteps):
        k1 = velocities(X, topo)[1]
        k2 = velocities(X + dt / 2 * k1, topo)[1]
        k3 = velocities(X + dt / 2 * k2, topo)[1]
        k4 = velocities(X + h3 * k3, topo)[1]
        X = X + (k1 * (dt / 6) + k2 * (dt / 3) + k3 * (dt / 3) + k4 * (dt / 6))
    return X
