"""Gradients through the GPU velocity and RK4 run, as torch autograd Functions.

The forward of each Function calls the HIP library (ov_gpu_velocities /
ov_gpu_rk4_run); its backward calls the hand-written adjoint
(ov_gpu_velocity_vjp / ov_gpu_rk4_run_vjp), which recomputes the trajectory
on the device with checkpoints.  Positions are (N, 3) float64 tensors
indexed like the tangle's vnodes; torch is only the autograd plumbing — every
derivative of the filament dynamics comes from the library.

The Kelvin-ring differentiation workload lives here too: a ring of radius r
whose points carry circularly polarised Kelvin waves with amplitudes A_m
(modes 2..8, seeded amplitudes and phases, as tangle.kelvin_ring), a loss of
the final state — squared distance of the centroid to a seeded target plus
the Kelvin-wave energy, defined here as the mean squared displacement of the
points from the ring's mean circle:
    E = mean_k[(rho_k - mean rho)^2 + (z_k - c_z)^2],
rho_k = |(x_k - c)_xy|, c the centroid, the ring axis along z.  Neither the
displacement split nor this energy is defined by the reference.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import tangle as T
from .gpu import Context


@dataclass
class Run:
    """What an RK4Run differentiates: nsteps steps of `tangle`'s topology
    (its vnodes are replaced by the input) under `params`."""
    ctx: Context
    tangle: T.Tangle
    params: T.Params
    t0: float
    dt: float
    nsteps: int
    checkpoint_every: int = 20
    enforce_boundaries: bool = False
    profile: bool = False   # per-kernel HIP-event timings of the adjoint (Context.kernel_stats)


def _with_positions(t: T.Tangle, X: torch.Tensor) -> T.Tangle:
    c = t.copy()
    c.vnodes[:] = X.detach().to("cpu", torch.float64).numpy()
    return c


class RK4Run(torch.autograd.Function):
    """X0 (N, 3) -> positions after run.nsteps RK4 steps."""

    @staticmethod
    def forward(fctx, X0: torch.Tensor, run: Run):
        t = _with_positions(run.tangle, X0)
        run.ctx.rk4_run(t, run.t0, run.dt, run.nsteps, run.params,
                        enforce_boundaries=run.enforce_boundaries)
        fctx.save_for_backward(X0)
        fctx.run = run
        return torch.from_numpy(t.vnodes.copy()).to(X0.device)

    @staticmethod
    def backward(fctx, g: torch.Tensor):
        (X0,) = fctx.saved_tensors
        run = fctx.run
        t = _with_positions(run.tangle, X0)
        gx = run.ctx.rk4_run_vjp(t, run.t0, run.dt, run.nsteps, run.params,
                                 g.detach().to("cpu", torch.float64).numpy(),
                                 checkpoint_every=run.checkpoint_every,
                                 enforce_boundaries=run.enforce_boundaries, profile=run.profile)
        return torch.from_numpy(gx).to(g.device), None


class Velocity(torch.autograd.Function):
    """X (N, 3) -> vels (N, 3) (update_tangle + update_velocities)."""

    @staticmethod
    def forward(fctx, X: torch.Tensor, ctx: Context, tangle: T.Tangle, params: T.Params,
                time: float = 0.0):
        t = _with_positions(tangle, X)
        ctx.velocities(t, time, params, geometry=True)
        fctx.save_for_backward(X)
        fctx.args = (ctx, tangle, params, time)
        return torch.from_numpy(t.vels.copy()).to(X.device)

    @staticmethod
    def backward(fctx, g: torch.Tensor):
        (X,) = fctx.saved_tensors
        ctx, tangle, params, time = fctx.args
        t = _with_positions(tangle, X)
        gx = ctx.velocity_vjp(t, time, params, g.detach().to("cpu", torch.float64).numpy())
        return torch.from_numpy(gx).to(g.device), None, None, None, None


def rk4_run(X0: torch.Tensor, run: Run) -> torch.Tensor:
    return RK4Run.apply(X0, run)


def velocity(X: torch.Tensor, ctx: Context, tangle: T.Tangle, params: T.Params,
             time: float = 0.0) -> torch.Tensor:
    return Velocity.apply(X, ctx, tangle, params, time)


# ---- Kelvin-ring differentiation workload ------------------------------------

@dataclass
class KelvinRing:
    npts: int = 4096
    r: float = 0.01
    modes: tuple = tuple(range(2, 9))
    amps: np.ndarray = None      # seeded U[0, amp_hi]
    phases: np.ndarray = None    # seeded U[0, 2 pi)
    target: np.ndarray = None    # seeded centroid target (cm)


def kelvin_setup(seed=1004, npts=4096, r=0.01, modes=tuple(range(2, 9)), amp_hi=2e-4) -> KelvinRing:
    """Amplitudes and phases drawn exactly as tangle.kelvin_ring draws them
    (SplitMix64(seed)); the centroid target comes from SplitMix64(seed + 1):
    the ring's self-induced drift direction (z) times U[0.5, 1.5] r, plus
    U[-1, 1] 0.1 r in x and y."""
    rng = T.SplitMix64(seed)
    amps = np.array([rng.uniform(0, amp_hi) for _ in modes])
    phases = np.array([rng.uniform(0, 2 * math.pi) for _ in modes])
    g = T.SplitMix64(seed + 1)
    target = np.array([g.uniform(-1, 1) * 0.1 * r, g.uniform(-1, 1) * 0.1 * r, g.uniform(0.5, 1.5) * r])
    return KelvinRing(npts, r, tuple(modes), amps, phases, target)


def kelvin_positions(k: KelvinRing, amps: torch.Tensor) -> torch.Tensor:
    """(npts, 3) positions of the ring for mode amplitudes `amps`, the same
    placement as tangle.kelvin_ring (differentiable in amps)."""
    phi = torch.arange(k.npts, dtype=torch.float64) * (2 * math.pi / k.npts)
    er = torch.stack([torch.sin(phi), -torch.cos(phi), torch.zeros_like(phi)], 1)
    ez = torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64)
    X = k.r * er
    for j, m in enumerate(k.modes):
        arg = m * phi + float(k.phases[j])
        X = X + amps[j] * (torch.cos(arg)[:, None] * er + torch.sin(arg)[:, None] * ez)
    return X


def kelvin_loss(X: torch.Tensor, target) -> torch.Tensor:
    """|centroid - target|^2 + Kelvin-wave energy (mean squared displacement
    from the mean circle), both in cm^2."""
    c = X.mean(0)
    tgt = torch.as_tensor(target, dtype=X.dtype)
    d = X - c
    rho = torch.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2)
    energy = ((rho - rho.mean()) ** 2 + d[:, 2] ** 2).mean()
    return ((c - tgt) ** 2).sum() + energy
