"""GPU adjoint (ov_gpu_velocity_vjp, ov_gpu_rk4_run_vjp) against
torch.autograd through the pinned torch restatement (tests/torch_ref.py) and
against central differences of the oracle and of the GPU forward.

Tolerances: the hand-written VJP and autograd differentiate the same FP64
function in different orders, so they agree to rounding: 1e-10 relative
(max over points of |g - g_ref| / rms |g_ref|).  Central differences
(h = 1e-9 cm) carry their own truncation and rounding: 1e-6 relative.
"""
import numpy as np
import pytest
import torch

from openvort_amd import adjoint as A
from openvort_amd import tangle as T
from openvort_amd.gpu import OV_E_UNSUPPORTED, OVError
from tests import adjoint_cases as AC
from tests import torch_ref as TR

pytestmark = pytest.mark.gpu

AD_TOL = 1e-10
FD_TOL = 1e-6


def _err(g, ref):
    g, ref = np.asarray(g).reshape(-1, 3), np.asarray(ref).reshape(-1, 3)
    return float(np.linalg.norm(g - ref, axis=1).max() / np.sqrt(np.mean(np.sum(ref * ref, 1))))


def _ref_velocity_vjp(t, p, w_act):
    topo = TR.Topo(t, p)
    X = torch.from_numpy(t.vnodes[t.active].copy()).requires_grad_(True)
    (TR.velocities(X, topo)[1] * torch.from_numpy(w_act)).sum().backward()
    return X.grad.numpy()


@pytest.mark.parametrize("name", AC.NAMES)
def test_velocity_vjp_matches_autograd(ctx, name):
    t, p = AC.case(name)
    print("case", name, "points", len(t.active), flush=True)
    w = AC.seeded((t.N, 3), 5)
    g = ctx.velocity_vjp(t, 0.0, p, w)
    ref = _ref_velocity_vjp(t, p, w[t.active])
    e = _err(g[t.active], ref)
    print(name, "vjp vs autograd", e, flush=True)
    assert e < AD_TOL
    empty = np.setdiff1d(np.arange(t.N), t.active)
    assert np.all(g[empty] == 0)


def test_velocity_vjp_many_tiles_matches_autograd(ctx):
    """1024 points: 16 tiles, 128 chunks per pass (the small cases fit one tile)."""
    t = T.kelvin_ring(npts=1024)
    p = T.default_params(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0,
                         vs=(0.0, 0.01, 0.0), vn=(0.2, 0.0, 0.0))
    w = AC.seeded((t.N, 3), 14)
    g = ctx.velocity_vjp(t, 0.0, p, w)
    e = _err(g, _ref_velocity_vjp(t, p, w))
    print("1024-point vjp vs autograd", e, flush=True)
    assert e < AD_TOL


@pytest.mark.parametrize("name", ["ring_mf_flows", "periodic_holes", "mirror", "small_loops"])
def test_segment_pass_forms_agree(ctx, name):
    """The fused segment pass (each pair once, target-tile partials) and the
    two-pass form (ov_ctx_set_adjoint_partial_limit 0) give the same VJPs to
    rounding (1e-13; they sum the segment terms in different orders), and the
    two-pass form matches autograd as the fused one does."""
    t, p = AC.case(name)
    w = AC.seeded((t.N, 3), 15)
    g1 = ctx.velocity_vjp(t, 0.0, p, w)
    r1 = ctx.rk4_run_vjp(t, 0.0, 2e-6, 3, p, w, checkpoint_every=2)
    ctx.set_adjoint_partial_limit(0)
    try:
        g2 = ctx.velocity_vjp(t, 0.0, p, w)
        r2 = ctx.rk4_run_vjp(t, 0.0, 2e-6, 3, p, w, checkpoint_every=2)
    finally:
        ctx.set_adjoint_partial_limit(128 << 20)
    e, er = _err(g2, g1), _err(r2, r1)
    ea = _err(g2[t.active], _ref_velocity_vjp(t, p, w[t.active]))
    print(name, "two-pass vs fused", e, er, "two-pass vs autograd", ea, flush=True)
    assert e < 1e-13 and er < 1e-13 and ea < AD_TOL


def test_fused_pass_after_a_large_tangle(ctx):
    """The fused pass's partial planes are strided by the tangle's own point
    count, so a context that held a large tangle before (buffers of its
    capacity) still takes the fused pass for a small one, with the same
    result as a fresh context's."""
    t, p = AC.case("ring_mf_flows")
    w = AC.seeded((t.N, 3), 16)
    g0 = ctx.velocity_vjp(t, 0.0, p, w)
    big = T.random_rings(1009, 0.1, n_points_target=60000, spacing=2e-4, r_lo=0.002, r_hi=0.02)
    ctx.velocities(big, 0.0, T.default_params(vortex_width=1.3e-8, hyperfriction=0))
    ctx.reset_stats()
    g1 = ctx.velocity_vjp(t, 0.0, p, w, profile=True)
    n_fused, _ = ctx.kernel_stats("ka_bs_fused")
    print("after", len(big.active), "points: fused launches", n_fused, flush=True)
    assert n_fused == 1
    assert np.array_equal(g1, g0)


@pytest.mark.parametrize("name", ["ring_mf_flows", "periodic_holes", "mirror"])
def test_velocity_vjp_matches_oracle_fd(ctx, oracle, name):
    t, p = AC.case(name)
    w = AC.seeded((t.N, 3), 6)
    g = ctx.velocity_vjp(t, 0.0, p, w)
    act = t.active
    h = 1e-9
    for s in range(2):
        d = np.zeros((t.N, 3))
        d[act] = AC.seeded((len(act), 3), 200 + s)
        a, b = t.copy(), t.copy()
        a.vnodes += h * d
        b.vnodes -= h * d
        oracle.update_tangle(a, 0.0, p)
        oracle.update_tangle(b, 0.0, p)
        fd = float(((a.vels - b.vels)[act] * w[act]).sum() / (2 * h))
        ad = float((g * d).sum())
        print(name, s, "fd", fd, "ad", ad, flush=True)
        assert abs(fd - ad) <= FD_TOL * abs(ad)


@pytest.mark.parametrize("variant", [T.RK_REFERENCE, T.RK_CLASSIC])
@pytest.mark.parametrize("name", ["ring_mf_flows", "periodic_holes", "small_loops"])
def test_rk4_run_vjp_matches_autograd(ctx, name, variant):
    """Three steps: GPU adjoint vs autograd through the torch RK4."""
    t, p = AC.case(name)
    p.rk_variant = variant
    dt, nsteps = 2e-6, 3
    w = AC.seeded((t.N, 3), 7)
    g = ctx.rk4_run_vjp(t, 0.0, dt, nsteps, p, w, checkpoint_every=2)
    topo = TR.Topo(t, p)
    X = torch.from_numpy(t.vnodes[t.active].copy()).requires_grad_(True)
    (TR.rk4_run(X, topo, dt, nsteps, variant) * torch.from_numpy(w[t.active])).sum().backward()
    e = _err(g[t.active], X.grad.numpy())
    print(name, variant, "rk4 vjp vs autograd", e, flush=True)
    assert e < AD_TOL
    empty = np.setdiff1d(np.arange(t.N), t.active)
    assert np.array_equal(g[empty], w[empty])   # EMPTY points do not move


def test_rk4_run_vjp_checkpointing_is_exact(ctx):
    """Every checkpoint spacing recomputes the same trajectory: identical bits."""
    t, p = AC.case("periodic_holes")
    w = AC.seeded((t.N, 3), 8)
    gs = [ctx.rk4_run_vjp(t, 0.0, 1e-6, 7, p, w, checkpoint_every=k) for k in (0, 1, 3, 7, 100)]
    for g in gs[1:]:
        assert np.array_equal(g, gs[0])


def test_rk4_run_vjp_matches_gpu_fd(ctx):
    """Ten steps with the periodic wrap: central differences of the GPU
    forward (h = 1e-9 cm) against the adjoint."""
    t, p = AC.case("periodic_holes")
    dt, nsteps = 2e-6, 10
    act = t.active
    w = np.zeros((t.N, 3))
    w[act] = AC.seeded((len(act), 3), 9)
    g = ctx.rk4_run_vjp(t, 0.0, dt, nsteps, p, w, checkpoint_every=4, enforce_boundaries=True)
    h = 1e-9
    for s in range(2):
        d = np.zeros((t.N, 3))
        d[act] = AC.seeded((len(act), 3), 300 + s)
        a, b = t.copy(), t.copy()
        a.vnodes += h * d
        b.vnodes -= h * d
        ctx.rk4_run(a, 0.0, dt, nsteps, p, enforce_boundaries=True)
        ctx.rk4_run(b, 0.0, dt, nsteps, p, enforce_boundaries=True)
        fd = float(((a.vnodes - b.vnodes)[act] * w[act]).sum() / (2 * h))
        ad = float((g * d).sum())
        print(s, "fd", fd, "ad", ad, flush=True)
        assert abs(fd - ad) <= FD_TOL * abs(ad)


def test_rk4_run_vjp_zero_steps_is_identity(ctx):
    t, p = AC.case("ring_mf_flows")
    w = AC.seeded((t.N, 3), 10)
    assert np.array_equal(ctx.rk4_run_vjp(t, 0.0, 1e-6, 0, p, w), w)


def test_adjoint_rejects_tree_and_caller_k1(ctx):
    t, p = AC.case("ring_mf_flows")
    w = AC.seeded((t.N, 3), 10)
    p.use_BH = 1
    with pytest.raises(OVError) as e:
        ctx.velocity_vjp(t, 0.0, p, w)
    assert e.value.code == OV_E_UNSUPPORTED
    p.use_BH = 0
    from openvort_amd.gpu import OV_USE_CALLER_K1
    import ctypes as C
    g = np.zeros((t.N, 3))
    ts = t.cstruct()
    rc = ctx.lib.ov_gpu_rk4_run_vjp(ctx.ptr, C.byref(ts), 0.0, 1e-6, 1, C.byref(p), OV_USE_CALLER_K1,
                                    0, w.ctypes.data, g.ctypes.data)
    assert rc == OV_E_UNSUPPORTED


def test_mixed_forward_gradient_close_to_fp64(ctx):
    """precision=MIXED changes only the forward sums entering the adjoint."""
    t, p = AC.case("periodic_holes")
    w = AC.seeded((t.N, 3), 12)
    g64 = ctx.rk4_run_vjp(t, 0.0, 2e-6, 3, p, w)
    p.precision = T.PREC_MIXED
    gm = ctx.rk4_run_vjp(t, 0.0, 2e-6, 3, p, w)
    e = _err(gm[t.active], g64[t.active])
    print("mixed vs fp64 gradient", e, flush=True)
    assert e < 1e-6


def test_autograd_function_kelvin_amplitudes(ctx):
    """d loss / d amplitudes through RK4Run (GPU forward + adjoint) against
    central differences of the GPU forward in the amplitudes."""
    k = A.kelvin_setup(npts=256)
    base = T.kelvin_ring(npts=256)
    p = T.default_params(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0)
    run = A.Run(ctx, base, p, 0.0, 1e-6, 6, checkpoint_every=4)
    amps = torch.tensor(k.amps, requires_grad=True)
    loss = A.kelvin_loss(A.rk4_run(A.kelvin_positions(k, amps), run), k.target)
    loss.backward()
    h = 1e-9
    for j in range(len(k.modes)):
        e = torch.zeros_like(amps)
        e[j] = h
        with torch.no_grad():
            lp = A.kelvin_loss(A.rk4_run(A.kelvin_positions(k, amps + e), run), k.target)
            lm = A.kelvin_loss(A.rk4_run(A.kelvin_positions(k, amps - e), run), k.target)
        fd = float((lp - lm) / (2 * h))
        ad = float(amps.grad[j])
        print("mode", k.modes[j], "fd", fd, "ad", ad, flush=True)
        assert abs(fd - ad) <= 1e-5 * max(abs(ad), 1e-3)


def test_velocity_function_backward(ctx):
    t, p = AC.case("ring_mf_flows")
    X = torch.from_numpy(t.vnodes.copy()).requires_grad_(True)
    w = torch.from_numpy(AC.seeded((t.N, 3), 13))
    (A.velocity(X, ctx, t, p) * w).sum().backward()
    g = ctx.velocity_vjp(t, 0.0, p, w.numpy())
    assert np.array_equal(X.grad.numpy(), g)


def test_kelvin_ring_full_size_gradient_and_fd_mechanism(ctx):
    """The differentiation workload's ring (4096 points) over 5 steps at
    dt = 1e-7: the GPU adjoint against autograd through the torch
    restatement run on the GPU (the FP64 reference gradient).  Central
    differences in a random direction are biased here by the reference's
    collinearity cut-off (|R.R'/(l l') - 1| < 1e-8 zeroes a pair,
    src/tangle.c:505-506): ~1e5 pairs sit near it, a perturbation flips
    O(h) of them and each flip jumps the velocity, so the bias does not
    shrink with h.  With the cut-off set to 0 in the restatement, its
    central differences match its autograd gradient."""
    dev = torch.device("cuda:0")
    t = T.kelvin_ring()
    p = T.default_params(vortex_width=1.3e-8, alpha=0.098, alpha_p=0.016, hyperfriction=0)
    dt, nsteps = 1e-7, 5
    w = AC.seeded((t.N, 3), 15)
    d = torch.from_numpy(AC.seeded((t.N, 3), 16)).to(dev)
    g = ctx.rk4_run_vjp(t, 0.0, dt, nsteps, p, w, checkpoint_every=2)
    wt = torch.from_numpy(w).to(dev)
    for tol in (1e-8, 0.0):
        topo = TR.Topo(t, p)
        topo.ndot_tol = tol
        X = torch.from_numpy(t.vnodes.copy()).to(dev).requires_grad_(True)
        (TR.rk4_run(X, topo, dt, nsteps) * wt).sum().backward()
        ref = X.grad.cpu().numpy()
        if tol > 0:
            e = _err(g, ref)
            print("4096-point adjoint vs autograd", e, flush=True)
            assert e < AD_TOL
        else:
            X0 = X.detach()
            ad = float((X.grad * d).sum())
            with torch.no_grad():
                f = lambda Y: float((TR.rk4_run(Y, topo, dt, nsteps) * wt).sum())
                h = 1e-9
                fd = (f(X0 + h * d) - f(X0 - h * d)) / (2 * h)
            print("cut-off off: fd", fd, "ad", ad, flush=True)
            assert abs(fd - ad) <= FD_TOL * abs(ad)
        del X
        torch.cuda.empty_cache()


@pytest.mark.parametrize("nsteps,every", [(7, 2), (8, 2), (5, 0)])
def test_rk4_run_vjp_graph_replay_matches_eager(ctx, nsteps, every):
    """Full checkpoint segments after the first are replayed HIP graphs;
    with profiling every launch is eager.  Every replay must give the eager
    bits (several replays per call, eager work between calls)."""
    t, p = AC.case("periodic_holes")
    w = AC.seeded((t.N, 3), 17)
    eager = ctx.rk4_run_vjp(t, 0.0, 1e-6, nsteps, p, w, checkpoint_every=every,
                            enforce_boundaries=True, profile=True)
    for _ in range(2):
        g = ctx.rk4_run_vjp(t, 0.0, 1e-6, nsteps, p, w, checkpoint_every=every,
                            enforce_boundaries=True)
        assert np.array_equal(g, eager)
