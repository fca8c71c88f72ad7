"""Pins the torch restatement used as the autograd reference of the adjoint:
its values against the oracle (itself pinned to the reference's C), its
gradients against central differences of the oracle.  CPU only."""
import numpy as np
import pytest
import torch

from tests import adjoint_cases as AC
from tests import torch_ref as TR
from tests.helpers import rel_err


def _oracle_vels(oracle, t, p, X_act=None):
    c = t.copy()
    if X_act is not None:
        c.vnodes[t.active] = X_act
    oracle.update_tangle(c, 0.0, p)
    return c.vels[t.active], c.vs[t.active]


@pytest.mark.parametrize("name", AC.NAMES)
def test_torch_ref_values_match_oracle(name, oracle):
    t, p = AC.case(name)
    topo = TR.Topo(t, p)
    X = torch.from_numpy(t.vnodes[t.active].copy())
    vs, vel = TR.velocities(X, topo)
    ov, ovs = _oracle_vels(oracle, t, p)
    e, es = rel_err(vel.numpy(), ov), rel_err(vs.numpy(), ovs)
    print(name, "vels", e, "vs", es)
    assert e[0] < 1e-11 and es[0] < 1e-11


@pytest.mark.parametrize("name", AC.NAMES)
def test_torch_ref_gradient_matches_oracle_fd(name, oracle):
    """d/dh w.V(X + h delta) by central differences of the oracle (h = 1e-9 cm)
    against autograd's (dV/dX)^T w . delta, three seeded directions."""
    t, p = AC.case(name)
    topo = TR.Topo(t, p)
    X0 = t.vnodes[t.active].copy()
    n = len(X0)
    w = AC.seeded((n, 3), 11)
    X = torch.from_numpy(X0.copy()).requires_grad_(True)
    (TR.velocities(X, topo)[1] * torch.from_numpy(w)).sum().backward()
    g = X.grad.numpy()
    h = 1e-9
    for s in range(3):
        d = AC.seeded((n, 3), 100 + s)
        vp, _ = _oracle_vels(oracle, t, p, X0 + h * d)
        vm, _ = _oracle_vels(oracle, t, p, X0 - h * d)
        fd = float(((vp - vm) * w).sum() / (2 * h))
        ad = float((g * d).sum())
        print(name, s, "fd", fd, "ad", ad, "rel", abs(fd - ad) / abs(ad))
        assert abs(fd - ad) <= 1e-6 * abs(ad)


def test_torch_ref_rk4_matches_oracle(oracle):
    """Two reference-variant RK4 steps of the restatement against the
    oracle's update_tangle + rk4_step."""
    t, p = AC.case("ring_mf_flows")
    topo = TR.Topo(t, p)
    X = TR.rk4_run(torch.from_numpy(t.vnodes[t.active].copy()), topo, 2e-6, 2)
    c = t.copy()
    for i in range(2):
        oracle.step(c, i * 2e-6, 2e-6, p)
    assert np.abs(X.numpy() - c.vnodes[t.active]).max() < 1e-16
