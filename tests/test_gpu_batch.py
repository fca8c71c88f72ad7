"""Ensembles in one launch per stage (ov_gpu_rk4_run_batch) against the same
tangles stepped one at a time through the single-tangle path, and one
member against the oracle."""
import numpy as np
import pytest

from openvort_amd import tangle as T
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _members():
    ens = T.ensemble(n_tangles=7, seed=77, n_lo=150, n_hi=2500)
    # vary the boxes and add an LIA-only member and an open-boundary member
    ens[2][0].box_hi = (0.06, 0.05, 0.07)
    ens[3][1]["LIA_only"] = 1
    b = T.TangleBuilder()
    b.add_ring((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.01, 256)
    ens.append((b.build(), dict(alpha=0.098, alpha_p=0.016, vn=(0.0, 0.0, 0.0), temperature=1.6)))
    return ens


def _params(kw, prec):
    kw = dict(kw)
    kw.pop("temperature", None)
    return T.default_params(vortex_width=1.3e-8, hyperfriction=0, precision=prec, **kw)


@pytest.mark.parametrize("prec", [T.PREC_FP64, T.PREC_MIXED])
def test_batch_matches_single(ctx, prec):
    ens = _members()
    tangles = [t.copy() for t, _ in ens]
    params = [_params(kw, prec) for _, kw in ens]
    ctx.rk4_run_batch(tangles, 0.0, 5e-6, 3, params, enforce_boundaries=True)
    print("tuning (unit, blocks, ms)", ctx.batch_tuning(), flush=True)
    for (t, kw), b, p in zip(ens, tangles, params):
        a = t.copy()
        ctx.rk4_run(a, 0.0, 5e-6, 3, p, enforce_boundaries=True)
        act = t.active
        assert np.abs(a.vnodes[act] - b.vnodes[act]).max() < 1e-15
        assert rel_err(b.vels[act], a.vels[act])[0] < 1e-11
        assert rel_err(b.vs[act], a.vs[act])[0] < 1e-11


def test_batch_member_matches_oracle(ctx, oracle):
    ens = _members()
    tangles = [t.copy() for t, _ in ens]
    params = [_params(kw, T.PREC_FP64) for _, kw in ens]
    ctx.rk4_run_batch(tangles, 0.0, 5e-6, 1, params)
    for i in (0, 4):
        a = ens[i][0].copy()
        oracle.step(a, 0.0, 5e-6, params[i])
        act = a.active
        assert np.abs(a.vnodes[act] - tangles[i].vnodes[act]).max() < 1e-16
        assert rel_err(tangles[i].vels[act], a.vels[act])[0] < 1e-11


def test_batch_rejects_mixed_precision_members(ctx):
    ens = _members()[:2]
    params = [_params(ens[0][1], T.PREC_FP64), _params(ens[1][1], T.PREC_MIXED)]
    from openvort_amd.gpu import OVError
    with pytest.raises(OVError):
        ctx.rk4_run_batch([t.copy() for t, _ in ens], 0.0, 5e-6, 1, params)
