"""The allocation-free entry point ov_gpu_velocities_noalloc against
ov_gpu_velocities (update_tangle semantics) on the same tangles.

Geometry comes from the same kernel on the same positions, so dxi, s' and
s'' agree bit for bit; the direct sum is split into at most 32 partial-sum
planes instead of up to 256, so velocities agree to FP64 summation order
(1e-12 relative).  Two calls in a row, the second on a smaller tangle: the
shared scratch keeps the first call's contents, and nothing of them may leak.
"""
import numpy as np
import pytest

from openvort_amd import tangle as T
from openvort_amd.gpu import OV_E_INVALID, OV_E_UNSUPPORTED, OVError
from tests.helpers import rel_err, tangle_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _noalloc(ctx, t, p, time):
    n = t.N
    assert np.all(t.status[:, 0] == T.FREE)
    dev = "cuda"
    pos = torch.tensor(np.ascontiguousarray(t.vnodes.T), device=dev)
    fwd = torch.tensor(t.connections[:, 0].astype(np.int32), device=dev)
    rev = torch.tensor(t.connections[:, 1].astype(np.int32), device=dev)
    out = {k: torch.full((3, n), np.nan, dtype=torch.float64, device=dev)
           for k in ("tangents", "normals", "vs", "vels")}
    dxi = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx.velocities_noalloc(pos.data_ptr(), fwd.data_ptr(), rev.data_ptr(), n, n, t, time, p,
                           out["tangents"].data_ptr(), out["normals"].data_ptr(), dxi.data_ptr(),
                           out["vs"].data_ptr(), out["vels"].data_ptr())
    res = {k: v.cpu().numpy().T.copy() for k, v in out.items()}
    res["dxi"] = dxi.cpu().numpy()
    return res


@pytest.mark.parametrize("prec", [T.PREC_FP64, T.PREC_MIXED])
def test_noalloc_matches_velocities(ctx, prec):
    p = tangle_params(precision=prec, near_radius=0.01)
    for seed, rings in ((1001, 6), (77, 3)):          # second call smaller
        t = T.random_rings(seed, 0.1, n_rings=rings)
        ref = t.copy()
        ctx.velocities(ref, 0.0, p, geometry=True)
        got = _noalloc(ctx, t, p, 0.0)
        print(prec, seed, t.N, "points", flush=True)
        np.testing.assert_array_equal(got["dxi"], ref.dxi)
        np.testing.assert_array_equal(got["tangents"], ref.tangents)
        np.testing.assert_array_equal(got["normals"], ref.normals)
        e_vs, e_v = rel_err(got["vs"], ref.vs), rel_err(got["vels"], ref.vels)
        print("vs", e_vs, "vels", e_v, flush=True)
        assert e_vs[0] < 1e-12 and e_v[0] < 1e-12


def test_noalloc_rejects_beyond_limits(ctx):
    t = T.random_rings(5, 0.1, n_rings=2)
    p = tangle_params()
    with pytest.raises(OVError) as e:
        ctx.velocities_noalloc(1, 1, 1, 20000, 20000, t, 0.0, p, 1, 1, 1, 1, 1)
    assert e.value.code == OV_E_INVALID
    with pytest.raises(OVError) as e:
        ctx.velocities_noalloc(1, 1, 1, 100, 100, t, 0.0, tangle_params(use_BH=1), 1, 1, 1, 1, 1)
    assert e.value.code == OV_E_UNSUPPORTED
