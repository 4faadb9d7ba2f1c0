"""Every supported patch order on the GPU against the oracle: each order has its
own lane grid in the register-resident inverses and its own edge-rule size."""
import numpy as np
import pytest

from qotential_amd import workload as W
from qotential_amd.engine import Engine

pytestmark = pytest.mark.gpu

# per-weight tolerance of RRQ pairs: bounded by 1e-16 times the condition number
# of the order's quaternion system (about 2.5e2 at p=4, 1.9e4 at p=6, 1.7e6 at p=8)
RRQ_TOL = {2: 1e-11, 3: 1e-11, 4: 1e-11, 5: 1e-10, 6: 1e-9, 7: 2e-8, 8: 2e-7}


@pytest.mark.parametrize("p", [2, 3, 4, 5, 6, 7, 8])
def test_order_vs_oracle(p, oracle):
    print(f"order {p}: VMEC surface 8x20, off-surface targets", flush=True)
    geo = W.geometry(8, 20, p, modes=W.VMEC_BENCH)
    tx, tri, dist = W.targets(geo, 600, -6, -1, sides=0, seed=20 + p)
    pos, q = W.point_charges(geo, seed=3)
    ub, gu = W.charge_field(geo["sx"], pos, q)
    ubn = (geo["snx"] * gu).sum(1)
    e = Engine(geo)
    e.set_targets(tx)
    e.nearlist()
    e.eval(exterior=True, side_tol=1e-10)
    w = e.weights()
    offs, idx = e.pairs()
    h = geo["h"]
    oracle.set_side_tol(1e-10)
    lab_all = w["label"]
    has_rrq = [k for k in range(len(offs) - 1) if (lab_all[offs[k]:offs[k + 1]] == 5).any()]
    assert has_rrq
    rng = np.random.default_rng(p)
    for k in rng.choice(has_rrq, min(4, len(has_rrq)), replace=False):
        sl = slice(k * h, (k + 1) * h)
        t = tx[idx[offs[k]:offs[k + 1]] - 1]
        r = oracle.rrq(p, t, geo["sx"][sl], geo["snx"][sl], geo["sw"][sl], geo["rts"][sl], geo["rps"][sl])
        lab = lab_all[offs[k]:offs[k + 1]]
        assert np.array_equal(lab, r["label"]), (p, k)
        for key in ("As", "Ad"):
            g = w[key][offs[k]:offs[k + 1]]
            err = np.abs(g - r[key]) / np.abs(r[key]).max(axis=1, keepdims=True)
            assert err[lab <= 4].max(initial=0.0) < 1e-13, (p, k, key)
            assert err[lab == 5].max(initial=0.0) < RRQ_TOL[p], (p, k, key, err[lab == 5].max())
        # the action on a smooth density agrees far below the per-weight bound
        vg = w["As"][offs[k]:offs[k + 1]] @ ubn[sl] - w["Ad"][offs[k]:offs[k + 1]] @ ub[sl]
        vo = r["As"] @ ubn[sl] - r["Ad"] @ ub[sl]
        scale = np.abs(r["As"]) @ np.abs(ubn[sl]) + np.abs(r["Ad"]) @ np.abs(ub[sl])
        assert np.max(np.abs(vg - vo) / scale) < 1e-10, (p, k)
