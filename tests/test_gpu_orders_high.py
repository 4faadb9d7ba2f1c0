"""Patch orders 9-12 on the GPU against the oracle.  These orders run kernels and
tile shapes that orders up to 8 do not: the matrix-core level geometry over 12-20
blocks of 4 panel nodes, direct kernels with 4 and 5 column tiles, and the
register-resident inverses spread over 3, 4, 6 and 8 waves with G finished by a
separate epilogue kernel (rows beyond 64 in the pivot mask).  The quaternion system's
condition number grows to about 2e10 at p = 12, so the per-weight bound is
looser, and so is the bound on the action on a smooth density."""
import numpy as np
import pytest

from qotential_amd import workload as W
from qotential_amd.engine import Engine

pytestmark = pytest.mark.gpu

# measured per-weight RRQ differences 9.3e-9, 7.5e-9, 2.3e-7, 1.4e-6 and density-action
# differences 2.7e-11, 2.2e-11, 5.6e-10, 3.9e-9 at p = 9..12; bounds about 10x those
RRQ_TOL = {9: 1e-7, 10: 1e-7, 11: 3e-6, 12: 2e-5}
ACT_TOL = {9: 1e-9, 10: 1e-9, 11: 1e-8, 12: 4e-8}


@pytest.mark.parametrize("p", [9, 10, 11, 12])
def test_high_order_vs_oracle(p, oracle):
    print(f"order {p}: VMEC surface 8x20, off-surface targets", flush=True)
    geo = W.geometry(8, 20, p, modes=W.VMEC_BENCH)
    tx, tri, dist = W.targets(geo, 400, -6, -1, sides=0, seed=20 + p)
    pos, q = W.point_charges(geo, seed=3)
    ub, gu = W.charge_field(geo["sx"], pos, q)
    ubn = (geo["snx"] * gu).sum(1)
    e = Engine(geo)
    e.set_targets(tx)
    e.nearlist()
    e.eval(exterior=True, side_tol=1e-10)
    w = e.weights()
    offs, idx = e.pairs()
    lg, bl = e.level_data()
    e.close()
    h = geo["h"]
    # level nodes and balls bit for bit (the matrix-core chain is the oracle's fma order)
    for k in (0, 123, 399):
        sl = slice(k * h, (k + 1) * h)
        L = oracle.patch_levels(p, geo["sx"][sl], geo["rts"][sl], geo["rps"][sl])
        assert np.array_equal(lg[k, 0:3].T, L["sx"]), (p, k)
        assert np.array_equal(lg[k, 6], L["sw"]), (p, k)
        assert np.array_equal(bl[k, :, 0:3], L["qpoint"]) and np.array_equal(bl[k, :, 3], L["qradii2"]), (p, k)
    h = geo["h"]
    oracle.set_side_tol(1e-10)
    lab_all = w["label"]
    has_rrq = [k for k in range(len(offs) - 1) if (lab_all[offs[k]:offs[k + 1]] == 5).any()]
    assert has_rrq
    rng = np.random.default_rng(p)
    worst = {"lev": 0.0, "rrq": 0.0, "act": 0.0}
    for k in rng.choice(has_rrq, min(3, len(has_rrq)), replace=False):
        sl = slice(k * h, (k + 1) * h)
        t = tx[idx[offs[k]:offs[k + 1]] - 1]
        r = oracle.rrq(p, t, geo["sx"][sl], geo["snx"][sl], geo["sw"][sl], geo["rts"][sl], geo["rps"][sl])
        lab = lab_all[offs[k]:offs[k + 1]]
        assert np.array_equal(lab, r["label"]), (p, k)
        for key in ("As", "Ad"):
            g = w[key][offs[k]:offs[k + 1]]
            err = np.abs(g - r[key]) / np.abs(r[key]).max(axis=1, keepdims=True)
            worst["lev"] = max(worst["lev"], err[lab <= 4].max(initial=0.0))
            worst["rrq"] = max(worst["rrq"], err[lab == 5].max(initial=0.0))
        vg = w["As"][offs[k]:offs[k + 1]] @ ubn[sl] - w["Ad"][offs[k]:offs[k + 1]] @ ub[sl]
        vo = r["As"] @ ubn[sl] - r["Ad"] @ ub[sl]
        scale = np.abs(r["As"]) @ np.abs(ubn[sl]) + np.abs(r["Ad"]) @ np.abs(ub[sl])
        worst["act"] = max(worst["act"], float(np.max(np.abs(vg - vo) / scale)))
    print(f"  order {p}: level weights {worst['lev']:.2e}, RRQ weights {worst['rrq']:.2e}, "
          f"density action {worst['act']:.2e}", flush=True)
    assert worst["lev"] < 1e-13
    assert worst["rrq"] < RRQ_TOL[p]
    assert worst["act"] < ACT_TOL[p]
