"""Edge cases of the engine and the drop-in on the GPU against the oracle: empty
inputs, pair lists holding one label only (the other stages then run on empty
lists), and one panel holding many more pairs than the others (several tiles of
every per-panel kernel next to panels with a single pair)."""
import numpy as np
import pytest

from qotential_amd import rrq as R
from qotential_amd import workload as W
from qotential_amd.engine import Engine

pytestmark = pytest.mark.gpu

P = 6
RRQ_TOL = 1e-9  # per-weight bound at p = 6 (condition number about 2e4)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def geo():
    return W.geometry(8, 20, P, modes=W.VMEC_BENCH)


def _density(geo):
    pos, q = W.point_charges(geo, seed=3)
    ub, gu = W.charge_field(geo["sx"], pos, q)
    return ub, (geo["snx"] * gu).sum(1)


def _own_panel_pairs(npan, tri):
    """pair list: every target against the panel it was generated on (1-based targets)."""
    order = np.lexsort((np.arange(tri.size), tri))
    offs = np.zeros(npan + 1, dtype=np.int64)
    np.add.at(offs, tri + 1, 1)
    return np.cumsum(offs), order + 1


def _check_panels(e, geo, tx, oracle, panels):
    w = e.weights()
    offs, idx = e.pairs()
    h = geo["h"]
    oracle.set_side_tol(1e-10)
    for k in panels:
        sl = slice(k * h, (k + 1) * h)
        t = tx[idx[offs[k]:offs[k + 1]] - 1]
        r = oracle.rrq(P, t, geo["sx"][sl], geo["snx"][sl], geo["sw"][sl], geo["rts"][sl], geo["rps"][sl])
        lab = w["label"][offs[k]:offs[k + 1]]
        assert np.array_equal(lab, r["label"]), k
        for key in ("As", "Ad"):
            g = w[key][offs[k]:offs[k + 1]]
            err = np.abs(g - r[key]) / np.abs(r[key]).max(axis=1, keepdims=True)
            assert err[lab <= 4].max(initial=0.0) < 1e-13, (k, key)
            assert err[lab == 5].max(initial=0.0) < RRQ_TOL, (k, key, err[lab == 5].max())
    return w


def test_empty_inputs(geo):
    say("no targets; targets with no near panel; drop-in with m = 0")
    ub, ubn = _density(geo)
    e = Engine(geo)
    e.set_targets(np.zeros((0, 3)))
    assert e.nearlist() == 0
    e.eval()
    u, val = e.apply(ub, ubn, np.zeros(0))
    assert u.size == 0 and val.size == 0
    far = np.array([[40.0, 0.0, 0.0], [0.0, -55.0, 3.0], [1.0, 2.0, 70.0]])
    e.set_targets(far)
    assert e.nearlist() == 0
    e.eval()
    assert e.weights()["As"].shape == (0, geo["h"])
    u0 = np.array([1.0, -2.0, 3.5])
    u, val = e.apply(ub, ubn, u0)
    assert np.array_equal(u, u0) and val.size == 0
    e.close()
    h = geo["h"]
    g = R.rrq(np.zeros((0, 3)), geo["sx"][:h], geo["snx"][:h], geo["sw"][:h], geo["rts"][:h], geo["rps"][:h], P)
    assert g["As"].shape == (0, h) and g["Omega"].shape == (4 * h, 0)


def test_rrq_pairs_only(geo, oracle):
    say("every pair RRQ: targets 1e-8..1e-7 h off their own panel, both sides")
    tx, tri, _ = W.targets(geo, 300, -8, -7, sides=0, seed=5)
    offs, idx1 = _own_panel_pairs(geo["ntri"], tri)
    e = Engine(geo)
    e.set_targets(tx)
    e.set_pairs(offs, idx1)
    e.eval(exterior=True, side_tol=1e-10)
    lab = e.weights()["label"]
    assert (lab == 5).all()
    panels = np.unique(tri)
    _check_panels(e, geo, tx, oracle, np.random.default_rng(1).choice(panels, 4, replace=False))
    e.close()


def test_level_pairs_only(geo, oracle):
    say("every pair level-served: targets 2..4 h off their own panel")
    tx, tri, _ = W.targets(geo, 300, np.log10(2.0), np.log10(4.0), sides=0, seed=6)
    offs, idx1 = _own_panel_pairs(geo["ntri"], tri)
    e = Engine(geo)
    e.set_targets(tx)
    e.set_pairs(offs, idx1)
    e.eval(exterior=True, side_tol=1e-10)
    lab = e.weights()["label"]
    assert (lab <= 4).all()
    _check_panels(e, geo, tx, oracle, np.random.default_rng(2).choice(np.unique(tri), 4, replace=False))
    e.close()


def test_one_crowded_panel(geo, oracle):
    say("one panel with 120 near targets at every distance, three panels with one target each")
    h = geo["h"]
    k0 = 57
    sl = slice(k0 * h, (k0 + 1) * h)
    sx, snx, sw = geo["sx"][sl], geo["snx"][sl], geo["sw"][sl]
    rng = np.random.default_rng(7)
    hd = np.sqrt(sw.sum())
    c = sx.mean(0)
    i = rng.integers(0, h, 120)
    d = hd * 10.0 ** rng.uniform(-7, 0.5, 120) * rng.choice([-1, 1], 120)
    crowd = sx[i] + 0.2 * (c - sx[i]) + d[:, None] * snx[i]
    ones, tri1, _ = W.targets(geo, 3, -6, -5, sides=0, seed=8)
    tx = np.vstack([crowd, ones])
    tri = np.concatenate([np.full(120, k0), tri1])
    offs, idx1 = _own_panel_pairs(geo["ntri"], tri)
    e = Engine(geo)
    e.set_targets(tx)
    e.set_pairs(offs, idx1)
    e.eval(exterior=True, side_tol=1e-10)
    lab = e.weights()["label"][offs[k0]:offs[k0 + 1]]
    say("labels on the crowded panel", np.bincount(lab, minlength=6)[1:])
    assert (lab == 5).sum() > 24 and (lab <= 4).sum() > 16  # several tiles of both kinds
    _check_panels(e, geo, tx, oracle, [k0] + [int(k) for k in tri1])
    # every target has one pair here, so the fixed-order per-target sum of the
    # corrections is that pair's value, bit for bit
    ub, ubn = _density(geo)
    u, val = e.apply(ub, ubn, np.zeros(tx.shape[0]))
    _, idx = e.pairs()
    assert np.all(np.isfinite(val))
    assert np.array_equal(u[idx - 1], val)
    e.close()


def test_engine_reuse_new_targets_same_nnz(geo):
    """One engine reused with a second target set whose pair list has the same
    number of pairs (targets permuted, plus targets with no pairs, so nt grows):
    the per-target sums must be rebuilt, giving a fresh engine's potential."""
    say("engine reuse: new targets, equal pair count")
    ub, ubn = _density(geo)
    tx, tri, _ = W.targets(geo, 300, -5.0, -1.0, sides=0, seed=11)
    e = Engine(geo)
    e.set_targets(tx)
    e.nearlist()
    e.eval(exterior=True, side_tol=1e-10)
    u_a, _ = e.apply(ub, ubn, np.zeros(len(tx)))
    offs, idx = e.pairs()
    # second set: the same points in another order, then 200 points with no pairs
    perm = np.random.default_rng(4).permutation(len(tx))
    inv = np.argsort(perm)
    far = tx.mean(0) + np.array([50.0, 0.0, 0.0]) + np.random.default_rng(5).standard_normal((200, 3))
    tx_b = np.concatenate([tx[perm], far])
    idx_b = np.concatenate([np.sort(inv[idx[offs[k]:offs[k + 1]] - 1]) + 1 for k in range(len(offs) - 1)])
    e.set_targets(tx_b)
    e.set_pairs(offs, idx_b)
    e.eval(exterior=True, side_tol=1e-10)
    u_b, _ = e.apply(ub, ubn, np.zeros(len(tx_b)))
    e.close()
    f = Engine(geo)
    f.set_targets(tx_b)
    f.set_pairs(offs, idx_b)
    f.eval(exterior=True, side_tol=1e-10)
    u_f, _ = f.apply(ub, ubn, np.zeros(len(tx_b)))
    f.close()
    assert np.array_equal(u_b, u_f)
    assert np.abs(u_b[:len(tx)] - u_a[perm]).max() <= 1e-12 * np.abs(u_a).max()
    assert not u_b[len(tx):].any()
