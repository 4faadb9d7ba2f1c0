"""The full benchmark workload (VMEC surface 48x120 charts, order 8, 14,250
triangles, 200,000 targets on both sides, 1.94M pairs) through the engine.

Against the oracle at full size: the near-list CSR and every pair's label bit
for bit, the order-8 level nodes and balls of a sample of panels bit for bit,
and that sample's weights.  Size-independent properties: the Green's
representation identity on every target, bit-identical repeated corrections,
and the one-stream schedule giving the three-stream result bit for bit."""
import os

import numpy as np
import pytest

from qotential_amd import workload as W
from qotential_amd.engine import Engine

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def bench_case():
    geo = W.geometry(48, 120, 8, modes=W.VMEC_BENCH)
    tx, tri, dist = W.targets(geo, 200000, -6.0, -1.0, sides=0, seed=1)
    pos, q = W.point_charges(geo, nchg=8, offset=0.4, seed=2)
    ub, gu = W.charge_field(geo["sx"], pos, q)
    ubn = (geo["snx"] * gu).sum(1)
    return geo, tx, dist, pos, q, ub, ubn


NTHREADS = max(1, min(16, os.cpu_count() or 1))


def test_bench_workload_lists_labels_levels_vs_oracle(bench_case, oracle):
    geo, tx, dist, pos, q, ub, ubn = bench_case
    h = geo["h"]
    e = Engine(geo)
    e.set_targets(tx)
    nnz = e.nearlist()
    offs, idx = e.pairs()
    say(f"near list on the GPU: {nnz} pairs; oracle with {NTHREADS} threads")
    o_offs, o_idx = oracle.nearlist(h, geo["sx"], geo["sw"], tx, nthreads=NTHREADS)
    assert np.array_equal(offs, o_offs) and np.array_equal(idx, o_idx)
    # the uniform-grid search: the same CSR
    import time
    e.sync()
    t0 = time.perf_counter()
    assert e.nearlist(grid=True) == nnz
    t_grid = time.perf_counter() - t0
    t0 = time.perf_counter()
    e.nearlist()
    t_brute = time.perf_counter() - t0
    say(f"near list: brute force {t_brute * 1e3:.1f} ms, uniform grid {t_grid * 1e3:.1f} ms (incl. host sync)")
    e.nearlist(grid=True)
    g_offs, g_idx = e.pairs()
    assert np.array_equal(g_offs, o_offs) and np.array_equal(g_idx, o_idx)
    e.eval(exterior=True, side_tol=1e-10)
    lab = e.weights()["label"]
    o_lab = oracle.labels(8, geo, tx, offs, idx, nthreads=NTHREADS)
    say("labels L1..L4, RRQ", np.bincount(lab, minlength=6)[1:].tolist())
    assert np.array_equal(lab, o_lab)
    # the Gauss-Newton projections stop before the 40-step cap (none of the oracle's
    # on 6,741 RRQ pairs of 300 sampled panels reaches it)
    ncap, step = e.projection_unconverged()
    say(f"capped projections: {ncap} of {(lab == 5).sum()} RRQ pairs, largest last step {step:.2e}")
    assert ncap <= 1e-4 * (lab == 5).sum() and step < 1e-12
    # order-8 level nodes and balls (the chained FP64 MFMA of the level kernel)
    lg, bl = e.level_data()
    rng = np.random.default_rng(5)
    sample = np.sort(rng.choice(geo["ntri"], 12, replace=False))
    for k in sample:
        sl = slice(k * h, (k + 1) * h)
        L = oracle.patch_levels(8, geo["sx"][sl], geo["rts"][sl], geo["rps"][sl])
        assert np.array_equal(lg[k, 0:3].T, L["sx"]), k
        assert np.array_equal(lg[k, 6], L["sw"]), k
        assert np.array_equal(bl[k, :, 0:3], L["qpoint"]) and np.array_equal(bl[k, :, 3], L["qradii2"]), k
        assert np.max(np.abs(lg[k, 3:6].T - L["snx"])) < 1e-15
    del lg, bl
    # the sample's weights: level pairs to 1e-13 of the row maximum; RRQ pairs
    # to the order-8 conditioning bound, and their action on the density to 1e-10
    w = e.weights()
    oracle.set_side_tol(1e-10)
    try:
        worst = [0.0, 0.0, 0.0]
        for k in sample[:6]:
            sl = slice(k * h, (k + 1) * h)
            a, b = offs[k], offs[k + 1]
            if b == a:
                continue
            r = oracle.rrq(8, tx[idx[a:b] - 1], geo["sx"][sl], geo["snx"][sl], geo["sw"][sl], geo["rts"][sl],
                           geo["rps"][sl], exterior=True)
            assert np.array_equal(w["label"][a:b], r["label"])
            lv = r["label"] <= 4
            for key, dens in (("As", ubn[sl]), ("Ad", ub[sl])):
                g = w[key][a:b]
                scale = np.abs(r[key]).max(axis=1, keepdims=True)
                err = np.abs(g - r[key]) / scale
                if lv.any():
                    worst[0] = max(worst[0], err[lv].max())
                if (~lv).any():
                    worst[1] = max(worst[1], err[~lv].max())
                    act = np.abs(g[~lv] @ dens - r[key][~lv] @ dens) / (np.abs(r[key][~lv]) @ np.abs(dens))
                    worst[2] = max(worst[2], act.max())
        say(f"sampled weights vs oracle: levels {worst[0]:.2e}, RRQ {worst[1]:.2e}, RRQ action {worst[2]:.2e}")
        assert worst[0] < 1e-13 and worst[1] < 2e-7 and worst[2] < 1e-10
    finally:
        oracle.set_side_tol(1e300)
    e.close()


def _run(geo, tx, ub, ubn):
    e = Engine(geo)
    e.set_targets(tx)
    nnz = e.nearlist()
    far = e.far(ub, ubn)
    e.eval(exterior=True, side_tol=1e-10)
    u, val = e.apply(ub, ubn, far)
    lab = np.bincount(e.weights()["label"], minlength=6)[1:]
    # a second step on the resident data repeats the per-pair corrections bit
    # for bit (apply_resident adds them into the resident potential once more)
    e.eval(exterior=True, side_tol=1e-10)
    e.apply_resident()
    val2, _ = e.values()
    e.close()
    return nnz, lab, u, val, val2


def test_bench_workload_identity_and_determinism(bench_case, monkeypatch):
    geo, tx, dist, pos, q, ub, ubn = bench_case
    say(f"full workload: {geo['ntri']} triangles, {tx.shape[0]} targets, three streams")
    nnz, lab, u, val, val2 = _run(geo, tx, ub, ubn)
    say("pairs", nnz, "labels L1..L4, RRQ", lab.tolist())
    assert nnz == 1939854 and lab.sum() == nnz and (lab > 0).all()
    # Green's identity on every target: the charges sit 0.4 along the outward
    # normal, so S[du/dn] - D[u] is u at targets inside the surface and 0 outside
    uex, _ = W.charge_field(tx, pos, q)
    err = np.abs(u - np.where(dist < 0.0, uex, 0.0)).max() / np.abs(uex).max()
    say(f"GRF max rel err {err:.3e}")
    assert err < 1e-7
    assert np.array_equal(val2, val)
    say("one stream")
    monkeypatch.setenv("QOL_ONE_STREAM", "1")
    nnz1, lab1, u1, val1, _ = _run(geo, tx, ub, ubn)
    monkeypatch.delenv("QOL_ONE_STREAM")
    assert nnz1 == nnz and np.array_equal(lab1, lab)
    assert np.array_equal(val1, val) and np.array_equal(u1, u)


def test_bench_workload_grf_decomposition(bench_case):
    """What sets the Green's-identity error of the bench workload (5.4e-8): the
    pairs the ball rule serves at level 1, farthest from the surface.  Serving
    every level pair one level finer halves the maximum; adding every target
    within 3 node radii of a panel to that panel's near list (so the smooth rule
    covers less) changes nothing."""
    from scipy.spatial import cKDTree
    geo, tx, dist, pos, q, ub, ubn = bench_case
    h, npan = geo["h"], geo["ntri"]
    uex, _ = W.charge_field(tx, pos, q)
    ref = np.where(dist < 0.0, uex, 0.0)
    scale = np.abs(uex).max()
    e = Engine(geo)
    e.set_targets(tx)
    e.nearlist()
    offs, idx = e.pairs()
    far = e.far(ub, ubn)

    def grf():
        e.eval(exterior=True, side_tol=1e-10)
        u, _ = e.apply(ub, ubn, far)
        return np.abs(u - ref) / scale

    err = grf()
    k = int(np.argmax(err))
    e.refine_levels(1)
    finer = grf().max()
    e.refine_levels(0)
    sx = geo["sx"].reshape(npan, h, 3)
    c = sx.mean(1)
    r = np.sqrt(((sx - c[:, None, :]) ** 2).sum(2).max(1))
    extra = cKDTree(tx).query_ball_point(c, 3.0 * r)
    lists = [np.union1d(idx[offs[j]:offs[j + 1]] - 1, np.asarray(extra[j], dtype=np.int64)) for j in range(npan)]
    w_offs = np.concatenate([[0], np.cumsum([len(x) for x in lists])]).astype(np.int64)
    e.set_pairs(w_offs, np.concatenate(lists) + 1)
    wide = grf().max()
    e.close()
    say(f"GRF max {err.max():.3e} (worst target at |d| = {abs(dist[k]):.1e}); one level finer {finer:.3e}; "
        f"near list to 3 node radii ({int(w_offs[-1])} pairs) {wide:.3e}")
    assert err.max() < 1e-7
    assert finer < 0.6 * err.max()
    assert abs(wide - err.max()) < 0.01 * err.max()

