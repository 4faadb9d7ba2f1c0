"""Drop-in entry points (qol_*_mex_) on the GPU against the oracle."""
import sys

import numpy as np
import pytest

from qotential_amd import rrq as R
from qotential_amd import workload as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _panel(geo, k):
    h = geo["h"]
    sl = slice(k * h, (k + 1) * h)
    return [geo[key][sl] for key in ("sx", "snx", "sw", "rts", "rps")]


# RRQ weights are the solution of the quaternion system M^T w = rhs, whose
# condition number grows with the patch order (about 2e2, 2e4, 2e6 at p = 4, 6, 8
# on these charts); two correct FP64 evaluations agree per weight to about
# cond * 1e-16 * 10, and on smooth densities to near machine precision.
RRQ_WEIGHT_TOL = {4: 1e-11, 6: 1e-9, 8: 2e-7}


def _targets(sx, snx, sw, n=40, seed=3):
    rng = np.random.default_rng(seed)
    c = sx.mean(0); nn = snx.mean(0); nn /= np.linalg.norm(nn)
    hd = np.sqrt(sw.sum())
    i = rng.integers(0, sx.shape[0], n)
    d = hd * 10.0 ** rng.uniform(-6, 0.3, n) * rng.choice([-1, 1], n)
    return sx[i] + 0.3 * (sx[i] - c) + d[:, None] * nn[None, :]


@pytest.mark.parametrize("p", [6, 8])
def test_rrq_mex_vs_oracle(p, oracle):
    say(f"qol_rrq_mex_ p={p}")
    geo = W.geometry(16, 40, p, modes=W.VMEC_BENCH)
    sx, snx, sw, rts, rps = _panel(geo, 123)
    tx = _targets(sx, snx, sw)
    g = R.rrq(tx, sx, snx, sw, rts, rps, p, exterior=True)
    oracle.set_side_tol(1e-10)
    o = oracle.rrq(p, tx, sx, snx, sw, rts, rps, exterior=True)
    lab = o["label"]
    assert (lab == 5).sum() > 5 and (lab <= 4).sum() > 5
    for key in ("As", "Ad"):
        scale = np.abs(o[key]).max(axis=1, keepdims=True)
        err = np.abs(g[key] - o[key]) / scale
        say(key, "level max rel", err[lab <= 4].max(), "rrq max rel", err[lab == 5].max())
        assert err[lab <= 4].max() < 1e-13
        assert err[lab == 5].max() < RRQ_WEIGHT_TOL[p]
    # action on a smooth density (point-charge field on the panel nodes)
    pos = sx.mean(0) + 2.0 * np.sqrt(sw.sum()) * np.array([[1.0, 0.5, 0.2], [-0.3, 1.0, 0.7]])
    f, gf = W.charge_field(sx, pos, np.array([1.0, -0.7]))
    fn = (snx * gf).sum(1)
    vg = g["As"] @ fn - g["Ad"] @ f
    vo = o["As"] @ fn - o["Ad"] @ f
    scale = np.abs(o["As"]) @ np.abs(fn) + np.abs(o["Ad"]) @ np.abs(f)
    say("smooth density max rel", np.max(np.abs(vg - vo) / scale))
    assert np.max(np.abs(vg - vo) / scale) < 1e-10
    assert np.max(np.abs(g["Ias"] - o["Ias"])) < 1e-11
    assert np.max(np.abs(g["Omega"] - o["Omega"])) < 1e-10 * np.abs(o["Omega"]).max()


def test_patch_levels_handles(oracle):
    say("qol_*_patch_levels_mex_")
    p = 8
    geo = W.geometry(16, 40, p, modes=W.VMEC_BENCH)
    sx, snx, sw, rts, rps = _panel(geo, 77)
    tx = _targets(sx, snx, sw, n=60, seed=5)
    lv, ier = R.build_patch_levels(sx, rts, rps, p - 1)
    assert ier == 0 and lv > 0
    lv2, ier2 = R.build_patch_levels(sx, rts, rps, p - 1)
    assert ier2 == 1  # one slot
    As, Ad, idxs, ier = R.lap3dsdlpmat_levels(tx, p * (p + 1) // 2, lv)
    assert ier == 0
    oA, oD, oidx, olab = oracle.levels_eval(p, tx, sx, rts, rps)
    assert np.array_equal(idxs, oidx)
    keep = olab <= 4
    assert np.max(np.abs(As[keep] - oA[keep])) <= 1e-13 * np.abs(oA[keep]).max()
    assert R.lap3dsdlpmat_levels(tx, 10, lv)[3] == 3
    assert R.clear_patch_levels(lv) == 0
    assert R.clear_patch_levels(lv) == 2
    assert R.lap3dsdlpmat_levels(tx, p * (p + 1) // 2, lv)[3] == 2
    assert R.clear_patch_levels_all() == 0


def test_getnearquad_triplets_vs_oracle(oracle):
    say("qol_getnearquad_lap_rrq_mex_")
    p = 4
    geo = W.geometry(12, 36, p)
    h = geo["h"]
    offs, idx = oracle.nearlist(h, geo["sx"], geo["sw"], geo["sx"])
    npan = geo["ntri"]
    sel = np.arange(0, npan, 37)  # a subset of panels keeps the oracle quick
    tcx_l, row_l, tcxi = [], [], [1]
    for k in range(npan):
        t = idx[offs[k]:offs[k + 1]] - 1 if k in set(sel) else np.zeros(0, dtype=np.int64)
        tcx_l.append(geo["sx"][t]); row_l.append(t + 1); tcxi.append(tcxi[-1] + len(t))
    tcx = np.concatenate(tcx_l); tcxrow = np.concatenate(row_l); tcxi = np.array(tcxi)
    g = R.getnearquad_lap_rrq(geo, tcx, tcxrow, tcxi, iside=30)
    oracle.set_side_tol(1e-10)
    o = oracle.getnearquad(p, geo, tcx, tcxrow, tcxi, iside=30)
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])
    for a, b in ((g[2], o[2]), (g[3], o[3])):
        assert np.max(np.abs(a - b)) < 1e-10 * np.abs(b).max()


def test_rrq_mex_from_concurrent_threads():
    """The drop-in is callable from several host threads at once (each call has
    its own engine and stream; per-order tables are built once under a lock):
    every thread's result equals the serial one bit for bit."""
    import threading
    say("qol_rrq_mex_ from 4 host threads at once")
    p = 6
    geo = W.geometry(16, 40, p, modes=W.VMEC_BENCH)
    jobs = []
    for j, k in enumerate((11, 222, 403, 777)):
        sx, snx, sw, rts, rps = _panel(geo, k)
        jobs.append((_targets(sx, snx, sw, n=30, seed=40 + j), sx, snx, sw, rts, rps))
    serial = [R.rrq(*a, p) for a in jobs]
    out = [None] * len(jobs)
    errs = []

    def run(i):
        try:
            for _ in range(3):
                out[i] = R.rrq(*jobs[i], p)
        except Exception as e:  # reported by the main thread
            errs.append(repr(e))

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(jobs))]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errs, errs
    assert all(not t.is_alive() for t in th)
    for a, b in zip(out, serial):
        for key in ("As", "Ad", "Omega", "Ias"):
            assert np.array_equal(a[key], b[key]), key
