"""The reference's triangle levels (PR_SQUARE = .false.) and adaptive walk
(PR_ADAPTIVE) on the GPU against the oracle in the same level form: level
nodes and balls bit for bit, labels bit for bit (1 level 1 whole, 2 a tree cut
of levels 2-3, 4 level 4, 5 RRQ in the adaptive form), weights to the level
and RRQ tolerances of the default form; the drop-in qol_rrq_mex_ under the
process-wide level mode; and the Green's identity on a VMEC surface in every
form."""
import numpy as np
import pytest

from qotential_amd import rrq as R
from qotential_amd import workload as W
from qotential_amd.engine import Engine

pytestmark = pytest.mark.gpu

RRQ_TOL = {4: 1e-11, 8: 2e-7}
FORMS = {"triangle": (False, False), "adaptive": (True, False)}


@pytest.fixture
def form(oracle, request):
    adaptive, square = FORMS[request.param]
    oracle.set_level_mode(adaptive, square)
    yield request.param, adaptive, square
    oracle.set_level_mode(False, True)
    R.set_level_mode(False, True)


@pytest.mark.parametrize("form", ["triangle", "adaptive"], indirect=True)
@pytest.mark.parametrize("p", [4, 8])
def test_level_forms_vs_oracle(oracle, form, p):
    name, adaptive, square = form
    print(f"{name} levels p={p}: VMEC surface 8x20, off-surface targets", flush=True)
    geo = W.geometry(8, 20, p, modes=W.VMEC_BENCH)
    h = geo["h"]
    tx, tri, dist = W.targets(geo, 3000, -7, -1, sides=0, seed=11)
    e = Engine(geo)
    e.set_level_mode(adaptive, square)
    e.set_targets(tx)
    nnz = e.nearlist()
    offs, idx = e.pairs()
    e.eval(exterior=True, side_tol=1e-10)
    w = e.weights()
    lab = w["label"]
    o_lab = oracle.labels(p, geo, tx, offs, idx)
    assert np.array_equal(lab, o_lab)
    counts = np.bincount(lab, minlength=6)[1:].tolist()
    print(f"  {nnz} pairs, labels 1..5 {counts}", flush=True)
    if adaptive:
        assert counts[1] > 0 and counts[2] == 0 and counts[4] > 0
    # level nodes and balls of sampled panels bit for bit
    lg, bl = e.level_data()
    rng = np.random.default_rng(3)
    sample = np.sort(rng.choice(geo["ntri"], 6, replace=False))
    for k in sample:
        sl = slice(k * h, (k + 1) * h)
        L = oracle.patch_levels(p, geo["sx"][sl], geo["rts"][sl], geo["rps"][sl])
        assert np.array_equal(lg[k, 0:3].T, L["sx"]) and np.array_equal(lg[k, 6], L["sw"]), k
        assert np.array_equal(bl[k, :, 0:3], L["qpoint"]) and np.array_equal(bl[k, :, 3], L["qradii2"]), k
    # weights of every pair of a few panels
    oracle.set_side_tol(1e-10)
    try:
        worst = {"lev": 0.0, "rrq": 0.0}
        busy = np.argsort(np.diff(offs))[-4:]
        for k in busy:
            a, b = offs[k], offs[k + 1]
            sl = slice(k * h, (k + 1) * h)
            r = oracle.rrq(p, tx[idx[a:b] - 1], geo["sx"][sl], geo["snx"][sl], geo["sw"][sl], geo["rts"][sl],
                           geo["rps"][sl], exterior=True)
            assert np.array_equal(r["label"], lab[a:b])
            lv = r["label"] <= 4
            for key in ("As", "Ad"):
                err = np.abs(w[key][a:b] - r[key]) / np.abs(r[key]).max(axis=1, keepdims=True)
                if lv.any():
                    worst["lev"] = max(worst["lev"], err[lv].max())
                if (~lv).any():
                    worst["rrq"] = max(worst["rrq"], err[~lv].max())
    finally:
        oracle.set_side_tol(1e300)
    print(f"  weights vs oracle: levels {worst['lev']:.2e}, RRQ {worst['rrq']:.2e}", flush=True)
    assert worst["lev"] < 1e-13 and worst["rrq"] < RRQ_TOL[p]
    e.close()


def test_dropin_under_each_level_mode(oracle):
    """qol_rrq_mex_ and qol_rrq_noalloc_ follow qol_set_level_mode (the Fortran
    module passes PR_ADAPTIVE / PR_SQUARE the same way); both set is refused"""
    p = 6
    geo = W.geometry(8, 20, p, modes=W.VMEC_BENCH)
    h = geo["h"]
    k = 41
    sl = slice(k * h, (k + 1) * h)
    rng = np.random.default_rng(2)
    nrm = geo["snx"][sl].mean(0)
    size = np.sqrt(geo["sw"][sl].sum())
    tx = geo["sx"][sl][rng.integers(0, h, 200)] + (size * 10 ** rng.uniform(-6, 0.3, 200))[:, None] * nrm
    pre = R.simplex_precomp(p + 2, p - 1)
    try:
        oracle.set_side_tol(1e-10)
        for name, (adaptive, square) in [("square", (False, True))] + list(FORMS.items()):
            R.set_level_mode(adaptive, square)
            oracle.set_level_mode(adaptive, square)
            r = oracle.rrq(p, tx, geo["sx"][sl], geo["snx"][sl], geo["sw"][sl], geo["rts"][sl], geo["rps"][sl],
                           exterior=True)
            for noalloc in (False, True):
                out = R.rrq(tx, geo["sx"][sl], geo["snx"][sl], geo["sw"][sl], geo["rts"][sl], geo["rps"][sl],
                            order=p, exterior=True, pre=pre, noalloc=noalloc)
                err = max(np.abs(out["As"] - r["As"]).max() / np.abs(r["As"]).max(),
                          np.abs(out["Ad"] - r["Ad"]).max() / np.abs(r["Ad"]).max())
                print(f"{'qol_rrq_noalloc_' if noalloc else 'qol_rrq_mex_'} {name}: labels "
                      f"{np.bincount(r['label'], minlength=6)[1:].tolist()}, max rel diff {err:.2e}", flush=True)
                assert err < 1e-9
        R.set_level_mode(True, True)
        with pytest.raises(Exception):
            R.rrq(tx, geo["sx"][sl], geo["snx"][sl], geo["sw"][sl], geo["rts"][sl], geo["rps"][sl], order=p,
                  exterior=True, pre=pre)
    finally:
        R.set_level_mode(False, True)
        oracle.set_level_mode(False, True)
        oracle.set_side_tol(1e300)


def test_grf_every_level_form():
    """the Green's identity on a VMEC surface (order 8, targets on both sides)
    in each level form: the three are close to each other"""
    geo = W.geometry(16, 40, 8, modes=W.VMEC_BENCH)
    tx, tri, dist = W.targets(geo, 20000, -6, -1, sides=0, seed=1)
    pos, q = W.point_charges(geo, nchg=8, offset=0.4, seed=2)
    ub, gu = W.charge_field(geo["sx"], pos, q)
    ubn = (geo["snx"] * gu).sum(1)
    uex, _ = W.charge_field(tx, pos, q)
    ref = np.where(dist < 0, uex, 0.0)
    errs = {}
    for name, (adaptive, square) in [("square", (False, True))] + list(FORMS.items()):
        e = Engine(geo)
        e.set_level_mode(adaptive, square)
        e.set_targets(tx)
        e.nearlist()
        far = e.far(ub, ubn)
        e.eval(exterior=True, side_tol=1e-10)
        u, _ = e.apply(ub, ubn, far)
        errs[name] = np.abs(u - ref).max() / np.abs(uex).max()
        print(f"{name}: GRF max rel err {errs[name]:.3e}, labels "
              f"{np.bincount(e.weights()['label'], minlength=6)[1:].tolist()}", flush=True)
        e.close()
    for name, v in errs.items():
        assert v < 10 * errs["square"] + 1e-12, name
