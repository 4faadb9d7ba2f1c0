"""GPU engine against the oracle (runs the HIP library on the device)."""
import sys

import numpy as np
import pytest

from qotential_amd import workload as W
from qotential_amd.engine import Engine

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng4(geo4):
    say("engine p=4 surface 12x36 (fixture geometry)")
    e = Engine(geo4)
    e.set_targets(geo4["sx"])
    return e


def test_nearlist_bit_exact(eng4, geo4, oracle):
    say("nearlist: surface nodes as targets")
    nnz = eng4.nearlist()
    assert nnz == 137259  # pair count of the reference's order-4 fixture run
    offs, idx = eng4.pairs()
    o_offs, o_idx = oracle.nearlist(geo4["h"], geo4["sx"], geo4["sw"], geo4["sx"])
    assert np.array_equal(offs, o_offs) and np.array_equal(idx, o_idx)
    # the uniform-grid search gives the same CSR
    assert eng4.nearlist(grid=True) == 137259
    g_offs, g_idx = eng4.pairs()
    assert np.array_equal(g_offs, o_offs) and np.array_equal(g_idx, o_idx)


def test_level_geometry_bit_exact(eng4, geo4, oracle):
    say("level geometry and balls")
    lg, bl = eng4.level_data()
    h = geo4["h"]
    for k in (0, 17, 555, 1109):
        sl = slice(k * h, (k + 1) * h)
        L = oracle.patch_levels(4, geo4["sx"][sl], geo4["rts"][sl], geo4["rps"][sl])
        assert np.array_equal(lg[k, 0:3].T, L["sx"]), k
        assert np.array_equal(lg[k, 6], L["sw"]), k
        assert np.array_equal(bl[k, :, 0:3], L["qpoint"]) and np.array_equal(bl[k, :, 3], L["qradii2"]), k
        assert np.max(np.abs(lg[k, 3:6].T - L["snx"])) < 1e-15


def test_weights_and_labels_vs_oracle(eng4, geo4, oracle):
    say("eval p=4, exterior limit for on-surface targets")
    eng4.eval(exterior=True, side_tol=1e300)
    w = eng4.weights()
    offs, idx = eng4.pairs()
    h = geo4["h"]
    oracle.set_side_tol(1e300)
    for k in (3, 100, 555, 900):
        sl = slice(k * h, (k + 1) * h)
        tx = geo4["sx"][idx[offs[k]:offs[k + 1]] - 1]
        r = oracle.rrq(4, tx, geo4["sx"][sl], geo4["snx"][sl], geo4["sw"][sl], geo4["rts"][sl], geo4["rps"][sl],
                       exterior=True)
        lab = w["label"][offs[k]:offs[k + 1]]
        assert np.array_equal(lab, r["label"]), k
        for key in ("As", "Ad"):
            g = w[key][offs[k]:offs[k + 1]]
            scale = np.abs(r[key]).max(axis=1, keepdims=True)
            err = np.abs(g - r[key]) / scale
            lv = lab <= 4
            assert err[lv].max() < 1e-13, (k, key, err[lv].max())
            if (~lv).any():
                assert err[~lv].max() < 1e-10, (k, key, err[~lv].max())


def test_grf_fixture_order4(eng4, geo4, fixture4, oracle):
    say("GRF order 4 fixture: far field + close correction on the GPU")
    ub, gu = W.harmonic_exp(geo4["sx"])
    ubn = (geo4["snx"] * gu).sum(1)
    eng4.eval(exterior=True, side_tol=1e300)
    far = eng4.far(ub, ubn)
    far_o = oracle.far_direct(geo4["sx"], geo4["sx"], geo4["snx"], geo4["sw"], ub, ubn, nthreads=8)
    assert np.max(np.abs(far - far_o)) < 1e-12 * np.abs(far_o).max()
    u, val = eng4.apply(ub, ubn, far)
    err = np.abs(u).max() / np.abs(ubn).max()
    say("GRF max rel err", err, "reference", float(fixture4["grf"]))
    assert err < 2e-3
    offs, idx = eng4.pairs()
    oracle.set_side_tol(1e300)
    uo = far_o.copy()
    val_o, lab_o = oracle.close_apply(4, geo4, geo4["sx"], offs, idx, ub, ubn, exterior=True, u=uo, nthreads=8)
    assert np.max(np.abs(u - uo)) < 1e-11 * np.abs(ubn).max()


def test_fused_corrections_match_apply(eng4, geo4):
    """With the density resident before eval, eval makes the per-pair corrections
    itself (level pairs right after the direct kernel, RRQ pairs after the
    weights); they equal a plain apply bit for bit, and so does the potential."""
    say("fused corrections vs plain apply")
    ub, gu = W.harmonic_exp(geo4["sx"])
    ubn = (geo4["snx"] * gu).sum(1)
    far = eng4.far(ub, ubn)              # density now resident, u = far field
    eng4.eval(exterior=True, side_tol=1e300)
    eng4.apply_resident()                # corrections came from eval
    val_f, u_f = eng4.values()
    u_ref, val_ref = eng4.apply(ub, ubn, far)
    assert np.array_equal(val_f, val_ref)
    assert np.array_equal(u_f, u_ref)
