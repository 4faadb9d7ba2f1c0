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


def test_config1_interior_targets_vs_oracle(oracle):
    """The CPU reference configuration on the GPU: VMEC surface 16x40 charts,
    order 4, 4,000 interior targets at h 10^U, U ~ U[-6,-1], point-charge
    density.  The potential equals the oracle's (near list, labels, close
    corrections, far field) to 1e-11 of max |u|, and the Green's identity holds
    to the order-4 level (u = the charges' field inside)."""
    say("config 1: 16x40 p=4, 4000 interior targets")
    geo = W.geometry(16, 40, 4, modes=W.VMEC_BENCH)
    tx, tri, dist = W.targets(geo, 4000, -6.0, -1.0, sides=-1, seed=1)
    assert (dist < 0).all()
    pos, q = W.point_charges(geo, nchg=8, offset=0.4, seed=2)
    ub, gu = W.charge_field(geo["sx"], pos, q)
    ubn = (geo["snx"] * gu).sum(1)
    e = Engine(geo)
    e.set_targets(tx)
    nnz = e.nearlist()
    offs, idx = e.pairs()
    o_offs, o_idx = oracle.nearlist(geo["h"], geo["sx"], geo["sw"], tx, nthreads=8)
    assert np.array_equal(offs, o_offs) and np.array_equal(idx, o_idx)
    far = e.far(ub, ubn)
    e.eval(exterior=True, side_tol=1e-10)
    u, _ = e.apply(ub, ubn, far)
    lab = e.weights()["label"]
    e.close()
    uo = oracle.far_direct(tx, geo["sx"], geo["snx"], geo["sw"], ub, ubn, nthreads=8)
    oracle.set_side_tol(1e-10)
    try:
        oracle.close_apply(4, geo, tx, o_offs, o_idx, ub, ubn, exterior=True, u=uo, nthreads=8)
    finally:
        oracle.set_side_tol(1e300)
    uex, _ = W.charge_field(tx, pos, q)
    diff = np.abs(u - uo).max() / np.abs(uo).max()
    grf = np.abs(u - uex).max() / np.abs(uex).max()
    say(f"  {nnz} pairs, labels L1..L4, RRQ {np.bincount(lab, minlength=6)[1:].tolist()}; "
        f"GPU vs oracle {diff:.2e}; GRF {grf:.3e}")
    assert nnz == 40357
    assert diff < 1e-11
    assert grf < 5e-2
