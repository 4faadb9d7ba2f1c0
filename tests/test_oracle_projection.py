"""The Gauss-Newton projection of RRQ targets onto their curved patch, for targets
placed beyond the patch's vertices and edges (oracle, CPU).

The projection (foot point, side, Dirac-string direction) takes at most 40 steps.
It stops when the step |du| + |dv| falls below 1e-15 or below 8 eps |x| sqrt((a11
+ a22) / det), eight times the step that rounding of the residual alone makes.
(With 1e-15 alone, 19% of the benchmark workload's RRQ projections cycled at that
floor, 5e-15 to 1.3e-14 on its small panels, until the cap.)  Projections still
moving at the cap keep their last iterate and are counted, with their largest last
step.  Checked here:
  - no projection of the targets around many panels of orders 4, 8 and 12, on the
    clamped boundary or off it, reaches the cap;
  - the result's dependence on rounding near the boundary: the oracle built with
    fused multiply-adds differs from the plain build, in the action on a smooth
    density, by less than 500 eps / (distance to the boundary point, in units of
    the patch size) and elsewhere by less than a fixed bound per order.  With the
    edge root started from a real guess near a vertex (the cancelling form of its
    imaginary part) the two builds differ there by 0.1-0.2;
  - RRQ on targets beyond the vertices and edges (the quarter points and
    midpoints of every edge, at 0.1 and 0.3 of the patch size beyond it or 0.1 off
    the surface) applied to smooth densities agrees with brute-force quadrature of
    the same patch to the accuracy RRQ has elsewhere at that order, with the
    solid angle to 1e-12.
"""
import numpy as np
import pytest

from qotential_amd import workload as W
from panel_points import action_bound, boundary_targets


@pytest.mark.parametrize("p", [4, 8, 12])
def test_no_projection_reaches_the_cap(oracle, p):
    geo = W.geometry(8, 20, p, modes=W.VMEC_BENCH)
    h = geo["h"]
    oracle.set_side_tol(1e-10)
    oracle.projection_unconverged(True)
    nrrq = 0
    try:
        for k in range(0, geo["ntri"], 73):
            sl = slice(k * h, (k + 1) * h)
            tx, _ = boundary_targets(oracle, p, geo["sx"][sl], np.sqrt(geo["sw"][sl].sum()))
            r = oracle.rrq(p, tx, geo["sx"][sl], geo["snx"][sl], geo["sw"][sl], geo["rts"][sl], geo["rps"][sl],
                           exterior=True)
            nrrq += int((r["label"] == 5).sum())
    finally:
        oracle.set_side_tol(1e300)
    n, step = oracle.projection_unconverged(True)
    print(f"p={p}: {nrrq} RRQ pairs beyond vertices and edges, {n} capped, largest last step {step:.2e}")
    assert nrrq > 1000 and n == 0


# smooth-density error relative to the row's absolute sum: measured 1.4e-3,
# 1.9e-5, 2.3e-7 at p = 4, 6, 8 (targets 0.1 beyond an edge are the worst);
# bounds 2-4x those
EDGE_BF_TOL = {4: 3e-3, 6: 5e-5, 8: 1e-6}


@pytest.mark.parametrize("p", [4, 6, 8])
def test_rrq_beyond_vertices_and_edges_vs_brute_force(oracle, p):
    geo = W.geometry(16, 40, p, modes=W.VMEC_BENCH)
    h = geo["h"]
    k = 123
    sl = slice(k * h, (k + 1) * h)
    sx, snx, sw, rts, rps = (geo[key][sl] for key in ("sx", "snx", "sw", "rts", "rps"))
    tx, ab = boundary_targets(oracle, p, sx, np.sqrt(sw.sum()))
    oracle.set_side_tol(1e-10)
    try:
        r = oracle.rrq(p, tx, sx, snx, sw, rts, rps, exterior=True)
    finally:
        oracle.set_side_tol(1e300)
    # the RRQ pairs brute force resolves: 0.1 off the surface or 0.1 beyond the boundary
    sel = np.where((r["label"] == 5) & ((np.abs(ab[:, 1]) >= 0.1) | (ab[:, 0] >= 0.1)))[0]
    assert len(sel) > 100
    xs = tx[sel]
    m = len(xs)
    As = np.zeros((m, h), order="F"); Ad = np.zeros((m, h), order="F"); om = np.zeros(m)
    oracle.lib().orc_brute(oracle.i64(m), oracle._p(oracle._f(xs)), oracle.i64(p), oracle._p(oracle._f(sx)),
                           oracle._p(oracle._f(snx)), oracle.i64(64), oracle.i64(20), oracle._p(As), oracle._p(Ad),
                           oracle._p(om))
    assert np.abs(r["Ias"][sel] - om).max() < 1e-12
    worst = 0.0
    for f in (np.ones(h), sx[:, 0], np.exp(sx[:, 0] / 2) * np.cos(sx[:, 1] / 2), sx[:, 2] ** 2):
        for i, j in enumerate(sel):
            es = abs(r["As"][j] @ f - As[i] @ f) / (np.abs(As[i]).sum() * np.abs(f).max())
            ed = abs(r["Ad"][j] @ f - Ad[i] @ f) / (np.abs(Ad[i]).sum() * np.abs(f).max())
            worst = max(worst, es, ed)
    print(f"p={p}: {m} targets beyond vertices and edges, smooth densities vs brute force {worst:.2e}")
    assert worst < EDGE_BF_TOL[p]


@pytest.mark.parametrize("p", [4, 8])
def test_rounding_spread_near_the_boundary(oracle, p):
    geo = W.geometry(8, 20, p, modes=W.VMEC_BENCH)
    h = geo["h"]
    pos, q = W.point_charges(geo, seed=3)
    ub, gu = W.charge_field(geo["sx"], pos, q)
    ubn = (geo["snx"] * gu).sum(1)
    oracle.set_side_tol(1e-10)
    worst, nrrq = 0.0, 0
    try:
        for k in range(0, geo["ntri"], 73):
            sl = slice(k * h, (k + 1) * h)
            args = [geo[key][sl] for key in ("sx", "snx", "sw", "rts", "rps")]
            tx, ab = boundary_targets(oracle, p, geo["sx"][sl], np.sqrt(geo["sw"][sl].sum()))
            r = oracle.rrq(p, tx, *args, exterior=True)
            rf = oracle.rrq(p, tx, *args, exterior=True, fma=True)
            assert np.array_equal(r["label"], rf["label"])
            rrq = r["label"] == 5
            nrrq += int(rrq.sum())
            act = np.abs((rf["As"] - r["As"]) @ ubn[sl] - (rf["Ad"] - r["Ad"]) @ ub[sl])
            act /= np.abs(r["As"]) @ np.abs(ubn[sl]) + np.abs(r["Ad"]) @ np.abs(ub[sl])
            ratio = act[rrq] / action_bound(p, ab[rrq])
            worst = max(worst, ratio.max())
    finally:
        oracle.set_side_tol(1e300)
    print(f"p={p}: {nrrq} RRQ pairs, largest action spread / bound {worst:.2e}")
    assert nrrq > 1000 and worst < 1.0
