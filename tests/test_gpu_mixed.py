"""Mixed-order surfaces on the GPU: one engine per order, potentials summed in
ascending order.  A single-order surface reproduces the plain engine bit for bit;
a mixed surface of orders 4, 6 and 8 satisfies Green's representation identity
about as well as the order-4 surface (its least accurate order)."""
import numpy as np
import pytest

from qotential_amd import grf
from qotential_amd import mixed as M
from qotential_amd import workload as W

pytestmark = pytest.mark.gpu


def _grf_err(u, tx, dist, pos, q):
    uex, _ = W.charge_field(tx, pos, q)
    return float(np.abs(u - np.where(dist < 0, uex, 0.0)).max() / np.abs(uex).max())


def test_single_order_equals_engine():
    print("mixed driver, one order (6) on the 16x40 surface, against the plain engine", flush=True)
    s = M.surface(16, 40, (6,), modes=W.VMEC_BENCH)
    g = W.geometry(16, 40, 6, modes=W.VMEC_BENCH)
    assert np.array_equal(s["parts"][6]["sx"], g["sx"])
    pos, q = W.point_charges(g, seed=2)
    tx, _, dist = W.targets(g, 2000, -8, -1, sides=0, seed=4)
    u_m, _ = M.layer_potential(s, tx, lambda x: W.charge_field(x, pos, q))
    ub, gu = W.charge_field(g["sx"], pos, q)
    u_e, eng = grf.layer_potential(g, tx, ub, (g["snx"] * gu).sum(1))
    eng.close()
    assert np.array_equal(u_m, u_e)


def test_mixed_orders_grf():
    """The mixed surface against uniform order-4 and order-8 surfaces on the same
    charts, each on its own targets, by distance band: its error lies between
    theirs (charts of different orders meet at seams that match only to the
    coarser order's interpolation error; at this size that does not dominate).
    Measured: mixed 1.5e-2, 1.6e-2, 1.5e-3; order 4 4.0e-2, 4.2e-2, 1.7e-3;
    order 8 2.2e-6, 2.2e-6, 6.7e-7 at U in [-8,-6], [-5,-3], [-2,-1]."""
    print("mixed orders 4, 6, 8 against uniform 4 and 8 on the 16x40 surface, by distance band", flush=True)
    g8 = W.geometry(16, 40, 8, modes=W.VMEC_BENCH)
    pos, q = W.point_charges(g8, seed=2)
    dens = lambda x: W.charge_field(x, pos, q)
    res = {}
    for name, orders in (("mixed", (4, 6, 8)), ("p4", (4,)), ("p8", (8,))):
        s = M.surface(16, 40, orders, modes=W.VMEC_BENCH, seed=1, chart_order=8)
        for lo, hi in ((-8, -6), (-5, -3), (-2, -1)):
            tx, dist, _ = M.targets(s, 3000, lo, hi, sides=0, seed=9)
            u, stats = M.layer_potential(s, tx, dens)
            res[(name, lo)] = _grf_err(u, tx, dist, pos, q)
        print("  %-5s GRF rel err  U[-8,-6] %.2e  U[-5,-3] %.2e  U[-2,-1] %.2e" %
              (name, res[(name, -8)], res[(name, -5)], res[(name, -2)]), flush=True)
    for lo in (-8, -5, -2):
        assert res[("p8", lo)] < res[("mixed", lo)] <= 1.2 * res[("p4", lo)], lo
    assert res[("p8", -8)] < 1e-5


# density action of the busiest panels' RRQ weights against the oracle, by order:
# measured 4.6e-13, 2.7e-12, 2.5e-11, 5.3e-10, 4.1e-8 at p = 4..12 (hundreds of
# targets per panel down to h = 1e-8; the quaternion system's condition number,
# about 2e10 at p = 12, times the FP64 unit roundoff is 2e-6 per weight); bounds 10x
MIXED_ACT_TOL = {4: 5e-12, 6: 3e-11, 8: 3e-10, 10: 5e-9, 12: 4e-7}


def _check_sample_vs_oracle(oracle, p, g, tx, e, pos, q):
    """labels, level nodes and weights of the three busiest panels of one order"""
    h = g["h"]
    offs, idx = e.pairs()
    w = e.weights()
    lg, bl = e.level_data()
    ub, gu = W.charge_field(g["sx"], pos, q)
    ubn = (g["snx"] * gu).sum(1)
    oracle.set_side_tol(1e-10)
    worst = [0.0, 0.0]
    try:
        for k in np.argsort(np.diff(offs))[-3:]:
            a, b = offs[k], offs[k + 1]
            sl = slice(k * h, (k + 1) * h)
            L = oracle.patch_levels(p, g["sx"][sl], g["rts"][sl], g["rps"][sl])
            assert np.array_equal(lg[k, 0:3].T, L["sx"]) and np.array_equal(bl[k, :, 3], L["qradii2"]), (p, k)
            r = oracle.rrq(p, tx[idx[a:b] - 1], g["sx"][sl], g["snx"][sl], g["sw"][sl], g["rts"][sl], g["rps"][sl],
                           exterior=True)
            assert np.array_equal(w["label"][a:b], r["label"]), (p, k)
            lv = r["label"] <= 4
            for key, dens in (("As", ubn[sl]), ("Ad", ub[sl])):
                gk = w[key][a:b]
                if lv.any():
                    err = np.abs(gk[lv] - r[key][lv]) / np.abs(r[key][lv]).max(axis=1, keepdims=True)
                    worst[0] = max(worst[0], err.max())
                if (~lv).any():
                    act = np.abs(gk[~lv] @ dens - r[key][~lv] @ dens) / (np.abs(r[key][~lv]) @ np.abs(dens))
                    worst[1] = max(worst[1], act.max())
    finally:
        oracle.set_side_tol(1e300)
    print(f"    vs oracle on 3 panels: level weights {worst[0]:.2e}, RRQ action {worst[1]:.2e}", flush=True)
    assert worst[0] < 1e-13 and worst[1] < MIXED_ACT_TOL[p]


def test_mixed_full_size(oracle):
    """The mixed-order workload at full size: 64x160 charts of the bench surface,
    orders 4, 6, 8, 10, 12 drawn per chart, 1,000,000 targets at h 10^U,
    U in [-8,-1], both sides.  Every order's near list, levels and RRQ run on the
    device; the per-order evaluation is timed (inputs resident), each order's
    labels, level nodes and weights are checked against the oracle on its three
    busiest panels (a check that sees quadrature errors far below the seams'
    effect on the identity), and the GRF error of the whole surface is bounded by
    the uniform order-4 figure measured on the 16x40 surface."""
    import time
    from qotential_amd.engine import Engine
    print("mixed orders 4..12, 64x160 charts, 1M targets", flush=True)
    s = M.surface(64, 160, (4, 6, 8, 10, 12), modes=W.VMEC_BENCH, seed=3)
    pos, q = W.point_charges(s["parts"][12], nchg=8, offset=0.4, seed=2)
    tx, dist, _ = M.targets(s, 1_000_000, -8, -1, sides=0, seed=11)
    u = np.zeros(tx.shape[0])
    npairs, t_eval = 0, 0.0
    engines = {}
    for p in sorted(s["parts"]):
        g = s["parts"][p]
        e = Engine(g)
        engines[p] = e
        e.set_targets(tx)
        nnz = e.nearlist()
        e.eval(exterior=True, side_tol=1e-10)  # first call: tables and buffers
        e.sync()
        t0 = time.perf_counter()
        e.eval(exterior=True, side_tol=1e-10)
        e.sync()
        dt = time.perf_counter() - t0
        npairs += nnz
        t_eval += dt
        print(f"  order {p}: {g['ntri']} triangles, {nnz} pairs, eval {1e3 * dt:.1f} ms", flush=True)
        _check_sample_vs_oracle(oracle, p, g, tx, e, pos, q)
    # every order launched before any is waited for: the orders overlap
    t0 = time.perf_counter()
    for e in engines.values():
        e.eval(exterior=True, side_tol=1e-10)
    for e in engines.values():
        e.sync()
    t_conc = time.perf_counter() - t0
    print(f"  all orders launched together: {1e3 * t_conc:.1f} ms ({npairs / t_conc:.3e} pairs/s)", flush=True)
    for p, e in engines.items():
        g = s["parts"][p]
        ub, gu = W.charge_field(g["sx"], pos, q)
        ubn = (g["snx"] * gu).sum(1)
        far = e.far(ub, ubn)
        up, _ = e.apply(ub, ubn, far)
        e.close()
        u += up
    err = _grf_err(u, tx, dist, pos, q)
    print(f"  total {npairs} pairs, eval {1e3 * t_eval:.1f} ms ({npairs / t_eval:.3e} pairs/s), "
          f"GRF rel err {err:.2e}", flush=True)
    assert npairs > 5_000_000
    assert err < 4.5e-2
