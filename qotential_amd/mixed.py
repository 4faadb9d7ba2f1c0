"""Mixed-order surfaces: every chart of a stellarator carries its own patch order
(for example 4, 6, 8, 10 and 12 drawn uniformly per chart), and both triangles of
a chart share it.

The close-evaluation path is per panel (rrq_r64, src/lap3d_close_mod.f90:622-881,
is called once per triangle by stellarator_grf_core_mod.f90:759-888 with that
triangle's own order), so the surface splits into one sub-surface per order. Each
goes to its own Engine, which keeps its panels, level tables and RRQ tables
resident; the near list is built per order over all the targets, so a target near
panels of several orders gets its pairs from each. The layer potential at a
target is the sum over the orders, in ascending order (a fixed order, so the result
is deterministic), of that order's far field plus its close corrections.
"""
import numpy as np

from . import workload as W
from .engine import Engine


def surface(mp, np_, orders, modes=None, nfp=5, seed=0, gpu=False, chart_order=None):
    """Chart list of the surface (aspect splits evaluated at chart_order, by
    default the highest order), an order per chart drawn uniformly from orders,
    and the triangulated sub-surface of each order.  Returns dict(parts={p: geo},
    chart_p, charts, mp, np, modes, nfp); geo["chart_index"] lists the charts of
    that part in the surface's chart order."""
    orders = sorted(set(int(p) for p in orders))
    pref = int(chart_order or max(orders))
    cl = W.geometry(mp, np_, pref, modes=modes, nfp=nfp)["charts"]
    chart_p = np.random.default_rng(seed).choice(orders, size=len(cl))
    parts = {}
    for p in orders:
        sel = np.nonzero(chart_p == p)[0]
        if sel.size == 0:
            continue
        g = W.geometry_of_charts(mp, np_, p, cl[sel], modes=modes, nfp=nfp, gpu=gpu)
        g["chart_index"] = sel
        parts[p] = g
    return dict(parts=parts, chart_p=chart_p, charts=cl, mp=mp, np=np_, modes=modes, nfp=nfp)


def targets(surf, nt, umin, umax, sides=0, seed=1):
    """near-surface targets (workload.targets) drawn from each order's triangles
    in proportion to its triangle count; returns tx, signed distance, order of
    the triangle each target was placed from."""
    parts = surf["parts"]
    ntri = np.array([parts[p]["ntri"] for p in parts], dtype=np.float64)
    counts = np.floor(nt * ntri / ntri.sum()).astype(np.int64)
    counts[-1] += nt - counts.sum()
    txs, ds, ps = [], [], []
    for i, (p, g) in enumerate(parts.items()):
        if counts[i] == 0:
            continue
        tx, _, dist = W.targets(g, int(counts[i]), umin, umax, sides=sides, seed=seed + 7919 * i)
        txs.append(tx); ds.append(dist); ps.append(np.full(counts[i], p))
    return np.vstack(txs), np.concatenate(ds), np.concatenate(ps)


def layer_potential(surf, tx, density, exterior=True, side_tol=1e-10):
    """S[du/dn] - D[u] at targets tx for the density u given by
    density(x) -> (u, grad u) at the surface nodes.  Every order's evaluation is
    launched before any is waited for (each engine works on its own streams, so
    the orders overlap on the device); far fields and corrections are then added
    in ascending order.  Returns u and per-order statistics {p: dict(pairs, labels)}."""
    orders = sorted(surf["parts"])
    engines, dens = {}, {}
    try:
        for p in orders:
            g = surf["parts"][p]
            ub, gu = density(g["sx"])
            dens[p] = (ub, (g["snx"] * gu).sum(1))
            e = Engine(g)
            engines[p] = e
            e.set_targets(tx)
            e.nearlist()
        for p in orders:
            engines[p].eval(exterior=exterior, side_tol=side_tol)
        u = np.zeros(tx.shape[0])
        stats = {}
        for p in orders:
            e = engines[p]
            ub, ubn = dens[p]
            far = e.far(ub, ubn)
            up, val = e.apply(ub, ubn, far)
            u += up
            lab = e.weights()["label"]
            stats[p] = dict(pairs=int(val.size), labels=np.bincount(lab, minlength=6)[1:].tolist())
    finally:
        for e in engines.values():
            e.close()
    return u, stats
