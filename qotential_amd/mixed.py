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
import ctypes as C

import numpy as np

from . import workload as W
from ._lib import check, dbl, i64, lib
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


class MixedBatch:
    """One step of every order of a mixed-order surface from one call.

    Builds one engine per order (targets, near list, far field, resident
    density, one eager evaluation) and a batch over them (qol_batch_create).
    step() launches eval + corrections + the per-target sums of all orders
    (qol_batch_launch); autotune() picks each order's schedule and solve
    kernels and the orders' launch sequence by timing (qol_batch_autotune).
    capture=True would record the step into one HIP graph (qol_batch_capture);
    the library refuses that for now (its first run ended the process with a
    host-side segmentation fault whose cause is not yet found).
    """

    def __init__(self, surf, tx, density, exterior=True, side_tol=1e-10, capture=False):
        self.orders = sorted(surf["parts"])
        self.engines, self.far, self.dens = {}, {}, {}
        for p in self.orders:
            g = surf["parts"][p]
            ub, gu = density(g["sx"])
            self.dens[p] = (ub, (g["snx"] * gu).sum(1))
            e = Engine(g)
            e.set_targets(tx)
            e.nearlist()
            self.far[p] = e.far(*self.dens[p])
            e.eval(exterior=exterior, side_tol=side_tol)
            e.apply(*self.dens[p], self.far[p])
            self.engines[p] = e
        arr = (C.c_void_p * len(self.orders))(*[self.engines[p]._h for p in self.orders])
        h = C.c_void_p()
        check(lib().qol_batch_create(i64(len(self.orders)), arr, C.byref(h)), "qol_batch_create")
        self._h = h
        if capture:
            check(lib().qol_batch_capture(h, i64(1 if exterior else 0), dbl(side_tol)), "qol_batch_capture")

    def autotune(self, reps=3):
        n = len(self.orders)
        ms = np.zeros(6 * n + 2)
        ch = np.zeros(2 * n + 1, dtype=np.int64)
        check(lib().qol_batch_autotune(self._h, i64(reps), ms.ctypes.data_as(C.c_void_p),
                                       ch.ctypes.data_as(C.c_void_p)), "qol_batch_autotune")
        table = {p: dict(ms=ms[6 * i:6 * i + 6].reshape(2, 3).tolist(), one_stream=int(ch[2 * i]),
                         solve_variant=int(ch[2 * i + 1])) for i, p in enumerate(self.orders)}
        table["graph_order_ms"] = dict(ascending=float(ms[6 * n]), descending=float(ms[6 * n + 1]))
        table["descending"] = bool(ch[2 * n])
        return table

    def step(self):
        check(lib().qol_batch_launch(self._h), "qol_batch_launch")

    def sync(self):
        check(lib().qol_batch_sync(self._h), "qol_batch_sync")

    def values(self):
        """each order's per-pair corrections of the last step, {p: val}"""
        self.sync()
        return {p: self.engines[p].values()[0] for p in self.orders}

    def potential(self):
        """S[du/dn] - D[u] at the targets from one batch step: each order's
        resident potential is reset to its far field, the step adds the
        corrections, and the orders are summed in ascending order"""
        self.sync()
        for p in self.orders:
            self.engines[p].set_potential(self.far[p])
        self.step()
        self.sync()
        u = np.zeros(self.engines[self.orders[0]].nt)
        for p in self.orders:
            u += self.engines[p].values()[1]
        return u

    def pairs(self):
        return sum(e.nnz for e in self.engines.values())

    def close(self):
        if getattr(self, "_h", None):
            lib().qol_batch_destroy(self._h)
            self._h = None
        for e in self.engines.values():
            e.close()
        self.engines = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
