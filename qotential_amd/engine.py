"""Python mirror of the library's engine API (batched close evaluation on one GPU).

An Engine holds one triangulated surface of patch order p in HBM; targets, the
near list (reference ball rule, or caller-given CSR), the per-pair weights from
the refinement levels or RRQ, and the near-correction apply all stay on the
device.  There is no CPU path: a missing library or device raises QolError.
"""
import ctypes as C

import numpy as np

from ._lib import QolError, check, dbl, i64, lib, ptr

vp = C.c_void_p


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def order_info(p):
    info = np.zeros(13, dtype=np.int64)
    check(lib().qol_order_info(i64(p), ptr(info)), "qol_order_info")
    return dict(h=int(info[0]), nq=int(info[1]), ncoef=int(info[2]), ntot=int(info[3]), nsubtot=int(info[4]),
                nl=info[5:9].copy(), nsub=info[9:13].copy())


def order_level(p, lev):
    """one refinement level's reference table: uv, wt, interp (n, h), suboff, subtyp"""
    info = order_info(p)
    n, ns, h = int(info["nl"][lev - 1]), int(info["nsub"][lev - 1]), info["h"]
    uv = np.zeros((n, 2)); wt = np.zeros(n); im = np.zeros((n, h), order="F")
    so = np.zeros(ns + 1, dtype=np.int64); st = np.zeros(ns, dtype=np.int64)
    check(lib().qol_order_level(i64(p), i64(lev), ptr(uv), ptr(wt), ptr(im), ptr(so), ptr(st)), "qol_order_level")
    return dict(uv=uv, wt=wt, interpmat=im, suboff=so, subtyp=st)


class Engine:
    """Close evaluation of S and D for a surface of npan triangles of order p.

    geo: dict with sx, snx, sw, rts, rps (nsrc x 3 / nsrc) as built by
    qotential_amd.workload.geometry; chart=(sxbd, rv) selects the exact chart
    boundary (ichart = 1).
    """

    def __init__(self, geo, chart=None):
        self.p = int(geo["p"])
        self.h = self.p * (self.p + 1) // 2
        self.npan = int(geo["sw"].shape[0] // self.h)
        self._keep = [_f(geo[k]) for k in ("sx", "snx", "sw", "rts", "rps")]
        sxbd = rv = None
        ichart = 0
        if chart is not None:
            sxbd, rv = _f(chart[0]), _f(chart[1])
            ichart = 1
            self._keep += [sxbd, rv]
        h = vp()
        check(lib().qol_engine_create(i64(self.p), i64(self.npan), *[ptr(a) for a in self._keep[:5]], i64(ichart),
                                      ptr(sxbd), ptr(rv), i64(0), C.byref(h)), "qol_engine_create")
        self._h = h
        self.nt = 0
        self.nnz = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().qol_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_targets(self, tx):
        tx = _f(tx)
        self.nt = tx.shape[0]
        check(lib().qol_engine_set_targets(self._h, i64(self.nt), ptr(tx), i64(0)), "qol_engine_set_targets")

    def set_near_radius(self, diam_factor=0.0):
        """radius of the next near list: 0 the reference's ball rule, f > 0 a
        ball of f times the panel's largest corner distance (f = 2: the 2h rule)"""
        check(lib().qol_engine_set_near_radius(self._h, dbl(diam_factor)), "qol_engine_set_near_radius")

    def nearlist(self, grid=False):
        """near list by the reference's ball rule, or the radius set_near_radius
        chose (brute force, or through a uniform grid of the targets: the same
        CSR)"""
        n = i64(0)
        fn = lib().qol_engine_nearlist_grid if grid else lib().qol_engine_nearlist
        check(fn(self._h, C.byref(n)), "qol_engine_nearlist")
        self.nnz = n.value
        return self.nnz

    def set_pairs(self, offs, idx1):
        offs = np.ascontiguousarray(offs, dtype=np.int64)
        idx1 = np.ascontiguousarray(idx1, dtype=np.int64)
        check(lib().qol_engine_set_pairs(self._h, ptr(offs), ptr(idx1)), "qol_engine_set_pairs")
        self.nnz = int(offs[-1])

    def pairs(self):
        offs = np.zeros(self.npan + 1, dtype=np.int64)
        idx = np.zeros(max(self.nnz, 1), dtype=np.int64)
        check(lib().qol_engine_get_pairs(self._h, ptr(offs), ptr(idx)), "qol_engine_get_pairs")
        return offs, idx[: self.nnz]

    def eval(self, exterior=True, side_tol=1e-10, want_omega=False):
        check(lib().qol_engine_eval(self._h, i64(1 if exterior else 0), dbl(side_tol), i64(1 if want_omega else 0)),
              "qol_engine_eval")

    def set_level_mode(self, adaptive=False, square=True):
        """the reference's PR_ADAPTIVE / PR_SQUARE for this engine: square levels
        (default), triangle levels (square=False), or the adaptive walk over
        triangle levels (adaptive=True, square=False); labels then read 1 level 1
        whole, 2 a tree cut of levels 2-3, 4 level 4, 5 RRQ"""
        check(lib().qol_engine_set_level_mode(self._h, i64(1 if adaptive else 0), i64(1 if square else 0)),
              "qol_engine_set_level_mode")

    def set_precision(self, split_levels):
        """levels 1..split_levels on the split-precision (FP32 matrix core) direct kernel; 0 = all FP64"""
        check(lib().qol_engine_set_precision(self._h, i64(split_levels)), "qol_engine_set_precision")

    def set_selector(self, selector, threshold=0.9):
        """label pairs with a learned selector (qotential_amd.selector.Selector; None: the ball rule)"""
        self._sel = selector
        check(lib().qol_engine_set_selector(self._h, selector.handle if selector is not None else None,
                                            dbl(threshold)), "qol_engine_set_selector")

    def force_label(self, label):
        """every pair served one way: 0 own nodes, 1..4 level, 5 RRQ; -1 the ball rule"""
        check(lib().qol_engine_force_label(self._h, i64(label)), "qol_engine_force_label")

    def refine_levels(self, shift):
        """the ball rule with level pairs served `shift` levels finer (at most level 4)"""
        check(lib().qol_engine_refine_levels(self._h, i64(shift)), "qol_engine_refine_levels")

    def pair_features(self):
        """the selector's 10 features of every pair, (nnz, 10) float32"""
        f = np.zeros((max(self.nnz, 1), 10), dtype=np.float32)
        check(lib().qol_engine_pair_features(self._h, ptr(f)), "qol_engine_pair_features")
        return f[: self.nnz]

    def sync(self):
        check(lib().qol_engine_sync(self._h), "qol_engine_sync")

    def stage_ms(self):
        ms = np.zeros(10)
        check(lib().qol_engine_stage_ms(self._h, ptr(ms)), "qol_engine_stage_ms")
        return dict(levels=ms[0], classify=ms[1], direct=ms[2], rrq_setup=ms[3], rrq_solve=ms[4], rrq_project=ms[5],
                    rrq_root_ssq=ms[6], rrq_moments=ms[7], rrq_solid_angle=ms[8], rrq_weights=ms[9])

    def projection_unconverged(self):
        """(count, largest last step): RRQ pairs of the last eval whose
        Gauss-Newton projection had not stopped by its 40th and last step (they
        keep that last iterate), and the largest such step |du| + |dv|"""
        n, st = C.c_int64(), C.c_double()
        check(lib().qol_engine_projection_unconverged(self._h, C.byref(n), C.byref(st)),
              "qol_engine_projection_unconverged")
        return n.value, st.value

    def labels(self):
        """the label of every pair (1-4 the serving level, 5 RRQ, 0 the panel's own
        nodes), without copying the weights"""
        lab = np.zeros(max(self.nnz, 1), dtype=np.int64)
        check(lib().qol_engine_get_weights(self._h, None, None, ptr(lab), None, None), "qol_engine_get_weights")
        return lab[:self.nnz]

    def weights(self, omega=False):
        n = max(self.nnz, 1)
        As = np.zeros((n, self.h), order="F"); Ad = np.zeros((n, self.h), order="F")
        lab = np.zeros(n, dtype=np.int64)
        Om = np.zeros((4 * self.h, n), order="F") if omega else None
        Ia = np.zeros(n) if omega else None
        check(lib().qol_engine_get_weights(self._h, ptr(As), ptr(Ad), ptr(lab), ptr(Om), ptr(Ia)),
              "qol_engine_get_weights")
        k = self.nnz
        out = dict(As=As[:k], Ad=Ad[:k], label=lab[:k])
        if omega:
            out["Omega"] = Om[:, :k]
            out["Ias"] = Ia[:k]
        return out

    def apply(self, ub, ubn, u0):
        u = _f(u0).copy()
        val = np.zeros(max(self.nnz, 1))
        check(lib().qol_engine_apply(self._h, ptr(_f(ub)), ptr(_f(ubn)), ptr(u), ptr(val)), "qol_engine_apply")
        return u, val[: self.nnz]

    def apply_resident(self):
        check(lib().qol_engine_apply_resident(self._h), "qol_engine_apply_resident")

    def values(self):
        """Per-pair corrections and the resident potential, copied back."""
        val = np.zeros(max(self.nnz, 1)); u = np.zeros(max(self.nt, 1))
        check(lib().qol_engine_get_values(self._h, ptr(val), ptr(u)), "qol_engine_get_values")
        return val[:self.nnz], u[:self.nt]

    def set_potential(self, u):
        """the resident potential the per-target sums add into"""
        check(lib().qol_engine_set_potential(self._h, ptr(_f(u))), "qol_engine_set_potential")

    def far(self, ub, ubn):
        u = np.zeros(max(self.nt, 1))
        check(lib().qol_engine_far(self._h, ptr(_f(ub)), ptr(_f(ubn)), ptr(u)), "qol_engine_far")
        return u[: self.nt]

    def level_info(self):
        """this engine's level sizes (they follow its level form)"""
        info = np.zeros(12, dtype=np.int64)
        check(lib().qol_engine_level_info(self._h, ptr(info)), "qol_engine_level_info")
        return dict(ntot=int(info[0]), nsubtot=int(info[1]), nl=info[2:6].copy(), nsub=info[6:10].copy(),
                    square=bool(info[10]), adaptive=bool(info[11]))

    def level_data(self):
        info = self.level_info()
        lg = np.zeros((self.npan, 7, info["ntot"]))
        bl = np.zeros((self.npan, info["nsubtot"], 4))
        check(lib().qol_engine_level_data(self._h, ptr(lg), ptr(bl)), "qol_engine_level_data")
        return lg, bl

    def triplets(self, trow):
        trow = np.ascontiguousarray(trow, dtype=np.int64)
        n = max(self.nnz * self.h, 1)
        si = np.zeros(n, dtype=np.int64); sj = np.zeros(n, dtype=np.int64)
        vs = np.zeros(n); vd = np.zeros(n)
        check(lib().qol_engine_triplets(self._h, ptr(trow), ptr(si), ptr(sj), ptr(vs), ptr(vd)), "qol_engine_triplets")
        k = self.nnz * self.h
        return si[:k], sj[:k], vs[:k], vd[:k]
