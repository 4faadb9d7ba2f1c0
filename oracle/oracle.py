"""ctypes wrapper around liboracle.so -- the CPU FP64 restatement of the RRQ path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
Arrays use the reference layout: x(3,n) is a C-contiguous (n,3) array; A(m,n)
is returned as an (m,n) numpy array (stored Fortran order internally).
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None
_LIB_FMA = None
_SIDE_TOL = 1e300          # the side tolerance and level form, applied to both builds
_LEVEL_MODE = (False, True)

i64 = C.c_int64
dbl = C.c_double
P = C.c_void_p


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build()
        _LIB = C.CDLL(str(path))
    return _LIB


def _p(a):
    return a.ctypes.data_as(P) if a is not None else None


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def lib_fma():
    """the oracle compiled with fused multiply-adds (liboracle_fma.so): the same
    algorithm rounded differently, for measuring a result's rounding sensitivity"""
    global _LIB_FMA
    if _LIB_FMA is None:
        path = _HERE / "liboracle_fma.so"
        if not path.exists():
            build()
        _LIB_FMA = C.CDLL(str(path))
        _LIB_FMA.orc_set_side_tol(dbl(_SIDE_TOL))
        _LIB_FMA.orc_set_level_mode(i64(1 if _LEVEL_MODE[0] else 0), i64(1 if _LEVEL_MODE[1] else 0))
    return _LIB_FMA


def set_side_tol(tol):
    global _SIDE_TOL
    _SIDE_TOL = tol
    lib().orc_set_side_tol(dbl(tol))
    if _LIB_FMA is not None:
        _LIB_FMA.orc_set_side_tol(dbl(tol))


def projection_unconverged(reset=True):
    """(count, largest last step): Gauss-Newton projections that had not
    stopped by their 40th step, since the last reset"""
    f = lib().orc_projection_unconverged
    f.restype = C.c_int64
    st = C.c_double()
    n = int(f(C.c_int64(1 if reset else 0), C.byref(st)))
    return n, st.value


def set_level_mode(adaptive=False, square=True):
    """PR_ADAPTIVE / PR_SQUARE for every later call (both true is refused)"""
    global _LEVEL_MODE
    lib().orc_set_level_mode.restype = i64
    if lib().orc_set_level_mode(i64(1 if adaptive else 0), i64(1 if square else 0)) != 0:
        raise ValueError("PR_SQUARE and PR_ADAPTIVE cannot be combined")
    _LEVEL_MODE = (adaptive, square)
    if _LIB_FMA is not None:
        _LIB_FMA.orc_set_level_mode(i64(1 if adaptive else 0), i64(1 if square else 0))


def gauss(n):
    t = np.zeros(n); w = np.zeros(n); D = np.zeros((n, n), order="F")
    lib().orc_gauss(i64(n), _p(t), _p(w), _p(D))
    return t, w, D


def vr_nodes(order):
    n = (order + 1) * (order + 2) // 2
    uv = np.zeros((n, 2)); w = np.zeros(n)
    lib().orc_vr_nodes(i64(order), _p(uv), _p(w))
    return uv, w


def koorn_pols(u, v, nmax):
    n = (nmax + 1) * (nmax + 2) // 2
    p = np.zeros(n); du = np.zeros(n); dv = np.zeros(n)
    lib().orc_koorn_pols(dbl(u), dbl(v), i64(nmax), _p(p), _p(du), _p(dv))
    return p, du, dv


def simplex_precomp(nquad, korder):
    kp = (korder + 1) * (korder + 2) // 2
    t = np.zeros(nquad); w = np.zeros(nquad); D = np.zeros((nquad, nquad), order="F")
    wb = np.zeros(nquad); L = np.zeros((nquad, nquad), order="F")
    um = np.zeros((kp, kp), order="F"); vm = np.zeros((kp, kp), order="F")
    lib().orc_simplex_precomp(i64(nquad), i64(korder), i64(kp), _p(t), _p(w), _p(D), _p(wb), _p(L), _p(um), _p(vm))
    return dict(tgl=t, wgl=w, Dgl=D, w_bclag=wb, Legmat=L, umatr=um, vmatr=vm)


def patch_ref_info(order):
    """level sizes for patch order p (korder=p-1, npols=hdim, orderff=p+2)."""
    nl = np.zeros(4, dtype=np.int64); ns = np.zeros(4, dtype=np.int64)
    h = order * (order + 1) // 2
    lib().orc_patch_ref_info(i64(order - 1), i64(h), i64(order + 2), _p(nl), _p(ns))
    return nl, ns


def patch_ref_level(order, lev):
    nl, ns = patch_ref_info(order)
    h = order * (order + 1) // 2
    n, s = int(nl[lev - 1]), int(ns[lev - 1])
    uv = np.zeros((n, 2)); wt = np.zeros(n); im = np.zeros((n, h), order="F")
    so = np.zeros(s + 1, dtype=np.int64); st = np.zeros(s, dtype=np.int64)
    lib().orc_patch_ref_level(i64(order - 1), i64(h), i64(order + 2), i64(lev), _p(uv), _p(wt), _p(im), _p(so), _p(st))
    return dict(uv=uv, wt=wt, interpmat=im, suboff=so, subtyp=st)


def patch_levels(order, sx, rts, rps, dist=1.4):
    nl, ns = patch_ref_info(order)
    h = order * (order + 1) // 2
    nt, nst = int(nl.sum()), int(ns.sum())
    lsx = np.zeros((nt, 3)); lsnx = np.zeros((nt, 3)); lsw = np.zeros(nt)
    qp = np.zeros((nst, 3)); qr = np.zeros(nst)
    lib().orc_patch_levels(i64(order - 1), i64(h), _p(_f(sx)), _p(_f(rts)), _p(_f(rps)), i64(order + 2), dbl(dist),
                           _p(lsx), _p(lsnx), _p(lsw), _p(qp), _p(qr))
    return dict(sx=lsx, snx=lsnx, sw=lsw, qpoint=qp, qradii2=qr, nl=nl, nsub=ns)


def lap3dsdlpmat(tx, sx, snx, sw):
    tx = _f(tx); m = tx.shape[0]; n = sw.shape[0]
    As = np.zeros((m, n), order="F"); Ad = np.zeros((m, n), order="F")
    lib().orc_lap3dsdlpmat(i64(m), _p(tx), i64(n), _p(_f(sx)), _p(_f(snx)), _p(_f(sw)), _p(As), _p(Ad))
    return As, Ad


def levels_eval(order, tx, sx, rts, rps, dist=1.4):
    tx = _f(tx); m = tx.shape[0]; h = order * (order + 1) // 2
    As = np.zeros((m, h), order="F"); Ad = np.zeros((m, h), order="F")
    idxs = np.zeros(m, dtype=np.int64); ms = np.zeros(1, dtype=np.int64); lab = np.zeros(m, dtype=np.int64)
    lib().orc_levels_eval(i64(m), _p(tx), i64(h), _p(_f(sx)), _p(_f(rts)), _p(_f(rps)), i64(order - 1),
                          i64(order + 2), dbl(dist), _p(As), _p(Ad), _p(idxs), _p(ms), _p(lab))
    return As, Ad, idxs[:ms[0]], lab


def rrq(order, tx, sx, snx, sw, rts, rps, exterior=True, ichart=0, sxbd_chart=None, rv_chart=None,
        timeinfo=None, fma=False):
    """rrq_r64 for m targets against one panel; fma=True runs the build with fused
    multiply-adds (lib_fma)"""
    tx = _f(tx); m = tx.shape[0]; h = order * (order + 1) // 2; nq = order + 2
    pre = simplex_precomp(nq, order - 1)
    As = np.zeros((m, h), order="F"); Ad = np.zeros((m, h), order="F")
    Om = np.zeros((4 * h, m), order="F"); Ia = np.zeros(m); lab = np.zeros(m, dtype=np.int64)
    ti = np.zeros(20) if timeinfo is None else timeinfo
    sxbd = np.zeros((3 * nq, 3)) if sxbd_chart is None else _f(sxbd_chart)
    rv = np.zeros((3, 3)) if rv_chart is None else _f(rv_chart)
    (lib_fma() if fma else lib()).orc_rrq(
                  i64(m), _p(tx), i64(h), _p(_f(sx)), _p(_f(snx)), _p(_f(sw)), _p(_f(rts)), _p(_f(rps)),
                  i64(order), i64(nq), i64(order + 2), dbl(1.4), i64(1 if exterior else 0), i64(ichart),
                  _p(sxbd), _p(rv), _p(pre["tgl"]), _p(pre["wgl"]), _p(pre["Dgl"]), _p(pre["w_bclag"]),
                  _p(pre["Legmat"]), _p(pre["umatr"]), _p(pre["vmatr"]), _p(As), _p(Ad), _p(Om), _p(Ia),
                  _p(ti), _p(lab))
    return dict(As=As, Ad=Ad, Omega=Om, Ias=Ia, label=lab, timeinfo=ti)


def set_near_radius(diam_factor=0.0):
    """radius rule of later nearlist calls: 0 the reference ball, f > 0 f times
    the panel's largest corner distance"""
    lib().orc_set_near_radius(dbl(diam_factor))


def nearlist(hdim, sx, sw, tx, nthreads=1):
    sx = _f(sx); sw = _f(sw); tx = _f(tx)
    npan = sw.shape[0] // hdim
    offs = np.zeros(npan + 1, dtype=np.int64)
    lib().orc_nearlist_count_mt(i64(npan), i64(hdim), _p(sx), _p(sw), i64(tx.shape[0]), _p(tx), _p(offs),
                                i64(nthreads))
    nnz = int(offs[-1])
    idx = np.zeros(nnz, dtype=np.int64)
    lib().orc_nearlist_fill_mt(i64(npan), i64(hdim), _p(sx), _p(sw), i64(tx.shape[0]), _p(tx), _p(offs), _p(idx),
                               i64(nthreads))
    return offs, idx


def labels(order, geo, tx, offs, idx, nthreads=1, cuts=False):
    """level label (1..4, or 5 for RRQ) of every pair of a CSR near list; with
    the adaptive level mode also the tree cut of each pair (cuts=True)"""
    h = order * (order + 1) // 2
    npan = geo["sw"].shape[0] // h
    lab = np.zeros(int(offs[-1]), dtype=np.int64)
    cut = np.zeros(int(offs[-1]), dtype=np.int64) if cuts else None
    lib().orc_labels(i64(npan), i64(order), _p(_f(geo["sx"])), _p(_f(geo["rts"])), _p(_f(geo["rps"])), _p(_f(tx)),
                     _p(np.ascontiguousarray(offs, dtype=np.int64)), _p(np.ascontiguousarray(idx, dtype=np.int64)),
                     _p(lab), _p(cut), i64(nthreads))
    return (lab, cut) if cuts else lab


def close_apply(order, geo, tx, offs, idx, ub, ubn, exterior=True, ichart=0, sxbd_all=None, rv_all=None,
                u=None, nthreads=1, timeinfo=None):
    h = order * (order + 1) // 2; nq = order + 2
    npan = geo["sw"].shape[0] // h
    nnz = int(offs[-1])
    val = np.zeros(nnz); lab = np.zeros(nnz, dtype=np.int64)
    ti = np.zeros(20) if timeinfo is None else timeinfo
    sxbd = np.zeros(1) if sxbd_all is None else _f(sxbd_all)
    rv = np.zeros(1) if rv_all is None else _f(rv_all)
    lib().orc_close_apply(i64(npan), i64(order), i64(h), i64(nq), _p(_f(geo["sx"])), _p(_f(geo["snx"])),
                          _p(_f(geo["sw"])), _p(_f(geo["rts"])), _p(_f(geo["rps"])), _p(_f(tx)),
                          _p(np.ascontiguousarray(offs, dtype=np.int64)), _p(np.ascontiguousarray(idx, dtype=np.int64)),
                          _p(_f(ub)), _p(_f(ubn)), i64(1 if exterior else 0), i64(ichart), _p(sxbd), _p(rv),
                          _p(val), _p(u), _p(lab), _p(ti), i64(nthreads))
    return val, lab


def far_direct(tx, sx, snx, sw, ub, ubn, nthreads=1):
    tx = _f(tx); u = np.zeros(tx.shape[0])
    lib().orc_far_direct(i64(tx.shape[0]), _p(tx), i64(sw.shape[0]), _p(_f(sx)), _p(_f(snx)), _p(_f(sw)),
                         _p(_f(ub)), _p(_f(ubn)), _p(u), i64(nthreads))
    return u


def getnearquad(order, geo, tcx, tcxrow, tcxi, iside=30):
    h = order * (order + 1) // 2; nq = order + 2
    npan = geo["sw"].shape[0] // h
    nnz = int((tcxi[-1] - 1) * h)
    si = np.zeros(nnz, dtype=np.int64); sj = np.zeros(nnz, dtype=np.int64)
    vs = np.zeros(nnz); vd = np.zeros(nnz); ti = np.zeros(20)
    lib().orc_getnearquad(i64(npan), i64(order), i64(h), i64(nq), i64(geo["sw"].shape[0]), _p(_f(geo["sx"])),
                          _p(_f(geo["snx"])), _p(_f(geo["sw"])), _p(_f(geo["rts"])), _p(_f(geo["rps"])),
                          i64(tcx.shape[0]), _p(_f(tcx)), _p(np.ascontiguousarray(tcxrow, dtype=np.int64)),
                          _p(np.ascontiguousarray(tcxi, dtype=np.int64)), i64(order + 2), dbl(1.4), i64(iside),
                          i64(nnz), _p(si), _p(sj), _p(vs), _p(vd), _p(ti))
    return si, sj, vs, vd


def closepanel_dlp(tx, sx, order, ref=1):
    """close DLP matrix Ac (m, order^2) of a tensor-product patch (closepanel.cpp)"""
    tx = _f(tx); m = tx.shape[0]
    Ac = np.zeros((m, order * order), order="F")
    lib().orc_closepanel_dlp(i64(m), _p(tx), i64(order), _p(_f(sx)), i64(ref), _p(Ac))
    return Ac
