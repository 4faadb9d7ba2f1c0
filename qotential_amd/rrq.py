"""Host-side mirror of qotential's close-evaluation entry points.

Same names, argument meaning and error behaviour as the reference's MEX/Fortran
interface (src/lap3d_close_mex.f90, src/patch_refine_mex.f90), implemented by
calling the drop-in symbols of libqotential_hip.so:

  simplex_precomp(nquad, korder)                     qol_simplex_precomp_mex
  rrq(tx, sx, snx, sw, rts, rps, order, ...)         qol_rrq_mex
  getnearquad_lap_rrq(npan, nterms, ...)             qol_getnearquad_lap_rrq_mex
  build_patch_levels / lap3dsdlpmat_levels /
  clear_patch_levels / clear_patch_levels_all        qol_*_patch_levels_mex
"""
import ctypes as C

import numpy as np

from ._lib import QolError, dbl, i64, lib, ptr


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _r(v):
    return C.byref(i64(v))


def _status(what):
    st = lib().qol_last_status()
    if st != 0:
        raise QolError(f"{what}: {lib().qol_last_error().decode()}")


def set_level_mode(adaptive=False, square=True):
    """PR_ADAPTIVE / PR_SQUARE (src/patch_refine_mod.f90:51-55) for every later
    drop-in call; both true is refused at the next evaluation, as the reference stops"""
    lib().qol_set_level_mode(i64(1 if adaptive else 0), i64(1 if square else 0))
    _status("qol_set_level_mode")


def level_mode():
    a, s = C.c_int64(), C.c_int64()
    lib().qol_get_level_mode(C.byref(a), C.byref(s))
    return bool(a.value), bool(s.value)


def simplex_precomp(nquad, korder):
    kp = (korder + 1) * (korder + 2) // 2
    out = dict(tgl=np.zeros(nquad), wgl=np.zeros(nquad), Dgl=np.zeros((nquad, nquad), order="F"),
               w_bclag=np.zeros(nquad), Legmat=np.zeros((nquad, nquad), order="F"),
               umatr=np.zeros((kp, kp), order="F"), vmatr=np.zeros((kp, kp), order="F"))
    lib().qol_simplex_precomp_mex_(_r(nquad), _r(korder), _r(kp), *[ptr(out[k]) for k in
                                   ("tgl", "wgl", "Dgl", "w_bclag", "Legmat", "umatr", "vmatr")])
    _status("qol_simplex_precomp_mex")
    return out


def rrq(tx, sx, snx, sw, rts, rps, order, exterior=True, ichart=0, sxbd_chart=None, rv_chart=None, pre=None,
        timeinfo=None, noalloc=False):
    """As, Ad (m x n), Omega (4h x m), IalphaAsvestas (m) for m targets against one panel.

    noalloc=True calls qol_rrq_noalloc_ (no device allocation per call: the
    library's fixed device scratch, which bounds m)."""
    tx = _f(tx)
    m = tx.shape[0]
    n = order * (order + 1) // 2
    nq = order + 2
    pre = pre or simplex_precomp(nq, order - 1)
    As = np.zeros((m, n), order="F"); Ad = np.zeros((m, n), order="F")
    Om = np.zeros((4 * n, m), order="F"); Ia = np.zeros(m)
    ti = np.zeros(20) if timeinfo is None else timeinfo
    sxbd = np.zeros((3 * nq, 3)) if sxbd_chart is None else _f(sxbd_chart)
    rv = np.zeros((3, 3)) if rv_chart is None else _f(rv_chart)
    fn = lib().qol_rrq_noalloc_ if noalloc else lib().qol_rrq_mex_
    fn(_r(m), ptr(tx), _r(n), ptr(_f(sx)), ptr(_f(snx)), ptr(_f(sw)), ptr(_f(rts)), ptr(_f(rps)),
                       _r(order), _r(nq), _r(order + 2), C.byref(dbl(1.4)), _r(1 if exterior else 0), _r(0),
                       _r(ichart), ptr(sxbd), ptr(rv), *[ptr(pre[k]) for k in
                       ("tgl", "wgl", "Dgl", "w_bclag", "Legmat", "umatr", "vmatr")],
                       ptr(As), ptr(Ad), ptr(Om), ptr(Ia), ptr(ti))
    _status("qol_rrq_noalloc" if noalloc else "qol_rrq_mex")
    st = C.c_double()
    n = lib().qol_last_projection_unconverged(C.byref(st))
    return dict(As=As, Ad=Ad, Omega=Om, Ias=Ia, timeinfo=ti, unconverged=(n, st.value))


def getnearquad_lap_rrq(geo, tcx, tcxrow, tcxi, iside=30):
    """COO triplets (sparsei, sparsej, sparsevs, sparsevd) of the near correction."""
    p = geo["p"]; h = p * (p + 1) // 2; nq = p + 2
    npan = geo["sw"].shape[0] // h
    tcx = _f(tcx)
    tcxrow = np.ascontiguousarray(tcxrow, dtype=np.int64)
    tcxi = np.ascontiguousarray(tcxi, dtype=np.int64)
    nnz = int((tcxi[-1] - 1) * h)
    si = np.zeros(max(nnz, 1), dtype=np.int64); sj = np.zeros(max(nnz, 1), dtype=np.int64)
    vs = np.zeros(max(nnz, 1)); vd = np.zeros(max(nnz, 1)); ti = np.zeros(20)
    lib().qol_getnearquad_lap_rrq_mex_(_r(npan), _r(p), _r(h), _r(nq), _r(npan * h), *[ptr(_f(geo[k])) for k in
                                       ("sx", "snx", "sw", "rts", "rps")], _r(tcx.shape[0]), ptr(tcx), ptr(tcxrow),
                                       ptr(tcxi), _r(p + 2), C.byref(dbl(1.4)), _r(iside), _r(0), _r(nnz), ptr(si),
                                       ptr(sj), ptr(vs), ptr(vd), ptr(ti))
    _status("qol_getnearquad_lap_rrq_mex")
    return si[:nnz], sj[:nnz], vs[:nnz], vd[:nnz]


def lap3ddlp_closepanel(tx, sx, order, ref=1, if_adapt=True):
    """close DLP matrix Ac (m x order^2) of a tensor-product patch (qol_lap3ddlp_closepanel_mex)"""
    tx = _f(tx)
    m = tx.shape[0]
    Ac = np.zeros((m, order * order), order="F")
    lib().qol_lap3ddlp_closepanel_mex_(_r(m), ptr(tx), _r(order * order), ptr(_f(sx)), _r(order), _r(ref),
                                       C.byref(dbl(1.0 if if_adapt else 0.0)), ptr(Ac))
    _status("qol_lap3ddlp_closepanel_mex")
    return Ac


def build_patch_levels(sx, rts, rps, korder, dist=1.4, nlevel=4):
    npols = (korder + 1) * (korder + 2) // 2
    lv = i64(0); ier = i64(0)
    lib().qol_build_patch_levels_mex_(_r(korder), _r(npols), ptr(_f(sx)), ptr(_f(rts)), ptr(_f(rps)),
                                      _r(korder + 3), C.byref(dbl(dist)), _r(nlevel), C.byref(lv), C.byref(ier))
    return lv.value, ier.value


def lap3dsdlpmat_levels(tx, npols, lv, As=None, Ad=None):
    tx = _f(tx)
    m = tx.shape[0]
    As = np.zeros((m, npols), order="F") if As is None else As
    Ad = np.zeros((m, npols), order="F") if Ad is None else Ad
    idxs = np.zeros(max(m, 1), dtype=np.int64); ms = i64(0); ier = i64(0)
    lib().qol_lap3dsdlpmat_levels_mex_(_r(m), ptr(tx), _r(npols), _r(lv), _r(0), ptr(As), ptr(Ad), ptr(idxs),
                                       C.byref(ms), C.byref(ier))
    return As, Ad, idxs[: ms.value], ier.value


def clear_patch_levels(lv):
    ier = i64(0)
    lib().qol_clear_patch_levels_mex_(_r(lv), C.byref(ier))
    return ier.value


def clear_patch_levels_all():
    ier = i64(0)
    lib().qol_clear_patch_levels_all_mex_(C.byref(ier))
    return ier.value
