"""Stellarator workloads: geometry, near-surface targets, point-charge densities.

Geometry and targets come from the library's host generator (a restatement of
utils/stellarator_geo_mex.c); densities are closed forms evaluated here.
"""
import numpy as np

from ._lib import check, dbl, i64, lib, ptr

# 5-period surface R = 1 + 0.3 cos th + 0.06 cos(th - 5 ph), Z = 0.3 sin th - 0.06 sin(th - 5 ph)
VMEC_BENCH = np.array([[0, 0, 1.0, 0.0], [1, 0, 0.3, 0.3], [1, 1, 0.06, -0.06]], dtype=np.float64)


def _modes(modes):
    nmode = 0 if modes is None else len(modes)
    mn = np.zeros(4) if modes is None else np.ascontiguousarray(modes, dtype=np.float64)
    return nmode, mn


def charts(mp, np_, p, modes=None, nfp=5, restol=0.0, cap=0, gpu=False):
    """The adaptive chart list (qol_geo_charts, stellarator_charts): (n, 6) rows
    a1 b1 a2 b2 jj ii.  restol <= 0: aspect splits only.  gpu=True evaluates the
    curvature tails on the GPU (qol_geo_charts_gpu)."""
    L = lib()
    nmode, mn = _modes(modes)
    fn = L.qol_geo_charts_gpu if gpu else L.qol_geo_charts
    n = fn(i64(mp), i64(np_), i64(p), i64(nfp), i64(nmode), ptr(mn), dbl(restol), i64(cap), None)
    if n == -1:
        raise RuntimeError("qol_geo_charts: the chart list would exceed cap; loosen restol")
    if n < 0:
        raise RuntimeError("qol_geo_charts: %s" % L.qol_last_error().decode())
    cl = np.zeros((n, 6))
    m = fn(i64(mp), i64(np_), i64(p), i64(nfp), i64(nmode), ptr(mn), dbl(restol), i64(cap), ptr(cl))
    assert m == n
    return cl


def curvature_tails_gpu(mp, np_, p, cl, modes=None, nfp=5):
    """panel_resolution of every chart of cl on the GPU (qol_geo_curvature_tails_gpu)."""
    nmode, mn = _modes(modes)
    cl = np.ascontiguousarray(cl, dtype=np.float64).reshape(-1, 6)
    out = np.zeros(len(cl))
    check(lib().qol_geo_curvature_tails_gpu(i64(mp), i64(np_), i64(p), i64(nfp), i64(nmode), ptr(mn), i64(len(cl)),
                                            ptr(cl), ptr(out)), "qol_geo_curvature_tails_gpu")
    return out


def curvature_tail(mp, np_, p, chart, modes=None, nfp=5):
    """panel_resolution of one chart (qol_geo_curvature_tail)."""
    nmode, mn = _modes(modes)
    c = np.ascontiguousarray(chart, dtype=np.float64)
    v = lib().qol_geo_curvature_tail(i64(mp), i64(np_), i64(p), i64(nfp), i64(nmode), ptr(mn), ptr(c))
    if v < 0:
        raise RuntimeError("qol_geo_curvature_tail failed")
    return v


def geometry(mp, np_, p, modes=None, nfp=5, iterate_aspect=None, gpu=False, restol=None, cap=0):
    """Triangulated surface: dict(sx, snx, sw, rts, rps (nsrc,3|nsrc), ntri, p, h).

    modes=None -> the reference's built-in 7-term surface (single aspect pass);
    modes=(nmode,4) array of (m, n, Rc, Zs) -> VMEC form with nfp periods.
    restol -> the adaptive chart list (aspect, curvature refinement for restol > 0,
    2:1 balance) of qol_geo_charts, built by qol_geo_build_charts[_gpu].
    gpu=True runs the per-chart work on the GPU (qol_geo_build_gpu).
    """
    L = lib()
    nmode, mn = _modes(modes)
    h = p * (p + 1) // 2
    if restol is not None:
        return geometry_of_charts(mp, np_, p, charts(mp, np_, p, modes, nfp, restol, cap, gpu=gpu), modes, nfp, gpu,
                                  restol=restol)
    it = (0 if modes is None else 1) if iterate_aspect is None else int(iterate_aspect)
    ntri = L.qol_geo_count(i64(mp), i64(np_), i64(p), i64(nfp), i64(nmode), ptr(mn), i64(it))
    if ntri <= 0:
        raise RuntimeError("geometry chart build failed")
    nsrc = ntri * h
    sx = np.zeros((nsrc, 3)); snx = np.zeros((nsrc, 3)); sw = np.zeros(nsrc)
    rts = np.zeros((nsrc, 3)); rps = np.zeros((nsrc, 3)); cl = np.zeros((ntri // 2, 6))
    build = L.qol_geo_build_gpu if gpu else L.qol_geo_build
    check(build(i64(mp), i64(np_), i64(p), i64(nfp), i64(nmode), ptr(mn), i64(it), ptr(sx), ptr(snx),
                ptr(sw), ptr(rts), ptr(rps), ptr(cl)), "qol_geo_build_gpu" if gpu else "qol_geo_build")
    return dict(sx=sx, snx=snx, sw=sw, rts=rts, rps=rps, ntri=ntri, p=p, h=h, mp=mp, np=np_, charts=cl,
                modes=modes, nfp=nfp, iterate_aspect=it, restol=None)


def geometry_of_charts(mp, np_, p, cl, modes=None, nfp=5, gpu=False, restol=0.0):
    """Triangulated surface of a given chart list (rows a1 b1 a2 b2 jj ii) at
    patch order p (qol_geo_build_charts[_gpu], stellarator_geo_charts)."""
    L = lib()
    nmode, mn = _modes(modes)
    cl = np.ascontiguousarray(cl, dtype=np.float64).reshape(-1, 6)
    h = p * (p + 1) // 2
    ntri = 2 * len(cl)
    nsrc = ntri * h
    sx = np.zeros((nsrc, 3)); snx = np.zeros((nsrc, 3)); sw = np.zeros(nsrc)
    rts = np.zeros((nsrc, 3)); rps = np.zeros((nsrc, 3))
    build = L.qol_geo_build_charts_gpu if gpu else L.qol_geo_build_charts
    check(build(i64(mp), i64(np_), i64(p), i64(nfp), i64(nmode), ptr(mn), i64(len(cl)), ptr(cl), ptr(sx),
                ptr(snx), ptr(sw), ptr(rts), ptr(rps)),
          "qol_geo_build_charts_gpu" if gpu else "qol_geo_build_charts")
    return dict(sx=sx, snx=snx, sw=sw, rts=rts, rps=rps, ntri=ntri, p=p, h=h, mp=mp, np=np_, charts=cl,
                modes=modes, nfp=nfp, iterate_aspect=1, restol=restol)


def chart_boundary(geo):
    """exact chart boundary at the edge Gauss-Legendre nodes (ichart=1 inputs)."""
    L = lib()
    p = geo["p"]; nq = p + 2; ntri = geo["ntri"]
    nmode, mn = _modes(geo["modes"])
    sxbd = np.zeros((ntri, 3 * nq, 3)); rv = np.zeros((ntri, 3, 3))
    if geo.get("restol") is not None:
        cl = np.ascontiguousarray(geo["charts"])
        check(L.qol_geo_chart_boundary_charts(i64(geo["mp"]), i64(geo["np"]), i64(p), i64(geo["nfp"]), i64(nmode),
                                              ptr(mn), i64(len(cl)), ptr(cl), i64(nq), ptr(sxbd), ptr(rv)),
              "qol_geo_chart_boundary_charts")
        return sxbd, rv
    check(L.qol_geo_chart_boundary(i64(geo["mp"]), i64(geo["np"]), i64(p), i64(geo["nfp"]), i64(nmode), ptr(mn),
                                   i64(geo["iterate_aspect"]), i64(nq), ptr(sxbd), ptr(rv)), "qol_geo_chart_boundary")
    return sxbd, rv


def targets(geo, nt, umin, umax, sides=0, seed=1):
    """near-surface targets: offset side*h*10^U along the patch normal."""
    tx = np.zeros((nt, 3)); tri = np.zeros(nt, dtype=np.int64); dist = np.zeros(nt)
    check(lib().qol_make_targets(i64(geo["p"]), i64(geo["ntri"]), ptr(geo["sx"]), ptr(geo["snx"]), i64(nt),
                                 dbl(umin), dbl(umax), i64(sides), C_u64(seed), ptr(tx), ptr(tri), ptr(dist)),
          "qol_make_targets")
    return tx, tri, dist


def C_u64(v):
    import ctypes
    return ctypes.c_uint64(v)


def point_charges(geo, nchg=8, offset=0.4, seed=2):
    """nchg charges q~N(0,1) at x_s + offset*n(x_s) for surface nodes x_s drawn uniformly."""
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, geo["sx"].shape[0], size=nchg)
    pos = geo["sx"][idx] + offset * geo["snx"][idx]
    q = rng.standard_normal(nchg)
    return pos, q


def charge_field(x, pos, q):
    """u = sum q/(4 pi |x-p|) and grad u."""
    d = x[:, None, :] - pos[None, :, :]
    r = np.sqrt((d * d).sum(-1))
    u = (q[None, :] / (4 * np.pi * r)).sum(1)
    g = -(q[None, :, None] * d / (4 * np.pi * r[..., None] ** 3)).sum(1)
    return u, g


def harmonic_exp(x, lam=0.5):
    """the reference GRF density exp(lam x) cos(lam y) and its gradient
    (stellarator_grf_core_mod.f90:1089-1103)."""
    e = np.exp(lam * x[:, 0])
    u = e * np.cos(lam * x[:, 1])
    g = np.stack([lam * e * np.cos(lam * x[:, 1]), -lam * e * np.sin(lam * x[:, 1]), np.zeros(len(x))], 1)
    return u, g
