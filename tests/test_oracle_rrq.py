"""Pins of the oracle's RRQ restatement (CPU only).

The RRQ helpers of the reference live in absent submodules (SSQ weights, solid
angle, omega moments), so no function-level golden vector exists for them.
These tests pin them by what the mathematics fixes:
  - singularity-swap line weights against adaptive quadrature (scipy) on a
    curved edge, and to roundoff on a straight edge where the rule is exact;
  - the solid angle and D[1] = Omega/4pi on flat triangles against the
    Van Oosterom-Strackee closed form;
  - Gauss's law (D[1] = -1 inside, 0 outside) for the direct kernel on a closed
    stellarator surface;
  - RRQ weights applied to smooth densities against brute-force quadrature of
    the same polynomial patch, with the spectral decay in p the method has
    (the reference's published GRF table: 1.8e-4, 2.1e-5, 2.7e-7 at p=4, 6, 8).
"""
import numpy as np
import pytest
from scipy.integrate import quad

from qotential_amd import workload as W


def _curve_coefs(oracle, nq, curve):
    pre = oracle.simplex_precomp(nq, 3)
    t, Leg = pre["tgl"], pre["Legmat"]
    vals = np.array([curve(s) for s in t])
    coef = np.concatenate([Leg @ vals[:, c] for c in range(3)])
    return t, coef


def _ssq(oracle, coef, nq, x):
    w = np.zeros(nq)
    so = np.zeros(4)
    oracle.lib().orc_debug_ssq(oracle._p(coef), oracle.i64(nq), oracle._p(np.ascontiguousarray(x, dtype=float)),
                               oracle._p(w), oracle._p(so))
    return w, so


@pytest.mark.parametrize("s0", [0.3, -0.95, 0.99])
def test_ssq_curved_edge_vs_adaptive(oracle, s0):
    nq = 10
    curve = lambda s: np.array([s, 0.1 * s * s + 0.05 * s, 0.05 * s ** 3 - 0.02 * s * s])
    dcurve = lambda s: np.array([1.0, 0.2 * s + 0.05, 0.15 * s * s - 0.04 * s])
    t, coef = _curve_coefs(oracle, nq, curve)
    f = lambda s: np.cos(s) + s * s
    for d in [1e-1, 1e-2, 1e-4, 1e-6]:
        n = np.cross(dcurve(s0), [0.3, 0.1, 1.0])
        x = curve(s0) + d * n / np.linalg.norm(n)
        w, so = _ssq(oracle, coef, nq, x)
        assert so[2] == 1.0 and abs(so[0] - s0) < 2 * d and 0 < abs(so[1]) < 2 * d
        R = lambda s: np.linalg.norm(curve(s) - x)
        approx = sum(w[i] * f(t[i]) / R(t[i]) for i in range(nq))
        exact = quad(lambda s: f(s) / R(s), -1, 1, points=[s0], epsabs=1e-15, epsrel=1e-14, limit=500)[0]
        # a 10-node rule on a curved edge: the smooth factor |s-s0|/R(s) is not a
        # polynomial, measured error 1e-10..4e-9
        assert abs(approx - exact) < 1e-8 * abs(exact), (d, approx, exact)


def test_ssq_straight_edge_exact(oracle):
    """On a straight edge |s - s0| / R(s) is constant, so the rule integrates
    polynomial densities of degree < nq exactly."""
    nq = 8
    a, b = np.array([0.2, -0.1, 0.4]), np.array([0.7, 0.3, -0.2])
    curve = lambda s: a + s * b
    t, coef = _curve_coefs(oracle, nq, curve)
    bl = np.linalg.norm(b)
    for s0, d in [(0.1, 1e-3), (-0.7, 1e-6), (0.95, 3e-2)]:
        n = np.cross(b, [0.0, 0.0, 1.0])
        x = curve(s0) + d * n / np.linalg.norm(n)
        w, _ = _ssq(oracle, coef, nq, x)
        for deg in range(nq):
            f = lambda s: (s - 0.2) ** deg
            R = lambda s: np.linalg.norm(curve(s) - x)
            approx = sum(w[i] * f(t[i]) / R(t[i]) for i in range(nq))
            # closed form: int (s-0.2)^deg / (|b| sqrt((s-s0)^2 + e^2)) ds by quad on the
            # smooth (after the substitution s = s0 + e sinh u) integrand
            e = d / bl
            g = lambda u: f(s0 + e * np.sinh(u)) / bl
            u0, u1 = np.arcsinh((-1 - s0) / e), np.arcsinh((1 - s0) / e)
            exact = quad(g, u0, u1, epsabs=1e-15, epsrel=1e-14, limit=500)[0]
            # the target's own rounding (eps |x|) fixes the root's imaginary part
            # d / |b| only to eps |x| / d relative, which moves the log term by
            # about 2 eps |x| / d
            cond = 10 * np.finfo(float).eps * np.linalg.norm(x) / d
            assert abs(approx - exact) < 1e-12 * max(1.0, abs(exact)) + cond, (s0, d, deg, approx, exact)


@pytest.mark.parametrize("x0,d", [(1.0 + 1e-8, 1e-9), (1.0, 1e-9), (-1.0 - 3e-9, 2e-9), (0.999999, 1e-10)])
def test_ssq_root_near_an_endpoint(oracle, x0, d):
    """A target 1e-9 off a straight edge's line near its endpoint: the root of
    |X(s) - x|^2 is x0 + i d / |b| exactly.  The chord start's imaginary part
    must come out positive (from the cross product: sqrt(|C|^2 |D|^2 - (C.D)^2)
    rounds to 0 here, and Newton from a real start never leaves the real axis,
    fails, and the edge loses its singularity swap)."""
    nq = 10
    a, b = np.array([0.2, -0.1, 0.4]), np.array([0.7, 0.3, -0.2])
    t, coef = _curve_coefs(oracle, nq, lambda s: a + s * b)
    n = np.cross(b, [0.0, 0.0, 1.0])
    x = a + x0 * b + d * n / np.linalg.norm(n)
    w, so = _ssq(oracle, coef, nq, x)
    bl = np.linalg.norm(b)
    assert so[2] == 1.0
    # the target itself carries rounding of about 1e-16 in each coordinate
    assert abs(so[0] - x0) < 1e-14 and abs(abs(so[1]) - d / bl) < 1e-15


def _flat_panel(oracle, p):
    h = p * (p + 1) // 2
    uv, w = oracle.vr_nodes(p - 1)
    v0, v1, v2 = np.array([0.1, -0.2, 0.3]), np.array([1.1, 0.1, 0.2]), np.array([0.3, 0.9, 0.5])
    sx = v0 + uv[:, :1] * (v1 - v0) + uv[:, 1:2] * (v2 - v0)
    nrm = np.cross(v1 - v0, v2 - v0)
    a2 = np.linalg.norm(nrm)
    geo = dict(sx=sx, snx=np.tile(nrm / a2, (h, 1)), sw=w * a2, rps=np.tile(v1 - v0, (h, 1)),
               rts=np.tile(v2 - v0, (h, 1)))
    return geo, (v0, v1, v2), nrm / a2


def _solid_angle(x, v):
    """Van Oosterom-Strackee: signed solid angle of triangle v seen from x."""
    a, b, c = (vi - x for vi in v)
    la, lb, lc = (np.linalg.norm(z) for z in (a, b, c))
    num = np.dot(a, np.cross(b, c))
    den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
    return 2.0 * np.arctan2(num, den)


@pytest.mark.parametrize("p", [4, 8])
def test_solid_angle_and_D1_flat_triangle(oracle, p):
    oracle.set_side_tol(1e-10)
    try:
        geo, v, n = _flat_panel(oracle, p)
        xs = []
        for bu, bv in [(0.3, 0.3), (0.1, 0.2), (0.45, 0.45), (0.02, 0.5)]:
            for d in [1e-2, -1e-3, 1e-5, -1e-7]:
                xs.append(v[0] + bu * (v[1] - v[0]) + bv * (v[2] - v[0]) + d * n)
        xs = np.array(xs)
        r = oracle.rrq(p, xs, geo["sx"], geo["snx"], geo["sw"], geo["rts"], geo["rps"], exterior=True)
        assert (r["label"] == 5).all()
        for i, x in enumerate(xs):
            om = _solid_angle(x, v)
            assert abs(r["Ias"][i] - om) < 1e-13
            # D[1](x) = int (x-y).n / (4 pi |x-y|^3) dS = -om / (4 pi) in this sign
            assert abs(4 * np.pi * r["Ad"][i].sum() + om) < 1e-11
    finally:
        oracle.set_side_tol(1e300)


def test_gauss_law_direct_kernel(oracle):
    """D[1] = -1 inside a closed surface and 0 outside (smooth quadrature,
    targets far from the surface)."""
    geo = W.geometry(16, 40, 6, modes=W.VMEC_BENCH)
    tx = np.array([[1.0, 0.0, 0.0], [0.0, 1.02, 0.0], [-0.98, 0.0, 0.01], [2.0, 0.0, 0.0], [0.0, 0.0, 1.0],
                   [0.0, 0.0, 0.0]])
    u = oracle.far_direct(tx, geo["sx"], geo["snx"], geo["sw"], np.ones(len(geo["sw"])), np.zeros(len(geo["sw"])),
                          nthreads=4)
    # far_direct returns S[ubn] - D[ub]; smooth-rule error at 0.3-0.6 from the
    # surface is 1e-9..1e-8 at this resolution
    assert np.abs(u[:3] - 1.0).max() < 1e-7
    assert np.abs(u[3:]).max() < 1e-7


def test_rrq_smooth_density_vs_brute_force(oracle):
    """RRQ applied to smooth densities against brute-force quadrature of the same
    degree-(p-1) patch; the error falls spectrally with p."""
    oracle.set_side_tol(1e-10)
    bound = {4: 1e-3, 6: 2e-5, 8: 5e-7}
    errs = {}
    try:
        for p in (4, 6, 8):
            h = p * (p + 1) // 2
            geo = W.geometry(16, 40, p, modes=W.VMEC_BENCH)
            worst = 0.0
            for k in (123, 777):
                sl = slice(k * h, (k + 1) * h)
                sx, snx, sw, rts, rps = (geo[key][sl] for key in ("sx", "snx", "sw", "rts", "rps"))
                Lp = max(np.linalg.norm(sx[i] - sx[j]) for i in range(h) for j in range(h))
                rng = np.random.default_rng(3)
                xs = []
                for i in range(8):
                    j = rng.integers(h)
                    d = 10 ** rng.uniform(-1.5, -0.8) * Lp * (1 if i % 2 == 0 else -1)
                    xs.append(sx[j] + d * snx[j])
                xs = np.array(xs)
                m = len(xs)
                r = oracle.rrq(p, xs, sx, snx, sw, rts, rps, exterior=True)
                assert (r["label"] == 5).all()
                As = np.zeros((m, h), order="F")
                Ad = np.zeros((m, h), order="F")
                om = np.zeros(m)
                oracle.lib().orc_brute(oracle.i64(m), oracle._p(oracle._f(xs)), oracle.i64(p),
                                       oracle._p(oracle._f(sx)), oracle._p(oracle._f(snx)), oracle.i64(64),
                                       oracle.i64(20), oracle._p(As), oracle._p(Ad), oracle._p(om))
                # the solid angle is exact on the curved patch as well
                assert np.abs(r["Ias"] - om).max() < 1e-9
                dens = [np.ones(h), sx[:, 0], np.exp(sx[:, 0] / 2) * np.cos(sx[:, 1] / 2), sx[:, 2] ** 2]
                for f in dens:
                    for i in range(m):
                        es = abs(r["As"][i] @ f - As[i] @ f) / (np.abs(As[i]).sum() * np.abs(f).max())
                        ed = abs(r["Ad"][i] @ f - Ad[i] @ f) / (np.abs(Ad[i]).sum() * np.abs(f).max())
                        worst = max(worst, es, ed)
            errs[p] = worst
            assert worst < bound[p], (p, worst)
    finally:
        oracle.set_side_tol(1e300)
    assert errs[8] < errs[6] < errs[4]


def test_nearlist_count_fixture(oracle, geo4):
    """The ball rule on the order-4 fixture surface (its nodes reproduced to
    1e-12), targets = the nodes: 137,259 pairs, each panel's list ascending."""
    offs, idx = oracle.nearlist(10, geo4["sx"], geo4["sw"], geo4["sx"])
    assert int(offs[-1]) == 137259
    for k in range(0, len(offs) - 1, 97):
        seg = idx[offs[k]:offs[k + 1]]
        assert (np.diff(seg) > 0).all()
