"""The oracle's tensor-patch close DLP (oracle/closepanel.cpp, restating what
src/lap3d_close_mod.f90:135-270 computes) pinned by closed forms and brute
force: on a flat square the weights' row sum is D[1] = -Omega/(4 pi) (Van
Oosterom-Strackee solid angle) down to 1e-6 from the plane; on the reference's
own curved smoke-test patch (test/test_lap3d_close.f90: a torus patch, order 8)
the weights applied to smooth densities match a brute-force tensor Gauss rule
at distances where that rule converges."""
import numpy as np
import pytest


def _solid_square(x, a):
    def tri(r1, r2, r3):
        R1, R2, R3 = r1 - x, r2 - x, r3 - x
        l1, l2, l3 = map(np.linalg.norm, (R1, R2, R3))
        num = np.dot(R1, np.cross(R2, R3))
        den = l1 * l2 * l3 + np.dot(R1, R2) * l3 + np.dot(R1, R3) * l2 + np.dot(R2, R3) * l1
        return 2 * np.arctan2(num, den)
    c = [np.array(v, dtype=float) for v in ([-a, -a, 0], [a, -a, 0], [a, a, 0], [-a, a, 0])]
    return tri(c[0], c[1], c[2]) + tri(c[0], c[2], c[3])


def flat_square(oracle, p, a):
    t, _, _ = oracle.gauss(p)
    return np.array([[a * t[i], a * t[j], 0.0] for j in range(p) for i in range(p)])


def torus_patch(oracle, p=8):
    """the patch of test/test_lap3d_close.f90:44-73 (a = 1, b = 0.5, ref = 1),
    centred on the origin"""
    t, _, _ = oracle.gauss(p)
    pts = []
    for i in range(p):
        for j in range(p):
            x1, x2 = 0.5 * t[j], 0.25 * t[i]
            pts.append([0.5 * np.sin(x1), (1 + 0.5 * np.cos(x1)) * np.sin(x2), (1 + 0.5 * np.cos(x1)) * np.cos(x2)])
    s = np.array(pts)
    return s - s.mean(0)


def test_flat_square_row_sums_are_solid_angles(oracle):
    p, a = 6, 0.1
    sx = flat_square(oracle, p, a)
    xs = np.array([[0.01, 0.02, 0.05], [0.03, -0.05, 1e-3], [0.0, 0.0, 1e-6], [0.09, 0.09, -1e-4],
                   [0.2, 0.0, 0.01], [0.099, -0.02, 2e-5]])
    Ac = oracle.closepanel_dlp(xs, sx, p)
    for i, x in enumerate(xs):
        assert abs(Ac[i].sum() + _solid_square(x, a) / (4 * np.pi)) < 2e-11, i


def brute_dlp(sx, p, xs, oracle, nsub=24, qq=20):
    """tensor Gauss qq x qq on nsub x nsub cells of the parameter square, with the
    patch's tensor interpolant, in numpy"""
    t, _, _ = oracle.gauss(p)
    g, w, _ = oracle.gauss(qq)
    edges = np.linspace(-1, 1, nsub + 1)
    us = np.concatenate([0.5 * (edges[k] + edges[k + 1]) + 0.5 * (edges[1] - edges[0]) * g for k in range(nsub)])
    ws = np.tile(w * 0.5 * (edges[1] - edges[0]), nsub)

    def lag(u):
        L = np.ones((len(u), p)); D = np.zeros((len(u), p))
        for j in range(p):
            for k in range(p):
                if k != j:
                    L[:, j] *= (u - t[k]) / (t[j] - t[k])
            for r in range(p):
                if r == j:
                    continue
                term = np.full(len(u), 1.0 / (t[j] - t[r]))
                for k in range(p):
                    if k != j and k != r:
                        term *= (u - t[k]) / (t[j] - t[k])
                D[:, j] += term
        return L, D
    Lu, Du = lag(us)
    S = sx.reshape(p, p, 3)  # [b, a, c]
    y = np.einsum("ua,vb,bac->vuc", Lu, Lu, S)
    yu = np.einsum("ua,vb,bac->vuc", Du, Lu, S)
    yv = np.einsum("ua,vb,bac->vuc", Lu, Du, S)
    nn = np.cross(yu, yv)
    W2 = np.outer(ws, ws)  # [v, u]
    out = np.zeros((len(xs), p * p))
    for k, x in enumerate(xs):
        d = x - y
        r = np.sqrt((d ** 2).sum(-1))
        kd = (d * nn).sum(-1) / r ** 3 / (4 * np.pi) * W2
        out[k] = np.einsum("vu,ua,vb->ba", kd, Lu, Lu).ravel()
    return out


def test_reference_smoke_patch_vs_brute_force(oracle):
    p = 8
    sx = torus_patch(oracle, p)
    # the reference smoke test's targets (test/test_lap3d_close.f90:76-81) and two nearer ones
    xs = np.array([[0.05 * k, 0.03 * k - 0.05, 0.02 * k + 0.01] for k in range(1, 6)])
    n = np.array([0.0, 0.0, 1.0])
    xs = np.vstack([xs, sx[27] + 0.02 * n, sx[40] - 0.03 * n])
    Ac = oracle.closepanel_dlp(xs, sx, p)
    Ab = brute_dlp(sx, p, xs, oracle)
    for f in (np.ones(p * p), sx[:, 0], np.exp(sx[:, 1]) * np.cos(sx[:, 2])):
        e = np.abs(Ac @ f - Ab @ f) / (np.abs(Ab) @ np.abs(f))
        assert e.max() < 1e-10, e
