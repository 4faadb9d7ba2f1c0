"""Points on and around one curved patch, from its node values: the patch map and
its tangents at any (u, v) of the reference triangle (the interpolant through the
Vioreanu-Rokhlin nodes in the oracle's Koornwinder basis), and targets placed
beyond its vertices and edges."""
import numpy as np

# (u, v) on the boundary of the reference triangle and the outward direction there
# in (u, v): the three vertices (outward along the bisector), edge midpoints and
# quarter points of each edge
S2 = 1 / np.sqrt(2)
BOUNDARY = [(0.0, 0.0, -S2, -S2), (1.0, 0.0, S2, -S2), (0.0, 1.0, -S2, S2),
            (0.5, 0.0, 0.0, -1.0), (0.25, 0.0, 0.0, -1.0), (0.0, 0.5, -1.0, 0.0), (0.0, 0.75, -1.0, 0.0),
            (0.5, 0.5, S2, S2), (0.75, 0.25, S2, S2)]
# offsets beyond the boundary (tangent plane) and along the normal, in units of the
# patch size sqrt(area)
ALONG = (0.0, 1e-8, 1e-4, 1e-2, 0.1, 0.3)
NORMAL = (1e-9, -1e-9, 1e-5, -1e-5, 1e-2, -1e-2, 0.1, -0.1)


def patch_eval(oracle, p, sx, u, v):
    """point and tangents of the order-p patch with node values sx (h x 3) at (u, v)"""
    uv, _ = oracle.vr_nodes(p - 1)
    V = np.array([oracle.koorn_pols(a, b, p - 1)[0] for a, b in uv])
    coef = np.linalg.solve(V, sx)
    P, Du, Dv = oracle.koorn_pols(u, v, p - 1)
    return P @ coef, Du @ coef, Dv @ coef


def boundary_targets(oracle, p, sx, size):
    """targets y + size (a t + b n) for every boundary point y of BOUNDARY, t the
    outward direction in the tangent plane, n the unit normal, a in ALONG and b in
    NORMAL; returns the targets and their (a, b)"""
    pts, ab = [], []
    for u, v, du, dv in BOUNDARY:
        y, yu, yv = patch_eval(oracle, p, sx, u, v)
        n = np.cross(yu, yv)
        n /= np.linalg.norm(n)
        t = du * yu + dv * yv
        t -= (t @ n) * n
        t /= np.linalg.norm(t)
        for a in ALONG:
            for b in NORMAL:
                pts.append(y + size * (a * t + b * n))
                ab.append((a, b))
    return np.array(pts), np.array(ab)


# action on a smooth density, plain build against the build with fused
# multiply-adds: at most 10 eps / distance near the boundary and 4e-13, 3.1e-12,
# 9e-10 elsewhere at p = 4, 8, 12 (a third rounding, clang with fused
# multiply-adds, at most 43 eps / distance); bounds about 10x
SPREAD_K = 500
SPREAD_BASE = {4: 1e-11, 8: 3e-11, 12: 1e-8}


def action_bound(p, ab, base=None):
    """bound on the density action's rounding spread for targets placed at (a, b):
    SPREAD_K eps / distance, at least the per-order base"""
    dist = np.maximum(np.hypot(ab[:, 0], ab[:, 1]), 1e-300)
    return np.maximum(SPREAD_BASE[p] if base is None else base, SPREAD_K * np.finfo(float).eps / dist)
