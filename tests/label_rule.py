"""Shared helpers for the label-rule tests: sub-patch balls restated in numpy
from a panel's level nodes, and targets placed exactly on a ball's boundary.

The rule (src/patch_refine_mod.f90:471-498 balls, :541-560 classification):
a sub-patch's ball has centre qpoint = sum(w x) / sum(w) and squared radius
qradii2 = dist^2 max(rmax^2, f sum(w)), with f = 1/2 for square (GL) cells and 1
for triangles, rmax the largest node distance from the centre and dist = 1.4;
a target is inside when |qpoint - x|^2 < qradii2 (strict) and is served at the
first level where it is inside no ball.
"""
import numpy as np

DIST2 = 1.4 * 1.4


def balls_numpy(x, w, suboff, subtyp):
    """(nsub, 4) array of centre and squared radius from nodes x (n, 3), weights w"""
    out = np.zeros((len(subtyp), 4))
    for q in range(len(subtyp)):
        a, b = int(suboff[q]), int(suboff[q + 1])
        sw = w[a:b].sum()
        qp = (x[a:b] * w[a:b, None]).sum(0) / sw
        rmax2 = ((x[a:b] - qp) ** 2).sum(1).max()
        f = 0.5 if subtyp[q] == 1 else 1.0
        out[q, :3] = qp
        out[q, 3] = DIST2 * max(rmax2, f * sw)
    return out


def d2(q, x):
    """the classification's squared distance, in its order of operations"""
    d0 = q[0] - x[0]
    d1 = q[1] - x[1]
    d2_ = q[2] - x[2]
    return d0 * d0 + d1 * d1 + d2_ * d2_


def boundary_targets(q, r2, n):
    """Targets along the unit direction n from the ball centre q: one whose
    squared distance (as the classification computes it) equals r2 exactly, and
    the nearest ones just inside and just outside.  None when no exact hit is
    found in the scan."""
    q = np.asarray(q, dtype=np.float64)
    s0 = np.sqrt(r2)
    ds = np.spacing(s0)
    exact = inside = outside = None
    for j in range(-4000, 4001):
        s = s0 + j * ds
        x = q + s * n
        v = d2(q, x)
        if v == r2 and exact is None:
            exact = x
        elif v < r2:
            inside = x
        elif v > r2 and outside is None:
            outside = x
        if exact is not None and inside is not None and outside is not None:
            break
    if exact is None:
        return None
    return exact, inside, outside


def boundary_cases(L, ref_level, p):
    """(level, cell type, [exact, inside, outside] targets) for one triangle and
    one square cell per level 2..4: exactly on that cell's ball, outside every
    other ball of the level and inside some ball of each coarser level."""
    off = np.concatenate([[0], np.cumsum(L["nl"])])
    soff = np.concatenate([[0], np.cumsum(L["nsub"])])
    B = np.concatenate([L["qpoint"], L["qradii2"][:, None]], 1)
    centre = L["sx"][: off[1]].mean(0)
    cases = []
    for lev in range(2, 5):
        R = ref_level(p, lev)
        nrm = L["snx"][off[lev - 1]:off[lev]]
        Bl = B[soff[lev - 1]:soff[lev]]
        for typ in (0, 1):
            found = False
            for q in np.where(R["subtyp"] == typ)[0]:
                a, b = R["suboff"][q], R["suboff"][q + 1]
                n = nrm[a:b].mean(0)
                n /= np.linalg.norm(n)
                t = Bl[q, :3] - centre
                t -= (t @ n) * n
                t /= max(np.linalg.norm(t), 1e-300)
                dirs = [sg * np.cos(al) * n + np.sin(al) * t for al in (0.0, 0.4, 0.8, 1.2) for sg in (1.0, -1.0)]
                for dv in dirs:
                    r = boundary_targets(Bl[q, :3], Bl[q, 3], dv / np.linalg.norm(dv))
                    if r is None:
                        continue
                    ex = r[0]
                    if min(d2(Bl[o, :3], ex) / Bl[o, 3] for o in range(len(Bl)) if o != q) <= 1 + 1e-6:
                        continue
                    if not all(any(d2(B[s, :3], ex) < B[s, 3] for s in range(soff[l], soff[l + 1]))
                               for l in range(lev - 1)):
                        continue
                    cases.append((lev, typ, np.array(r)))
                    found = True
                    break
                if found:
                    break
    return cases
