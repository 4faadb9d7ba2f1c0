"""The reference's other two level forms in the oracle (PR_SQUARE = .false. and
PR_ADAPTIVE, src/patch_refine_mod.f90:51-55, selected by the bench driver's
iadap = 3 and 1, test/stellarator_bench.f90:46-47), pinned without the
reference:

- triangle levels (:229-240): the level sizes n1, 4 n1, 16 n1 and 43 n1, each
  level integrating the polynomials its rule is exact for, and the library's
  host tables equal to the oracle's;
- the adaptive walk (:740-771): level-3 cells 4q..4q+3 lie inside level-2 cell
  q (the child index the walk uses), labels and tree cuts against a numpy
  statement of the walk, including targets exactly on a ball (the walk's leaf
  test is strict >, so such a target is opened; level 4's inside test is strict
  <, so there it is outside);
- the weights of a tree cut: the whole-level cuts equal direct evaluation on
  levels 2 and 3, and every cut integrates a smooth density like brute-force
  quadrature.
"""
import numpy as np
import pytest

from qotential_amd import engine as E
from qotential_amd import rrq as R
from qotential_amd import workload as W
from label_rule import boundary_targets, d2


@pytest.fixture
def triangle(oracle):
    oracle.set_level_mode(False, False)
    R.set_level_mode(False, False)
    yield oracle
    oracle.set_level_mode(False, True)
    R.set_level_mode(False, True)


@pytest.fixture
def adaptive(oracle):
    oracle.set_level_mode(True, False)
    yield oracle
    oracle.set_level_mode(False, True)


# the cuts' error on a smooth density, relative to the row's absolute sum:
# measured 1.0e-6 at p = 6 (the level rules' accuracy at the walk's distances,
# between the square levels' 3e-5 at p = 4 and 1e-9 at p = 8); bound 10x
ADAPT_FTOL = 1e-5


def _simplex_monomial(a, b):
    from math import factorial
    return factorial(a) * factorial(b) / factorial(a + b + 2)


@pytest.mark.parametrize("p", [4, 8])
def test_triangle_levels_sizes_exactness_and_library_tables(triangle, p):
    oracle = triangle
    nl, ns = oracle.patch_ref_info(p)
    n1 = (p + 2) * (p + 3) // 2
    assert nl.tolist() == [n1, 4 * n1, 16 * n1, 43 * n1] and ns.tolist() == [1, 4, 16, 43]
    info = E.order_info(p)
    assert info["nl"].tolist() == nl.tolist() and info["nsub"].tolist() == ns.tolist()
    for lev in range(1, 5):
        L = oracle.patch_ref_level(p, lev)
        assert (L["subtyp"] == 0).all() and np.array_equal(L["suboff"], n1 * np.arange(ns[lev - 1] + 1))
        u, v = L["uv"][:, 0], L["uv"][:, 1]
        # the VR rule of order p + 1 and its 4-way splits are exact to degree p + 1
        for a in range(p + 2):
            for b in range(p + 2 - a):
                assert abs((L["wt"] * u ** a * v ** b).sum() - _simplex_monomial(a, b)) < 1e-14
        G = E.order_level(p, lev)
        for key in ("uv", "wt", "interpmat", "suboff", "subtyp"):
            assert np.array_equal(G[key], L[key]), (lev, key)


def _affine_fit(src, dst):
    A = np.hstack([src, np.ones((len(src), 1))])
    M, *_ = np.linalg.lstsq(A, dst, rcond=None)
    return M


def test_walk_children_lie_in_their_parent(triangle):
    """cell s of level 3 is the image of a level-1 node set under an affine map;
    its parent in the walk, level-2 cell s // 4, contains it"""
    oracle = triangle
    p = 6
    n1 = (p + 2) * (p + 3) // 2
    L1, L2, L3 = (oracle.patch_ref_level(p, lev) for lev in (1, 2, 3))
    tri = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    for q in range(4):
        M = _affine_fit(L1["uv"], L2["uv"][q * n1:(q + 1) * n1])
        verts = np.hstack([tri, np.ones((3, 1))]) @ M
        T = np.array([verts[1] - verts[0], verts[2] - verts[0]]).T
        for s in range(16):
            lam = np.linalg.solve(T, (L3["uv"][s * n1:(s + 1) * n1] - verts[0]).T)
            inside = (lam > -1e-12).all() and (lam.sum(0) < 1 + 1e-12).all()
            assert inside == (s // 4 == q), (q, s)


def _walk_numpy(B, soff, x):
    """label, cut and leaves of target x from the balls B (centre, r^2)"""
    def outside_leaf(s):  # the walk's leaf test, d^2 > r^2
        return d2(B[s, :3], x) > B[s, 3]

    if outside_leaf(soff[0]):
        return 1, 0
    opened = 0
    for q in range(4):
        if outside_leaf(soff[1] + q):
            continue
        opened |= 1 << q
        for c in range(4):
            if not outside_leaf(soff[2] + 4 * q + c):
                for s in range(soff[3], soff[4]):  # one_level: inside is d^2 < r^2
                    if d2(B[s, :3], x) < B[s, 3]:
                        return 5, 0
                return 4, 0
    return 2, 1 + opened


@pytest.mark.parametrize("p,k", [(4, 77), (8, 300)])
def test_walk_labels_and_cuts_vs_numpy(adaptive, p, k):
    oracle = adaptive
    geo = W.geometry(16, 40, p, modes=W.VMEC_BENCH)
    h = geo["h"]
    sl = slice(k * h, (k + 1) * h)
    L = oracle.patch_levels(p, geo["sx"][sl], geo["rts"][sl], geo["rps"][sl])
    B = np.concatenate([L["qpoint"], L["qradii2"][:, None]], 1)
    soff = np.concatenate([[0], np.cumsum(L["nsub"])])
    # targets around the panel at many distances, and on balls of levels 1-4
    rng = np.random.default_rng(4)
    c = geo["sx"][sl].mean(0)
    size = np.sqrt(geo["sw"][sl].sum())
    tx = [c + size * 10 ** rng.uniform(-3, 0.5) * d / np.linalg.norm(d) for d in rng.normal(size=(400, 3))]
    exact = []
    for s in (0, soff[1] + 1, soff[2] + 6, soff[3] + 20):
        for _ in range(40):
            nrm = rng.normal(size=3)
            r = boundary_targets(B[s, :3], B[s, 3], nrm / np.linalg.norm(nrm))
            if r is not None:
                exact.append(r[0])
                tx.extend(r)
                break
    assert len(exact) == 4
    tx = np.array(tx)
    sub = {key: geo[key][sl] for key in ("sx", "rts", "rps", "sw")}
    lab, cut = oracle.labels(p, sub, tx, np.array([0, len(tx)]), np.arange(1, len(tx) + 1), cuts=True)
    for j, x in enumerate(tx):
        assert (lab[j], cut[j] if lab[j] == 2 else 0) == _walk_numpy(B, soff, x), j
    assert set(np.unique(lab)) >= {1, 2, 5}
    # exactly on the level-1 ball: opened (not a leaf); on a level-4 ball: outside it
    j1 = next(i for i, x in enumerate(tx) if np.array_equal(x, exact[0]))
    assert lab[j1] != 1
    assert d2(B[soff[3] + 20, :3], exact[3]) == B[soff[3] + 20, 3]


def _direct_level(oracle, L, ref, lev, x):
    off = np.concatenate([[0], np.cumsum(L["nl"])])
    a, b = off[lev - 1], off[lev]
    As, Ad = oracle.lap3dsdlpmat(x[None, :], L["sx"][a:b], L["snx"][a:b], L["sw"][a:b])
    return As @ ref["interpmat"], Ad @ ref["interpmat"]


def test_cut_weights_whole_levels_and_brute_force(adaptive):
    oracle = adaptive
    p = 6
    geo = W.geometry(16, 40, p, modes=W.VMEC_BENCH)
    h = geo["h"]
    k = 200
    sl = slice(k * h, (k + 1) * h)
    L = oracle.patch_levels(p, geo["sx"][sl], geo["rts"][sl], geo["rps"][sl])
    refs = {lev: oracle.patch_ref_level(p, lev) for lev in (2, 3)}
    rng = np.random.default_rng(9)
    c = geo["sx"][sl].mean(0)
    nrm = geo["snx"][sl].mean(0)
    nrm /= np.linalg.norm(nrm)
    size = np.sqrt(geo["sw"][sl].sum())
    base = geo["sx"][sl][rng.integers(0, h, 600)]
    tx = base + (size * 10 ** rng.uniform(-1.3, 0.2, 600) * rng.choice([-1, 1], 600))[:, None] * nrm
    sub = {key: geo[key][sl] for key in ("sx", "snx", "sw", "rts", "rps")}
    lab, cut = oracle.labels(p, sub, tx, np.array([0, len(tx)]), np.arange(1, len(tx) + 1), cuts=True)
    r = oracle.rrq(p, tx, sub["sx"], sub["snx"], sub["sw"], sub["rts"], sub["rps"], exterior=True)
    assert np.array_equal(r["label"], lab)
    whole = {1: 2, 16: 3}  # no level-2 cell opened: level 2; all opened: level 3
    seen = set()
    for j in np.where(lab == 2)[0]:
        if cut[j] in whole:
            lev = whole[cut[j]]
            s, d = _direct_level(oracle, L, refs[lev], lev, tx[j])
            scale = np.abs(s).max()
            assert np.abs(r["As"][j] - s[0]).max() < 1e-13 * scale and np.abs(r["Ad"][j] - d[0]).max() < 1e-13 * np.abs(d).max()
            seen.add(cut[j])
    assert seen == {1, 16}, seen
    # every cut against brute-force quadrature of the interpolated patch on a
    # smooth density (48^2 sub-triangles, VR order 20)
    tree = np.where(lab == 2)[0]
    xs = tx[tree]
    m = len(xs)
    Ab = np.zeros((m, h), order="F"); Db = np.zeros((m, h), order="F"); om = np.zeros(m)
    oracle.lib().orc_brute(oracle.i64(m), oracle._p(oracle._f(xs)), oracle.i64(p), oracle._p(oracle._f(sub["sx"])),
                           oracle._p(oracle._f(sub["snx"])), oracle.i64(48), oracle.i64(20), oracle._p(Ab),
                           oracle._p(Db), oracle._p(om))
    f = np.exp(sub["sx"][:, 0] / 2) * np.cos(sub["sx"][:, 1] / 2)
    worst = 0.0
    for i, j in enumerate(tree):
        for A, Bm in ((r["As"][j], Ab[i]), (r["Ad"][j], Db[i])):
            worst = max(worst, abs((A - Bm) @ f) / np.abs(Bm).sum())
    print(f"tree cuts {sorted(set(cut[tree].tolist()))}: smooth-density error vs brute force {worst:.2e}")
    assert worst < ADAPT_FTOL
    assert len(set(cut[tree])) >= 6
