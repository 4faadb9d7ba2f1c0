"""Oracle per-order tables against known answers and invariants."""
import numpy as np


def test_gauss_legendre_known_answer(oracle):
    # GL-8 nodes/weights hard-coded in the reference's test/test_lap3d_close.f90:35-42
    t, w, D = oracle.gauss(8)
    ref_t = [-0.9602898564975363, -0.7966664774136267, -0.5255324099163290, -0.1834346424956498]
    ref_w = [0.1012285362903763, 0.2223810344533745, 0.3137066458778873, 0.3626837833783620]
    assert np.max(np.abs(t[:4] - ref_t)) < 2e-16 and np.max(np.abs(t[4:] + np.array(ref_t)[::-1])) < 2e-16
    assert np.max(np.abs(w[:4] - ref_w)) < 2e-16
    assert abs(w.sum() - 2.0) < 1e-15
    # differentiation matrix is exact on polynomials of degree < 8
    f = t ** 7 - 3 * t ** 2
    assert np.max(np.abs(D @ f - (7 * t ** 6 - 6 * t))) < 1e-12


def test_vr_rules_exact(oracle):
    for order in (3, 7, 9, 11, 20):
        uv, w = oracle.vr_nodes(order)
        assert abs(w.sum() - 0.5) < 1e-14
        # int_T u^a v^b = a! b! / (a+b+2)!
        from math import factorial
        for a, b in [(order, 0), (order // 2, order - order // 2), (1, order - 1)]:
            ex = factorial(a) * factorial(b) / factorial(a + b + 2)
            assert abs((w * uv[:, 0] ** a * uv[:, 1] ** b).sum() - ex) < 1e-14


def test_level_counts_per_order(oracle):
    # node and sub-patch counts per level (patch_refine_mod.f90:100-421 restated)
    expect = {4: [21, 78, 300, 825], 6: [36, 136, 528, 1444], 8: [55, 174, 604, 1767],
              10: [78, 256, 912, 2626], 12: [105, 354, 1284, 3657]}
    for p, nl in expect.items():
        n, ns = oracle.patch_ref_info(p)
        assert list(n) == nl
        assert list(ns) == [1, 3, 10, 30]


def test_interpolation_reproduces_polynomials(oracle):
    p = 8
    uvn, _ = oracle.vr_nodes(p - 1)
    f = lambda u, v: (1 + u - 2 * v) ** 7 + u ** 3 * v ** 4 - 0.3 * v ** 6
    for lev in (1, 2, 3, 4):
        L = oracle.patch_ref_level(p, lev)
        err = np.max(np.abs(L["interpmat"] @ f(uvn[:, 0], uvn[:, 1]) - f(L["uv"][:, 0], L["uv"][:, 1])))
        assert err < 1e-12
        # level weights integrate the reference triangle exactly
        assert abs(L["wt"].sum() - 0.5) < 1e-14
        assert L["suboff"][-1] == L["uv"].shape[0]


def test_product_tables_bitwise_equal_oracle(oracle):
    """The library's per-order level tables equal the oracle's bit for bit, so the
    device's level balls (and hence the level labels) can be compared exactly."""
    from qotential_amd._lib import lib, i64, ptr
    for p in (4, 8, 12):
        info = np.zeros(13, dtype=np.int64)
        assert lib().qol_order_info(i64(p), ptr(info)) == 0
        h = int(info[0])
        for lev in (1, 2, 3, 4):
            n, ns = int(info[4 + lev]), int(info[8 + lev])
            uv = np.zeros((n, 2)); wt = np.zeros(n); im = np.zeros((n, h), order="F")
            so = np.zeros(ns + 1, dtype=np.int64); st = np.zeros(ns, dtype=np.int64)
            assert lib().qol_order_level(i64(p), i64(lev), ptr(uv), ptr(wt), ptr(im), ptr(so), ptr(st)) == 0
            L = oracle.patch_ref_level(p, lev)
            assert np.array_equal(uv, L["uv"]) and np.array_equal(wt, L["wt"])
            assert np.array_equal(so, L["suboff"]) and np.array_equal(st, L["subtyp"])
            assert np.array_equal(im, L["interpmat"]), (p, lev, np.abs(im - L["interpmat"]).max())


def test_level_rules_exact_on_monomials(oracle):
    """Every refinement level is a composite rule over sub-patches that tile the
    reference triangle (VR rules of order orderff-1 on sub-triangles, Gauss
    squares on merged pairs), so each integrates u^a v^b exactly up to degree
    p+1.  A missing, doubled or misplaced sub-patch breaks this."""
    from math import factorial
    for p in (4, 8, 12):
        for lev in (1, 2, 3, 4):
            L = oracle.patch_ref_level(p, lev)
            u, v, w = L["uv"][:, 0], L["uv"][:, 1], L["wt"]
            for a in range(p + 2):
                for b in range(p + 2 - a):
                    ex = factorial(a) * factorial(b) / factorial(a + b + 2)
                    assert abs((w * u ** a * v ** b).sum() - ex) < 1e-14, (p, lev, a, b)
            # nodes inside the triangle
            assert (u >= -1e-15).all() and (v >= -1e-15).all() and (u + v <= 1 + 1e-15).all()
