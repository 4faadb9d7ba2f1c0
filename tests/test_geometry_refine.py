"""Adaptive chart list (qol_geo_charts: aspect splits, curvature refinement by
the mean curvature's Legendre tail, 2:1 balance; utils/stellarator_geo_mex.c:
359-411, 427-546) against the reference's W7-X fixture count, the circular
torus's closed-form mean curvature, and the properties the refinement promises."""
import os

import numpy as np
import pytest
from numpy.polynomial import legendre

from qotential_amd import workload as W

HERE = os.path.dirname(__file__)


@pytest.fixture(scope="module")
def w7x():
    return np.loadtxt(os.path.join(HERE, "golden", "w7x_modes.txt"))


@pytest.fixture(scope="module")
def w7x_charts(w7x):
    return W.charts(12, 36, 6, modes=w7x, nfp=5, restol=0.1)


def test_w7x_fixture_triangle_count(w7x_charts):
    # web/stellarator/fixtures/native/w7x-order6-direct.json: mp 12, np 36,
    # order 6, restol 0.1 -> ntri 5500, nsrc 115500
    assert 2 * len(w7x_charts) == 5500
    assert 2 * len(w7x_charts) * 21 == 115500


def _boxes(cl, mp, np_):
    ts1, ts2 = np.pi / np_, np.pi / mp
    a1, b1, a2, b2, jj, ii = cl.T
    return np.stack([((b1 - a1) + 2 * jj) * ts1, ((b1 + a1) + 2 * jj) * ts1,
                     ((b2 - a2) + 2 * ii) * ts2, ((b2 + a2) + 2 * ii) * ts2], 1)


def test_refined_list_tiles_the_surface(w7x_charts):
    cl = w7x_charts
    # every base cell is covered once: the parameter areas sum to mp * np ...
    assert abs((cl[:, 0] * cl[:, 2]).sum() - 12 * 36) < 1e-9
    # ... and no two charts overlap
    bx = _boxes(cl, 12, 36)
    ov1 = np.minimum(bx[:, None, 1], bx[None, :, 1]) - np.maximum(bx[:, None, 0], bx[None, :, 0])
    ov2 = np.minimum(bx[:, None, 3], bx[None, :, 3]) - np.maximum(bx[:, None, 2], bx[None, :, 2])
    both = (ov1 > 1e-12) & (ov2 > 1e-12)
    np.fill_diagonal(both, False)
    assert not both.any()


def test_refined_list_is_2to1_balanced(w7x_charts):
    bx = _boxes(w7x_charts, 12, 36)
    e1, e2 = 1e-9 * np.pi / 36, 1e-9 * np.pi / 12
    # faces at constant t1 shared over a positive t2 length: t2 extents within 2x
    f1 = (np.abs(bx[:, None, 1] - bx[None, :, 0]) < e1) | (np.abs(bx[:, None, 0] - bx[None, :, 1]) < e1)
    f1 &= np.minimum(bx[:, None, 3], bx[None, :, 3]) - np.maximum(bx[:, None, 2], bx[None, :, 2]) > e2
    ext2 = bx[:, 3] - bx[:, 2]
    assert f1.sum() > 0
    assert not (f1 & (ext2[:, None] > 2 * ext2[None, :] + e2)).any()
    f2 = (np.abs(bx[:, None, 3] - bx[None, :, 2]) < e2) | (np.abs(bx[:, None, 2] - bx[None, :, 3]) < e2)
    f2 &= np.minimum(bx[:, None, 1], bx[None, :, 1]) - np.maximum(bx[:, None, 0], bx[None, :, 0]) > e1
    ext1 = bx[:, 1] - bx[:, 0]
    assert f2.sum() > 0
    assert not (f2 & (ext1[:, None] > 2 * ext1[None, :] + e1)).any()


def test_restol_zero_and_cap(w7x):
    cl0 = W.charts(12, 36, 6, modes=w7x, nfp=5, restol=0.0)
    g = W.geometry(12, 36, 6, modes=w7x, nfp=5)  # iterated aspect splits
    assert np.array_equal(cl0, g["charts"])
    with pytest.raises(RuntimeError, match="cap"):
        W.charts(12, 36, 6, modes=w7x, nfp=5, restol=0.1, cap=2000)


@pytest.mark.parametrize("p,mp", [(14, 2), (14, 3), (14, 4)])
def test_curvature_tail_circular_torus(p, mp):
    """Torus R = R0 + r cos th, Z = r sin th: H(th) = (R0 + 2 r cos th) / (2 r (R0 + r cos th))
    depends on the poloidal angle only, so the tensor spectrum is the 1-D Legendre
    spectrum of H over the chart's poloidal range (degree 0 in the toroidal one).
    The library differentiates the chart map spectrally on its p x p grid; at
    these orders and chart sizes (a half to a quarter of the poloidal circle) that
    is resolved, and the tails agree to 1% (at p = 6 and 8 they differ by up to
    several times: there the tail measures the grid's own error)."""
    R0, r = 1.0, 0.3
    modes = np.array([[0, 0, R0, 0.0], [1, 0, r, r]])
    x0, _ = legendre.leggauss(p)
    V = legendre.legvander(x0, p - 1)
    ts2 = np.pi / mp
    for chart in ([1.0, 0.0, 1.0, 0.0, 3.0, 1.0], [1.0, 0.0, 1.0, 0.0, 1.0, 2.0], [0.5, 0.5, 1.0, 0.0, 7.0, 1.0]):
        th = ((chart[2] * x0 + chart[3]) + 2 * chart[5]) * ts2
        H = (R0 + 2 * r * np.cos(th)) / (2 * r * (R0 + r * np.cos(th)))
        c = np.linalg.solve(V, H)
        want = np.sqrt((c[p - 2:] ** 2).sum() / (c ** 2).sum())
        got = W.curvature_tail(mp, 10, p, np.array(chart), modes=modes, nfp=5)
        assert want > 1e-10
        assert abs(got - want) < 0.01 * want, (chart, got, want)


def test_chart_boundary_of_a_chart_list(w7x):
    """qol_geo_chart_boundary_charts on the aspect-only list equals
    qol_geo_chart_boundary bit for bit; on the refined list every edge node is
    finite and each triangle's vertices are distinct and within its edges' span."""
    g_it = W.geometry(12, 36, 6, modes=w7x, nfp=5)
    g_cl = W.geometry(12, 36, 6, modes=w7x, nfp=5, restol=0.0)
    assert np.array_equal(g_it["charts"], g_cl["charts"])
    b_it, v_it = W.chart_boundary(g_it)
    b_cl, v_cl = W.chart_boundary(g_cl)
    assert np.array_equal(b_it, b_cl) and np.array_equal(v_it, v_cl)
    g_rf = W.geometry(12, 36, 6, modes=w7x, nfp=5, restol=0.1)
    b_rf, v_rf = W.chart_boundary(g_rf)
    assert b_rf.shape == (5500, 3 * 8, 3) and np.all(np.isfinite(b_rf))
    # each triangle's three vertices are distinct and lie within its edge nodes' span
    span = np.ptp(b_rf, axis=1).max(axis=1)
    d = np.linalg.norm(v_rf[:, 0] - v_rf[:, 1], axis=1)
    assert np.all(d > 0) and np.all(d < 2 * span)
