"""Workload generator against the reference's native fixtures."""
import numpy as np
import pytest

from qotential_amd import workload as W


@pytest.mark.parametrize("order", [4, 6])
def test_builtin_surface_matches_fixture(order, fixture4, fixture6):
    fx = fixture4 if order == 4 else fixture6
    g = W.geometry(int(fx["mp"]), int(fx["np_"]), order)
    assert g["ntri"] == int(fx["ntri"]) and g["sx"].shape[0] == int(fx["nsrc"])
    sel = fx["sel"]
    # the reference's cross-toolchain gate: sx, snx, sw within 1e-12
    assert np.max(np.abs(g["sx"][sel] - fx["sx"])) < 1e-12
    assert np.max(np.abs(g["snx"][sel] - fx["snx"])) < 1e-12
    assert np.max(np.abs(g["sw"][sel] - fx["sw"])) < 1e-12
    ub, gu = W.harmonic_exp(g["sx"][sel])
    assert np.max(np.abs(ub - fx["ub"])) < 1e-11
    assert np.max(np.abs((g["snx"][sel] * gu).sum(1) - fx["ubn"])) < 1e-11


def test_vmec_chart_counts():
    # the benchmark surface: 16x40 -> 1580 triangles at p=4, 48x120 -> 14250 at p=8
    assert W.geometry(16, 40, 4, modes=W.VMEC_BENCH)["ntri"] == 1580
    from qotential_amd._lib import lib, i64, ptr
    mn = np.ascontiguousarray(W.VMEC_BENCH)
    assert lib().qol_geo_count(i64(48), i64(120), i64(8), i64(5), i64(3), ptr(mn), i64(1)) == 14250


def test_targets_offsets():
    g = W.geometry(16, 40, 4, modes=W.VMEC_BENCH)
    tx, tri, dist = W.targets(g, 200, -6, -1, sides=-1, seed=1)
    assert np.all(dist < 0) and tx.shape == (200, 3)
    h = g["h"]
    # the offset is along the panel normal: |x - patch point| == |dist|
    for j in range(5):
        k = tri[j]
        assert np.linalg.norm(tx[j] - g["sx"][k * h:(k + 1) * h].mean(0)) < 0.5
