"""The reference's W7-X GRF case end to end on the GPU: the W7-X surface (288 modes,
nfp 5) on 12x36 charts at order 6 with curvature refinement (restol 0.1: 5,500
triangles, 115,500 nodes), charts and nodes built on the GPU, every node a
target (exterior limit), density exp(x/2) cos(y/2) (stellarator_grf_core_mod.f90:
909-928, 1089-1103).  The reference's fixture (web/stellarator/fixtures/native/
w7x-order6-direct.json) reports GRF max rel err 2.3159e-3 for this case; its
binary arrays are not in the reference, so the comparison is of that number.
Our SSQ and solid-angle quadratures are re-derived (the reference's helpers are
in absent submodules), so the two errors are not equal: measured here 5.0e-4 with
the exact chart boundary the reference's GRF loop uses (ichart = 1), and 8.4e-3
with the polynomial boundary (ichart = 0).  The test holds the error at or below
the reference's."""
import os

import numpy as np
import pytest

from qotential_amd import workload as W
from qotential_amd.engine import Engine

pytestmark = pytest.mark.gpu


def test_grf_w7x_order6():
    print("W7-X 12x36 order 6 restol 0.1, GRF with the surface nodes as targets", flush=True)
    modes = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "w7x_modes.txt"))
    geo = W.geometry(12, 36, 6, modes=modes, nfp=5, restol=0.1, gpu=True)
    assert geo["ntri"] == 5500 and geo["sx"].shape[0] == 115500
    # the reference's GRF loop passes the exact chart boundary (ichart = 1 by default,
    # stellarator_grf_core_mod.f90:78, 714-737)
    e = Engine(geo, chart=W.chart_boundary(geo))
    e.set_targets(geo["sx"])
    nnz = e.nearlist()
    ub, gu = W.harmonic_exp(geo["sx"])
    ubn = (geo["snx"] * gu).sum(1)
    e.eval(exterior=True, side_tol=1e300)
    far = e.far(ub, ubn)
    u, val = e.apply(ub, ubn, far)
    lab = np.bincount(e.weights()["label"], minlength=6)[1:]
    e.close()
    err = float(np.abs(u).max() / np.abs(ubn).max())
    print(f"  {nnz} pairs, labels {lab.tolist()}, GRF max rel err {err:.4e} (reference 2.3159e-3)", flush=True)
    assert np.all(np.isfinite(u))
    assert err <= 2.3159e-3
