"""The "radius 2h" near list on the GPU (radius twice the panel's largest corner
distance about the mean of its nodes): the brute-force and uniform-grid kernels
give the oracle's CSR bit for bit, the list holds the reference ball's, and the
Green's identity with the wider list is no worse than with the ball rule's."""
import numpy as np
import pytest

from qotential_amd import workload as W
from qotential_amd.engine import Engine

pytestmark = pytest.mark.gpu


def test_radius_2h_near_list_vs_oracle_and_grf(oracle):
    print("radius-2h near list: VMEC surface 16x40, order 8, 20,000 targets on both sides", flush=True)
    geo = W.geometry(16, 40, 8, modes=W.VMEC_BENCH)
    tx, tri, dist = W.targets(geo, 20000, -6, -1, sides=0, seed=1)
    pos, q = W.point_charges(geo, nchg=8, offset=0.4, seed=2)
    ub, gu = W.charge_field(geo["sx"], pos, q)
    ubn = (geo["snx"] * gu).sum(1)
    uex, _ = W.charge_field(tx, pos, q)
    ref = np.where(dist < 0, uex, 0.0)
    grf = {}
    for f in (0.0, 2.0):
        e = Engine(geo)
        e.set_targets(tx)
        e.set_near_radius(f)
        nnz = e.nearlist()
        offs, idx = e.pairs()
        assert e.nearlist(grid=True) == nnz
        g_offs, g_idx = e.pairs()
        oracle.set_near_radius(f)
        try:
            o_offs, o_idx = oracle.nearlist(geo["h"], geo["sx"], geo["sw"], tx, nthreads=8)
        finally:
            oracle.set_near_radius(0.0)
        assert np.array_equal(offs, o_offs) and np.array_equal(idx, o_idx), f
        assert np.array_equal(g_offs, o_offs) and np.array_equal(g_idx, o_idx), f
        if f == 0.0:
            ball = (offs, idx)
        else:
            for k in range(len(offs) - 1):
                assert set(ball[1][ball[0][k]:ball[0][k + 1]]) <= set(idx[offs[k]:offs[k + 1]])
        far = e.far(ub, ubn)
        e.eval(exterior=True, side_tol=1e-10)
        u, _ = e.apply(ub, ubn, far)
        grf[f] = np.abs(u - ref).max() / np.abs(uex).max()
        print(f"  factor {f}: {nnz} pairs, labels {np.bincount(e.weights()['label'], minlength=6)[1:].tolist()}, "
              f"GRF {grf[f]:.3e}", flush=True)
        e.close()
    assert grf[2.0] <= 1.01 * grf[0.0]
