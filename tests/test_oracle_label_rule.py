"""The oracle's level label rule against an independent numpy statement of the
ball formula (src/patch_refine_mod.f90:471-498) and targets placed exactly on a
ball's boundary (strict < at :541-560).  A dropped 1/2 square-cell factor, dist
in place of dist^2, or <= in place of < each fails here."""
import numpy as np
import pytest

from qotential_amd import workload as W
from label_rule import balls_numpy, boundary_cases


def _levels(oracle, p, k):
    geo = W.geometry(16, 40, p, modes=W.VMEC_BENCH)
    h = geo["h"]
    sl = slice(k * h, (k + 1) * h)
    L = oracle.patch_levels(p, geo["sx"][sl], geo["rts"][sl], geo["rps"][sl])
    sub = {key: geo[key][sl] for key in ("sx", "rts", "rps", "sw")}
    return L, sub


@pytest.mark.parametrize("p,k", [(4, 77), (8, 300)])
def test_oracle_balls_match_numpy_formula(oracle, p, k):
    L, _ = _levels(oracle, p, k)
    off = np.concatenate([[0], np.cumsum(L["nl"])])
    soff = np.concatenate([[0], np.cumsum(L["nsub"])])
    nsq = 0
    for lev in range(1, 5):
        R = oracle.patch_ref_level(p, lev)
        a, b = off[lev - 1], off[lev]
        B = balls_numpy(L["sx"][a:b], L["sw"][a:b], R["suboff"], R["subtyp"])
        qp = L["qpoint"][soff[lev - 1]:soff[lev]]
        qr = L["qradii2"][soff[lev - 1]:soff[lev]]
        assert np.abs(B[:, :3] - qp).max() < 1e-14 * np.abs(qp).max()
        assert np.abs(B[:, 3] - qr).max() < 1e-13 * qr.max()
        nsq += int((R["subtyp"] == 1).sum())
    assert nsq > 0  # square cells are present, so their factor is exercised


@pytest.mark.parametrize("p,k", [(4, 77), (8, 300)])
def test_oracle_labels_on_ball_boundary(oracle, p, k):
    L, sub = _levels(oracle, p, k)
    cases = boundary_cases(L, oracle.patch_ref_level, p)
    assert len(cases) >= 4 and {c[1] for c in cases} == {0, 1}
    for lev, typ, tx in cases:
        lab = oracle.labels(p, sub, tx, np.array([0, 3]), np.array([1, 2, 3]))
        # on the boundary counts as outside (strict <); just inside is refined
        assert lab[0] == lev and lab[2] == lev and lab[1] > lev, (lev, typ, lab)
