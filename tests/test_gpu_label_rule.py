"""The GPU's level label rule against an independent numpy statement of the
ball formula (src/patch_refine_mod.f90:471-498), computed from the GPU's own
level nodes, and targets placed exactly on a ball's boundary (strict < at
:541-560), for triangle and square cells at orders 4 and 8."""
import numpy as np
import pytest

from qotential_amd import workload as W
from qotential_amd.engine import Engine, order_info, order_level
from label_rule import balls_numpy, boundary_cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("p,k", [(4, 77), (8, 300)])
def test_gpu_balls_and_boundary_labels(p, k):
    say(f"label rule p={p} panel {k}")
    geo = W.geometry(16, 40, p, modes=W.VMEC_BENCH)
    h = geo["h"]
    sl = slice(k * h, (k + 1) * h)
    one = {key: geo[key][sl] for key in ("sx", "snx", "sw", "rts", "rps")}
    one["p"] = p
    e = Engine(one)
    lg, bl = e.level_data()
    info = order_info(p)
    off = np.concatenate([[0], np.cumsum(info["nl"])])
    soff = np.concatenate([[0], np.cumsum(info["nsub"])])
    x, nrm, w = lg[0, 0:3].T, lg[0, 3:6].T, lg[0, 6]
    for lev in range(1, 5):
        R = order_level(p, lev)
        a, b = off[lev - 1], off[lev]
        B = balls_numpy(x[a:b], w[a:b], R["suboff"], R["subtyp"])
        G = bl[0, soff[lev - 1]:soff[lev]]
        assert np.abs(B[:, :3] - G[:, :3]).max() < 1e-14 * np.abs(G[:, :3]).max(), lev
        assert np.abs(B[:, 3] - G[:, 3]).max() < 1e-13 * G[:, 3].max(), lev
    L = dict(sx=x, snx=nrm, sw=w, qpoint=bl[0, :, :3], qradii2=bl[0, :, 3], nl=info["nl"], nsub=info["nsub"])
    cases = boundary_cases(L, order_level, p)
    assert len(cases) >= 4 and {c[1] for c in cases} == {0, 1}
    tx = np.concatenate([c[2] for c in cases])
    e.set_targets(tx)
    e.set_pairs(np.array([0, len(tx)]), np.arange(1, len(tx) + 1))
    e.eval(exterior=True, side_tol=1e-10)
    lab = e.weights()["label"].reshape(-1, 3)
    e.close()
    for (lev, typ, _), lb in zip(cases, lab):
        say("level", lev, "cell type", typ, "labels (on, inside, outside)", lb.tolist())
        assert lb[0] == lev and lb[2] == lev and lb[1] > lev, (lev, typ, lb)
