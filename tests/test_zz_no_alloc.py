"""qol_rrq_noalloc_ (no device allocation per call, fixed device scratch)
against qol_rrq_mex_ on the same inputs: two calls in a row, the second with
fewer targets, then a call too large for the scratch, which is refused before
any device work with the outputs left untouched."""
import numpy as np
import pytest

from qotential_amd import rrq as R
from qotential_amd import workload as W
from qotential_amd._lib import QolError, lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _panel_and_targets(p, k, m, seed):
    geo = W.geometry(16, 40, p, modes=W.VMEC_BENCH)
    h = geo["h"]
    sl = slice(k * h, (k + 1) * h)
    sx, snx, sw, rts, rps = (geo[key][sl] for key in ("sx", "snx", "sw", "rts", "rps"))
    rng = np.random.default_rng(seed)
    c = sx.mean(0)
    n = snx.mean(0)
    n /= np.linalg.norm(n)
    hd = np.sqrt(sw.sum())
    base = sx[rng.integers(0, h, m)] + 0.3 * hd * rng.standard_normal((m, 3))
    tx = base + (hd * 10.0 ** rng.uniform(-6, 0.3, m) * rng.choice([-1, 1], m))[:, None] * n
    return (sx, snx, sw, rts, rps), tx


@pytest.mark.parametrize("p", [8, 6])
def test_noalloc_matches_rrq_mex(p):
    for call, (k, m, seed) in enumerate([(300, 3000, 1), (41, 257, 2)]):
        say(f"p={p} call {call + 1}: panel {k}, {m} targets")
        panel, tx = _panel_and_targets(p, k, m, seed)
        a = R.rrq(tx, *panel, p, exterior=True, noalloc=True)
        b = R.rrq(tx, *panel, p, exterior=True)
        for key in ("As", "Ad", "Omega", "Ias"):
            assert np.array_equal(a[key], b[key]), (p, call, key)
        assert a["timeinfo"][0] > 0.0


def test_noalloc_refuses_oversize_call():
    p = 8
    panel, tx = _panel_and_targets(p, 300, 40000, 3)
    with pytest.raises(QolError, match="device scratch"):
        R.rrq(tx, *panel, p, exterior=True, noalloc=True)
    # the library stays usable after the refusal
    panel, tx = _panel_and_targets(p, 300, 100, 4)
    a = R.rrq(tx, *panel, p, exterior=True, noalloc=True)
    b = R.rrq(tx, *panel, p, exterior=True)
    assert np.array_equal(a["As"], b["As"]) and np.array_equal(a["Ad"], b["Ad"])
