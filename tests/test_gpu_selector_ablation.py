"""What the trained selector (qotential_amd/selector_model.npz, trained by
qotential_amd/train_selector.py on 1.99M pairs of 32 random stellarators with
this library's HIP kernels) gains over the method without it, on a stellarator
it has not seen: the ball rule alone, the ball rule one level finer everywhere,
and the selector with the ball rule below confidence 0.9, evaluated on
1,000,000 targets at h 10^U, U ~ U[-6,-1], both sides.  Measured: the
selector serves about 10% of the pairs with the panel's own nodes and cuts the
direct level FLOPs by 8% at the same Green's-identity error; its per-pair
inference costs more wall time than that saves."""
import numpy as np
import pytest

from qotential_amd import selector as S
from qotential_amd import workload as W

pytestmark = pytest.mark.gpu


def test_selector_ablation_unseen_surface():
    modes, nfp = S.random_stellarator(np.random.default_rng(12345))
    geo = W.geometry(16, 16 * nfp, 8, modes=modes, nfp=nfp)
    sel = S.Selector.load()
    res = S.ablation(sel, geo, nt=1_000_000)
    sel.close()
    for k, v in res.items():
        print(k, v, flush=True)
    rule, finer, learned = res["rule"], res["rule+1"], res["selector"]
    # the same accuracy as the method without it ...
    assert abs(learned["grf_max"] - rule["grf_max"]) <= 1e-9
    assert learned["grf_p999"] <= rule["grf_p999"] * (1 + 1e-6) + 1e-12
    # ... with fewer direct FLOPs: pairs moved to their own nodes or coarser levels
    assert learned["labels"][0] > 0.05 * sum(learned["labels"])
    assert learned["direct_gflop"] < 0.95 * rule["direct_gflop"]
    # uniform refinement costs far more for the same error on this surface
    assert finer["direct_gflop"] > 2 * rule["direct_gflop"]
