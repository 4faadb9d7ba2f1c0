"""Host logic of the learned selector: labels from serving errors, class
weights, feature normalisation, the random-stellarator draw."""
import numpy as np

from qotential_amd import selector as S


def test_classes_from_errors():
    tol = 1e-11
    big, small = 1.0, 1e-13
    err = np.array([
        [small, small, small, small, small, 0.0],   # own nodes suffice
        [big, big, small, 0.0, 0.0, 0.0],           # level 2 is the coarsest within tol
        [big, big, big, big, 0.0, 0.0],             # level 4 (rule label 4)
        [big, big, big, big, small, 0.0],           # level 4 where the rule says RRQ
        [big, big, big, big, big, 0.0],             # RRQ
        [big, tol, 0.0, 0.0, 0.0, 0.0],             # tolerance is inclusive
    ])
    rule = np.array([3, 3, 4, 5, 5, 2])
    assert S.classes_from_errors(err, rule, tol).tolist() == [0, 2, 4, 6, 5, 1]


def test_serve_rank_orders_by_cost():
    assert S.serve_rank([0, 1, 2, 3, 4, 6, 5]).tolist() == [0, 1, 2, 3, 4, 4, 5]


def test_class_weights_balance():
    lab = np.array([0] * 900 + [3] * 100)
    w = S.class_weights(lab)
    assert w.shape == (S.N_CLASS,) and w[1] == 0.0
    assert abs(w[lab].mean() - 1.0) < 1e-6
    assert np.isclose(w[3] / w[0], 3.0)


def test_feature_norm_and_init():
    rng = np.random.default_rng(0)
    f = rng.standard_normal((5000, S.N_FEAT)) * 3 + 2
    f[:, 8] = 8 / 12  # constant feature: std floored
    n = S.feature_norm(f)
    z = (f - n[:S.N_FEAT]) * n[S.N_FEAT:]
    assert np.abs(z.mean(0)).max() < 1e-4
    assert np.abs(z[:, :8].std(0) - 1).max() < 1e-4 and np.all(np.isfinite(z))
    p = S.init_params(1)
    assert p.shape == (S.N_PARAM,) and p.dtype == np.float32 and not p[S.N_W:].any()


def test_random_stellarator_ranges():
    rng = np.random.default_rng(3)
    for _ in range(50):
        modes, nfp = S.random_stellarator(rng)
        assert nfp in (3, 4, 5) and modes.shape == (5, 4)
        assert modes[0].tolist() == [0, 0, 1.0, 0.0]
        assert 0.2 <= modes[1, 2] <= 0.35 and modes[1, 2] == modes[1, 3]
        assert np.abs(modes[2:, 2:]).max() <= 0.05
        assert len({(int(m), int(n)) for m, n in modes[:, :2]}) == 5
