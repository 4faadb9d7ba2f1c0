"""Mixed-order batches: every order's step launched from one call.

The batch step's potential equals the per-order evaluation bit for bit, on
every step, with other eager work and synchronisations between steps; after
the autotuner has picked each order's schedule and solve kernels, the
potential still agrees with the untuned one to the conditioning of the solve
variants.  Capturing the step into a HIP graph is switched off in the library
(its first run ended the test process with a host-side segmentation fault
whose cause is not yet found); the graph test is skipped for that reason."""
import time

import numpy as np
import pytest

from qotential_amd import mixed as M
from qotential_amd import workload as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.skip(reason="graph capture switched off: its first run ended the process with a host-side "
                         "segmentation fault whose cause is not yet found")
def test_graph_capture():
    pass


def test_batch_step_equals_eager_on_every_step():
    say("batch step vs per-order evaluation: 16x40 surface, orders 4, 6, 8, 5000 targets")
    s = M.surface(16, 40, (4, 6, 8), modes=W.VMEC_BENCH, seed=1, chart_order=8)
    g8 = W.geometry(16, 40, 8, modes=W.VMEC_BENCH)
    pos, q = W.point_charges(g8, seed=2)
    dens = lambda x: W.charge_field(x, pos, q)
    tx, dist, _ = M.targets(s, 5000, -8, -1, sides=0, seed=3)
    u_e, stats = M.layer_potential(s, tx, dens)
    b = M.MixedBatch(s, tx, dens)
    try:
        for rep in range(3):
            u_g = b.potential()
            assert np.array_equal(u_g, u_e), rep
            # eager work between replays: one order's engine evaluated directly
            e = b.engines[6]
            e.eval()
            e.sync()
        table = b.autotune(reps=2)
        say("autotune:", table)
        u_t = b.potential()
        assert np.abs(u_t - u_e).max() <= 1e-10 * np.abs(u_e).max()
        assert np.array_equal(b.potential(), u_t)
    finally:
        b.close()


def test_batch_mixed_full_size():
    """The mixed-order workload: 64x160 charts of the bench surface, orders 4,
    6, 8, 10, 12 per chart, 1,000,000 targets at h 10^U, U in [-8, -1], both
    sides; one batch call per step, autotuned; GRF as the per-order path."""
    s = M.surface(64, 160, (4, 6, 8, 10, 12), modes=W.VMEC_BENCH, seed=5)
    g = W.geometry(16, 40, 8, modes=W.VMEC_BENCH)
    pos, q = W.point_charges(g, seed=2)
    dens = lambda x: W.charge_field(x, pos, q)
    tx, dist, _ = M.targets(s, 1000000, -8, -1, sides=0, seed=7)
    say("full-size mixed batch: building engines")
    b = M.MixedBatch(s, tx, dens)
    try:
        u0 = b.potential()
        nnz = b.pairs()
        b.sync()
        t0 = time.perf_counter()
        for _ in range(5):
            b.step()
        b.sync()
        t_graph = (time.perf_counter() - t0) / 5
        say(f"{nnz} pairs; one batch call per step: {t_graph * 1e3:.1f} ms ({nnz / t_graph:.3e} pairs/s)")
        table = b.autotune(reps=2)
        say("autotune:", {k: v for k, v in table.items()})
        b.sync()
        t0 = time.perf_counter()
        for _ in range(5):
            b.step()
        b.sync()
        t_tuned = (time.perf_counter() - t0) / 5
        u1 = b.potential()
        uex, _ = W.charge_field(tx, pos, q)
        err = np.abs(u1 - np.where(dist < 0, uex, 0.0)).max() / np.abs(uex).max()
        say(f"autotuned: {t_tuned * 1e3:.1f} ms ({nnz / t_tuned:.3e} pairs/s); GRF {err:.3e}; "
            f"tuned vs default {np.abs(u1 - u0).max() / np.abs(u0).max():.1e}")
        assert err < 4.5e-2
        assert np.abs(u1 - u0).max() <= 1e-8 * np.abs(u0).max()
    finally:
        b.close()
