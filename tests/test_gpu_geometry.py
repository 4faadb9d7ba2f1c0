"""The GPU geometry generator (qol_geo_build_gpu) against the host generator:
same charts and triangulation; nodes, tangents, normals and weights equal up to
the rounding of the device sin/cos (every sum runs in the host's order)."""
import time

import numpy as np
import pytest

from qotential_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mp,np_,p,vmec", [(12, 36, 6, False), (16, 40, 4, True), (48, 120, 8, True)])
def test_geometry_gpu_vs_host(mp, np_, p, vmec):
    print(f"geometry {mp}x{np_} p={p} {'VMEC' if vmec else 'built-in'} surface", flush=True)
    modes = W.VMEC_BENCH if vmec else None
    t0 = time.perf_counter()
    host = W.geometry(mp, np_, p, modes=modes)
    t1 = time.perf_counter()
    gpu = W.geometry(mp, np_, p, modes=modes, gpu=True)
    t2 = time.perf_counter()
    print(f"  {host['ntri']} triangles: host {t1 - t0:.3f} s, GPU call {t2 - t1:.3f} s", flush=True)
    assert gpu["ntri"] == host["ntri"]
    assert np.array_equal(gpu["charts"], host["charts"])
    # node positions: a few ulp of the surface scale
    assert np.abs(gpu["sx"] - host["sx"]).max() < 1e-14 * np.abs(host["sx"]).max()
    # tangents: the spectral derivative scales the rounding by the grid's
    # differentiation matrix (entries up to about p^2)
    for key in ("rts", "rps"):
        assert np.abs(gpu[key] - host[key]).max() < 1e-12 * np.abs(host[key]).max(), key
    assert np.abs(gpu["snx"] - host["snx"]).max() < 1e-12
    assert np.abs(gpu["sw"] / host["sw"] - 1.0).max() < 1e-12
    # triangle counts of the benchmark surfaces (aspect splits of utils/stellarator_geo_mex.c:427-479)
    if vmec:
        assert host["ntri"] == {(16, 40): 1580, (48, 120): 14250}[(mp, np_)]


def test_geometry_gpu_vs_host_w7x_refined():
    """W7-X (288 modes, nfp 5), 12x36 charts at order 6 with curvature refinement
    (restol 0.1): the refined chart list built on the GPU against the host."""
    print("geometry W7-X 12x36 p=6 restol 0.1", flush=True)
    import os
    modes = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "w7x_modes.txt"))
    host = W.geometry(12, 36, 6, modes=modes, restol=0.1)
    gpu = W.geometry(12, 36, 6, modes=modes, restol=0.1, gpu=True)
    assert host["ntri"] == gpu["ntri"] == 5500
    assert np.array_equal(gpu["charts"], host["charts"])
    assert np.abs(gpu["sx"] - host["sx"]).max() < 1e-14 * np.abs(host["sx"]).max()
    for key in ("rts", "rps"):
        assert np.abs(gpu[key] - host[key]).max() < 1e-12 * np.abs(host[key]).max(), key
    assert np.abs(gpu["snx"] - host["snx"]).max() < 1e-12
    assert np.abs(gpu["sw"] / host["sw"] - 1.0).max() < 1e-12


def test_curvature_tails_gpu_vs_host():
    """panel_resolution on the GPU (one wave per chart) against the host for 300
    W7-X charts, and the refined chart list built with it."""
    print("curvature tails on the GPU: W7-X 12x36 p=6", flush=True)
    import os
    modes = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "w7x_modes.txt"))
    cl = W.charts(12, 36, 6, modes=modes, nfp=5, restol=0.1)
    sel = cl[np.random.default_rng(1).choice(len(cl), 300, replace=False)]
    tg = W.curvature_tails_gpu(12, 36, 6, sel, modes=modes, nfp=5)
    th = np.array([W.curvature_tail(12, 36, 6, c, modes=modes, nfp=5) for c in sel])
    # the device sin/cos rounding passes through second spectral derivatives:
    # measured up to 1.7e-12 on tails of 1e-3 to 1e-1
    assert np.all(np.abs(tg - th) <= 1e-8 * th + 1e-13), np.abs(tg - th).max()
    clg = W.charts(12, 36, 6, modes=modes, nfp=5, restol=0.1, gpu=True)
    assert np.array_equal(clg, cl) and 2 * len(clg) == 5500
