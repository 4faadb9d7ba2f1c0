"""The Fortran boundary: the reference's benchmark driver test/stellarator_bench.f90
compiles unchanged against this library's Fortran modules
(qotential_amd/fortran/qotential_modules.f90: quatapproximation_mod,
stellarator_mesh_mod, lap3d_close_mod, patch_refine_mod) and links to
libqotential_hip.so; on a GPU the driver built by oracle/fortran_bench.mk runs
its first case through the drop-in with each of its three level forms."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF_BENCH = Path("/root/reference/test/stellarator_bench.f90")
FLANG = "/opt/rocm/bin/amdflang"
BUILT = ROOT / "oracle" / "_ref" / "stellarator_bench"


@pytest.mark.skipif(not REF_BENCH.exists() or not Path(FLANG).exists(),
                    reason="needs the reference checkout and amdflang")
def test_reference_bench_compiles_unchanged(tmp_path):
    mods = ROOT / "qotential_amd" / "fortran" / "qotential_modules.f90"
    subprocess.run([FLANG, "-O2", "-module-dir", str(tmp_path), "-c", str(mods), "-o", str(tmp_path / "m.o")],
                   check=True)
    subprocess.run([FLANG, "-O2", "-I", str(tmp_path), str(REF_BENCH), str(tmp_path / "m.o"),
                    "-L", str(ROOT / "qotential_amd"), "-lqotential_hip", "-o", str(tmp_path / "bench")], check=True)
    assert (tmp_path / "bench").exists()


@pytest.mark.gpu
@pytest.mark.skipif(not BUILT.exists(), reason="oracle/_ref/stellarator_bench not built (no reference checkout)")
@pytest.mark.parametrize("iadap", [0, 3, 1])
def test_reference_bench_runs_on_the_gpu(iadap):
    """case 1 (order 4, 24x72 charts) with the default levels (iadap = 0),
    triangle levels (3) and the adaptive walk (1): the driver sets PR_SQUARE and
    PR_ADAPTIVE from iadap (test/stellarator_bench.f90:46-47)"""
    print(f"running the reference's stellarator_bench (case 1: order 4, 24x72, iadap {iadap}) through the drop-in",
          flush=True)
    r = subprocess.run([str(BUILT), "1", "0", str(iadap)], capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    out = r.stdout
    assert "nterms =   4" in out and "target pps" in out and f"iadap = {iadap}" in out
    tt = float(out.split("target time")[1].split()[0])
    assert tt > 0.0
