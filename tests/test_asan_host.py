"""AddressSanitizer + UndefinedBehaviorSanitizer run of the host code: the
library's geometry generator, per-order tables and workload C-ABI
(qotential_amd/csrc/geometry.cpp, host_tables.cpp, workload_capi.cpp,
status.cpp) and the oracle (oracle/*.cpp), built with g++ into one host
executable (tests/asan_driver.cpp) that exercises them on small surfaces.  Any
sanitizer report aborts the run."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_asan_ubsan(tmp_path):
    src = [ROOT / "tests" / "asan_driver.cpp"]
    src += [ROOT / "qotential_amd" / "csrc" / f for f in ("geometry.cpp", "host_tables.cpp", "workload_capi.cpp",
                                                          "status.cpp")]
    src += [ROOT / "oracle" / f for f in ("tables.cpp", "levels.cpp", "rrq.cpp", "closepanel.cpp", "capi.cpp")]
    exe = tmp_path / "asan_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fopenmp", "-ffp-contract=off", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", f"-I{ROOT / 'include'}",
           f"-I{ROOT / 'qotential_amd' / 'csrc'}", f"-I{ROOT / 'oracle'}", *map(str, src), "-o", str(exe)]
    subprocess.run(cmd, check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=900, env=env)
    print(r.stdout[-3000:], r.stderr[-6000:])
    assert r.returncode == 0 and "asan driver done" in r.stdout
