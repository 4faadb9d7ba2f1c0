import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def fixture4():
    return dict(np.load(GOLDEN / "stellarator_order4.npz"))


@pytest.fixture(scope="session")
def fixture6():
    return dict(np.load(GOLDEN / "stellarator_order6.npz"))


@pytest.fixture(scope="session")
def geo4():
    from qotential_amd import workload as W
    return W.geometry(12, 36, 4)


@pytest.fixture(scope="session")
def geo8():
    from qotential_amd import workload as W
    return W.geometry(12, 36, 8)
