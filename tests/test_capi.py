"""The C-ABI library loads and exports every symbol include/qotential_hip.h declares."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    txt = (ROOT / "include" / "qotential_hip.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qol_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_declared_symbols():
    from qotential_amd._lib import LIB_PATH
    lib = ctypes.CDLL(str(LIB_PATH))
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
