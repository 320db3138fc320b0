"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(pmx_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "pmx_ow_assemble" in names and "pmx_ow_solve" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    from pymrst_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.fail("pmx library not built (python -m pymrst_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers exactly the declared entry points
    assert set(_lib.SIGNATURES) == set(declared_symbols())


def test_solver_defaults_are_the_reference_defaults():
    from pymrst_amd import _lib
    o = _lib.solver_defaults()
    assert o.rtol == 1e-8 and o.restart == 50 and o.maxiter == 200 and o.precision == 64
