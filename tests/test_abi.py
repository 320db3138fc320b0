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


def test_autotune_table_file_round_trip(tmp_path):
    """The launch-shape table written by pmx_autotune_save reads back entry
    for entry (host only: no GPU call), a reload replaces entries of the same
    class, and a malformed file is refused."""
    from pymrst_amd import _lib
    src = tmp_path / "shapes.txt"
    rows = ["1 2 80 11 0 4", "2 4 72 11 5 0", "2 4 64 12 5 6", "3 8 84 18 0 4"]
    src.write_text("\n".join(rows) + "\n")
    lib = _lib.load()
    _lib.check(lib.pmx_autotune_reset(), "pmx_autotune_reset")
    assert _lib.autotune_load(src) == 4
    out = tmp_path / "again.txt"
    assert _lib.autotune_save(out) == 4
    assert sorted(out.read_text().split("\n")[:-1]) == sorted(rows)
    (tmp_path / "same.txt").write_text("2 4 72 11 5 3\n")
    assert _lib.autotune_load(tmp_path / "same.txt") == 1
    n = ctypes.c_int32()
    _lib.check(lib.pmx_autotune_entries(ctypes.byref(n)), "pmx_autotune_entries")
    assert n.value == 4
    _lib.autotune_save(out)
    assert "2 4 72 11 5 3" in out.read_text()
    (tmp_path / "bad.txt").write_text("1 2 3\n")
    with pytest.raises(_lib.PmxError):
        _lib.autotune_load(tmp_path / "bad.txt")
    _lib.check(lib.pmx_autotune_reset(), "pmx_autotune_reset")
