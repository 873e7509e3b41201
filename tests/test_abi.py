"""The C-ABI library loads and exports every entry point include/amr_hip.h
declares (no device work: this runs without a GPU)."""
import ctypes
import re

from amrlib import ROOT, load_package


def declared_entry_points():
    text = (ROOT / "include" / "amr_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|const char\*)\s+(amr_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_entry_points()
    assert "amr_tree_create" in names and "amr_solver_advance_batch_async" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    pkg = load_package()
    lib = ctypes.CDLL(str(pkg.LIB_PATH))
    missing = [n for n in declared_entry_points() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    pkg = load_package()
    assert set(declared_entry_points()) == set(pkg.exported_symbols())


def test_struct_layouts():
    pkg = load_package()
    assert ctypes.sizeof(pkg.TreeConfig) == 4 * 4 + 8 + 3 * 8
    assert ctypes.sizeof(pkg.Criterion) == 4 * 4 + 2 * 8
    assert pkg.HALO_META_DTYPE.itemsize == 36
