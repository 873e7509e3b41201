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
    assert ctypes.sizeof(pkg.TreeConfig) == 4 * 4 + 8 + 3 * 8 + 3 * 4 + 4
    assert ctypes.sizeof(pkg.Criterion) == 4 * 4 + 2 * 8
    assert pkg.HALO_META_DTYPE.itemsize == 36


def test_patch_shape_validation_without_device():
    """Patch shapes are validated before any device work: odd extents, a
    halo deeper than half a patch, a missing extent and a patch_size that
    disagrees with the extents return AMR_ERR_ARGUMENT."""
    pkg = load_package()
    lib = pkg.load()
    for ext, h, size in (((15, 8, 0), 1, 0), ((8, 2, 0), 2, 0), ((8, 0, 0), 1, 0), ((16, 8, 0), 1, 16),
                         ((16, 8, 4), 1, 0)):
        cfg = pkg.TreeConfig()
        cfg.rank, cfg.max_depth, cfg.halo_width, cfg.capacity, cfg.patch_size = 2, 7, h, 64, size
        for d in range(3):
            cfg.patch_extents[d] = ext[d]
            cfg.domain_length[d] = 1.0
        out = ctypes.c_void_p()
        assert lib.amr_tree_create(ctypes.byref(cfg), ctypes.byref(out)) == pkg.AMR_ERR_ARGUMENT, (ext, h, size)
        assert not out.value
        assert lib.amr_last_error()
