"""The C ABI on the CPU: the in-tree library loads (no GPU is touched) and
exports every function include/dplus_mi355x.h declares, and the ctypes binding
declares exactly that set."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dplus_mi355x.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(dplus_[a-z0-9_]+)\s*\(", text))


def test_header_declares_the_api():
    names = declared()
    for must in ("dplus_grid_create", "dplus_orientation_average", "dplus_backend_start_generate",
                 "dplus_engine_submit", "dplus_session_evaluate", "dplus_amp_read", "dplus_dock_grid_space_filling"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import dplus_mi355x._lib as L
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    import dplus_mi355x._lib as L
    assert set(L.SIGNATURES) == declared()
    L.lib()  # every signature binds


def test_build_shape_table_text_round_trip():
    """The autotuner's shape table (host-only state): load, save, clear, and
    rejection of malformed or out-of-range lines with ERROR_INVALID_ARGS."""
    import dplus_mi355x as D
    from dplus_mi355x._lib import DplusError
    D.build_shapes_clear()
    try:
        assert D.build_shapes_save() == ""
        D.build_shapes_load("# grid_shells tile_class wg slices\n28 20 256 4\n53 27 512 2\n")
        assert D.build_shapes_save() == "28 20 256 4\n53 27 512 2\n"
        D.build_shapes_load("28 20 128 3")  # replaces the (28, 20) entry
        assert D.build_shapes_save() == "28 20 128 3\n53 27 512 2\n"
        for bad in ("28 20 96 4", "28 20 256 9", "0 1 256 1", "28 20 256"):
            try:
                D.build_shapes_load(bad)
            except DplusError as e:
                assert e.error_code == 21
            else:
                raise AssertionError(bad)
        assert D.build_shapes_save() == "28 20 128 3\n53 27 512 2\n"
    finally:
        D.build_shapes_clear()
