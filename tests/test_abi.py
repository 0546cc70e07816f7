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
