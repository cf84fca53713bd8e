"""The C-ABI library loads on the CPU and exports every entry point include/cutcell_hip.h declares
(no compute calls: this container has no GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    src = open(os.path.join(ROOT, "include", "cutcell_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(cc_\w+)\s*\(", src, re.M)))


def test_header_declares_the_expected_surface():
    names = declared_entry_points()
    for n in ("cc_cache_create", "cc_cache_destroy", "cc_update", "cc_update_from_phi",
              "cc_column", "cc_copy_to_host", "cc_generate_mesh_ms", "cc_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from cutcell import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_entry_points() if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= set(declared_entry_points())


def test_version_string_without_a_device():
    from cutcell import _lib
    assert b"gfx950" in _lib.lib().cc_version()
