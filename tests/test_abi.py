"""The C-ABI library loads on a machine without a GPU and exports every entry
point include/fts_hip.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "langevin_fts_amd", "libfts_hip.so")


def declared():
    src = open(os.path.join(REPO, "include", "fts_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fts_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "fts_solver_compute_propagators" in names
    assert "fts_am_calculate_new_fields" in names
    assert len(names) >= 25


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfts_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.fts_abi_version.restype = ctypes.c_int
    assert lib.fts_abi_version() == 1
    lib.fts_last_error.restype = ctypes.c_char_p
    assert lib.fts_last_error() == b""


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfts_hip.so not built")
def test_null_arguments_are_errors_not_crashes():
    lib = ctypes.CDLL(LIB)
    lib.fts_last_error.restype = ctypes.c_char_p
    assert lib.fts_grid_set_lx(None, None) != 0
    assert b"null" in lib.fts_last_error()
    assert lib.fts_solver_compute_concentrations(None) != 0
