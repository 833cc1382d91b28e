"""The C-ABI library loads and exports every function include/cryoim.h declares
(no compute call: runs without a GPU)."""
import ctypes
import os

import pytest

from cryoim_amd import lib


def test_header_declares_entry_points():
    names = lib.header_functions()
    assert "cim_init" in names and "cim_rec_insert" in names and "cim_search_global" in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("HIP library not built")
    L = ctypes.CDLL(lib.LIB_PATH)
    missing = [n for n in lib.header_functions() if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"
