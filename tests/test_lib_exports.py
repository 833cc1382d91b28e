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


def test_local_symmetry_operators_without_gpu():
    """cim_sym_ops is host-only: the local programs' generators (icos2f_sym_matc.f90)."""
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("HIP library not built")
    import numpy as np
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lref_units.npz"))
    for sym, key in (("c5", "lsym_0_5"), ("i2", "lsym_2_60"), ("i5", "lsym_3_60")):
        assert np.abs(lib.sym_ops(sym) - g[key]).max() < 2e-6
    assert lib.sym_ops("icos").shape == (60, 3, 3)
    with pytest.raises(lib.CimError):
        lib.sym_ops("d7")


def test_icosprocess_local_operators_without_gpu(O32):
    """cim_icosprocess_sym_ops (host-only): the mode-6 local groups, against the
    oracle's operators, which reproduce the reference's mode-6 map."""
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("HIP library not built")
    import numpy as np
    for sym, (N, _) in O32.LOCAL_SYM.items():
        got = lib.icosprocess_sym_ops(sym)
        assert got.shape == (N, 3, 3)
        assert np.abs(got - O32.icosprocess_local_ops(sym)).max() < 1e-7, sym
    for bad in ("c9", "icos", "i2", ""):
        with pytest.raises(lib.CimError):
            lib.icosprocess_sym_ops(bad)


def test_local_reconstruct_operators_without_gpu(O32):
    """cim_local_reconstruct_sym_ops (host-only) against the oracle's float32
    restatement of the same generators."""
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("HIP library not built")
    import numpy as np
    for sym in ("i2", "I3", "i5", "c1", "c5", "C7", "C_12", "c15"):
        got = lib.local_reconstruct_sym_ops(sym)
        ref = O32.local_reconstruct_ops(sym)
        assert got.shape == ref.shape, sym
        assert np.abs(got - ref).max() < 1e-7, sym
    with pytest.raises(lib.CimError):
        lib.local_reconstruct_sym_ops("c16")


def test_local_box_operators_without_gpu(O32):
    """cim_local_box_sym_ops (host-only): local_box's spellings, including the
    default of 60 C_matrix operators for a name it does not know."""
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("HIP library not built")
    import numpy as np
    for sym in ("i2", "I3", "i5", "c1", "C5", "c10", "C12", "c15", "C_12", "d7"):
        got = lib.local_box_sym_ops(sym)
        ref = O32.local_box_ops(sym)
        assert got.shape == ref.shape, sym
        assert np.abs(got - ref).max() < 1e-7, sym
    assert lib.local_box_sym_ops("C_12").shape == (60, 3, 3)
