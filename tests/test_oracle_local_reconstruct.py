"""Oracle of local_reconstruct (its own operator generators, the periodic
local box cut, the insert and its substep) against the reference program's
outputs (tests/golden/lrec.npz: the program built with the one line flang
rejects removed) and closed forms of the generators."""
import numpy as np
import pytest

import pipelines
from conftest import rel_l2

SYMS = [("c5", "c5"), ("i2", "i2"), ("i3", "I3"), ("i5", "i5")]


@pytest.fixture(scope="module")
def lr():
    return np.load(pipelines.golden_path("lrec.npz"))


def _base(lr):
    box = int(lr["lrec_box"][0])
    return (lr["lrec_images"], lr["lrec_cen"], lr["lrec_angles"], lr["lrec_shifts"], lr["lrec_df"],
            float(lr["lrec_apix"][0]), box, 10)        # maxR = int(2 * 32 / 6)


@pytest.mark.parametrize("tag,sym", SYMS)
@pytest.mark.parametrize("bits,tol", [(32, 2e-6), (64, 1e-5)])
def test_local_map_matches_reference(bits, tol, O32, O64, lr, tag, sym):
    O = O32 if bits == 32 else O64
    v = pipelines.local_reconstruct_volume(O, *_base(lr), sym)
    err = rel_l2(O.fold_and_map(v, 10), lr[f"lrec_{tag}_map"])
    print(f"FP{bits} local_reconstruct {sym}: rel-L2 {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("bits,tol", [(32, 2e-6), (64, 1e-5)])
def test_zhigh_defocus_matches_reference(bits, tol, O32, O64, lr):
    O = O32 if bits == 32 else O64
    v = pipelines.local_reconstruct_volume(O, *_base(lr), "c1", z=lr["lrec_z"], sign="-")
    assert rel_l2(O.fold_and_map(v, 10), lr["lrec_zhigh_map"]) < tol


@pytest.mark.parametrize("bits,tol", [(32, 2e-6), (64, 1e-5)])
def test_local_substep_matches_reference(bits, tol, O32, O64, lr):
    """reconstruct_substep.f90: mask (radius imgmask/apix = 12) before the
    norm, rcut maxR-1, compact layout, weights from 0.1."""
    O = O32 if bits == 32 else O64
    assert list(lr["lrecsub_hdr"]) == [32, 10, 0, 0]
    assert np.allclose(lr["lrecsub_hdrf"], [2.0, 12.0])
    v = pipelines.local_reconstruct_volume(O, *_base(lr), "c5", compact=True, mask_r=12.0)
    for o, k in zip(v, ("fa", "fb", "wg")):
        assert rel_l2(o, lr[f"lrecsub_{k}"]) < tol


def _closed(ops, tol):
    for a in ops:
        for b in ops:
            c = a @ b
            assert np.abs(ops - c).reshape(len(ops), -1).max(1).min() < tol


def test_icosahedral_generators(O64):
    """I2 is the icosahedral group of icos2f_sym_mat (same 60 rotations);
    I3 and I5 are groups with a 3-fold / 5-fold about z."""
    ref = O64.icos_ops()
    for sym in ("i2", "I3", "i5"):
        ops = O64.local_reconstruct_ops(sym)
        assert ops.shape == (60, 3, 3)
        assert np.allclose(np.einsum("nij,nkj->nik", ops, ops), np.eye(3), atol=1e-12)
        _closed(ops, 1e-12)
    i2 = O64.local_reconstruct_ops("i2")
    for r in ref:
        assert np.abs(i2 - r).reshape(60, -1).max(1).min() < 1e-6
    for sym, n in (("I3", 3), ("i5", 5)):
        ops = O64.local_reconstruct_ops(sym)
        a = 2 * np.pi / n
        rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
        assert min(np.abs(ops - rz).reshape(60, -1).max(1).min(), np.abs(ops - rz.T).reshape(60, -1).max(1).min()) < 1e-6


def test_c_matrix_order(O64):
    """C_matrix: the first operator is the rotation by 360/N, the last the identity."""
    for n in (1, 2, 5, 7, 12):
        ops = O64.local_reconstruct_ops(f"c{n}")
        a = 2 * np.pi / n
        assert np.allclose(ops[0], [[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]], atol=1e-12)
        assert np.allclose(ops[-1], np.eye(3), atol=1e-9)
    assert O64.local_reconstruct_sym("C_12") == (0, 12) and O64.local_reconstruct_sym("C12") == (0, 1)
