"""The icosahedral operators against the symmetry-matrix data files the
reference ships in local_box/SRC (tests/golden/symfiles.npz, read column by
column as read_sym_matrix.f90:8-12 reads them): the 60 operators of the
hard-coded table (icos2f_sym_mat.f90) and the I2 / I3 / I5 generators of the
local programs (icos_matrix_i2/i3/i5.f90) must contain every matrix of them,
and the vertex / face files' axes must be icosahedral directions."""
import numpy as np
import pytest

import pipelines


@pytest.fixture(scope="module")
def g():
    return np.load(pipelines.golden_path("symfiles.npz"))


def _index_in(A, G, tol):
    """For each matrix of A, the index of the element of G within tol (max abs), or -1."""
    d = np.abs(A[:, None] - G[None]).max(axis=(-1, -2))
    i = d.argmin(axis=1)
    return np.where(d[np.arange(len(A)), i] < tol, i, -1)


def test_icos2f_table_is_the_group_of_the_file(O64, g):
    """icos2f_sym_matrix.txt holds the same 60 operators as the table (another order)."""
    A, T = g["sym_icos2f_sym_matrix"], O64.icos_ops()
    ia = _index_in(A, T, 1e-7)
    assert (ia >= 0).all() and len(set(ia.tolist())) == 60
    assert (_index_in(T, A, 1e-7) >= 0).all()


@pytest.mark.parametrize("name,sym,n", [("icos2f_12vertex_matrix", "i2", 12), ("icos2f_30edgecenter", "i2", 30),
                                        ("icos5f_12vertex_matrix", "i5", 12), ("icos3f_20facecenter", "i3", 20)])
def test_generators_contain_the_file_operators(O32, O64, g, name, sym, n):
    """Every matrix of the vertex / edge / face files is an element of the group
    generated in the file's frame (float32 files: 1e-6)."""
    A = g["sym_" + name]
    assert A.shape == (n, 3, 3)
    for O in (O64, O32):
        G = O.local_reconstruct_ops(sym)
        ia = _index_in(A, G, 2e-6)
        assert (ia >= 0).all(), (name, np.nonzero(ia < 0))
        assert len(set(ia.tolist())) == n


@pytest.mark.parametrize("name,n,dots,repeats", [
    ("icos5f_12vertex_matrix", 12, (1 / np.sqrt(5), -1 / np.sqrt(5), -1.0), [(1, 9), (2, 4)]),
    ("icos3f_20facecenter", 20, (np.sqrt(5) / 3, -np.sqrt(5) / 3, 1 / 3, -1 / 3, -1.0), [])])
def test_file_axes(g, name, n, dots, repeats):
    """With the file's own axis along z (5-fold for the vertex file, 3-fold for
    the face file), S^T z runs over vertices (face centres): unit vectors whose
    pairwise cosines are those of the icosahedron (dodecahedron for the faces).
    The 20 face matrices pick 20 distinct centres; the 12 vertex matrices, as
    shipped, pick only 10 vertices -- rows 2 and 10, and rows 3 and 5 (1-based)
    are different rotations about the same vertex."""
    A = g["sym_" + name]
    v = np.array([a.T @ np.array([0.0, 0.0, 1.0]) for a in A])
    c = v @ v.T
    assert np.abs(np.linalg.norm(v, axis=1) - 1).max() < 1e-6
    same = [tuple(ij) for ij in np.argwhere(np.triu(c > 1 - 1e-3, 1)).tolist()]
    assert same == repeats
    off = c[~np.eye(n, dtype=bool) & (c < 1 - 1e-3)]
    assert np.abs(off[:, None] - np.array(dots)[None]).min(axis=1).max() < 1e-5


def test_sym_matrix_53_file_has_one_matrix_that_is_no_rotation(g):
    """icos2f_sym_matrix_53.txt: 29 of its 30 matrices are icosahedral operators;
    one is not a rotation at all (determinant -0.309) -- data as shipped, which
    a sym_matrix run of local_box would use as given."""
    A = g["sym_icos2f_sym_matrix_53"]
    det = np.linalg.det(A)
    bad = np.nonzero(np.abs(det - 1) > 1e-6)[0]
    assert len(bad) == 1 and abs(det[bad[0]] + 0.309017) < 1e-5
    from oracle import Oracle
    ia = _index_in(np.delete(A, bad, axis=0), Oracle(64).icos_ops(), 1e-7)
    assert (ia >= 0).all()
