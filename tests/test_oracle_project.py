"""The oracle's project_template2d (icosprocess/SRC/project_template2d.f90:3-80)
against the reference's own projections (tests/golden/proj.npz: the unit
driver's project_template2d, and mode 0's project2d template stack), and
against what the geometry fixes: at the identity orientation a projection is
the plain column sum of the voxels inside the radius cut, and the projection
of the Gaussian-blob phantom approaches its analytic line integrals."""
import numpy as np
import pytest

from conftest import rel_l2
from cryoim_amd import synth
import pipelines


@pytest.fixture(scope="module")
def pj():
    return np.load(pipelines.golden_path("proj.npz"))


@pytest.mark.parametrize("bits,tol", [(32, 1e-6), (64, 1e-5)])
def test_projection_matches_reference(bits, tol, O32, O64, pj):
    O = O32 if bits == 32 else O64
    for ang, ref in zip(pj["proj_angles"], pj["proj_out"]):
        got = O.project_template2d(pj["proj_model"], ang)
        err = rel_l2(got, ref)
        print(f"FP{bits} projection at {ang}: rel-L2 {err:.2e} vs reference")
        assert err < tol


def test_mode0_stack_is_the_projection(pj):
    """project2d (mode 0) and project_template2d are the same arithmetic."""
    assert np.array_equal(pj["mode0_stack"], pj["proj_out"][:3])


def test_identity_projection_is_the_column_sum(O64):
    rng = np.random.default_rng(3)
    n = 20
    vol = rng.standard_normal((n, n, n))
    got = O64.project_template2d(vol, (0.0, 0.0, 0.0))
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    inside = np.sqrt((x - n // 2) ** 2 + (y - n // 2) ** 2 + (z - n // 2) ** 2) < n // 2 - 1
    assert np.allclose(got, np.where(inside, vol, 0.0).sum(axis=0), rtol=1e-12, atol=1e-12)


def test_projection_of_blobs_matches_line_integrals(O64):
    box = 64
    ph = synth.make_phantom(box, radius=(14, 20), sigma=(2.0, 3.0), seed=12)
    vol = synth.phantom_map(ph).astype(np.float64)
    for ang in ((75.0, 10.0, 30.0), (40.0, -70.0, 250.0)):
        got = O64.project_template2d(vol, ang)
        want = synth.project(ph, np.array([ang]))[0]
        err = rel_l2(got, want)
        print(f"projection at {ang} vs analytic line integrals: rel-L2 {err:.2e}")
        assert err < 0.02
