"""Oracle of the real-space icosahedral average (realspaceavg=y): icosprocess's
img3d_icosavg and merge_icosprocess_subrec's, against the reference programs'
own outputs (tests/golden/icosavg.npz) and against closed forms."""
import numpy as np
import pytest

import pipelines
from conftest import rel_l2

TNF = ("fa", "fb", "wg")


@pytest.fixture(scope="module")
def a():
    return np.load(pipelines.golden_path("icosavg.npz"))


@pytest.fixture(scope="module")
def g():
    return np.load(pipelines.golden_path("fsc.npz"))


def _tnf(g, key):
    return [g[f"{key}_{k}"] for k in TNF]


@pytest.mark.parametrize("bits,tol", [(32, 1e-6), (64, 1e-6)])
def test_full_average_matches_reference(bits, tol, O32, O64, a):
    """icosprocess mode 1 with realspaceavg=y: the averaged map from the
    unaveraged one (both written by the reference for the same particles)."""
    O = O32 if bits == 32 else O64
    out = O.icosavg_full(a["avg_input_map"], 20.0)
    err = rel_l2(out, a["avg_map"])
    print(f"FP{bits} img3d_icosavg vs reference: rel-L2 {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("bits", [32, 64])
def test_merge_averages_match_reference(bits, O32, O64, g, a):
    """nicosavg = 1 in the tnf headers: merge_reconstruct (average on the full
    map, then the crop) and merge_reconstruct_FSC (average of the first
    (box-8)^3 values as a cube), and icosprocess's reconstruct_FSC."""
    O = O32 if bits == 32 else O64
    assert int(a["avg_tnf_hdr"][0]) == 1
    box, maxR, mask = 48, int(g["s1_hdr"][1]), float(g["s1_hdrf"][1])
    m = O.merge_map([_tnf(g, "s1"), _tnf(g, "s2")], box, maxR, imgmask=mask, icosavg=True)
    assert rel_l2(m, a["mavg_map"]) < 1e-5
    odd = [x.astype(np.float32) + y for x, y in zip(_tnf(g, "p1_odd"), _tnf(g, "p2_odd"))]
    eve = [x.astype(np.float32) + y for x, y in zip(_tnf(g, "p1_eve"), _tnf(g, "p2_eve"))]
    _, m = O.fsc_map(odd, eve, box, maxR, imgmask=mask, merge=True, icosavg=True)
    err = rel_l2(m, a["mfscavg_map"])
    print(f"FP{bits} merge FSC + average vs reference: rel-L2 {err:.2e}")
    assert err < 1e-5
    _, m = O.fsc_map(_tnf(g, "half_odd"), _tnf(g, "half_eve"), box, maxR, imgmask=mask, icosavg=True)
    assert rel_l2(m, a["fscavg_map"]) < 1e-5


@pytest.mark.parametrize("variant", ["full", "merge"])
def test_constant_map_stays_constant_inside_r0(O64, variant):
    """Trilinear weights sum to 1 and every tap of a voxel within r0 < n/2-1
    lies inside the cube: a constant map is kept inside r0, zeroed outside."""
    n, r0 = 24, 9.5
    v = np.full((n, n, n), 2.5)
    out = O64.icosavg_full(v, r0) if variant == "full" else O64.icosavg_merge(v, r0)
    z, y, x = np.meshgrid(*(np.arange(n) - n // 2,) * 3, indexing="ij")
    r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    assert np.abs(out[r <= r0] - 2.5).max() < 1e-12
    assert np.all(out[r > r0] == 0.0)


def test_merge_copies_equal_the_full_average(O64):
    """The icosahedral mean is invariant under the 2-folds about x, y and z,
    so merge's quarter-and-copy equals the full mean for r0 < n/2 - 1."""
    n = 28
    v = np.random.default_rng(4).normal(size=(n, n, n))
    full = O64.icosavg_full(v, 12.0)
    merged = O64.icosavg_merge(v, 12.0)
    assert np.abs(full - merged).max() < 1e-12 * np.abs(full).max() + 1e-12


def test_merge_differs_only_on_the_low_faces_when_r0_reaches_n_half(O64):
    """With r0 >= n/2 merge's loops (ky, lz <= n/2-1) leave the faces at
    -n/2 to its copies, and the hx = -n/2 copies land on x = 0 of the next
    row (the reference's flat index): the result differs from the full mean
    only where some index is 0, i.e. some centred coordinate is -n/2."""
    n = 16
    v = np.random.default_rng(5).normal(size=(n, n, n))
    merged = O64.icosavg_merge(v, 8.0)
    full = O64.icosavg_full(v, 8.0)
    d = np.abs(merged - full) > 1e-9
    assert d.any()
    z, y, x = np.nonzero(d)
    assert np.all((x == 0) | (y == 0) | (z == 0))
