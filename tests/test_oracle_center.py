"""Oracle of icosprocess mode 3 (centre finding: calc_center.f90 with a model,
calc_center_selfrot180.f90 without) pinned against closed forms and against
the reference program's own outputs (tests/golden/center.npz)."""
import numpy as np
import pytest

import pipelines

N = 32


def _rot180(a):
    """x -> (n - x) mod n on both axes: what one img2d_FFT round trip does,
    since img2d_FFT stores b = -Im in both directions."""
    return np.roll(a[::-1, ::-1], 1, axis=(0, 1))


@pytest.fixture(scope="module")
def cen():
    return np.load(pipelines.golden_path("center.npz"))


def _grid(n=N):
    return np.meshgrid(*(np.arange(n) - n // 2,) * 3, indexing="ij")


def test_gauss_regression_recovers_a_gaussian(O64, O32):
    """log of a Gaussian is the quadric the fit assumes: exact offset and peak."""
    for O, tol in ((O64, 1e-12), (O32, 2e-5)):
        for dx0, dy0, A, s in ((0.3, -0.2, 0.8, 1.5), (-0.45, 0.1, 0.3, 0.9), (0.0, 0.0, 1.0, 2.0)):
            z = np.array([[A * np.exp(-((i - dx0) ** 2 + (j - dy0) ** 2) / (2 * s * s)) for i in (-1, 0, 1)]
                          for j in (-1, 0, 1)])
            dx, dy, cc = O.gauss_regression(z)
            assert abs(dx - dx0) < tol and abs(dy - dy0) < tol and abs(cc - A) < tol


def test_lowpass_round_trip_rotates_and_scales(O64):
    """img2d_lowpass with H -> 1 returns the image rotated by 180 degrees about
    pixel (n/2, n/2) and scaled by 1/n^2 (forward 1/n^2, then H/n^2)."""
    img = np.random.default_rng(0).normal(size=(N, N))
    assert np.abs(O64.lowpass(img, 10 ** 6) - _rot180(img) / N ** 2).max() < 1e-9


def test_apply_ctf_of_one_rotates(O64):
    img = np.random.default_rng(1).normal(size=(N, N))
    assert np.abs(O64.apply_ctf(img, np.ones((N, N))) - _rot180(img)).max() < 1e-12


def test_cross_correlation_autocorrelation_peak(O64):
    """CC of an image with itself peaks at (n/2, n/2) with mean(a^2)."""
    img = np.random.default_rng(2).normal(size=(N, N))
    cc = O64.cross_correlation(img, img)
    k = np.unravel_index(np.argmax(cc), cc.shape)
    assert k == (N // 2, N // 2)
    assert abs(cc[k] - (img ** 2).mean()) < 1e-12


def test_projection_is_central_slice(O64):
    """At theta = phi = omega = 0 the interpolated central slice lands on grid
    points: the projection is the z-sum / n, rotated by the backward FFT."""
    z, y, x = _grid()
    vol = np.exp(-((x - 3) ** 2 + (y + 2) ** 2 + (z - 1) ** 2) / (2 * 2.0 ** 2))
    fr, fi = O64.center_model(vol)
    p = O64.project2d_ft(fr, fi, 0.0, 0.0, 0.0)
    ref = _rot180(vol.sum(0) / N)
    assert np.abs(p - ref).max() < 1e-6 * ref.max()


def test_model_mask_keeps_the_inside(O64):
    z, y, x = _grid()
    vol = np.random.default_rng(3).normal(size=(N, N, N))
    m = O64.img3d_mask(vol, 8.0)
    r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    assert np.array_equal(m[r <= 5.0], vol[r <= 5.0])
    assert np.allclose(m[r > 11.0], m[r > 11.0].flat[0])


def test_selfrot180_finds_the_shift(O64):
    """A centrosymmetric object at (1.5, -1): the 180-degree self-correlation
    peaks at twice the shift, which is halved."""
    _, y, x = _grid()
    xs, ys = x[0], y[0]
    r2 = (xs - 1.5) ** 2 + (ys + 1.0) ** 2
    b = np.exp(-r2 / (2 * 1.5 ** 2)) + 0.3 * np.exp(-r2 / (2 * 4.0 ** 2))
    cx, cy, cc = O64.center_selfrot180_one(b, None, 12)
    assert abs(cx - 1.5) < 1e-3 and abs(cy + 1.0) < 1e-3 and cc > 0.95


def test_center_peak_clamps_large_shifts(O64):
    """|shift| >= n/16 is replaced by 0 (calc_center.f90:84-85)."""
    yy, xx = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    ccmap = np.exp(-((xx - N // 2 - 2.2) ** 2 + (yy - N // 2 + 0.6) ** 2) / 4.0) + 1e-3
    cx, cy, _ = O64.center_peak(ccmap)
    assert cx == 0.0 and abs(cy + 0.6) < 0.05
    cx, cy, _ = O64.center_peak(ccmap, halve=True)
    assert abs(cx - 1.1) < 0.05 and abs(cy + 0.3) < 0.05


def center_inputs(O, cen):
    box, apix = cen["cen_images"].shape[-1], 2.0
    d = pipelines.derived(box, apix, maxRes=6.0, imgmask_A=20)
    fr, fi = O.center_model(cen["cen_model"], d["mask"])
    ctf = [O.ctf2d(box, cen["cen_df"][i, 0], cen["cen_df"][i, 1], 0.0, apix) for i in range(len(cen["cen_df"]))]
    return d, fr, fi, ctf


@pytest.mark.parametrize("bits,tol", [(32, 2e-5), (64, 1e-4)])
def test_calc_center_matches_reference(bits, tol, O32, O64, cen):
    O = O32 if bits == 32 else O64
    d, fr, fi, ctf = center_inputs(O, cen)
    out = np.array([O.calc_center_one(fr, fi, cen["cen_images"][i], *cen["cen_guess"][i], cen["cen_angles"][i],
                                      ctf[i], d["mask"], d["maxR"]) for i in range(len(ctf))])
    err = np.abs(out - cen["cen_model_out"]).max()
    print(f"FP{bits} calc_center vs reference: max |d(x, y, cc)| {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("bits,tol", [(32, 2e-5), (64, 1e-4)])
def test_selfrot180_matches_reference(bits, tol, O32, O64, cen):
    O = O32 if bits == 32 else O64
    d = pipelines.derived(cen["cen_images"].shape[-1], 2.0, maxRes=6.0, imgmask_A=20)
    out = np.array([O.center_selfrot180_one(im, d["mask"], d["maxR"]) for im in cen["cen_images"]])
    err = np.abs(out - cen["cen_self_out"]).max()
    print(f"FP{bits} calc_center_selfrot180 vs reference: max |d(x, y, cc)| {err:.2e}")
    assert err < tol


def test_output_ort_pr_is_arccos_of_cc(cen):
    """The ortfile written after mode 3 holds PR = acos(cc) in degrees."""
    for key in ("cen_model_out", "cen_self_out"):
        pr = np.degrees(np.arccos(np.clip(cen[key][:, 2], -1, 1)))
        assert np.abs(pr - cen[key + "_ort"][:, 2]).max() < 2e-4
        assert np.abs(cen[key][:, :2] - cen[key + "_ort"][:, :2]).max() < 1e-4
