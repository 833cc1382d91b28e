"""HIP centre finding (icosprocess mode 3) against the reference program's
outputs (tests/golden/center.npz), the FP64 oracle at box 128, and the
shift of noise-free projections at box 384."""
import numpy as np
import pytest

import pipelines
from cryoim_amd import lib, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cen():
    return np.load(pipelines.golden_path("center.npz"))


def _records(angles, xy, df):
    P = np.zeros((len(angles), lib.PART_STRIDE), np.float32)
    P[:, 0:3] = angles
    P[:, 3:5] = xy
    if df is not None:
        P[:, 6:8] = df
    P[:, 11] = 1.0
    return P


def _golden_params(cen):
    box, apix = cen["cen_images"].shape[-1], 2.0
    d = pipelines.derived(box, apix, maxRes=6.0, imgmask_A=20)
    return d, lib.make_params(box, 0, d["maxR"], apply_ctf=True, mask_on=True, apix=apix, imgmask=d["mask"])


@pytest.mark.parametrize("method", [lib.CENTER_MODEL, lib.CENTER_SELFROT180])
def test_center_matches_reference(ctx, cen, method):
    """calc_center / calc_center_selfrot180 on the reference's own inputs:
    x, y, the fitted peak and PR = acos(cc) of its output ortfile."""
    d, p = _golden_params(cen)
    key = "cen_model_out" if method == lib.CENTER_MODEL else "cen_self_out"
    if method == lib.CENTER_MODEL:
        ctx.center_model(p, cen["cen_model"])
    parts = _records(cen["cen_angles"], cen["cen_guess"], cen["cen_df"])
    out, cc = ctx.center(p, method, cen["cen_images"], parts)
    got = np.column_stack([out[:, 3], out[:, 4], cc])
    err = np.abs(got - cen[key]).max(0)
    print(f"{key}: max |dx| {err[0]:.2e} |dy| {err[1]:.2e} |dcc| {err[2]:.2e}")
    assert err.max() < 2e-4
    assert np.abs(out[:, 5] - cen[key + "_ort"][:, 2]).max() < 2e-2      # PR (degrees), ortfile has 4 decimals
    assert np.array_equal(out[:, :3], parts[:, :3])                       # angles untouched


@pytest.mark.parametrize("method", [lib.CENTER_MODEL, lib.CENTER_SELFROT180])
def test_center_vs_fp64_oracle_box128(ctx, O64, method):
    box, apix, n = 128, 1.5, 12
    ph = synth.make_phantom(box, radius=(30, 45), sigma=(2.0, 3.5), seed=21)
    vol = synth.phantom_map(ph).astype(np.float32)
    ps = synth.make_particles(ph, n, seed=22, shift_max=3.0, snr=0.5, ctf=True, apix=apix, df_range=(1.0, 2.5))
    rng = np.random.default_rng(23)
    guess = (ps.shifts + rng.uniform(-0.5, 0.5, ps.shifts.shape)).astype(np.float32)
    maxR, mask = 24, 56.0
    p = lib.make_params(box, 0, maxR, apply_ctf=True, mask_on=True, apix=apix, imgmask=mask)
    parts = _records(ps.angles, guess, ps.df)
    if method == lib.CENTER_MODEL:
        ctx.center_model(p, vol)
        fr, fi = O64.center_model(vol, mask)
        ref = np.array([O64.calc_center_one(fr, fi, ps.images[i], *guess[i], ps.angles[i],
                                            O64.ctf2d(box, ps.df[i, 0], ps.df[i, 1], 0.0, apix), mask, maxR)
                        for i in range(n)])
    else:
        ref = np.array([O64.center_selfrot180_one(ps.images[i], mask, maxR) for i in range(n)])
    out, cc = ctx.center(p, method, ps.images, parts)
    got = np.column_stack([out[:, 3], out[:, 4], cc])
    err = np.abs(got - ref).max(0)
    print(f"method {method} box {box}: max |dx| {err[0]:.2e} |dy| {err[1]:.2e} |dcc| {err[2]:.2e}")
    assert err.max() < 1e-3


@pytest.fixture(scope="module")
def box384():
    box, n = 384, 8
    ph = synth.make_phantom(box, radius=(120, 165), sigma=(6.0, 12.0), seed=31)
    vol = synth.phantom_map(ph).astype(np.float32)
    rng = np.random.default_rng(32)
    ang = synth.random_au_orientations(n, rng)
    sh = rng.uniform(-10.0, 10.0, (n, 2))
    imgs = synth.project(ph, ang, sh).astype(np.float32)
    return vol, ang, sh, imgs


def test_center_recovers_projection_shifts_box384(ctx, box384):
    """Noise-free projections shifted by known amounts: both methods find the
    shift to within a pixel (the 3x3 log-quadric fit of a broad, non-Gaussian
    peak is biased by a few tenths of a pixel; the FP64 oracle shows the same
    bias on these inputs)."""
    vol, ang, sh, imgs = box384
    n = len(ang)
    p = lib.make_params(384, 0, 72, apply_ctf=False, mask_on=True, apix=1.5, imgmask=180.0)
    ctx.center_model(p, vol)
    for method in (lib.CENTER_MODEL, lib.CENTER_SELFROT180):
        out, cc = ctx.center(p, method, imgs, _records(ang, np.zeros((n, 2)), None))
        err = np.abs(out[:, 3:5] - sh).max()
        print(f"box 384 method {method}: max shift error {err:.3f} px, cc min {cc.min():.3f}")
        assert err < 1.0 and cc.min() > 0.5


def test_center_integer_shift_equivariance_box384(ctx, box384):
    """Rolling a particle image by whole pixels, with its mask centre moved by
    the same amount, moves the correlation map by exactly that much: the
    fitted centre moves by the roll and the peak value is unchanged."""
    vol, ang, sh, imgs = box384
    n = 4
    p = lib.make_params(384, 0, 72, apply_ctf=True, mask_on=True, apix=1.5, imgmask=150.0)
    ctx.center_model(p, vol)
    df = np.full((n, 2), 1.8, np.float32)
    # images carry the CTF the projection is given, so the peak is the true
    # one (unclamped) rather than a side lobe beyond box/16
    c = synth.ctf2d(384, 1.8, 1.8, 0.0, 1.5)
    imgs = np.stack([synth.apply_ctf(imgs[i], c) for i in range(n)]).astype(np.float32)
    base, cc0 = ctx.center(p, lib.CENTER_MODEL, imgs[:n], _records(ang[:n], sh[:n], df))
    roll = np.array([[3, -2], [-5, 1], [0, 4], [2, 2]])
    rolled = np.stack([np.roll(imgs[i], (roll[i, 1], roll[i, 0]), axis=(0, 1)) for i in range(n)])
    out, cc1 = ctx.center(p, lib.CENTER_MODEL, rolled, _records(ang[:n], sh[:n] + roll, df))
    d = out[:, 3:5] - base[:, 3:5] - roll
    print(f"equivariance: max |d centre| {np.abs(d).max():.2e}, max |d cc| {np.abs(cc1 - cc0).max():.2e}")
    assert np.abs(base[:, 3:5] - sh[:n]).max() < 1.0 and cc0.min() > 0.5
    assert np.abs(d).max() < 2e-3 and np.abs(cc1 - cc0).max() < 1e-4


def test_center_chunks_are_independent(ctx, cen):
    """More particles than one chunk holds: each result equals the
    single-particle one (particles never interact)."""
    d, p = _golden_params(cen)
    ctx.center_model(p, cen["cen_model"])
    idx = np.arange(1100) % len(cen["cen_images"])
    parts = _records(cen["cen_angles"][idx], cen["cen_guess"][idx], cen["cen_df"][idx])
    out, cc = ctx.center(p, lib.CENTER_MODEL, cen["cen_images"][idx], parts)
    one, cc1 = ctx.center(p, lib.CENTER_MODEL, cen["cen_images"],
                          _records(cen["cen_angles"], cen["cen_guess"], cen["cen_df"]))
    assert np.array_equal(out, one[idx]) and np.array_equal(cc, cc1[idx])


def test_center_errors(cen):
    d, p = _golden_params(cen)
    c = lib.Context(0)
    try:
        parts = _records(cen["cen_angles"], cen["cen_guess"], cen["cen_df"])
        with pytest.raises(lib.CimError, match="cim_center_model"):
            c.center(p, lib.CENTER_MODEL, cen["cen_images"], parts)
        with pytest.raises(lib.CimError, match="method"):
            c.center(p, 7, cen["cen_images"], parts)
        c.center_model(p, cen["cen_model"])
        q = lib.make_params(64, 0, 10, apix=2.0, imgmask=20.0)
        with pytest.raises(lib.CimError, match="box 64"):
            c.center(q, lib.CENTER_MODEL, np.zeros((1, 64, 64), np.float32), parts[:1])
    finally:
        c.close()
