"""The C oracle pinned against golden vectors recorded from the reference's
own Fortran (tests/golden/make_golden.py) and against mathematical
invariants.  CPU only."""
import os

import numpy as np
import pytest

import pipelines
from conftest import rel_l2

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def units():
    return np.load(os.path.join(G, "units.npz"))


@pytest.fixture(scope="module")
def prv():
    return np.load(os.path.join(G, "pr.npz"))


@pytest.fixture(scope="module")
def lref():
    return np.load(os.path.join(G, "lref_units.npz"))


# ---- rotations ------------------------------------------------------------

def test_symops_match_reference_table(O64, units):
    assert np.abs(O64.icos_ops() - units["symops"]).max() < 1e-6


def test_symops_form_the_icosahedral_group(O64):
    ops = O64.icos_ops()
    for S in ops:
        assert abs(np.linalg.det(S) - 1) < 1e-12
        assert np.abs(S @ S.T - np.eye(3)).max() < 1e-12
    flat = ops.reshape(60, 9)
    for A in ops[::7]:
        for B in ops[::5]:
            C = (A @ B).reshape(9)
            assert np.abs(flat - C).max(axis=1).min() < 1e-12       # closure
    d = np.abs(flat[:, None, :] - flat[None, :, :]).max(axis=2)
    assert (d + np.eye(60) * 9).min() > 0.1                           # 60 distinct


def test_euler(O64, O32, units):
    for a, R in zip(units["euler_in"], units["euler_out"]):
        assert np.abs(O32.euler_mat(*a) - R).max() < 2e-6
        M = O64.euler_mat(*a)
        assert np.abs(M - R).max() < 2e-6
        assert abs(np.linalg.det(M) - 1) < 1e-12


def test_euler_identity_closed_form(O64):
    assert np.abs(O64.euler_mat(0, 0, 0) - np.eye(3)).max() < 1e-15
    # theta only: rotation about y
    t = np.deg2rad(30.0)
    Ry = np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])
    assert np.abs(O64.euler_mat(30, 0, 0) - Ry).max() < 1e-15


def test_crosslin_angles(O32, O64, units):
    for pair, out in zip(units["crosslin_in"], units["crosslin_out"]):
        a1, a2 = O32.crosslin_angle(pair[:3], pair[3:])
        skip = out[0] >= 999
        assert np.array_equal(a1 >= 999, skip)
        d1 = (a1[~skip] - out[0][~skip] + 180) % 360 - 180
        d2 = (a2[~skip] - out[1][~skip] + 180) % 360 - 180
        assert np.abs(d1).max() < 2e-3 and np.abs(d2).max() < 2e-3


def test_crosslin_common_line_lies_in_both_planes(O64):
    """v = z1 x S z2 lies in both central planes: its image-frame directions
    (cos a1, sin a1, 0) and (cos a2, sin a2, 0) map to the same model vector."""
    ops = O64.icos_ops()
    a1, a2 = O64.crosslin_angle((75.0, 12.0, 40.0), (82.0, -7.0, 210.0))
    E1, E2 = O64.euler_mat(75, 12, 40), O64.euler_mat(82, -7, 210)
    for j in range(60):
        if a1[j] >= 999:
            continue
        u1 = E1 @ np.array([np.cos(np.deg2rad(a1[j])), np.sin(np.deg2rad(a1[j])), 0])
        u2 = ops[j] @ E2 @ np.array([np.cos(np.deg2rad(a2[j])), np.sin(np.deg2rad(a2[j])), 0])
        assert np.abs(u1 - u2).max() < 1e-9


def test_return_au(O32, units):
    for a, out in zip(units["au_in"], units["au_out"]):
        r = np.array(O32.return_au(*a))
        d = (r - out + 180) % 360 - 180
        assert np.abs(d).max() < 2e-3, (a, r, out)


# ---- per-image operations ------------------------------------------------------

def test_ctf(O32, O64, units):
    for p, out in zip(units["ctf_in"], units["ctf_out"]):
        df1, df2, ast, apix, vol, cs, amp = p
        c32 = O32.ctf2d(40, df1, df2, ast, apix, vol, cs, amp)
        assert np.abs(c32 - out).max() < 1e-4        # REAL*4 phase of ~100 rad: ~3e-5 rounding
        c64 = O64.ctf2d(40, df1, df2, ast, apix, vol, cs, amp)
        assert np.abs(c64 - out).max() < 1e-4
        assert abs(c64[20, 20] + amp) < 1e-6                  # s = 0: ctf = -A


def test_mask_and_norm(O32, units):
    for img, p, out in zip(units["mask_in"], units["mask_par"], units["mask_out"]):
        assert np.abs(O32.mask(img, *p) - out).max() < 1e-5
    assert np.abs(O32.norm(units["norm_in"]) - units["norm_out"]).max() < 1e-5


def test_fft_convention(O64, units):
    a, b = units["fft2_in"]
    fa, fb = O64.fft2(a, b, 1, True)
    assert rel_l2(fa + 1j * fb, units["fft2_out"][0] + 1j * units["fft2_out"][1]) < 1e-6
    # the reference's backward call (conj_out applies to img2d_FFT's b = -aimag)
    ba, bb = O64.fft2(units["fft2_out"][0], units["fft2_out"][1], -1, True)
    assert rel_l2(ba + 1j * bb, units["fft2_out"][2] + 1j * units["fft2_out"][3]) < 1e-6


@pytest.mark.parametrize("n", [8, 12, 15])
def test_own_fft_against_brute_force_dft(O64, n):
    rng = np.random.default_rng(n)
    re, im = rng.standard_normal((2, n, n))
    br, bi = O64.dft2_brute(re, im, -1)
    fa, fb = O64.fft2(re, im, 1, False)
    assert rel_l2(fa + 1j * fb, (br + 1j * bi) / n ** 2) < 1e-12
    # Parseval with 1/N forward scaling
    assert abs(np.sum(fa ** 2 + fb ** 2) * n * n - np.sum(re ** 2 + im ** 2)) < 1e-9 * np.sum(re ** 2 + im ** 2)


def test_centshift(O32, units):
    a, b = units["shift_in"]
    x, y = units["shift_xy"]
    sa, sb = O32.ft_shift(a, b, x, y, True)
    assert np.abs(sa - units["shift_out"][0]).max() < 2e-5
    assert np.abs(sb - units["shift_out"][1]).max() < 2e-5


def test_shift_theorem(O64):
    """An integer real-space roll equals the phase ramp of img2d_FT_centshift."""
    rng = np.random.default_rng(4)
    n = 32
    img = rng.standard_normal((n, n))
    a, b = O64.fft2(O64.checker(img), np.zeros((n, n)), 1, True)
    rolled = np.roll(img, (2, 3), axis=(0, 1))                 # x += 3, y += 2
    ra, rb = O64.fft2(O64.checker(rolled), np.zeros((n, n)), 1, True)
    sa, sb = O64.ft_shift(ra, rb, 3.0, 2.0, False)
    assert rel_l2(sa + 1j * sb, a + 1j * b) < 1e-10


def test_amp_expfit(O32, O64, units):
    a, b = units["expfit_in"]
    mn, mx = units["expfit_r"]
    assert np.abs(O32.amp_expfit(a, b, mn, mx) / units["expfit_out"] - 1).max() < 1e-5
    assert np.abs(O64.amp_expfit(a, b, mn, mx) / units["expfit_out"] - 1).max() < 1e-4


def test_amp_expfit_recovers_exponential(O64):
    n, mn, mx = 64, 2, 24
    ky, hx = np.meshgrid(np.arange(n) - n // 2, np.arange(n) - n // 2, indexing="ij")
    r = np.sqrt(hx ** 2 + ky ** 2)
    amp = np.exp(-0.05 * r + 1.5)
    f = O64.amp_expfit(amp, np.zeros_like(amp), mn, mx)
    # ring means of exp(a r + b) are close to exp(a <r> + b): the fit returns 0.8x that
    rr = np.array([r[(np.floor(r + 0.5) == k) & (ky >= 0) & (np.abs(hx) <= mx)].mean() for k in range(mn, mx + 1)])
    assert np.abs(f / (0.8 * np.exp(-0.05 * rr + 1.5)) - 1).max() < 2e-3


# ---- phase residuals -------------------------------------------------------------

def _pr_inputs(O, g, actf):
    box, mn, mx, T = g["pr_box"]
    ta, tb, tf = g["pr_ta"], g["pr_tb"], g["pr_tfcut"]
    if actf:
        ta, tb, tf = pipelines.template_inputs(O, ta, tb, g["pr_ctf"], mn, mx)
    return mn, mx, ta, tb, tf


@pytest.mark.parametrize("bits", [32, 64])
def test_pr_omega_matches_reference(bits, O32, O64, prv):
    O = O32 if bits == 32 else O64
    step = float(prv["pr_step"][0])
    for ic, actf in enumerate((1, 0)):
        mn, mx, ta, tb, tf = _pr_inputs(O, prv, actf)
        for k, (th, ph, om, x, y) in enumerate(prv["pr_omega_cases"]):
            out = O.pr_omega(prv["pr_pa"], prv["pr_pb"], prv["pr_fc0"], ta, tb, tf, prv["pr_tang"], mn, mx, th,
                             ph, step)
            ref = prv["pr_omega_out"][ic, k]
            tol = 2e-3 if bits == 32 else 0.05
            assert np.abs(out - ref).max() < tol, (bits, actf, k, np.abs(out - ref).max())


def test_pr_identical_inputs_give_zero(O64, prv):
    """PR of a line set against itself is 0 and against its negation 180."""
    a, b = prv["pr_pa"], prv["pr_pb"]
    assert O64.pr_ft(a, b, a, b, -1, 20) < 1e-9
    assert abs(O64.pr_ft(a, b, -a, -b, -1, 20) - 180) < 1e-9
    rng = np.random.default_rng(0)
    ph = rng.uniform(-np.pi, np.pi, a.shape)
    amp = np.hypot(a, b)
    pr = O64.pr_ft(a, b, amp * np.cos(ph), amp * np.sin(ph), -1, 20)
    assert 80 < pr < 100                                       # random phases ~ 90
    z = np.zeros_like(a)
    assert O64.pr_ft(z, z, z, z, -1, 20) == 1000.0               # empty sentinel


@pytest.mark.parametrize("bits", [32, 64])
def test_crosslin_pr_matches_reference(bits, O32, O64, prv):
    O = O32 if bits == 32 else O64
    for ic, actf in enumerate((1, 0)):
        mn, mx, ta, tb, tf = _pr_inputs(O, prv, actf)
        for k, (th, ph, om, x, y) in enumerate(prv["pr_cases"]):
            out = O.crosslin_pr(prv["pr_a0"], prv["pr_b0"], prv["pr_fc0"], ta, tb, tf, prv["pr_tang"], mn, mx, th,
                                ph, om, x, y)
            tol = 2e-3 if bits == 32 else 0.05
            assert abs(out - prv["pr_out"][ic, k]) < tol


# ---- local_refine kernels ------------------------------------------------------------

def test_fft3d_slice(O32, O64, lref):
    fr, fi = lref["slice_in"]
    for ang, out in zip(lref["slice_ang"], lref["slice_out"]):
        a, b = O32.fft3d_slice(fr, fi, *ang)
        assert np.abs(a - out[0]).max() < 1e-5 and np.abs(b - out[1]).max() < 1e-5


def test_fft3d_slice_identity_is_central_plane(O64):
    rng = np.random.default_rng(1)
    n = 16
    fr, fi = rng.standard_normal((2, n, n, n))
    a, b = O64.fft3d_slice(fr, fi, 0, 0, 0)
    ky, hx = np.meshgrid(np.arange(n) - n // 2, np.arange(n) - n // 2, indexing="ij")
    inside = np.sqrt(hx ** 2 + ky ** 2) <= n // 2 - 2
    assert np.abs(a[inside] - fr[n // 2][inside]).max() < 1e-12
    assert np.all(a[~inside] == 0)


def test_pr_ft_and_ftphase(O32, lref):
    a1, b1, a2, b2 = lref["prft_in"]
    for r, out in zip(lref["prft_r"], lref["prft_out"]):
        assert abs(O32.pr_ft(a1, b1, a2, b2, *r) - out) < 2e-3
    sa, sb = O32.ft_shift(a1, b1, *lref["ftphase_xy"], recentre=False)
    assert np.abs(sa - lref["ftphase_out"][0]).max() < 2e-5 and np.abs(sb - lref["ftphase_out"][1]).max() < 2e-5


def test_softmask_and_fft3(O32, O64, lref):
    v = O32.softmask3d(lref["softmask_in"], lref["softmask_r"][0])
    assert np.abs(v - lref["softmask_out"]).max() < 1e-4
    a, b = O64.fft3(lref["fft3_in"][0], lref["fft3_in"][1], 1, False)
    assert rel_l2(a + 1j * b, lref["fft3_out"][0] + 1j * lref["fft3_out"][1]) < 1e-6


# ---- whole programs -------------------------------------------------------------------

def _prog_setup(prog):
    box = int(prog["prog_box"][0])
    apix = float(prog["prog_apix"][0])
    d = pipelines.derived(box, apix, maxRes=6.0, minRes=100.0, imgmask_A=40.0)
    return box, apix, d


@pytest.mark.parametrize("bits,tol", [(32, 1e-5), (64, 1e-4)])
def test_mode1_map_matches_reference(bits, tol, O32, O64, prog):
    O = O32 if bits == 32 else O64
    box, apix, d = _prog_setup(prog)
    vol = pipelines.mode1_volume(O, prog["prog_images"], prog["prog_angles"], prog["prog_shifts"], prog["prog_df"],
                                 apix, d["maxR"], d["mask"])
    m = O.fold_and_map(vol, d["maxR"])
    err = rel_l2(m, prog["mode1_map"])
    print(f"FP{bits} oracle vs reference mode-1 map: rel-L2 {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("bits,tol", [(32, 1e-5), (64, 1e-4)])
def test_mode1_substep_tnf_matches_reference(bits, tol, O32, O64, prog):
    O = O32 if bits == 32 else O64
    box, apix, d = _prog_setup(prog)
    assert list(prog["mode1sub_hdr"]) == [box, d["maxR"]]
    vol = pipelines.mode1_volume(O, prog["prog_images"], prog["prog_angles"], prog["prog_shifts"], prog["prog_df"],
                                 apix, d["maxR"], d["mask"], compact=True)
    for o, r in zip(vol, (prog["mode1sub_fa"], prog["mode1sub_fb"], prog["mode1sub_wg"])):
        assert rel_l2(o, r) < tol


def _mode4_oracle(O, prog, i, step=6.0):
    box, apix, d = _prog_setup(prog)
    ta, tb = pipelines.templates(O, prog["prog_timgs"], np.zeros((3, 2)), d["mask"])
    ctf = O.ctf2d(box, prog["prog_df"][i, 0], prog["prog_df"][i, 1], 0.0, apix)
    ta, tb, tf = pipelines.template_inputs(O, ta, tb, ctf, d["minR"], d["maxR"])
    x, y = prog["prog_shifts"][i]
    (a0, b0), (pa, pb), fc0 = pipelines.particle_inputs(O, prog["prog_images"][i], x, y, d["mask"], d["minR"],
                                                        d["maxR"])
    return d, (a0, b0), (pa, pb), fc0, ta, tb, tf


def test_mode4_search_matches_reference(O32, prog):
    ref = prog["mode4_out"]
    for i in range(len(ref)):
        d, _, (pa, pb), fc0, ta, tb, tf = _mode4_oracle(O32, prog, i)
        o4 = O32.globesearch(pa, pb, fc0, ta, tb, tf, prog["prog_tang"], d["minR"], d["maxR"], 6.0)
        assert np.allclose(o4[:3], ref[i, :3], atol=1e-3), (i, o4, ref[i])
        assert abs(o4[3] - ref[i, 5]) < 2e-3


def test_mode5_refinement_matches_reference(O32, prog):
    ref = prog["mode5_out"]
    for i in range(len(ref)):
        d, (a0, b0), _, fc0, ta, tb, tf = _mode4_oracle(O32, prog, i)
        start = np.array([*prog["prog_angles"][i], *prog["prog_shifts"][i]], np.float32)
        io5, pr = O32.refine_particle(a0, b0, fc0, ta, tb, tf, prog["prog_tang"], d["minR"], d["maxR"], 3,
                                      d["dang"], d["dtran"], start)
        assert np.allclose(io5, ref[i, :5], atol=2e-3), (i, io5, ref[i])
        assert abs(pr - ref[i, 5]) < 2e-3


# ---- local_refine program -----------------------------------------------------------

@pytest.fixture(scope="module")
def lprog():
    return np.load(os.path.join(G, "lref_program.npz"))


def test_lref_symmetry_generators(O32, O64, lref):
    for kind, n in ((0, 5), (1, 4), (2, 60), (3, 60)):
        ref = lref[f"lsym_{kind}_{n}"]
        assert np.abs(O32.lref_symops(kind, n) - ref).max() < 2e-6
        ops = O64.lref_symops(kind, n)
        for S in ops:
            assert abs(np.linalg.det(S) - 1) < 1e-12
    # I2 is the icosahedral group of icos2f_sym_mat in another element order
    a = O64.lref_symops(2, 60).reshape(60, 9)
    b = O64.icos_ops().reshape(60, 9)
    assert all(np.abs(b - x).max(axis=1).min() < 1e-9 for x in a)


def test_icos_equalort(O32, lref):
    th, ph, om = lref["equalort_in"]
    for kind, n in ((0, 5), (2, 60), (3, 60)):
        out = O32.icos_equalort(th, ph, om, O32.lref_symops(kind, n))
        d = (out - lref[f"equalort_{kind}_{n}"] + 180) % 360 - 180
        assert np.abs(d).max() < 2e-3


@pytest.mark.parametrize("tag", ["c1", "c5"])
def test_local_refine_program_matches_reference(tag, O32, lprog):
    box, box0 = (int(v) for v in lprog["lprog_box"])
    apix = float(lprog["lprog_apix"][0])
    d = pipelines.lref_derived(box, apix, maxRes=8.0, imgmask_A=24.0)
    fr, fi = pipelines.lref_model_ft(O32, lprog["lprog_model"], d["mask"])
    ops = O32.lref_symops(0, 5) if tag == "c5" else None
    ref = lprog[f"lprog_{tag}_out"]
    for i in range(len(ref)):
        x0, y0 = (round(float(v) * 0.5, 4) for v in lprog["lprog_shifts"][i])
        hx0, ky0 = lprog["lprog_cen"][i]
        a20, b20, w = pipelines.lref_particle(O32, lprog["lprog_images"][i], box, hx0, ky0, x0, y0, d["mask"],
                                              d["maxres"])
        df = float(lprog["lprog_df"][i])
        ctf = O32.ctf2d(box, df, df, 0.0, apix, vol=200.0, df_scale=1.0)
        start = np.array([*lprog["lprog_start"][i], x0, y0], np.float32)
        start[:3] = np.round(start[:3], 4)
        io5, pr = O32.lref_refine(fr, fi, a20, b20, ctf, w, ops, 3, d["refine_angle"], d["refine_shift"], 0.8,
                                  d["maxres"], start)
        da = (io5[:3] - ref[i, :3] + 180) % 360 - 180
        assert np.abs(da).max() < 2e-3 and np.abs(io5[3:5] - ref[i, 3:5]).max() < 2e-3, (i, io5, ref[i])
        assert abs(pr - ref[i, 6]) < 2e-3
