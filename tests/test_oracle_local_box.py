"""Oracle of local_box (coordinate and stack modes) against the reference
program's own outputs (tests/golden/lbox.npz: the 30 particles of the
48-pixel synthetic set, three of them above PR_threshold), and against
closed forms: the FFT shift is a circular roll for integer shifts and moves
a band-limited cosine by any shift; equalort's angles give back S R; the
centre of an operator-free box is R^T x0 in closed form."""
import numpy as np
import pytest

import pipelines

X0_I5, X0_V12, X0_STACK = (0.0, 0.0, 18.0), (1.5, -2.0, 17.5), (0.0, 0.0, 14.0)


@pytest.fixture(scope="module")
def g():
    return np.load(pipelines.golden_path("lbox.npz"))


def good(g, thr=90.0, bound=100.0):
    """load_ort.f90:29-33: PR <= PR_threshold, |x|, |y| <= boundX."""
    pr, sh = g["lbox_pr"], g["lbox_shifts"]
    return [i for i in range(len(pr)) if pr[i] <= thr and abs(sh[i, 0]) <= bound and abs(sh[i, 1]) <= bound]


def angdiff(a, b):
    d = np.abs(np.asarray(a, np.float64) - b) % 360.0
    return np.minimum(d, 360.0 - d)


def coords(O, g, ops, x0):
    return np.concatenate([O.local_box_coords(g["lbox_angles"][i], g["lbox_shifts"][i], ops, x0, 48, 2.0)
                           for i in good(g)])


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("case", ["i5", "v12"])
def test_coordinates_match_reference(bits, case, O32, O64, g):
    """loc_box_coordinate: the record columns after the stack name are
    particleID theta phi omega x y PR imgID imgparticleID df1e4 df2e4 astig |
    ii hx0 ky0 lth lph lom fx fy z | PR (format 900, 4 decimals)."""
    O = O32 if bits == 32 else O64
    ops, x0 = (O.local_box_ops("i5"), X0_I5) if case == "i5" else (g["lbox_v12_ops"], X0_V12)
    ref = g[f"lbox_coords_{case}"]
    nops = len(ops)
    c = coords(O, g, ops, x0)
    assert ref.shape == (len(good(g)) * nops, 22)
    ids = np.repeat(np.array(good(g)) + 1, nops)
    assert np.array_equal(ref[:, 0], ids) and np.array_equal(ref[:, 12], np.tile(np.arange(1, nops + 1), len(good(g))))
    assert np.array_equal(c[:, 3:5], ref[:, 13:15])
    assert angdiff(c[:, 0:3], ref[:, 15:18]).max() < 1.2e-4
    assert np.abs(c[:, 5:8] - ref[:, 18:21]).max() < 1.0e-4
    assert np.abs(ref[:, 9] - 1e4 * g["lbox_df"][good(g)][:, 0].repeat(nops)).max() < 0.06


@pytest.mark.parametrize("bits,tol", [(32, 1e-6), (64, 1e-6)])
def test_stack_matches_reference(bits, tol, O32, O64, g):
    """loc_box: the per-operator boxes (24 of a 48-pixel image) and the short
    ortfile columns Nparticle lth lph lom fx fy PR imgID ii df1 df2 astig z."""
    O = O32 if bits == 32 else O64
    ops = g["lbox_v12_ops"]
    B, M = [], []
    for i in good(g):
        b, m = O.local_box_stack(g["lbox_images"][i], g["lbox_angles"][i], g["lbox_shifts"][i], ops, X0_STACK, 24,
                                 2.0)
        B.append(b)
        M.append(m)
    B, M = np.concatenate(B), np.concatenate(M)
    ref, st = g["lbox_stack_ort"], g["lbox_stack"]
    assert np.array_equal(ref[:, 0], np.arange(1, len(ref) + 1))
    assert angdiff(M[:, 0:3], ref[:, 1:4]).max() < 1.2e-4
    assert np.abs(M[:, 3:5] - ref[:, 4:6]).max() < 1e-4
    assert np.abs(M[:, 5] - ref[:, 12]).max() < 1e-4
    err = np.linalg.norm(B - st) / np.linalg.norm(st)
    print(f"FP{bits} local_box stack vs reference: rel-L2 {err:.2e}")
    assert err < tol


def test_header_counts_every_particle(g):
    """loc_box.f90:24: nz = Nsym * totalparticle (all 30, not the 27 written)."""
    h = g["lbox_stack_hdr"]
    assert list(h[:3]) == [24, 24, 12 * 30] and g["lbox_stack"].shape[0] == 12 * len(good(g))


def test_operator_sets(O32, O64):
    """local_box's spellings: C10..C15 upper-case without '_', and an unknown
    name keeps Nsym = 60 for C_matrix; the generated sets are groups."""
    assert O64.local_box_sym("C12") == (0, 12) and O64.local_box_sym("C_12") == (0, 60)
    assert O64.local_box_sym("I3") == (3, 60) and O64.local_box_sym("d5") == (0, 60)
    for sym in ("i2", "i3", "i5", "c7", "x"):
        ops = O64.local_box_ops(sym)
        n = len(ops)
        prods = np.einsum("aij,bjk->abik", ops, ops).reshape(-1, 9)
        flat = ops.reshape(n, 9)
        d = np.abs(prods[:, None, :] - flat[None]).max(-1).min(-1)
        assert d.max() < 1e-9, sym


@pytest.mark.parametrize("n", [48, 30])
def test_realshift_integer_is_a_roll(O64, n):
    a = np.random.default_rng(n).normal(size=(n, n))
    out = O64.img2d_realshift(a, 3.0, -5.0)
    assert np.abs(out - np.roll(a, (-5, 3), axis=(0, 1))).max() < 1e-10


def test_realshift_moves_a_band_limited_cosine(O64):
    n, kx, ky, sx, sy = 48, 5, -3, 0.37, -1.81
    y, x = np.mgrid[0:n, 0:n]
    a = np.cos(2 * np.pi * (kx * x + ky * y) / n)
    out = O64.img2d_realshift(a, sx, sy)
    assert np.abs(out - np.cos(2 * np.pi * (kx * (x - sx) + ky * (y - sy)) / n)).max() < 1e-10


def test_equalort_angles_rebuild_the_product(O64):
    """E(equalort(S, angles)) = S E(angles) (theta in (0, 180))."""
    rng = np.random.default_rng(3)
    ops = O64.local_box_ops("i2")
    for _ in range(5):
        ang = (rng.uniform(5, 175), rng.uniform(-180, 180), rng.uniform(-180, 180))
        c = O64.local_box_coords(ang, (0.0, 0.0), ops, (0.0, 0.0, 0.0), 48, 1.0)
        E = O64.euler_mat(*ang)
        for k in range(60):
            assert np.abs(O64.euler_mat(*c[k, 0:3]) - ops[k] @ E).max() < 1e-9


def test_identity_centre_closed_form(O64):
    """S = 1: c = E^T x0 + box/2 + (x, y); for x0 = (0, 0, z) the centre is
    box/2 + x - z sin(theta) cos(omega), box/2 + y + z sin(theta) sin(omega),
    and the height z cos(theta)."""
    th, ph, om, x, y, z, n, apix = 63.0, 21.0, 144.0, 1.25, -0.75, 17.0, 48, 1.7
    c = O64.local_box_coords((th, ph, om), (x, y), np.eye(3)[None], (0.0, 0.0, z), n, apix)[0]
    t, o = np.radians(th), np.radians(om)
    cx = n / 2 + x - z * np.sin(t) * np.cos(o)
    cy = n / 2 + y + z * np.sin(t) * np.sin(o)
    assert c[3] == int(cx) and c[4] == int(cy)
    assert abs(c[5] - (cx - int(cx))) < 1e-9 and abs(c[6] - (cy - int(cy))) < 1e-9
    assert abs(c[7] - z * np.cos(t) * apix) < 1e-9


def test_stack_box_is_the_normalised_centre_cut(O64):
    """S = 1, x0 = 0, no shift: the box is the centred sub x sub cut of the
    image, normalised to mean 0 and rms 1."""
    n, sub = 32, 12
    a = np.random.default_rng(9).normal(size=(n, n)) * 3 + 2
    b, m = O64.local_box_stack(a, (40.0, 10.0, 20.0), (0.0, 0.0), np.eye(3)[None], (0.0, 0.0, 0.0), sub, 1.0)
    cut = a[n // 2 - sub // 2:n // 2 + sub // 2, n // 2 - sub // 2:n // 2 + sub // 2]
    ref = (cut - cut.mean()) / cut.std()
    assert np.abs(b[0] - ref).max() < 1e-9
