"""Oracle of the half-map / FSC / tnf-merge path against the reference's own
outputs (tests/golden/fsc.npz: reconstruct_half + reconstruct_FSC run by
icosprocess with FSC=y, and merge_icosprocess_subrec on its half and partial
tnf files), plus properties the mathematics fixes."""
import numpy as np
import pytest

import pipelines
from conftest import rel_l2

TNF = ("fa", "fb", "wg")


@pytest.fixture(scope="module")
def g():
    return np.load(pipelines.golden_path("fsc.npz"))


def _tnf(g, key):
    return [g[f"{key}_{k}"] for k in TNF]


def _setup(prog):
    box = int(prog["prog_box"][0])
    apix = float(prog["prog_apix"][0])
    return box, apix, pipelines.derived(box, apix, maxRes=6.0, imgmask_A=40)


@pytest.mark.parametrize("bits,tol", [(32, 1e-5), (64, 2e-5)])
def test_half_maps_match_reference(bits, tol, O32, O64, prog, g):
    """reconstruct_half.f90: compact sums, rcut = maxR, particles first, first+2, ..
    then first+1, ..; mask at particle(first) for both halves."""
    O = O32 if bits == 32 else O64
    box, apix, d = _setup(prog)
    assert list(g["half_odd_hdr"]) == [box, d["maxR"], 0, 0]
    for par, key in ((0, "half_odd"), (1, "half_eve")):
        vol = pipelines.half_volume(O, prog["prog_images"], prog["prog_angles"], prog["prog_shifts"],
                                    prog["prog_df"], apix, d["maxR"], d["mask"], par)
        for o, r in zip(vol, _tnf(g, key)):
            assert rel_l2(o, r) < tol


@pytest.mark.parametrize("bits", [32, 64])
def test_fsc_and_map_match_reference(bits, O32, O64, g):
    O = O32 if bits == 32 else O64
    box, maxR = int(g["half_odd_hdr"][0]), int(g["half_odd_hdr"][1])
    f, m = O.fsc_map(_tnf(g, "half_odd"), _tnf(g, "half_eve"), box, maxR)
    ref = g["fsc_txt"]
    assert np.array_equal(ref[:, 0], np.arange(maxR + 1))
    assert np.abs(f - ref[:, 1]).max() < 2e-6
    err = rel_l2(m, g["fsc_map"])
    print(f"FP{bits} oracle reconstruct_FSC map vs reference: rel-L2 {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("bits", [32, 64])
def test_merge_fsc_matches_reference(bits, O32, O64, g):
    O = O32 if bits == 32 else O64
    box, maxR = int(g["p1_odd_hdr"][0]), int(g["p1_odd_hdr"][1])
    assert int(g["p1_odd_hdr"][2]) == 1          # subrec=y halves carry nFSC = 1
    odd = [a.astype(np.float32) + b for a, b in zip(_tnf(g, "p1_odd"), _tnf(g, "p2_odd"))]
    eve = [a.astype(np.float32) + b for a, b in zip(_tnf(g, "p1_eve"), _tnf(g, "p2_eve"))]
    f, m = O.fsc_map(odd, eve, box, maxR, imgmask=float(g["p1_odd_hdrf"][1]), merge=True)
    assert np.abs(f - g["mfsc_txt"][:, 1]).max() < 2e-6
    assert m.shape == (box - 8,) * 3
    err = rel_l2(m, g["mfsc_map"])
    print(f"FP{bits} oracle merge_reconstruct_FSC map vs reference: rel-L2 {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("bits", [32, 64])
def test_merge_plain_matches_reference(bits, O32, O64, g):
    O = O32 if bits == 32 else O64
    box, maxR = int(g["s1_hdr"][0]), int(g["s1_hdr"][1])
    m = O.merge_map([_tnf(g, "s1"), _tnf(g, "s2")], box, maxR)
    err = rel_l2(m, g["merge_map"])
    print(f"FP{bits} oracle merge_reconstruct map vs reference: rel-L2 {err:.2e}")
    assert err < 1e-5


def _random_halves(rng, maxR, rho=0.5):
    n = (2 * maxR + 1) ** 2 * (maxR + 1)
    s = [rng.standard_normal(n) for _ in range(2)]
    w = [rng.uniform(0.5, 3.0, n) for _ in range(2)]
    odd = [s[0] + rho * rng.standard_normal(n), s[1] + rho * rng.standard_normal(n), w[0]]
    eve = [s[0] + rho * rng.standard_normal(n), s[1] + rho * rng.standard_normal(n), w[1]]
    return odd, eve


def test_fsc_identical_and_opposite_halves(O64):
    """FSC = 1 for identical halves and -1 for negated ones (weights then 0)."""
    rng = np.random.default_rng(1)
    maxR = 6
    odd, _ = _random_halves(rng, maxR)
    f, fa, fb, wg = O64.tnf_fsc(odd, odd, maxR)
    assert np.allclose(f, 1.0, atol=1e-9)
    f, fa, fb, wg = O64.tnf_fsc(odd, [-odd[0], -odd[1], odd[2]], maxR)
    assert np.allclose(f, -1.0, atol=1e-9)
    T = 2 * maxR + 1
    ir = _shells(maxR)
    inside = ir <= maxR
    assert np.all(fa.reshape(maxR + 1, T, T)[inside] == 0)          # weight 0 in every shell
    assert np.allclose(wg.reshape(maxR + 1, T, T)[inside], 1.0)     # (w1 + w2 - 2) * 0 + 1


def _shells(maxR):
    lz, ky, hx = np.meshgrid(np.arange(maxR + 1), np.arange(-maxR, maxR + 1), np.arange(-maxR, maxR + 1),
                             indexing="ij")
    return np.sqrt((hx * hx + ky * ky + lz * lz).astype(np.float32)).astype(np.int32)


def test_fsc_weighting_zeroes_run_ends(O64):
    """reconstruct_FSC.f90:82-84 zeroes a, b at the stale index: exactly the
    in-sphere voxels whose successor in (lz, ky, hx) order is outside."""
    rng = np.random.default_rng(2)
    maxR = 7
    odd, eve = _random_halves(rng, maxR, rho=0.2)
    f, fa, fb, wg = O64.tnf_fsc(odd, eve, maxR)
    assert np.all(f > 0.5)
    inside = (_shells(maxR) <= maxR).reshape(-1)
    succ_out = np.append(~inside[1:], True)
    zeroed = inside & succ_out
    assert zeroed.sum() > 0
    assert np.all(fa[zeroed] == 0) and np.all(fb[zeroed] == 0)
    assert np.all(fa[inside & ~succ_out] != 0)
    # the weights are kept there
    assert np.all(wg[zeroed] > 1.0)


def test_merge_crop_takes_next_plane_for_plane_ends(O64):
    nx = 14
    M = nx - 8
    m = np.random.default_rng(3).standard_normal((nx, nx, nx))
    out = np.zeros(M ** 3)
    O64.L.o_merge_crop(O64._p(m), nx, O64._p(out))
    out = out.reshape(M, M, M)
    straight = m[4:4 + M, 4:4 + M, 4:4 + M]
    diff = out != straight
    assert diff.sum() == M                          # one voxel per plane
    for z in range(M):
        assert out[z, M - 1, M - 1] == m[z + 5, M + 3, M + 3]


def test_integer_shell_fold_keeps_the_outer_shell(O64):
    """merge_reconstruct.f90:42 (integer dr) keeps voxels with maxR < r < maxR+1
    that reconstruct.f90's real test drops."""
    nx, maxR = 16, 5
    rng = np.random.default_rng(4)
    vol = (rng.standard_normal((nx,) * 3), rng.standard_normal((nx,) * 3), rng.uniform(1, 2, (nx,) * 3))
    a = O64.fold_and_map(vol, maxR, int_radius=False)
    b = O64.fold_and_map(vol, maxR, int_radius=True)
    fa = np.fft.fftn(b) - np.fft.fftn(a)
    assert rel_l2(a, b) > 1e-3 and np.abs(fa).max() > 0


# ---- icosprocess mode 6 (local reconstruction) --------------------------------

@pytest.fixture(scope="module")
def loc():
    return np.load(pipelines.golden_path("local.npz"))


def _local_setup(prog, loc):
    box, apix = int(prog["prog_box"][0]), float(prog["prog_apix"][0])
    lb = int(loc["loc_box"][0])
    d = pipelines.derived(lb, apix, maxRes=6.0, imgmask_A=24)   # calc_initialvalues uses loc_box in mode 6
    return lb, apix, d


def test_local_symmetry_operators(O64):
    """icos2f_sym_matc*: closed groups of z rotations of order N."""
    for sym, (N, step) in O64.LOCAL_SYM.items():
        ops = O64.icosprocess_local_ops(sym)
        assert ops.shape == (N, 3, 3)
        for k in range(N):
            a = np.deg2rad(k * 360.0 / N)
            assert np.allclose(ops[k], [[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]], atol=2e-7)


@pytest.mark.parametrize("bits,tol", [(32, 1e-5), (64, 1e-4)])
def test_local_map_matches_reference(bits, tol, O32, O64, prog, loc):
    O = O32 if bits == 32 else O64
    lb, apix, d = _local_setup(prog, loc)
    assert int(loc["locsub_hdr"][1]) == d["maxR"]
    vol = pipelines.local_volume(O, prog["prog_images"], prog["prog_angles"], prog["prog_shifts"], prog["prog_df"],
                                 loc["loc_x0y0"], apix, d["maxR"], d["mask"], lb, "c5a")
    m = O.fold_and_map(vol, d["maxR"])
    err = rel_l2(m, loc["loc_map"])
    print(f"FP{bits} oracle vs reference mode-6 map: rel-L2 {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("bits,tol", [(32, 1e-5), (64, 1e-4)])
def test_local_substep_tnf_matches_reference(bits, tol, O32, O64, prog, loc):
    O = O32 if bits == 32 else O64
    lb, apix, d = _local_setup(prog, loc)
    vol = pipelines.local_volume(O, prog["prog_images"], prog["prog_angles"], prog["prog_shifts"], prog["prog_df"],
                                 loc["loc_x0y0"], apix, d["maxR"], d["mask"], lb, "c5a", compact=True)
    for o, k in zip(vol, TNF):
        assert rel_l2(o, loc[f"locsub_{k}"]) < tol
