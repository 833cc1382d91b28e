"""Oracle of reconstruct_corrCTF (mode 1 with a defocus series) against the
reference program's outputs (tests/golden/corrctf.npz: the key-patched
icosprocess with corrCTF=0.2,2) and against what the arithmetic fixes."""
import numpy as np
import pytest

import pipelines
from conftest import rel_l2

TNF = ("fa", "fb", "wg")


@pytest.fixture(scope="module")
def cc():
    return np.load(pipelines.golden_path("corrctf.npz"))


def _setup(prog, cc):
    box, apix = int(prog["prog_box"][0]), float(prog["prog_apix"][0])
    step, cyc = float(cc["corr_step_cycles"][0]), int(cc["corr_step_cycles"][1])
    return box, apix, pipelines.derived(box, apix, maxRes=6.0, imgmask_A=40), step, cyc


@pytest.mark.parametrize("bits,tol", [(32, 1e-5), (64, 1e-4)])
def test_corrctf_map_matches_reference(bits, tol, O32, O64, prog, cc):
    O = O32 if bits == 32 else O64
    box, apix, d, step, cyc = _setup(prog, cc)
    vol = pipelines.corrctf_volume(O, prog["prog_images"], prog["prog_angles"], prog["prog_shifts"],
                                   prog["prog_df"], apix, d["maxR"], d["mask"], step, cyc)
    m = O.fold_and_map(vol, d["maxR"])
    err = rel_l2(m, cc["corr_map"])
    print(f"FP{bits} corrCTF map vs reference: rel-L2 {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("bits,tol", [(32, 1e-5), (64, 1e-4)])
def test_corrctf_substep_matches_reference(bits, tol, O32, O64, prog, cc):
    O = O32 if bits == 32 else O64
    box, apix, d, step, cyc = _setup(prog, cc)
    assert list(cc["corrsub_hdr"]) == [box, d["maxR"], 0, 0]
    vol = pipelines.corrctf_volume(O, prog["prog_images"], prog["prog_angles"], prog["prog_shifts"],
                                   prog["prog_df"], apix, d["maxR"], d["mask"], step, cyc, compact=True)
    for o, k in zip(vol, TNF):
        err = rel_l2(o, cc[f"corrsub_{k}"])
        print(f"FP{bits} corrCTF substep {k}: rel-L2 {err:.2e}")
        assert err < tol


def test_zero_cycles_is_the_plain_insert(O64, prog):
    """cycles = 0: one defocus, and the weight's ctf2d is that same CTF."""
    box, apix = int(prog["prog_box"][0]), float(prog["prog_apix"][0])
    d = pipelines.derived(box, apix, maxRes=6.0, imgmask_A=40)
    idx = slice(0, 4)
    a = pipelines.corrctf_volume(O64, prog["prog_images"][idx], prog["prog_angles"][idx], prog["prog_shifts"][idx],
                                 prog["prog_df"][idx], apix, d["maxR"], d["mask"], 0.3, 0, compact=True)
    vol = O64.new_volume(box, d["maxR"], compact=True, wg_init=0.0)
    for i in range(4):
        x, y = prog["prog_shifts"][i]
        fa, fb = O64.particle_ft(prog["prog_images"][i], x, y, mask_r=d["mask"],
                                 mask_xy=tuple(prog["prog_shifts"][0]), norm=True)
        c = O64.ctf2d(box, prog["prog_df"][i, 0], prog["prog_df"][i, 1], 0.0, apix)
        O64.insert(vol, fa, fb, c, prog["prog_angles"][i], d["maxR"], rcut=d["maxR"], compact=True)
    for x, y in zip(a, vol):
        assert np.abs(x - y).max() <= 1e-12 * max(1.0, np.abs(y).max())


def test_series_sums_are_linear_in_the_ctf(O64, prog):
    """fa, fb take sum_k ctf_k per sample and wg takes (2C+1) ctf_last^2:
    the series insert equals the plain insert with those two CTF images."""
    box, apix = int(prog["prog_box"][0]), float(prog["prog_apix"][0])
    d = pipelines.derived(box, apix, maxRes=6.0, imgmask_A=40)
    i = 3
    x, y = prog["prog_shifts"][i]
    fa, fb = O64.particle_ft(prog["prog_images"][i], x, y, mask_r=d["mask"], norm=True)
    c3, last = pipelines.corrctf_ctfs(O64, box, prog["prog_df"][i, 0], prog["prog_df"][i, 1], apix, 0.2, 2)
    s = O64.new_volume(box, d["maxR"], compact=True, wg_init=0.0)
    O64.insert_series(s, fa, fb, c3, last, prog["prog_angles"][i], d["maxR"], compact=True)
    p = O64.new_volume(box, d["maxR"], compact=True, wg_init=0.0)
    O64.insert(p, fa, fb, c3.sum(0), prog["prog_angles"][i], d["maxR"], compact=True)
    w = O64.new_volume(box, d["maxR"], compact=True, wg_init=0.0)
    O64.insert(w, fa, fb, np.sqrt(5.0) * np.abs(last), prog["prog_angles"][i], d["maxR"], compact=True)
    assert rel_l2(s[0], p[0]) < 1e-12 and rel_l2(s[1], p[1]) < 1e-12
    assert rel_l2(s[2], w[2]) < 1e-12
