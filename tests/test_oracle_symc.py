"""Oracle of icosprocess mode 1 with a cyclic sym (reconstruct_C1,
reconstruct_C1_substep, and img3d_icosavg over the group) against the
reference program's outputs (tests/golden/symc.npz)."""
import numpy as np
import pytest

import pipelines
from conftest import rel_l2


@pytest.fixture(scope="module")
def sc():
    return np.load(pipelines.golden_path("symc.npz"))


def _d(prog):
    box, apix = int(prog["prog_box"][0]), float(prog["prog_apix"][0])
    return box, apix, pipelines.derived(box, apix, maxRes=6.0, imgmask_A=40)


def _args(prog):
    return prog["prog_images"], prog["prog_angles"], prog["prog_shifts"], prog["prog_df"]


def test_c7_operators_are_the_closed_form(O32):
    """icos2f_sym_matc7: rotations about z by k 360/7 (float32 angles)."""
    ops = O32.icosprocess_local_ops("c7")
    for k in range(7):
        a = np.deg2rad(k * 360.0 / 7.0)
        assert np.allclose(ops[k], [[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]], atol=3e-7)


@pytest.mark.parametrize("sym,key", [("c5a", "c5a_map"), ("c7", "c7_map")])
@pytest.mark.parametrize("bits,tol", [(32, 1e-5), (64, 1e-4)])
def test_c1_map_matches_reference(bits, tol, O32, O64, prog, sc, sym, key):
    O = O32 if bits == 32 else O64
    box, apix, d = _d(prog)
    vol = pipelines.c1_volume(O, *_args(prog), apix, d["maxR"], d["mask"], sym)
    err = rel_l2(O.fold_and_map(vol, d["maxR"]), sc[key])
    print(f"FP{bits} reconstruct_C1 {sym} map vs reference: rel-L2 {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("bits,tol", [(32, 1e-5), (64, 1e-4)])
def test_c1_substep_matches_reference(bits, tol, O32, O64, prog, sc):
    O = O32 if bits == 32 else O64
    box, apix, d = _d(prog)
    mcs = float(sc["c5asub_maxcentshift"][0])
    kept = np.sqrt((prog["prog_shifts"].astype(np.float32) ** 2).sum(1)) <= np.float32(mcs)
    assert 0 < kept.sum() < len(kept)          # the filter drops some particles
    vol = pipelines.c1_volume(O, *_args(prog), apix, d["maxR"], d["mask"], "c5a", substep=True, maxcentshift=mcs)
    for o, k in zip(vol, ("fa", "fb", "wg")):
        assert rel_l2(o, sc[f"c5asub_{k}"]) < tol


@pytest.mark.parametrize("bits,tol", [(32, 1e-5), (64, 1e-4)])
def test_group_average_matches_reference(bits, tol, O32, O64, sc):
    """reconstruct_C1.f90:286: img3d_icosavg with sym = c5a averages over
    that group's 5 operators (radius imgmask)."""
    O = O32 if bits == 32 else O64
    out = O.icosavg_full(sc["c5a_map"], 20.0, ops=O.icosprocess_local_ops("c5a"))
    err = rel_l2(out, sc["c5a_avg_map"])
    print(f"FP{bits} c5a real-space average vs reference: rel-L2 {err:.2e}")
    assert err < tol
