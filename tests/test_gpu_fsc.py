"""HIP half maps, FSC and tnf merges against the reference's own outputs
(tests/golden/fsc.npz) and against the FP64 oracle at a larger size."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pipelines
from conftest import rel_l2
from cryoim_amd import formats, lib

pytestmark = pytest.mark.gpu
TNF = ("fa", "fb", "wg")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MERGE_EXE = os.path.join(ROOT, "cryoim_amd", "bin", "merge_icosprocess_subrec_hip")


@pytest.fixture(scope="module")
def g():
    return np.load(pipelines.golden_path("fsc.npz"))


def _tnf(g, key):
    return [g[f"{key}_{k}"] for k in TNF]


def _records(prog, idx, mask_xy):
    P = np.zeros((len(idx), lib.PART_STRIDE), np.float32)
    P[:, 0:3] = prog["prog_angles"][idx]
    P[:, 3:5] = prog["prog_shifts"][idx]
    P[:, 6:8] = prog["prog_df"][idx]
    P[:, 9:11] = mask_xy
    P[:, 11] = 1.0
    return P


def test_half_maps_match_reference(ctx, prog, g):
    """reconstruct_half.f90: compact layout with rcut = maxR (its flat-index
    wrap included), odd then even particles, mask at the first particle."""
    box, apix = int(prog["prog_box"][0]), float(prog["prog_apix"][0])
    d = pipelines.derived(box, apix, maxRes=6.0, imgmask_A=40)
    maxR = d["maxR"]
    p = lib.make_params(box, 0, maxR, apply_ctf=True, mask_on=True, norm_on=True, apix=apix, imgmask=d["mask"])
    n = len(prog["prog_images"])
    for par, key in ((0, "half_odd"), (1, "half_eve")):
        idx = np.arange(par, n, 2)
        ctx.rec_begin(box, maxR, layout=1, rcut=maxR, wg_init=0.0)
        ctx.rec_insert(p, prog["prog_images"][idx], _records(prog, idx, prog["prog_shifts"][0]))
        out = ctx.rec_get_tnf(maxR)
        for o, r in zip(out, _tnf(g, key)):
            err = rel_l2(o, r)
            print(f"{key}: rel-L2 {err:.2e}")
            assert err < 2e-5


def test_fsc_map_reconstruct_variant(ctx, g):
    box, maxR = int(g["half_odd_hdr"][0]), int(g["half_odd_hdr"][1])
    f, m = ctx.fsc_map(box, maxR, _tnf(g, "half_odd"), _tnf(g, "half_eve"))
    assert np.abs(f - g["fsc_txt"][:, 1]).max() < 2e-6
    err = rel_l2(m, g["fsc_map"])
    print(f"reconstruct_FSC map vs reference: rel-L2 {err:.2e}")
    assert err < 1e-5


def test_fsc_map_merge_variant(ctx, g):
    box, maxR = int(g["p1_odd_hdr"][0]), int(g["p1_odd_hdr"][1])
    odd = [a.astype(np.float32) + b for a, b in zip(_tnf(g, "p1_odd"), _tnf(g, "p2_odd"))]
    eve = [a.astype(np.float32) + b for a, b in zip(_tnf(g, "p1_eve"), _tnf(g, "p2_eve"))]
    f, m = ctx.fsc_map(box, maxR, odd, eve, imgmask=float(g["p1_odd_hdrf"][1]), merge=True)
    assert np.abs(f - g["mfsc_txt"][:, 1]).max() < 2e-6
    err = rel_l2(m, g["mfsc_map"])
    print(f"merge_reconstruct_FSC map vs reference: rel-L2 {err:.2e}")
    assert err < 1e-5


def test_merge_map_matches_reference(ctx, g):
    box, maxR = int(g["s1_hdr"][0]), int(g["s1_hdr"][1])
    ctx.rec_begin(box, maxR, layout=1, rcut=maxR, wg_init=0.01)
    for key in ("s1", "s2"):
        ctx.rec_add_tnf(*_tnf(g, key))
    m = ctx.merge_map(box)
    err = rel_l2(m, g["merge_map"])
    print(f"merge_reconstruct map vs reference: rel-L2 {err:.2e}")
    assert err < 1e-5


@pytest.fixture(scope="module")
def avg():
    return np.load(pipelines.golden_path("icosavg.npz"))


def test_icosavg_full_matches_reference(ctx, avg):
    """icosprocess mode 1 with realspaceavg=y (reconstruct.f90:253-255)."""
    out = ctx.icosavg(avg["avg_input_map"], 20.0, lib.ICOSAVG_FULL)
    err = rel_l2(out, avg["avg_map"])
    print(f"img3d_icosavg vs reference: rel-L2 {err:.2e}")
    assert err < 1e-5


def test_group_average_matches_reference(ctx):
    """reconstruct_C1.f90:286: img3d_icosavg over the c5a group's 5 operators."""
    sc = np.load(pipelines.golden_path("symc.npz"))
    out = ctx.icosavg(sc["c5a_map"], 20.0, lib.ICOSAVG_FULL, ops=lib.icosprocess_sym_ops("c5a"))
    err = rel_l2(out, sc["c5a_avg_map"])
    print(f"c5a real-space average vs reference: rel-L2 {err:.2e}")
    assert err < 1e-5


def test_merges_with_icosavg_match_reference(ctx, g, avg):
    """nicosavg = 1: merge_reconstruct, merge_reconstruct_FSC (average of the
    first (box-8)^3 values as a cube) and icosprocess's reconstruct_FSC."""
    box, maxR, mask = int(g["s1_hdr"][0]), int(g["s1_hdr"][1]), float(g["s1_hdrf"][1])
    ctx.rec_begin(box, maxR, layout=1, rcut=maxR, wg_init=0.01)
    for key in ("s1", "s2"):
        ctx.rec_add_tnf(*_tnf(g, key))
    m = ctx.merge_map(box, icosavg=True, imgmask=mask)
    e1 = rel_l2(m, avg["mavg_map"])
    odd = [a.astype(np.float32) + b for a, b in zip(_tnf(g, "p1_odd"), _tnf(g, "p2_odd"))]
    eve = [a.astype(np.float32) + b for a, b in zip(_tnf(g, "p1_eve"), _tnf(g, "p2_eve"))]
    _, m = ctx.fsc_map(box, maxR, odd, eve, imgmask=mask, merge=True, icosavg=True)
    e2 = rel_l2(m, avg["mfscavg_map"])
    _, m = ctx.fsc_map(box, maxR, _tnf(g, "half_odd"), _tnf(g, "half_eve"), imgmask=mask, icosavg=True)
    e3 = rel_l2(m, avg["fscavg_map"])
    print(f"averaged merge {e1:.2e}, merge FSC {e2:.2e}, reconstruct_FSC {e3:.2e}")
    assert max(e1, e2, e3) < 1e-5


@pytest.mark.parametrize("variant", [lib.ICOSAVG_FULL, lib.ICOSAVG_MERGE])
def test_icosavg_vs_fp64_oracle_box128(ctx, O64, variant):
    """Box 128.  The merge variant runs as merge_reconstruct_FSC calls it: a
    120^3 cube over the first values of the 128^3 map, radius 62 >= 60, so its
    taps read past the cube and its hx = -60 copies wrap to the next row."""
    box = 128
    rng = np.random.default_rng(41)
    v = rng.normal(size=(box, box, box)).astype(np.float32)
    if variant == lib.ICOSAVG_FULL:
        n, r0 = box, 60.0
        ref = O64.icosavg_full(v, r0)
    else:
        n, r0 = box - 8, 62.0
        ref = O64.icosavg_merge(v, r0, n=n)
    out = ctx.icosavg(v, r0, variant, n=n)
    m3 = n ** 3
    err = np.abs(out.reshape(-1)[:m3] - ref.reshape(-1)[:m3]).max()
    print(f"variant {variant}: max |d| {err:.2e} on a unit-variance map")
    assert err < 2e-5
    assert np.array_equal(out.reshape(-1)[m3:], v.reshape(-1)[m3:])    # values past n^3 untouched


@pytest.mark.parametrize("merge", [False, True])
def test_fsc_map_vs_fp64_oracle_box96(ctx, O64, merge):
    box, maxR = 96, 40
    rng = np.random.default_rng(7)
    n = (2 * maxR + 1) ** 2 * (maxR + 1)
    # signal decaying with radius, independent noise per half: FSC falls across the shells
    lz, ky, hx = np.meshgrid(np.arange(maxR + 1), np.arange(-maxR, maxR + 1), np.arange(-maxR, maxR + 1),
                             indexing="ij")
    r = np.sqrt(hx ** 2 + ky ** 2 + lz ** 2).reshape(-1)
    sig = np.exp(-r / 12.0)
    s = [sig * rng.standard_normal(n) for _ in range(2)]
    halves = []
    for _ in range(2):
        w = rng.uniform(1.0, 4.0, n)
        halves.append([((s[0] + 0.05 * rng.standard_normal(n)) * w).astype(np.float32),
                       ((s[1] + 0.05 * rng.standard_normal(n)) * w).astype(np.float32), w.astype(np.float32)])
    f, m = ctx.fsc_map(box, maxR, halves[0], halves[1], imgmask=40.0, merge=merge)
    fo, mo = O64.fsc_map(halves[0], halves[1], box, maxR, imgmask=40.0, merge=merge)
    assert f.max() > 0.9 and f.min() < 0.5
    assert np.abs(f - fo).max() < 1e-5
    err = rel_l2(m, mo)
    print(f"box 96 FSC map ({'merge' if merge else 'reconstruct'}) vs FP64 oracle: rel-L2 {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("compact", [False, True])
def test_local_reconstruction_matches_reference(ctx, prog, compact):
    """icosprocess mode 6 through the C ABI: loc_box windows cut around
    (x0, y0), the c5a local operators, weights starting at 0.01, mask centred on
    each particle's own shift; full loc_box^3 map (reconstruction_local.f90) or
    compact tnf with rcut maxR-1 (reconstruction_local_substep.f90)."""
    loc = np.load(pipelines.golden_path("local.npz"))
    apix, lb = float(prog["prog_apix"][0]), int(loc["loc_box"][0])
    d = pipelines.derived(lb, apix, maxRes=6.0, imgmask_A=24)
    maxR = d["maxR"]
    ops = lib.icosprocess_sym_ops("c5a")
    assert ops.shape == (5, 3, 3)
    wins = pipelines.local_windows(prog["prog_images"], loc["loc_x0y0"], lb)
    p = lib.make_params(lb, 0, maxR, apply_ctf=True, mask_on=True, norm_on=True, apix=apix, imgmask=d["mask"])
    n = len(wins)
    ctx.rec_begin(lb, maxR, layout=int(compact), rcut=maxR - 1 if compact else maxR, wg_init=0.01, ops=ops)
    ctx.rec_insert(p, wins, _records(prog, np.arange(n), prog["prog_shifts"]))
    if compact:
        for o, k in zip(ctx.rec_get_tnf(maxR), TNF):
            err = rel_l2(o, loc[f"locsub_{k}"])
            print(f"local substep {k}: rel-L2 {err:.2e}")
            assert err < 2e-5
    else:
        m = ctx.rec_finish(lb)
        err = rel_l2(m, loc["loc_map"])
        print(f"local map: rel-L2 {err:.2e}")
        assert err < 2e-5


def test_rec_begin_rejects_non_rotation_ops(ctx):
    ops = np.eye(3)[None] * 1.02
    with pytest.raises(lib.CimError):
        ctx.rec_begin(64, 20, layout=1, rcut=20, wg_init=0.0, ops=ops)


def _run(exe, d, *args):
    r = subprocess.run([exe, *args], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    return r.stdout


def _write_tnf(path, g, key, nicosavg):
    h, hf = g[key + "_hdr"], g[key + "_hdrf"]
    formats.write_tnf(path, int(h[0]), int(h[1]), float(hf[0]), float(hf[1]), *_tnf(g, key), fsc=int(h[2]),
                      realspaceavg=nicosavg)


def test_fortran_merge_driver_with_icosavg(g, avg):
    """realspaceavg=y tnf input (nicosavg = 1) through the merge driver."""
    if not os.path.exists(MERGE_EXE):
        pytest.skip("Fortran merge driver not built (amdflang missing)")
    d = tempfile.mkdtemp(prefix="cimavg_")
    for key, name in (("s1", "s1.tnf"), ("s2", "s2.tnf")):
        _write_tnf(os.path.join(d, name), g, key, nicosavg=1)
    _run(MERGE_EXE, d, "s1.tnf", "s2.tnf", "m0.mrc")
    m = np.fromfile(os.path.join(d, "m0.mrc"), dtype="<f4", offset=1024).reshape((40,) * 3)
    assert rel_l2(m, avg["mavg_map"]) < 1e-5
    for p in ("p1", "p2"):
        for h in ("odd", "eve"):
            _write_tnf(os.path.join(d, f"{p}{h}tnf"), g, f"{p}_{h}", nicosavg=1)
    _run(MERGE_EXE, d, "p1", "p2", "mfsc.mrc")
    m = np.fromfile(os.path.join(d, "mfsc.mrc"), dtype="<f4", offset=1024).reshape((40,) * 3)
    assert rel_l2(m, avg["mfscavg_map"]) < 1e-5


def test_fortran_merge_driver(g):
    """merge_icosprocess_subrec_hip with the reference program's command lines
    on the reference's own tnf files (plain and FSC-weighted merges)."""
    if not os.path.exists(MERGE_EXE):
        pytest.skip("Fortran merge driver not built (amdflang missing)")
    box, maxR = int(g["s1_hdr"][0]), int(g["s1_hdr"][1])
    with tempfile.TemporaryDirectory() as d:
        for key, name in (("s1", "s1.tnf"), ("s2", "s2.tnf"), ("p1_odd", "p1oddtnf"), ("p1_eve", "p1evetnf"),
                          ("p2_odd", "p2oddtnf"), ("p2_eve", "p2evetnf")):
            h, hf = g[key + "_hdr"], g[key + "_hdrf"]
            formats.write_tnf(os.path.join(d, name), int(h[0]), int(h[1]), float(hf[0]), float(hf[1]),
                              *_tnf(g, key), fsc=int(h[2]), realspaceavg=int(h[3]))
        _run(MERGE_EXE, d, "s1.tnf", "s2.tnf", "m0.mrc")
        raw = open(os.path.join(d, "m0.mrc"), "rb").read()
        assert np.array_equal(np.frombuffer(raw[:1024], "<i4"), g["merge_hdr"])
        m0 = np.frombuffer(raw[1024:], "<f4").reshape((box - 8,) * 3)
        assert rel_l2(m0, g["merge_map"]) < 1e-5
        _run(MERGE_EXE, d, "p1", "p2", "mfsc.mrc")
        m1 = np.fromfile(os.path.join(d, "mfsc.mrc"), "<f4", offset=1024).reshape((box - 8,) * 3)
        assert rel_l2(m1, g["mfsc_map"]) < 1e-5
        rows = np.array([[float(v) for v in line.split()] for line in open(os.path.join(d, "FSC.txt"))
                         if line.strip()])
        assert rows.shape == (maxR + 1, 3)
        assert np.abs(rows[:, 1] - g["mfsc_txt"][:, 1]).max() < 2e-6
        ok = np.isfinite(g["mfsc_txt"][:, 2])
        assert np.array_equal(np.isfinite(rows[:, 2]), ok)
