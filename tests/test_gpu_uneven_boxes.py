"""Full-size cases of the uneven-box workload: boxes 256, 384 and 512 with
phantoms of outer radius 100, 150 and 200 px (20 seeded Gaussian blobs, 60-fold
expanded), CTF-modulated particles (defocus U(0.8, 3.5) um, 300 kV, Cs 2.7 mm,
A 0.1, 1.5 A/px).

* Search (icosprocess mode 4, ort_globesearch.f90:2-141) on the real 0.5 deg AU
  grid x 720 omega at 8 A (maxR 48 / 72 / 96): the tabulated search with its
  exact re-score gives the plain exhaustive search's (theta, phi, omega, PR)
  bit for bit on every particle.
* Insert (reconstruct_substep.f90:101-145, compact tnf accumulator) at
  Nyquist-2 (maxR 126 / 190 / 254, rcut maxR-2): the whole mode-1 particle
  pipeline (mask, norm, FFT, centre shift, CTF, 60-fold trilinear insert)
  within rel-L2 1e-4 of the FP64 oracle, plain and compensated sums; the
  queue-packed brick kernel and the row-by-row brick kernel give the same
  bits; the weights obey the partition of unity, sum(wg) = 60 x sum(ctf^2)
  over the inserted samples, which no oracle enters.
"""
import os
import sys

import numpy as np
import pytest

from conftest import rel_l2
from cryoim_amd import lib, synth

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pipelines  # noqa: E402

pytestmark = pytest.mark.gpu

APIX = 1.5
OUTER = {256: 100.0, 384: 150.0, 512: 200.0}
COMPENSATED = 16   # CIM_REC_COMPENSATED


def _phantom(box, seed):
    s = OUTER[box] / 55.0
    return synth.make_phantom(box, nblobs=20, sigma=(2.0 * s, 4.0 * s), radius=(40.0 * s, 55.0 * s), seed=seed)


def _parts(ps, mask_first=False):
    P = np.zeros((len(ps.images), lib.PART_STRIDE), np.float32)
    P[:, 0:3] = ps.angles
    P[:, 3:5] = ps.shifts
    P[:, 6:8] = ps.df
    P[:, 8] = ps.astig
    # reconstruct_substep.f90 masks every particle around the first one's shift
    P[:, 9:11] = ps.shifts[0] if mask_first else ps.shifts
    P[:, 11] = 1.0
    return P


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("box", [256, 384, 512])
def test_tabulated_search_equals_plain_full_grid(ctx, box):
    ph = _phantom(box, seed=box)
    rng = np.random.default_rng(box + 1)
    T, n = 4, 12
    tang = synth.random_au_orientations(T, rng)
    timgs = synth.project(ph, tang)
    ps = synth.make_particles(ph, n, seed=box + 2, shift_max=3.0, df_range=(0.8, 3.5), snr=0.05, apix=APIX,
                              mask_radius=box / 2 - 4)
    d = pipelines.derived(box, APIX, maxRes=8.0, minRes=300.0, imgmask_A=(box / 2 - 4) * APIX)
    assert d["maxR"] == {256: 48, 384: 72, 512: 96}[box]
    common = dict(apply_ctf=True, mask_on=True, apix=APIX, imgmask=d["mask"], searchstep=0.5)
    p_plain = lib.make_params(box, d["minR"], d["maxR"], search_method=lib.SEARCH_PLAIN, **common)
    p_tab = lib.make_params(box, d["minR"], d["maxR"], search_method=lib.SEARCH_TABLES, search_topk=32, **common)
    _say(f"search box {box}: maxR {d['maxR']}, {n} particles, T {T}, 0.5 deg grid")
    ctx.search_templates(p_plain, timgs, np.c_[tang, np.zeros((T, 2))])
    plain = ctx.search_global(p_plain, ps.images, _parts(ps))
    ctx.counter_reset()
    tab = ctx.search_global(p_tab, ps.images, _parts(ps))
    same = np.all(tab[:, [0, 1, 2, 5]] == plain[:, [0, 1, 2, 5]], axis=1)
    _say(f"box {box}: identical {same.sum()} of {n}, fallbacks {ctx.counter('cl_fallback_particles')}")
    assert np.all(plain[:, 5] < 999.0)
    assert same.all()


@pytest.mark.parametrize("box", [256, 384, 512])
def test_insert_pipeline_matches_fp64_oracle_at_nyquist(ctx, O64, box):
    maxR = box // 2 - 2
    rcut = maxR - 2.0
    n = 3
    ph = _phantom(box, seed=box + 10)
    ps = synth.make_particles(ph, n, seed=box + 11, shift_max=3.0, snr=0.5, ctf=True, apix=APIX,
                              df_range=(0.8, 3.5))
    mask_r = box / 2 - 4
    vol = O64.new_volume(box, maxR, compact=True, wg_init=0.0)
    for i in range(n):
        x, y = ps.shifts[i]
        a, b = O64.particle_ft(ps.images[i], x, y, mask_r=mask_r, mask_xy=tuple(ps.shifts[0]), norm=True)
        c = O64.ctf2d(box, ps.df[i, 0], ps.df[i, 1], 0.0, APIX)
        O64.insert(vol, a, b, c, ps.angles[i], maxR, rcut=rcut, compact=True)
    p = lib.make_params(box, 0, maxR, apply_ctf=True, mask_on=True, norm_on=True, apix=APIX, imgmask=mask_r)
    got = {}
    for name, layout, kern in (("plain", 1, "brick"), ("compensated", 1 | COMPENSATED, "brick"),
                               ("rows", 1, "brickrows")):
        _say(f"insert box {box}: maxR {maxR}, {name} ({kern}), {n} particles")
        os.environ["CIM_INSERT"] = kern
        try:
            ctx.rec_begin(box, maxR, layout=layout, rcut=rcut, wg_init=0.0)
            ctx.rec_insert(p, ps.images, _parts(ps, mask_first=True))
            got[name] = ctx.rec_get_tnf(maxR)
        finally:
            os.environ.pop("CIM_INSERT", None)
    for name in ("plain", "compensated"):
        errs = [rel_l2(g, o) for g, o in zip(got[name], vol)]
        _say(f"box {box} {name}: rel-L2 vs FP64 oracle fa {errs[0]:.2e} fb {errs[1]:.2e} wg {errs[2]:.2e}")
        assert max(errs) < 1e-4
    for a, b in zip(got["plain"], got["rows"]):
        assert np.array_equal(a, b)
    # the coset insert (full layout, the default there): half-volume sums for z >= 1 and the
    # z = 0 plane after the fold's Friedel combination, against the same FP64 sums (rcut maxR-2
    # keeps every corner off the compact array's wrapped edge, so compact and full sums agree)
    _say(f"insert box {box}: coset insert, compensated, {n} particles")
    ctx.rec_begin(box, maxR, layout=COMPENSATED, rcut=rcut, wg_init=0.0)
    ctx.rec_insert(p, ps.images, _parts(ps, mask_first=True))
    cs = ctx.rec_get_tnf(maxR)

    def plane0(t):
        fa, fb, wg = (np.asarray(v, np.float64)[0] for v in t)
        return fa + fa[::-1, ::-1], fb - fb[::-1, ::-1], wg + wg[::-1, ::-1]

    # (the folded plane is real up to rounding: the 2-fold about z maps x to -x there)
    g0, o0 = plane0(cs), plane0(vol)
    errs = [rel_l2(g[1:], o[1:]) for g, o in zip(cs, vol)] + [rel_l2(g0[0] + 1j * g0[1], o0[0] + 1j * o0[1]),
                                                               rel_l2(g0[2], o0[2])]
    _say(f"box {box} cosets: rel-L2 vs FP64 oracle z >= 1: fa {errs[0]:.2e} fb {errs[1]:.2e} wg {errs[2]:.2e}; "
         f"folded z = 0: fa + i fb {errs[3]:.2e} wg {errs[4]:.2e}; |fb| there {np.abs(g0[1]).max():.1e} of "
         f"|fa| {np.abs(g0[0]).max():.1e}")
    assert max(errs) < 1e-4


@pytest.mark.parametrize("box", [256, 384, 512])
def test_partition_of_unity_at_nyquist(ctx, box):
    maxR = box // 2 - 2
    rcut = maxR - 2.0
    rng = np.random.default_rng(box + 20)
    n = 2
    re = rng.standard_normal((n, box, box)).astype(np.float32)
    im = rng.standard_normal((n, box, box)).astype(np.float32)
    ctf = rng.uniform(-1, 1, (n, box, box)).astype(np.float32)
    eul = np.array([[70.5, 10.25, 30.0], [88.0, -17.5, 211.5]], np.float32)
    _say(f"partition of unity box {box}: maxR {maxR}, rcut {rcut}")
    ctx.rec_begin(box, maxR, layout=1, rcut=rcut, wg_init=0.0)
    ctx.rec_insert_ft(re, im, eul, ctf)
    wg = ctx.rec_get_tnf(maxR)[2]
    c = box // 2
    ky, hx = np.meshgrid(np.arange(-maxR, maxR + 1), np.arange(0, maxR + 1), indexing="ij")
    keep = np.sqrt((hx * hx + ky * ky).astype(np.float32)) <= np.float32(rcut)
    tot = float(np.sum(ctf[:, ky[keep] + c, hx[keep] + c].astype(np.float64) ** 2))
    err = abs(wg.sum(dtype=np.float64) - 60 * tot) / (60 * tot)
    _say(f"box {box}: sum(wg) relative to 60 sum(ctf^2): {err:.2e}")
    assert err < 1e-5
