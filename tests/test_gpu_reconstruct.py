"""HIP direct Fourier inversion against the FP64 oracle of
icosprocess/SRC/reconstruct.f90 (full box, mode 1) and
reconstruct_substep.f90 (compact tnf accumulator).  Every test runs with
each insert kernel: the coset insert (the default for the full layout with
the 60 icosahedral operators; other cases fall back to the brick gather),
the 60-plane brick gather (CIM_INSERT=brick) and the atomic scatter
(CIM_INSERT=scatter)."""
import os

import numpy as np
import pytest

from conftest import rel_l2
from cryoim_amd import lib, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["cosets", "brick", "scatter"])
def insert_kernel(request):
    old = os.environ.get("CIM_INSERT")
    os.environ["CIM_INSERT"] = request.param
    yield request.param
    if old is None:
        os.environ.pop("CIM_INSERT", None)
    else:
        os.environ["CIM_INSERT"] = old


def _data(box=64, n=24, seed=11, ctf=True):
    ph = synth.make_phantom(box, radius=(box * 0.28, box * 0.4), sigma=(1.5, 2.5), seed=seed)
    ps = synth.make_particles(ph, n, seed=seed + 1, shift_max=2.0, snr=0.5, ctf=ctf, apix=1.5,
                              df_range=(1.0, 2.5))
    return ps


def _parts(ps, mask_first=True):
    n = len(ps.images)
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3] = ps.angles
    P[:, 3:5] = ps.shifts
    P[:, 6:8] = ps.df
    P[:, 8] = ps.astig
    if mask_first:                     # reconstruct.f90:92 masks around the first particle's shift
        P[:, 9:11] = ps.shifts[0]
    else:
        P[:, 9:11] = ps.shifts
    P[:, 11] = 1.0
    return P


def _oracle_volume(O, ps, maxR, mask_r, apix, compact=False, rcut=None, wg_init=0.01, ctf=True):
    box = ps.images.shape[-1]
    vol = O.new_volume(box, maxR, compact=compact, wg_init=wg_init)
    for i in range(len(ps.images)):
        x, y = ps.shifts[i]
        a, b = O.particle_ft(ps.images[i], x, y, mask_r=mask_r, mask_xy=tuple(ps.shifts[0]), norm=True)
        c = O.ctf2d(box, ps.df[i, 0], ps.df[i, 1], 0.0, apix) if ctf else None
        O.insert(vol, a, b, c, ps.angles[i], maxR, rcut=rcut, compact=compact)
    return vol


@pytest.mark.parametrize("box,maxR", [(64, 24), (96, 30)])
def test_reconstruct_map_matches_fp64_oracle(ctx, O64, insert_kernel, box, maxR):
    ps = _data(box=box, n=20)
    apix, mask_r = 1.5, box / 2 - 4
    p = lib.make_params(box, 0, maxR, apply_ctf=True, mask_on=True, norm_on=True, apix=apix, imgmask=mask_r)
    ctx.rec_begin(box, maxR, layout=0, rcut=maxR, wg_init=0.01)
    ctx.rec_insert(p, ps.images, _parts(ps))
    gmap = ctx.rec_finish(box)
    vol = _oracle_volume(O64, ps, maxR, mask_r, apix)
    omap = O64.fold_and_map(vol, maxR)
    err = rel_l2(gmap, omap)
    print(f"box {box} maxR {maxR}: map rel-L2 vs FP64 oracle = {err:.3e}")
    assert err < 1e-4
    # the accumulated sums themselves (tnf block of the full volume); the coset insert holds
    # half of the fold's Friedel combination on the z = 0 plane (the fold gives the same map)
    got = ctx.rec_get_tnf(maxR)
    c = box // 2
    sl = (slice(c, c + maxR + 1), slice(c - maxR, c + maxR + 1), slice(c - maxR, c + maxR + 1))
    want = tuple(v[sl] for v in vol)
    if insert_kernel == "cosets":
        for g, o in zip(got, want):
            assert rel_l2(g[1:], o[1:]) < 1e-4
        g0, o0 = _fold_plane0(got), _fold_plane0(want)
        assert rel_l2(g0[0] + 1j * g0[1], o0[0] + 1j * o0[1]) < 1e-4 and rel_l2(g0[2], o0[2]) < 1e-4
    else:
        for g, o in zip(got, want):
            assert rel_l2(g, o) < 1e-4


def test_reconstruct_substep_tnf(ctx, O64):
    box, maxR = 64, 24
    ps = _data(box=box, n=16, seed=5)
    apix, mask_r = 1.5, box / 2 - 4
    p = lib.make_params(box, 0, maxR, apply_ctf=True, mask_on=True, norm_on=True, apix=apix, imgmask=mask_r)
    ctx.rec_begin(box, maxR, layout=1, rcut=maxR - 2, wg_init=0.0)
    ctx.rec_insert(p, ps.images, _parts(ps))
    fa, fb, wg = ctx.rec_get_tnf(maxR)
    vol = _oracle_volume(O64, ps, maxR, mask_r, apix, compact=True, rcut=maxR - 2, wg_init=0.0)
    for g, o in ((fa, vol[0]), (fb, vol[1]), (wg, vol[2])):
        assert rel_l2(g, o) < 1e-4


def test_partition_of_unity_weights(ctx):
    """sum(wg) = 60 * sum(ctf^2) over inserted samples (+ init), independent of the oracle."""
    box, maxR = 64, 20
    rng = np.random.default_rng(3)
    n = 5
    re = rng.standard_normal((n, box, box)).astype(np.float32)
    im = rng.standard_normal((n, box, box)).astype(np.float32)
    ctf = rng.uniform(-1, 1, (n, box, box)).astype(np.float32)
    eul = np.array([[70, 10, 30], [80, -5, 200], [75, 3, 100], [88, 20, 5], [71, -15, 300]], np.float32)
    ctx.rec_begin(box, maxR, layout=1, rcut=maxR - 2, wg_init=0.0)
    ctx.rec_insert_ft(re, im, eul, ctf)
    fa, fb, wg = ctx.rec_get_tnf(maxR)
    c = box // 2
    tot = 0.0
    for ky in range(-maxR, maxR + 1):
        for hx in range(0, maxR + 1):
            if np.sqrt(np.float32(hx * hx + ky * ky)) <= maxR - 2:
                tot += float(np.sum(ctf[:, ky + c, hx + c].astype(np.float64) ** 2))
    assert abs(wg.sum(dtype=np.float64) - 60 * tot) / (60 * tot) < 1e-5


def test_unused_particles_are_skipped(ctx):
    box, maxR = 64, 20
    ps = _data(box=box, n=6, seed=9)
    p = lib.make_params(box, 0, maxR, apply_ctf=True, norm_on=True, apix=1.5)
    P = _parts(ps)
    P[:, 11] = 0.0
    ctx.rec_begin(box, maxR, layout=1, rcut=maxR - 2, wg_init=0.0)
    ctx.rec_insert(p, ps.images, P)
    fa, fb, wg = ctx.rec_get_tnf(maxR)
    assert not fa.any() and not wg.any()


@pytest.mark.parametrize("nops,n", [(1, 3), (7, 40), (60, 19)])
def test_brick_gather_equals_scatter(ctx, insert_kernel, nops, n):
    """The two insert kernels place every contribution identically; only the
    float summation order differs.  Random operators, n*nops across the
    gather's 1024-plane chunks, both layouts."""
    if insert_kernel != "brick":
        pytest.skip("compares the two kernels once")
    box, maxR = 80, 33
    rng = np.random.default_rng(nops)
    re = rng.standard_normal((n, box, box)).astype(np.float32)
    im = rng.standard_normal((n, box, box)).astype(np.float32)
    ctf = rng.uniform(-1, 1, (n, box, box)).astype(np.float32)
    eul = np.stack([rng.uniform(0, 180, n), rng.uniform(0, 360, n), rng.uniform(0, 360, n)], 1).astype(np.float32)
    q = np.linalg.qr(rng.standard_normal((nops, 3, 3)))[0]
    q *= np.sign(np.linalg.det(q))[:, None, None]
    ops = None if nops == 60 else q
    res = {}
    for layout, rcut in ((0, float(maxR)), (1, maxR - 2.0)):
        for kern in ("brick", "scatter"):
            os.environ["CIM_INSERT"] = kern
            ctx.rec_begin(box, maxR, layout=layout, rcut=rcut, wg_init=0.0, ops=ops)
            ctx.rec_insert_ft(re, im, eul, ctf)
            ctx.rec_insert_ft(re[: n // 2 + 1], im[: n // 2 + 1], eul[: n // 2 + 1], ctf[: n // 2 + 1])
            res[kern] = ctx.rec_get_tnf(maxR)
        # float32 sums of up to ~1e3 terms per voxel in different orders (the brick sums a
        # batch exactly, so fa, fb that cancel come out 0 where the scatter leaves noise)
        for a, b in zip(res["brick"], res["scatter"]):
            scale = float(np.abs(b).max())
            err_max = float(np.abs(a - b).max())
            err_l2 = rel_l2(a, b)
            print(f"nops {nops} layout {layout}: rel-L2 {err_l2:.2e}, max {err_max / scale:.2e} of max")
            assert err_l2 < 1e-5
            assert err_max <= 1e-4 * scale
        # weights are sums of non-negative terms: the same voxels are reached
        wa, wb = res["brick"][2], res["scatter"][2]
        assert float(((wa != 0) == (wb != 0)).mean()) > 0.9999


COMPENSATED = 16   # CIM_REC_COMPENSATED


@pytest.mark.parametrize("nops,layout", [(60, 0), (60, 1 | COMPENSATED), (7, 1), (60, 1)])
def test_queue_kernel_equals_row_kernel(ctx, insert_kernel, nops, layout):
    """The brick gather packed through per-wave queues (default) and the
    row-by-row brick gather (CIM_INSERT=brickrows) add the same fixed-point
    terms into exact sums: the same bits, over three batches (the second
    spans two 1024-plane chunks), for every brick shape of the queue kernel
    (CIM_BRICKQ: 16x16x8, 16^3, 8^3, 16x8x8 at 512 and 384 threads; the 16-voxel bricks overhang the
    118 x 118 x 60 accumulator) and in raster as well as central-first brick
    order."""
    if insert_kernel != "brick":
        pytest.skip("compares the two brick kernels once")
    box, maxR = 128, 58
    rng = np.random.default_rng(100 + nops + layout)
    n = 24
    re = rng.standard_normal((n, box, box)).astype(np.float32)
    im = rng.standard_normal((n, box, box)).astype(np.float32)
    ctf = rng.uniform(-1, 1, (n, box, box)).astype(np.float32)
    eul = np.stack([rng.uniform(0, 180, n), rng.uniform(0, 360, n), rng.uniform(0, 360, n)], 1).astype(np.float32)
    q = np.linalg.qr(rng.standard_normal((nops, 3, 3)))[0]
    q *= np.sign(np.linalg.det(q))[:, None, None]
    ops = None if nops == 60 else q
    res = {}
    runs = [("brickrows", "0", "central")] + [("brick", s, o) for s in ("0", "1", "2", "3", "4")
                                                for o in ("central", "raster")]
    try:
        for kern, shape, order in runs:
            os.environ["CIM_INSERT"] = kern
            os.environ["CIM_BRICKQ"] = shape
            os.environ["CIM_BRICK_ORDER"] = order
            ctx.rec_begin(box, maxR, layout=layout, rcut=maxR - 2.0, wg_init=0.0, ops=ops)
            ctx.rec_insert_ft(re[:3], im[:3], eul[:3], ctf[:3])
            ctx.rec_insert_ft(re[3:], im[3:], eul[3:], ctf[3:])
            ctx.rec_insert_ft(re[:1], im[:1], eul[:1], ctf[:1])
            res[(kern, shape, order)] = ctx.rec_get_tnf(maxR)
    finally:
        os.environ["CIM_INSERT"] = "brick"
        os.environ.pop("CIM_BRICKQ", None)
        os.environ.pop("CIM_BRICK_ORDER", None)
    ref = res[runs[0]]
    assert np.abs(ref[2]).max() > 0
    for key in runs[1:]:
        for a, b in zip(res[key], ref):
            assert np.array_equal(a, b), key


def test_compensated_sums_do_not_drift(ctx, insert_kernel):
    """256 one-particle batches of the same particle: the compensated sums stay at
    256 x the single insert (float64 reference), the plain FP32 sums drift."""
    if insert_kernel != "brick":
        pytest.skip("compensation is a brick-insert feature")
    box, maxR, reps = 64, 20, 256
    rng = np.random.default_rng(8)
    re = rng.standard_normal((1, box, box)).astype(np.float32)
    im = rng.standard_normal((1, box, box)).astype(np.float32)
    ctf = rng.uniform(0.2, 1.0, (1, box, box)).astype(np.float32)
    eul = np.array([[73.0, 11.0, 41.0]], np.float32)
    ctx.rec_begin(box, maxR, layout=1, rcut=maxR - 2, wg_init=0.0)
    ctx.rec_insert_ft(re, im, eul, ctf)
    one = [x.astype(np.float64) * reps for x in ctx.rec_get_tnf(maxR)]
    errs = {}
    for layout in (1, 1 | COMPENSATED):
        ctx.rec_begin(box, maxR, layout=layout, rcut=maxR - 2, wg_init=0.0)
        for _ in range(reps):
            ctx.rec_insert_ft(re, im, eul, ctf)
        got = ctx.rec_get_tnf(maxR)
        errs[layout] = max(rel_l2(g, o) for g, o in zip(got, one))
    print(f"after {reps} batches: plain FP32 rel-L2 {errs[1]:.2e}, compensated {errs[1 | COMPENSATED]:.2e}")
    assert errs[1 | COMPENSATED] < 2e-7
    assert errs[1 | COMPENSATED] < errs[1] / 4


def test_compensated_map_matches_fp64_oracle(ctx, O64, insert_kernel):
    if insert_kernel == "scatter":
        pytest.skip("compensation is a feature of the brick and coset inserts")
    box, maxR = 64, 24
    ps = _data(box=box, n=20)
    apix, mask_r = 1.5, box / 2 - 4
    p = lib.make_params(box, 0, maxR, apply_ctf=True, mask_on=True, norm_on=True, apix=apix, imgmask=mask_r)
    ctx.rec_begin(box, maxR, layout=COMPENSATED, rcut=maxR, wg_init=0.01)
    for i in range(0, 20, 5):      # four batches
        ctx.rec_insert(p, ps.images[i:i + 5], _parts(ps)[i:i + 5])
    gmap = ctx.rec_finish(box)
    vol = _oracle_volume(O64, ps, maxR, mask_r, apix)
    err = rel_l2(gmap, O64.fold_and_map(vol, maxR))
    print(f"compensated map vs FP64 oracle: rel-L2 {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("compact", [False, True])
def test_corrctf_matches_reference(ctx, O64, prog, insert_kernel, compact):
    """reconstruct_corrCTF / reconstruct_corr_CTF_substep (corrCTF=0.2,2): the
    insert with a defocus series, against the key-patched reference's map and
    tnf (tests/golden/corrctf.npz) and the FP64 oracle.  The reference's
    float32 sums (5 defoci x 60 operators x 8 corners per sample, in order)
    drift 3.4e-5 from the FP64 map on this case, as the FP32 oracle shows by
    replaying that order; the brick kernel's per-batch sums are exact, so its
    map sits with the FP64 one (the scatter kernel's float atomics add in
    arbitrary order, the reference's error class)."""
    import pipelines
    cc = np.load(pipelines.golden_path("corrctf.npz"))
    box, apix = int(prog["prog_box"][0]), float(prog["prog_apix"][0])
    d = pipelines.derived(box, apix, maxRes=6.0, imgmask_A=40)
    maxR = d["maxR"]
    step, cyc = float(cc["corr_step_cycles"][0]), int(cc["corr_step_cycles"][1])
    p = lib.make_params(box, 0, maxR, apply_ctf=True, mask_on=True, norm_on=True, apix=apix, imgmask=d["mask"],
                        corr_step=step, corr_cycles=cyc)
    n = len(prog["prog_images"])
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3] = prog["prog_angles"]
    P[:, 3:5] = prog["prog_shifts"]
    P[:, 6:8] = prog["prog_df"]
    P[:, 9:11] = prog["prog_shifts"][0]      # reconstruct_corrCTF.f90:95 masks at the first particle
    P[:, 11] = 1.0
    vol = pipelines.corrctf_volume(O64, prog["prog_images"], prog["prog_angles"], prog["prog_shifts"],
                                   prog["prog_df"], apix, maxR, d["mask"], step, cyc, compact=compact)
    tol64 = 5e-6 if insert_kernel == "brick" else 3e-5
    if compact:
        ctx.rec_begin(box, maxR, layout=1, rcut=maxR, wg_init=0.0)
        ctx.rec_insert(p, prog["prog_images"], P)
        for o, k, v in zip(ctx.rec_get_tnf(maxR), ("fa", "fb", "wg"), vol):
            e_ref, e_64 = rel_l2(o, cc[f"corrsub_{k}"]), rel_l2(o, v)
            print(f"corrCTF substep {k}: rel-L2 {e_ref:.2e} vs reference, {e_64:.2e} vs FP64 oracle")
            assert e_ref < 1e-4 and e_64 < tol64
    else:
        ctx.rec_begin(box, maxR, layout=0, rcut=maxR, wg_init=0.01)
        ctx.rec_insert(p, prog["prog_images"], P)
        m = ctx.rec_finish(box)
        e_ref, e_64 = rel_l2(m, cc["corr_map"]), rel_l2(m, O64.fold_and_map(vol, maxR))
        print(f"corrCTF map: rel-L2 {e_ref:.2e} vs reference, {e_64:.2e} vs FP64 oracle")
        assert e_ref < 1e-4 and e_64 < tol64


def test_corrctf_needs_the_ctf(ctx):
    p = lib.make_params(64, 0, 20, apply_ctf=False, apix=1.5, corr_step=0.1, corr_cycles=1)
    ctx.rec_begin(64, 20, layout=1, rcut=18, wg_init=0.0)
    with pytest.raises(lib.CimError, match="apply_ctf"):
        ctx.rec_insert(p, np.zeros((1, 64, 64), np.float32), np.zeros((1, lib.PART_STRIDE), np.float32))


def _fold_plane0(t):
    """The z = 0 plane of tnf-order sums (fa, fb, wg) combined with its Friedel
    mate as the fold does (reconstruct.f90:151-237): fa(x) + fa(-x),
    fb(x) - fb(-x), wg(x) + wg(-x)."""
    fa, fb, wg = (np.asarray(v, np.float64)[0] for v in t)
    return fa + fa[::-1, ::-1], fb - fb[::-1, ::-1], wg + wg[::-1, ::-1]


@pytest.mark.parametrize("layout", [0, COMPENSATED])
def test_coset_insert_matches_fp64_oracle(ctx, O64, insert_kernel, layout):
    """The coset insert (5 coset representatives of the signed-permutation
    subgroup into a full cube, symmetrised when read) and the 60-plane brick
    gather on the same random batches, both against the FP64 oracle's
    60-operator insert (reconstruct.f90:103-147): the half-volume sums for
    z >= 1, the z = 0 plane after the fold's Friedel combination, and the
    maps.  The sums are flushed once per read: reading twice, then inserting
    more, gives the sums of every batch once."""
    if insert_kernel != "cosets":
        pytest.skip("compares the two kernels once")
    box, maxR = 128, 58
    rng = np.random.default_rng(300 + layout)
    n = 24
    re = rng.standard_normal((n, box, box)).astype(np.float32)
    im = rng.standard_normal((n, box, box)).astype(np.float32)
    ctf = rng.uniform(-1, 1, (n, box, box)).astype(np.float32)
    eul = np.stack([rng.uniform(0, 180, n), rng.uniform(0, 360, n), rng.uniform(0, 360, n)], 1).astype(np.float32)
    vol = O64.new_volume(box, maxR, compact=False, wg_init=0.01)
    for i in range(n):
        O64.insert(vol, re[i], im[i], ctf[i], eul[i], maxR)
    c = box // 2
    sl = (slice(c, c + maxR + 1), slice(c - maxR, c + maxR + 1), slice(c - maxR, c + maxR + 1))
    want = tuple(v[sl] for v in vol)
    omap = O64.fold_and_map(vol, maxR)
    res = {}
    try:
        for kern in ("brick", "cosets"):
            os.environ["CIM_INSERT"] = kern
            ctx.counter_reset()
            ctx.rec_begin(box, maxR, layout=layout, rcut=float(maxR), wg_init=0.01)
            ctx.rec_insert_ft(re[:10], im[:10], eul[:10], ctf[:10])
            first = ctx.rec_get_tnf(maxR)
            again = ctx.rec_get_tnf(maxR)
            for a, b in zip(first, again):
                assert np.array_equal(a, b), kern
            ctx.rec_insert_ft(re[10:], im[10:], eul[10:], ctf[10:])
            res[kern] = (ctx.rec_get_tnf(maxR), ctx.rec_finish(box))
            assert ctx.counter("insert_coset_particles") == (n if kern == "cosets" else 0), kern
    finally:
        os.environ["CIM_INSERT"] = "cosets"
    o0 = _fold_plane0(want)
    errs = {}
    for kern, (t, m) in res.items():
        g0 = _fold_plane0(t)
        errs[kern] = ([rel_l2(a[1:], b[1:]) for a, b in zip(t, want)] +
                      [rel_l2(g0[0] + 1j * g0[1], o0[0] + 1j * o0[1]), rel_l2(g0[2], o0[2]), rel_l2(m, omap)])
        print(f"layout {layout} {kern}: vs FP64 oracle z >= 1 fa {errs[kern][0]:.2e} fb {errs[kern][1]:.2e} "
              f"wg {errs[kern][2]:.2e}; folded z = 0 {errs[kern][3]:.2e} / {errs[kern][4]:.2e}; "
              f"map {errs[kern][5]:.2e}")
    # both place samples in float32 from a float32 S R (the reference's REAL*4 arithmetic); the
    # coset kernel's fractions x - floor(x) are exact, while the 60-plane kernel adds box/2 before
    # truncating, as the reference does, and rounds them to that magnitude's ulp: measured 2.5-3.1e-6
    # against 6-7e-6 on these random (cancelling) sums
    assert max(errs["cosets"]) < 1e-5
    assert max(errs["cosets"]) <= max(errs["brick"])
    assert max(errs["brick"]) < 1e-4
