"""HIP common-lines search (icosprocess modes 4 and 5) against the FP64
oracle restatement of ort_globesearch / cal_crosslin_PR_omega /
ort_refinement on the same seeded inputs."""
import numpy as np
import pytest

from cryoim_amd import geometry, lib, synth
import pipelines

pytestmark = pytest.mark.gpu


def _setup(box=64, maxR=12, minR=1, T=3, n=2, seed=21, apix=2.0, ctf=True):
    ph = synth.make_phantom(box, radius=(box * 0.28, box * 0.4), sigma=(1.5, 2.5), seed=seed)
    rng = np.random.default_rng(seed)
    tang = synth.random_au_orientations(T, rng)
    tsh = rng.uniform(-1, 1, (T, 2))
    timgs = synth.project(ph, tang, tsh)
    ps = synth.make_particles(ph, n, seed=seed + 1, shift_max=2.0, snr=1.0, ctf=ctf, apix=apix,
                              df_range=(1.0, 2.0))
    return ph, tang, tsh, timgs, ps


def _oracle_inputs(O, tang, tsh, timgs, ps, i, minR, maxR, mask_r, apix, ctf=True):
    box = timgs.shape[-1]
    ta, tb, tfc = [], [], []
    c = O.ctf2d(box, ps.df[i, 0], ps.df[i, 1], 0.0, apix) if ctf else None
    for t in range(len(tang)):
        a, b = O.particle_ft(timgs[t], tsh[t, 0], tsh[t, 1], mask_r=mask_r)
        if ctf:
            a, b = a * c, b * c
        ta.append(a)
        tb.append(b)
        tfc.append(O.amp_expfit(a, b, minR, maxR))
    x, y = ps.shifts[i]
    a0, b0 = O.particle_ft(ps.images[i], x, y, mask_r=mask_r, shift=False)
    fc0 = O.amp_expfit(a0, b0, minR, maxR)
    pa, pb = O.ft_shift(a0, b0, x, y, True)
    return (a0, b0), (pa, pb), fc0, np.array(ta), np.array(tb), np.array(tfc)


def _parts(ps):
    P = np.zeros((len(ps.images), lib.PART_STRIDE), np.float32)
    P[:, 0:3] = ps.angles
    P[:, 3:5] = ps.shifts
    P[:, 6:8] = ps.df
    P[:, 11] = 1
    return P


VARIANTS = {"tuned": {}, "no_pair": {"CIM_SEARCH_NOPAIR": "1"}, "lds": {"CIM_SEARCH_LDS": "1"},
            "cache_vpt2_nopair": {"CIM_SEARCH_LDS": "0", "CIM_SEARCH_VPT": "2", "CIM_SEARCH_NOPAIR": "1"}}


@pytest.fixture(params=sorted(VARIANTS))
def variant(request, monkeypatch):
    """Every launch variant of the search kernel (the autotuner picks one per shape)."""
    for k, v in VARIANTS[request.param].items():
        monkeypatch.setenv(k, v)
    return request.param


@pytest.mark.parametrize("ctf", [True, False])
def test_pr_array_matches_oracle(ctx, O64, ctf, variant):
    box, minR, maxR, step, apix = 64, 1, 12, 3.0, 2.0
    mask_r = box / 2 - 4
    ph, tang, tsh, timgs, ps = _setup(box, maxR, minR, T=3, n=1, ctf=ctf, apix=apix)
    p = lib.make_params(box, minR, maxR, apply_ctf=ctf, mask_on=True, apix=apix, imgmask=mask_r,
                        searchstep=step)
    ctx.search_templates(p, timgs, np.concatenate([tang, tsh], axis=1))
    Nt, Np, No = geometry.search_grid(step)
    g = ctx.search_pr_array(p, ps.images[0], _parts(ps)[0], (Nt + 1, 2 * Np + 1, No))
    _, (pa, pb), fc0, ta, tb, tfc = _oracle_inputs(O64, tang, tsh, timgs, ps, 0, minR, maxR, mask_r, apix, ctf)
    o4, arr = O64.globesearch(pa, pb, fc0, ta, tb, tfc, tang, minR, maxR, step, want_array=True)
    inside = arr < 999.5
    assert (g < 999.5).sum() == inside.sum()
    assert np.array_equal(g >= 999.5, ~inside)
    err = np.abs(g[inside] - arr[inside])
    # the FP32-faithful restatement differs from FP64 in the same way: a
    # sample whose |F| sits within rounding of Fcut is gated differently
    # (cal_crosslin_PR_omega.f90:52,56), which moves one PR by ~0.1 deg at
    # this small radius; everything else agrees to ~1e-5 deg
    o32 = pytest.importorskip("oracle").Oracle(32)
    _, (qa, qb), qfc0, qta, qtb, qtfc = _oracle_inputs(o32, tang, tsh, timgs, ps, 0, minR, maxR, mask_r, apix, ctf)
    _, arr32 = o32.globesearch(qa, qb, qfc0, qta, qtb, qtfc, tang, minR, maxR, step, want_array=True)
    err32 = np.abs(arr32[inside] - arr[inside])
    print(f"PR |GPU - FP64 oracle|: max {err.max():.4f} deg, p99 {np.quantile(err, 0.99):.2e}, "
          f"mean {err.mean():.2e}; FP32 oracle vs FP64: max {err32.max():.4f}, p99 {np.quantile(err32, 0.99):.2e}")
    assert np.quantile(err, 0.99) < 2e-3
    assert err.mean() < 1e-3
    assert err.max() < 0.5


@pytest.mark.parametrize("pair,step", [(True, 1.0), (False, 0.5)])
def test_launch_shapes_give_the_same_bits(ctx, monkeypatch, pair, step):
    """Every launch shape of the exact kernel (columns per lane 1-4, particle
    taps from LDS or through the caches) computes each column with the same
    arithmetic in the same order, so the PR arrays are bitwise equal: the
    tabulated path re-scores with one shape and the plain path uses whichever
    the autotuner picks, and their winners must agree exactly.  The shapes
    tried are those the autotuner offers for the grid (a column count per lane
    that gives the same workgroup size as one less is not offered)."""
    box, minR, maxR, apix = 64, 1, 12, 2.0
    mask_r = box / 2 - 4
    ph, tang, tsh, timgs, ps = _setup(box, maxR, minR, T=3, n=1, apix=apix, seed=44)
    p = lib.make_params(box, minR, maxR, apply_ctf=True, mask_on=True, apix=apix, imgmask=mask_r,
                        searchstep=step)
    Nt, Np, No = geometry.search_grid(step)
    if not pair:
        monkeypatch.setenv("CIM_SEARCH_NOPAIR", "1")
    ncol = No // 2 if pair else No
    block = lambda v: (-(-(-(-ncol // v)) // 64)) * 64
    vpts = [v for v in range(1, 5) if v == 1 or block(v) != block(v - 1)]
    res = {}
    for vpt in vpts:
        for lds in ("0", "1"):
            monkeypatch.setenv("CIM_SEARCH_VPT", str(vpt))
            monkeypatch.setenv("CIM_SEARCH_LDS", lds)
            ctx.search_templates(p, timgs, np.concatenate([tang, tsh], axis=1))
            res[(vpt, lds)] = ctx.search_pr_array(p, ps.images[0], _parts(ps)[0], (Nt + 1, 2 * Np + 1, No))
    assert len(res) >= 6
    ref = res[(1, "0")]
    assert (ref < 999.5).any()
    for k, g in res.items():
        assert np.array_equal(g, ref), (k, int((g != ref).sum()))


def test_global_search_argmin(ctx, O64):
    box, minR, maxR, step, apix = 64, 1, 12, 3.0, 2.0
    mask_r = box / 2 - 4
    ph, tang, tsh, timgs, ps = _setup(box, maxR, minR, T=3, n=3, seed=33, apix=apix)
    p = lib.make_params(box, minR, maxR, apply_ctf=True, mask_on=True, apix=apix, imgmask=mask_r,
                        searchstep=step)
    ctx.search_templates(p, timgs, np.concatenate([tang, tsh], axis=1))
    out = ctx.search_global(p, ps.images, _parts(ps))
    for i in range(3):
        _, (pa, pb), fc0, ta, tb, tfc = _oracle_inputs(O64, tang, tsh, timgs, ps, i, minR, maxR, mask_r, apix)
        o4, arr = O64.globesearch(pa, pb, fc0, ta, tb, tfc, tang, minR, maxR, step, want_array=True)
        # decisions agree unless the oracle's best and runner-up are within the float noise
        srt = np.sort(arr.ravel())
        if srt[1] - srt[0] > 0.02:
            assert np.allclose(out[i, 0:3], o4[:3], atol=1e-3), (out[i, 0:3], o4[:3])
        assert abs(out[i, 5] - o4[3]) < 0.05


def test_local_refinement_matches_oracle(ctx, O64):
    box, minR, maxR, apix = 64, 1, 12, 2.0
    mask_r = box / 2 - 4
    ph, tang, tsh, timgs, ps = _setup(box, maxR, minR, T=3, n=2, seed=41, apix=apix)
    dang = 360.0 / (maxR * np.sqrt(5.0) * 3.14159)
    dtran = box / (maxR * np.sqrt(5.0))
    p = lib.make_params(box, minR, maxR, apply_ctf=True, mask_on=True, apix=apix, imgmask=mask_r,
                        ncycle=3, delta_angle=dang, delta_trans=dtran)
    ctx.search_templates(p, timgs, np.concatenate([tang, tsh], axis=1))
    P = _parts(ps)
    P[:, 0:3] += np.array([1.5, -1.0, 2.0], np.float32)      # start near the truth
    out = ctx.refine_local(p, ps.images, P)
    for i in range(2):
        (a0, b0), _, fc0, ta, tb, tfc = _oracle_inputs(O64, tang, tsh, timgs, ps, i, minR, maxR, mask_r, apix)
        io5, pr = O64.refine_particle(a0, b0, fc0, ta, tb, tfc, tang, minR, maxR, 3, dang, dtran,
                                      np.array([*P[i, 0:5]], np.float64))
        assert np.allclose(out[i, 0:5], io5, atol=2e-3), (out[i, 0:5], io5)
        assert abs(out[i, 5] - pr) < 0.05


def _search_pair(ctx, method, topk, box=96, maxR=20, minR=2, T=4, n=24, step=2.0, seed=31, monkeypatch=None):
    ph, tang, tsh, timgs, ps = _setup(box=box, maxR=maxR, minR=minR, T=T, n=n, seed=seed, apix=2.0)
    p = lib.make_params(box, minR, maxR, apply_ctf=True, mask_on=True, apix=2.0, imgmask=box / 2 - 4,
                        searchstep=step, search_method=method, search_topk=topk)
    ctx.search_templates(p, timgs, np.c_[tang, tsh])
    ctx.counter_reset()
    out = ctx.search_global(p, ps.images, _parts(ps))
    return out, ps


@pytest.mark.parametrize("variant_name", ["tuned", "no_pair", "lds"])
def test_tabulated_search_equals_plain(ctx, monkeypatch, variant_name):
    """CIM_SEARCH_TABLES re-scores its candidates with the exact kernel, so every
    particle whose plain minimum is among them gets the plain result bit for bit;
    on this set that is every particle (K = 32 of 198 AU points, 2 deg grid)."""
    for k, v in VARIANTS[variant_name].items():
        monkeypatch.setenv(k, v)
    plain, ps = _search_pair(ctx, lib.SEARCH_PLAIN, 0)
    tab, _ = _search_pair(ctx, lib.SEARCH_TABLES, 0)
    same = np.all(tab[:, [0, 1, 2, 5]] == plain[:, [0, 1, 2, 5]], axis=1)
    print(f"{variant_name}: identical {same.sum()} of {len(same)}, fallbacks {ctx.counter('cl_fallback_particles')}")
    assert same.all()


@pytest.mark.parametrize("coarse", ["8", "24"])
def test_tabulated_search_coarse_to_fine_equals_plain(ctx, monkeypatch, coarse):
    """The coarse-to-fine table scores (every other theta and phi index first,
    then the 3x3 neighbours of each particle's M best): on this set every
    particle still gets the plain result bit for bit (the 2 deg grid has 198 AU
    points; M = 8 and 24 of its coarse ones are refined)."""
    monkeypatch.setenv("CIM_TABLE_COARSE", coarse)
    plain, ps = _search_pair(ctx, lib.SEARCH_PLAIN, 0)
    ctx.counter_reset()
    tab, _ = _search_pair(ctx, lib.SEARCH_TABLES, 0)
    same = np.all(tab[:, [0, 1, 2, 5]] == plain[:, [0, 1, 2, 5]], axis=1)
    scored = ctx.counter("cl_gather_points")
    print(f"coarse M={coarse}: identical {same.sum()} of {len(same)}, scored (point, omega) {scored}")
    assert same.all()
    monkeypatch.setenv("CIM_TABLE_COARSE", "0")
    ctx.counter_reset()
    _search_pair(ctx, lib.SEARCH_TABLES, 0)
    assert scored < ctx.counter("cl_gather_points")     # fewer points scored than the full gather


def test_tabulated_search_exhaustive_candidates(ctx, monkeypatch):
    """With every AU point a candidate and full omega windows the tabulated search
    is the plain search: identical output whatever the tables rank."""
    monkeypatch.setenv("CIM_TABLE_WINDOW", "4096")
    monkeypatch.setenv("CIM_TABLE_FALLBACK", "0")
    plain, _ = _search_pair(ctx, lib.SEARCH_PLAIN, 0, n=6, step=3.0)
    tab, _ = _search_pair(ctx, lib.SEARCH_TABLES, 256, n=6, step=3.0)
    assert np.array_equal(tab[:, [0, 1, 2, 5]], plain[:, [0, 1, 2, 5]])


@pytest.mark.parametrize("step", [3.0, 2.0])
def test_tabulated_search_fallback_is_plain(ctx, monkeypatch, step):
    """The fallback path (forced for every particle) is the exhaustive search of
    the plain method, particle by particle: identical output.  At 2 deg the
    re-score uses omega windows (64 of 90 paired columns) and the fallback must
    not inherit them."""
    monkeypatch.setenv("CIM_TABLE_FALLBACK", "-1")
    plain, _ = _search_pair(ctx, lib.SEARCH_PLAIN, 0, n=8, step=step)
    tab, _ = _search_pair(ctx, lib.SEARCH_TABLES, 2, n=8, step=step)
    assert ctx.counter("cl_fallback_particles") == 8
    assert np.array_equal(tab[:, [0, 1, 2, 5]], plain[:, [0, 1, 2, 5]])


def test_tabulated_search_second_windows(ctx, monkeypatch):
    """Second windows (search_second) add candidates after the topk AU points:
    each re-scored column is the plain search's column bit for bit, so with the
    fallback off a particle's result can only move to a lower PR, every
    particle identical to the plain search without them stays identical, and
    every result's PR is at least the plain minimum.  On 2,000 particles of the
    box-128 recipe (1 deg grid, maxR 24) some exact minima lie in a second
    window."""
    import torch
    from cryoim_amd import synth
    box, apix, step = 128, 1.5, 1.0
    ph = synth.make_phantom(box, nblobs=20, sigma=(2.0, 4.0), radius=(40.0, 55.0), seed=101)
    tang = synth.random_au_orientations(4, np.random.default_rng(102))
    timgs = synth.project(ph, tang)
    ps = synth.make_particles_torch(ph, 2000, seed=103, shift_max=3.0, df_range=(1.0, 3.0), snr=0.1, apix=apix,
                                    mask_radius=box / 2 - 4)
    d = pipelines.derived(box, apix, maxRes=8.0, minRes=300.0, imgmask_A=(box / 2 - 4) * apix)
    P = np.zeros((2000, lib.PART_STRIDE), np.float32)
    P[:, 3:5] = ps.shifts
    P[:, 6:8] = ps.df
    P[:, 11] = 1.0
    out = {}
    # fallback off (results can only improve with more windows), then on (ranks counted)
    for name, method, S, fb in (("plain", lib.SEARCH_PLAIN, 0, "0"), ("s0", lib.SEARCH_TABLES, 0, "0"),
                                ("s8", lib.SEARCH_TABLES, 8, "0"), ("s8fb", lib.SEARCH_TABLES, 8, None)):
        if fb is None:
            monkeypatch.delenv("CIM_TABLE_FALLBACK", raising=False)
        else:
            monkeypatch.setenv("CIM_TABLE_FALLBACK", fb)
        p = lib.make_params(box, d["minR"], d["maxR"], apply_ctf=True, mask_on=True, apix=apix, imgmask=d["mask"],
                            searchstep=step, search_method=method, search_second=S)
        ctx.search_templates(p, timgs, np.c_[tang, np.zeros((4, 2))])
        ctx.counter_reset()
        out[name] = ctx.search_global(p, ps.images, P)
        out[name, "second"] = ctx.counter("cl_rank_second")
        torch.cuda.synchronize()
    plain, s0, s8 = out["plain"], out["s0"], out["s8"]
    same = {k: np.all(out[k][:, [0, 1, 2, 5]] == plain[:, [0, 1, 2, 5]], axis=1) for k in ("s0", "s8", "s8fb")}
    moved = ~np.all(s8[:, [0, 1, 2, 5]] == s0[:, [0, 1, 2, 5]], axis=1)
    print(f"identical to plain: {same['s0'].sum()} without, {same['s8'].sum()} with second windows, "
          f"{same['s8fb'].sum()} with the fallback too; {moved.sum()} results moved; "
          f"{out['s8fb', 'second']} minima found in a second window")
    assert np.all(s8[:, 5] <= s0[:, 5])
    assert np.all(same["s8"][same["s0"]])
    assert np.all(s8[:, 5] >= plain[:, 5])
    assert moved.sum() > 0 and same["s8"].sum() > same["s0"].sum()
    assert out["s8fb", "second"] > 0
    assert same["s8fb"].sum() >= same["s8"].sum()


def test_tabulated_search_needs_closed_grid(ctx):
    with pytest.raises(lib.CimError):
        _search_pair(ctx, lib.SEARCH_TABLES, 0, n=2, step=0.7)


def test_fused_preparation_same_bits(ctx, monkeypatch):
    """The pruned transforms of the search and the insert prepare the images as
    the row pass reads them (the default) or from a prepared copy
    (CIM_PREP_FUSED=0): the search's records and the inserted maps are the same
    bits either way, with the normalisation before the mask, after it, and off."""
    import torch
    from cryoim_amd import synth
    box, apix, step, n = 128, 1.5, 1.0, 96
    ph = synth.make_phantom(box, nblobs=20, sigma=(2.0, 4.0), radius=(40.0, 55.0), seed=111)
    tang = synth.random_au_orientations(4, np.random.default_rng(112))
    timgs = synth.project(ph, tang)
    ps = synth.make_particles_torch(ph, n, seed=113, shift_max=3.0, df_range=(1.0, 3.0), snr=0.1, apix=apix,
                                    mask_radius=box / 2 - 4)
    d = pipelines.derived(box, apix, maxRes=8.0, minRes=300.0, imgmask_A=(box / 2 - 4) * apix)
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 3:5] = ps.shifts
    P[:, 6:8] = ps.df
    P[:, 11] = 1.0
    maxR = box // 2 - 6
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("CIM_PREP_FUSED", fused)
        p = lib.make_params(box, d["minR"], d["maxR"], apply_ctf=True, mask_on=True, apix=apix, imgmask=d["mask"],
                            searchstep=step, search_method=lib.SEARCH_TABLES)
        ctx.search_templates(p, timgs, np.c_[tang, np.zeros((4, 2))])
        res = ctx.search_global(p, ps.images, P)
        out["search", fused] = res
        Q = P.copy()
        Q[:, 0:3] = res[:, 0:3]
        for norm in (0, 1, 2):
            pi_ = lib.make_params(box, 0, maxR, apply_ctf=True, mask_on=True, norm_on=norm, apix=apix,
                                  imgmask=d["mask"])
            ctx.rec_begin(box, maxR, layout=16, rcut=maxR, wg_init=0.01)   # 16: CIM_REC_COMPENSATED
            ctx.rec_insert(pi_, ps.images, Q)
            m = torch.empty((box, box, box), dtype=torch.float32, device="cuda:0")
            ctx.rec_finish(box, m)
            torch.cuda.synchronize()
            out["map", norm, fused] = m.cpu().numpy()
    assert np.array_equal(out["search", "1"], out["search", "0"])
    for norm in (0, 1, 2):
        a, b = out["map", norm, "1"], out["map", norm, "0"]
        assert np.isfinite(a).all() and np.abs(a).max() > 0
        assert np.array_equal(a, b), f"norm_on={norm}: max |diff| {np.abs(a - b).max()}"
