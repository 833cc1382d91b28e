"""Learned proposals, first slice (cryoim_amd/proposals.py with the network
trained on renders of the config-1 map, cryoim_amd/data/proposal_c1.npz)
against the plain and the tabulated search on 400 config-1 particles it was
not trained on (box 128, 1 deg grid, SNR 0.1): particles whose best
hypothesis confidence reaches the threshold are scored exactly over the AU
points within 3 deg of their 4 hypotheses (windows of 16 omega columns
around each hypothesis's omega), the others take the tabulated search.
Reports the proposal errors, the miss rate against the plain search and the
speed against the tabulated search; asserts what the construction fixes:
a particle whose cones hold the plain minimum gets the plain result bit for
bit, and a particle that falls back gets the tabulated result."""
import time

import numpy as np
import pytest

import pipelines
from cryoim_amd import geometry, lib, proposals

pytestmark = pytest.mark.gpu
CONE, THRESHOLD = 3.0, 0.5


def test_learned_proposals_miss_rate_and_speed(ctx):
    import torch
    n = 400
    w = pipelines.config1_workload(n)
    d = w["d"]
    step = w["step"]
    common = dict(apply_ctf=True, mask_on=True, apix=w["apix"], imgmask=d["mask"], searchstep=step)
    pp = lib.make_params(w["box"], d["minR"], d["maxR"], search_method=lib.SEARCH_PLAIN, **common)
    pt = lib.make_params(w["box"], d["minR"], d["maxR"], search_method=lib.SEARCH_TABLES, **common)
    T = len(w["tang"])
    ctx.search_templates(pt, w["timgs"], np.c_[w["tang"], np.zeros((T, 2))])
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 3:5] = w["ps"].shifts
    P[:, 6:8] = w["ps"].df
    P[:, 11] = 1.0
    images = torch.from_numpy(np.ascontiguousarray(w["ps"].images)).cuda()
    params = proposals.load_weights()

    def timed(f):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = f()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    plain, t_plain = timed(lambda: ctx.search_global(pp, images, P))
    ctx.search_global(pt, images, P)
    tab, t_tab = timed(lambda: ctx.search_global(pt, images, P))
    proposals.propose(ctx, params, images[:8])
    (ang, conf), t_cnn = timed(lambda: proposals.propose(ctx, params, images))
    t0 = time.perf_counter()
    cand, omega = proposals.cone_candidates(ang, conf, step, cone=CONE, min_conf=THRESHOLD)
    t_cones = time.perf_counter() - t0
    sure = conf.max(axis=1) >= THRESHOLD
    si, fi = np.nonzero(sure)[0], np.nonzero(~sure)[0]
    out = P.copy()
    t_cand = t_fb = 0.0
    if len(si):
        out[si], t_cand = timed(lambda: ctx.search_candidates(pt, images[si].contiguous(), P[si], cand[si], omega[si]))
    if len(fi):
        out[fi], t_fb = timed(lambda: ctx.search_global(pt, images[fi].contiguous(), P[fi]))

    best_err = np.stack([geometry.sym_distance_deg(ang[:, h], w["ps"].angles) for h in range(4)], 1).min(1)
    same = np.all(out[:, [0, 1, 2, 5]] == plain[:, [0, 1, 2, 5]], axis=1)
    steps_off = pipelines.angle_steps(out[:, :3], plain[:, :3], step)
    view = geometry.sym_distance_deg(out[:, :3], plain[:, :3])
    # does a confident particle's candidate list hold the plain minimum's point with its omega in a window?
    pts = proposals.grid_points(step)
    No = int(round(360 / step))
    held = np.zeros(n, bool)
    for i in si:
        q = np.nonzero(np.isclose(pts[:, 0], plain[i, 0]) & np.isclose(pts[:, 1], plain[i, 1]))[0]
        if len(q):
            kc = int(round(plain[i, 2] / step)) % No
            k = omega[i][cand[i] == q[0]]
            held[i] = np.any(np.minimum(np.abs(k - kc) % (No // 2), (No // 2) - np.abs(k - kc) % (No // 2)) <= 6)
    t_learned = t_cnn + t_cones + t_cand + t_fb
    print(f"proposal error (best of 4, symmetry-aware): median {np.median(best_err):.2f} deg, p90 "
          f"{np.percentile(best_err, 90):.2f}, within {CONE} deg {np.mean(best_err <= CONE):.3f}")
    print(f"confident (>= {THRESHOLD}): {len(si)}/{n}; their cones hold the plain minimum: {int(held[si].sum())}/"
          f"{len(si)}; fallback to the tables: {len(fi)}")
    print(f"learned vs plain: {int(same.sum())}/{n} identical, {int((steps_off <= 1.0 + 1e-4).sum())} within one "
          f"step, misses {int((steps_off > 1.0 + 1e-4).sum())} ({np.mean(steps_off > 1.0 + 1e-4):.3f}); views within "
          f"1 deg up to symmetry {int((view <= 1.0).sum())}")
    print(f"time for {n} particles: plain {t_plain * 1e3:.1f} ms, tables {t_tab * 1e3:.1f} ms, learned "
          f"{t_learned * 1e3:.1f} ms (network {t_cnn * 1e3:.1f}, cones on the host {t_cones * 1e3:.1f}, exact "
          f"cone scores {t_cand * 1e3:.1f}, tables fallback {t_fb * 1e3:.1f}); learned / tables speed "
          f"{t_tab / t_learned:.2f}")
    assert np.all(same[held]), "a cone that holds the plain minimum must give the plain result"
    assert np.array_equal(out[fi][:, [0, 1, 2, 5]], tab[fi][:, [0, 1, 2, 5]])
    assert np.median(best_err) < 10.0
