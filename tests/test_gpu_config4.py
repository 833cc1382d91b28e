"""The uneven-box and local-sub-volume workload (config 4) end to end.

* Boxes 256, 384 and 512 (outer radii 100 / 150 / 200 px) interleaved: each
  box's search + insert step captured as one HIP graph on its own context
  (stream), the three graphs replayed in turn twice; every replay's
  orientations and PR equal the step run call by call, and each map equals
  the eager one's.
* The vertex chain of the local programs on 128-pixel particles of a 60-fold
  map: local_box's coordinate mode (loc_box_coordinate.f90:16-41) with the 12
  five-fold vertex operators of icos5f_12vertex_matrix.txt gives each
  sub-particle's local angles, cut centre and fraction; local_refine
  (refine_PR.f90:3-260) refines them on 32-pixel cuts against the vertex
  sub-volume in steps of 0.25 deg for 12 cycles (up to +-3 deg); and
  local_reconstruct's substep (reconstruct_substep.f90, norm after mask) inserts the
  refined sub-particles into a 32^3 volume with the C5 operators.  Every
  stage against the FP64 oracle on the same inputs.
"""
import numpy as np
import pytest

import pipelines
from conftest import rel_l2
from cryoim_amd import lib, synth

pytestmark = pytest.mark.gpu
COMPENSATED = 16
APIX = 1.5
OUTER = {256: 100.0, 384: 150.0, 512: 200.0}


def _records(ps, mask_first=True):
    P = np.zeros((len(ps.images), lib.PART_STRIDE), np.float32)
    P[:, 3:5] = ps.shifts
    P[:, 6:8] = ps.df
    P[:, 9:11] = ps.shifts[0] if mask_first else ps.shifts
    P[:, 11] = 1.0
    return P


def test_uneven_boxes_interleaved_graphs():
    import torch
    torch.cuda.init()
    n = 16
    runs = []
    for box in (256, 384, 512):
        s = OUTER[box] / 55.0
        ph = synth.make_phantom(box, nblobs=20, sigma=(2.0 * s, 4.0 * s), radius=(40.0 * s, 55.0 * s), seed=box)
        rng = np.random.default_rng(box + 1)
        tang = synth.random_au_orientations(4, rng)
        timgs = synth.project(ph, tang)
        ps = synth.make_particles(ph, n, seed=box + 2, shift_max=3.0, df_range=(0.8, 3.5), snr=0.05, apix=APIX,
                                  mask_radius=box / 2 - 4)
        d = pipelines.derived(box, APIX, maxRes=8.0, minRes=300.0, imgmask_A=(box / 2 - 4) * APIX)
        maxR = box // 2 - 2
        psr = lib.make_params(box, d["minR"], d["maxR"], search_method=lib.SEARCH_TABLES, apply_ctf=True,
                              mask_on=True, apix=APIX, imgmask=d["mask"], searchstep=0.5)
        pin = lib.make_params(box, 0, maxR, apply_ctf=True, mask_on=True, norm_on=True, apix=APIX, imgmask=d["mask"])
        images = torch.from_numpy(np.ascontiguousarray(ps.images)).cuda()
        P = _records(ps)
        # eager: search, then two inserts at its orientations
        ce = lib.Context(0)
        ce.search_templates(psr, timgs, np.c_[tang, np.zeros((4, 2))])
        out = ce.search_global(psr, images, P)
        Q = P.copy()
        Q[:, 0:3] = out[:, 0:3]
        ce.rec_begin(box, maxR, layout=COMPENSATED, rcut=maxR, wg_init=0.01)
        ce.rec_insert(pin, images, Q)
        ce.rec_insert(pin, images, Q)
        ref_map = ce.rec_finish(box)
        ce.close()
        # graph on its own context
        cg = lib.Context(0)
        cg.search_templates(psr, timgs, np.c_[tang, np.zeros((4, 2))])
        cg.search_global(psr, images, P)
        cg.rec_begin(box, maxR, layout=COMPENSATED, rcut=maxR, wg_init=0.01)
        parts = torch.from_numpy(P.copy()).cuda()
        cg.step_capture(psr, pin, images, parts)
        runs.append(dict(box=box, ctx=cg, parts=parts, P=P, out=out, ref_map=ref_map))
        print(f"box {box}: graph captured", flush=True)
    for rep in range(2):
        for r in runs:
            r["parts"].copy_(torch.from_numpy(r["P"]))
            r["ctx"].step_replay()
            got = r["parts"].cpu().numpy()
            assert np.array_equal(got[:, [0, 1, 2, 5]], r["out"][:, [0, 1, 2, 5]]), (r["box"], rep)
    for r in runs:
        m = r["ctx"].rec_finish(r["box"])
        err = rel_l2(m, r["ref_map"])
        print(f"box {r['box']}: map after 2 interleaved replays vs 2 eager inserts: rel-L2 {err:.2e}")
        assert err < 1e-6
        r["ctx"].step_release()
        r["ctx"].close()


def test_vertex_chain_local_box_refine_reconstruct_vs_fp64_oracle(ctx, O64):
    g = np.load(pipelines.golden_path("lbox.npz"))
    ops = g["lbox_v12_ops"].reshape(-1, 3, 3)
    assert len(ops) == 12
    box, sub, n = 128, 32, 2
    ph = synth.make_phantom(box, nblobs=20, sigma=(2.0, 4.0), radius=(40.0, 55.0), seed=101)
    rng = np.random.default_rng(7)
    ang = synth.random_au_orientations(n, rng)
    sh = rng.uniform(-2, 2, (n, 2))
    df = rng.uniform(9000, 18000, n)            # Angstrom, as the local programs take it
    imgs = synth.project(ph, ang, sh, box=box)
    for i in range(n):
        c = synth.ctf2d(box, df[i], df[i], 0.0, APIX, vol=200.0, df_scale=1.0)
        imgs[i] = synth.apply_ctf(imgs[i], c) + 0.2 * rng.standard_normal((box, box)).astype(np.float32)
    imgs = imgs.astype(np.float32)
    x0 = np.array([0.0, 30.0, 42.0], np.float32)
    # 1. local_box coordinates (GPU against the oracle)
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3], P[:, 3:5] = ang, sh
    rec = ctx.local_box_coords(P, ops, x0, box, APIX)                      # (n, 12, 8)
    for i in range(n):
        o = O64.local_box_coords(ang[i], sh[i], ops, x0, box, APIX)
        assert np.array_equal(rec[i, :, 3:5], o[:, 3:5])
        assert np.abs(rec[i, :, 5:7] - o[:, 5:7]).max() < 2e-3
    rec = rec.reshape(-1, 8)
    m = len(rec)
    # 2. local_refine of the 24 sub-particles, started 1.5 deg off their local angles
    lx = np.repeat(np.arange(n), 12)
    model = synth.phantom_map(synth.Phantom(ph.centres - x0.astype(np.float64), ph.sigmas, ph.amps, sub))
    d = pipelines.lref_derived(sub, APIX, maxRes=6.0, imgmask_A=20.0)
    step, ncycle = 0.25, 12
    p = lib.make_params(sub, 0, sub // 2 - 2, apply_ctf=True, mask_on=True, conj_out=False, apix=APIX,
                        imgmask=d["mask"], vol=200.0, df_scale=1.0, ncycle=ncycle, delta_angle=step,
                        delta_trans=d["refine_shift"])
    L = np.zeros((m, lib.PART_STRIDE), np.float32)
    L[:, 0:3] = rec[:, 0:3] + rng.uniform(-1.5, 1.5, (m, 3)).astype(np.float32)
    L[:, 3:5] = rec[:, 5:7]
    L[:, lib.P_DF1] = df[lx]
    L[:, lib.P_DF2] = df[lx]
    L[:, lib.P_BOXX:lib.P_BOXY + 1] = rec[:, 3:5]
    ctx.lref_model(p, model)
    full = imgs[lx]
    out = ctx.lref_refine(p, full, L, sym="c1", convF=0.8, maxres=d["maxres"])
    fr, fi = pipelines.lref_model_ft(O64, model, d["mask"])
    same = 0
    for k in range(m):
        a20, b20, w = pipelines.lref_particle(O64, full[k], sub, int(L[k, lib.P_BOXX]), int(L[k, lib.P_BOXY]),
                                              float(L[k, 3]), float(L[k, 4]), d["mask"], d["maxres"])
        ctf = O64.ctf2d(sub, float(L[k, lib.P_DF1]), float(L[k, lib.P_DF2]), 0.0, APIX, vol=200.0, df_scale=1.0)
        io5, pr = O64.lref_refine(fr, fi, a20, b20, ctf, w, None, ncycle, step, d["refine_shift"], 0.8, d["maxres"],
                                  L[k, 0:5].astype(np.float64))
        da = (out[k, 0:3] - io5[:3] + 180) % 360 - 180
        same += np.abs(da).max() < 2e-3 and np.abs(out[k, 3:5] - io5[3:5]).max() < 2e-3
        assert abs(out[k, 5] - pr) < 0.05
    print(f"local_refine: {same}/{m} sub-particles on the FP64 oracle's path")
    assert same >= m - 2
    # 3. local_reconstruct substep (mask, then norm; C5) of the refined sub-particles
    maxR = sub // 2 - 2
    mask = d["mask"]
    cuts = np.stack([np.roll(np.roll(full[k], sub // 2 - int(L[k, lib.P_BOXY]), axis=0),
                             sub // 2 - int(L[k, lib.P_BOXX]), axis=1)[:sub, :sub] for k in range(m)])
    R = np.zeros((m, lib.PART_STRIDE), np.float32)
    R[:, 0:3], R[:, 3:5] = out[:, 0:3], out[:, 3:5]
    R[:, lib.P_DF1], R[:, lib.P_DF2] = df[lx], df[lx]
    R[:, 9:11], R[:, 11] = out[:, 3:5], 1.0
    pr_ = lib.make_params(sub, 0, maxR, apply_ctf=True, mask_on=True, norm_on=2, apix=APIX, imgmask=mask, vol=200.0,
                          df_scale=1.0)
    ctx.rec_begin(sub, maxR, layout=1, rcut=maxR - 1, wg_init=0.1, ops=lib.local_reconstruct_sym_ops("c5"))
    ctx.rec_insert(pr_, cuts, R)
    vol = pipelines.local_reconstruct_volume(O64, full, rec[:, 3:5].astype(int), out[:, 0:3], out[:, 3:5], df[lx],
                                             APIX, sub, maxR, "c5", compact=True, mask_r=mask)
    for o, v, k in zip(ctx.rec_get_tnf(maxR), vol, ("fa", "fb", "wg")):
        err = rel_l2(o, v)
        print(f"local_reconstruct {k}: rel-L2 {err:.2e} vs FP64 oracle")
        assert err < 1e-5
