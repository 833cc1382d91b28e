"""Throughput of the icosahedral search + insert hot path on one MI355X:
the chosen path against the plain FP32 path, on the same particles.

Workload: box 384 particles of a seeded 60-fold Gaussian-blob phantom,
CTF-modulated (defocus U(0.8, 3.5) um, 300 kV, Cs 2.7 mm, A 0.1, 1.5 A/px)
with Gaussian noise at SNR 0.05.  One step takes a batch of particles through

  * the icosprocess global search (mode 4: mask, FFT, Fcut fit, centre shift,
    common-lines phase residual against T templates over the 0.5 deg AU grid
    x 720 in-plane angles, first-minimum argmin) at maxR 72 (8 A), and
  * the 60-fold direct-Fourier-inversion insert (mode 1 pipeline: norm, CTF,
    mask, FFT, centre shift, trilinear insert) at maxR 190 (Nyquist-2) at
    the orientations the search returned.

Chosen path (value): the tabulated search (CIM_SEARCH_TABLES: per-particle
residual tables over ray pairs, (AU point, omega) scored from them coarse to
fine, the best 32 AU points re-scored exactly, exhaustive fallback) and the
compensated FP32 insert (CIM_REC_COMPENSATED).  Its steps are pipelined: the
search of batch k runs on one library context (stream) while the insert of
batch k-1 runs on another; the timed region holds exactly --steps batches
through both stages, and the same steps timed one after the other are
reported as extra.sequential_particles_per_s.  Plain path: the exhaustive
search of the reference's method and the plain FP32 insert.  The line
reports both rates, how many particles get a different orientation, and the
FSC between the two maps.

`--stage center` measures only the centre finding of icosprocess mode 3
(calc_center against the phantom map, and calc_center_selfrot180) on 224
of those particles; the default run reports its rates under extra.
`--stage lbox` measures local_box's stack mode on 256 of them: per particle
the 96-pixel boxes around its 12 five-fold vertices (local_z 140 px).
`--stage lref` measures local_refine on those 3,072 vertex sub-particles:
the C5 candidates, then 12 cycles of 0.25 deg angle and shift steps.

Inputs are resident in HBM before the timed region.  The roofline block
describes the kernel that takes most of a chosen step (its HIP-event time in
the timed steps); the cpu_baseline block times the C oracle port (FP32, one
process per host core) on a bounded sample of the same workload, before any
GPU use, and quotes the reference's own Fortran rate measured where its
sources were built (profiles/cpu_reference_box384.json).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BOX, APIX, MAXRES, MINRES = 384, 1.5, 8.0, 300.0
INSERT_MAXR = 190
STEP = 0.5
# MI355X published peaks: FP32 (vector and matrix alike), dense BF16 matrix, HBM3E
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
COMPENSATED = 16          # CIM_REC_COMPENSATED
# FP32 operations of one call of the exact score's sample (search.hip score_sample, paired: one
# interpolation feeds a column and its Friedel column, counted as two samples): position 6,
# bilinear weights 6, two 4-tap blends 14, |F| 4 (gate-failing samples stop here: 30), weight 2,
# four products 4, per column two sums, the folded phase difference (atan2_upper: reciprocal,
# 7 FMAs of the polynomial and 5 more: 20) and the two accumulations 3 (25 each): 86
FLOPS_PAIR_CALL_PASS = 86
FLOPS_PAIR_CALL_GATED = 30
# residual tables on the matrix cores: K features per (template ray, particle ray) entry and radius
# (2 for D, 4 NH = 16 for N at NH = 4), 2 flops per multiply-add
TABLE_FEATURES_PER_RADIUS = 2 + 4 * 4
# where each particle's exact minimum ranked among its table candidates (counted by the library)
RANK_KEYS = ("cl_rank_lt4", "cl_rank_lt8", "cl_rank_lt16", "cl_rank_lt24", "cl_rank_ge24", "cl_rank_second")
METRIC = "particles/s search+insert, box 384: chosen vs plain FP32 GPU at FSC>=0.999"


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=512, help="particles per step")
    ap.add_argument("--templates", type=int, default=4, help="template bank size T")
    ap.add_argument("--topk", type=int, default=32, help="AU points re-scored exactly per particle")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="time the chosen path step after step (search, then insert) on one stream")
    ap.add_argument("--stage", choices=("main", "center", "lbox", "lref", "missrate", "c1"), default="main",
                    help="center: only the centre finding of icosprocess mode 3 at box 384; "
                         "c1: the box-128 workload (1 deg grid, 2,000 particles per step), chosen vs plain; "
                         "lbox: only local_box's stack mode (12 vertex boxes of 96 pixels per particle); "
                         "lref: local_refine of those vertex sub-particles (C5, 12 cycles of 0.25 deg); "
                         "missrate: chosen vs plain search on --particles distinct particles")
    ap.add_argument("--particles", type=int, default=5120,
                    help="missrate: distinct particles (generated on the GPU in batches of --batch); "
                         "c1: particles per step (default 2000)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 1 GiB device copy first (HBM counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--chosen-only", action="store_true",
                    help="time only the chosen path, step after step (for counter runs under rocprofv3 --pmc): "
                         "no pipelined leg, plain path, maps or centre stage")
    return ap.parse_args()


def make_workload(batch: int, T: int, seed: int = 1):
    from cryoim_amd import synth
    s = BOX / 128.0
    ph = synth.make_phantom(BOX, nblobs=20, sigma=(2.0 * s, 4.0 * s), radius=(40.0 * s, 55.0 * s), seed=seed)
    rng = np.random.default_rng(seed + 1)
    tang = synth.random_au_orientations(T, rng)
    timgs = synth.project(ph, tang)
    ps = synth.make_particles(ph, batch, seed=seed + 2, shift_max=3.0, df_range=(0.8, 3.5), snr=0.05, apix=APIX,
                              mask_radius=BOX / 2 - 4)
    return ph, tang, timgs, ps


def derived():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pipelines
    return pipelines.derived(BOX, APIX, maxRes=MAXRES, minRes=MINRES, imgmask_A=(BOX / 2 - 4) * APIX)


def _fsc(a, b, rmax):
    """Shell FSC (ir = int(r)) of two real maps on the GPU, shells 0..rmax."""
    import torch
    fa, fb = torch.fft.rfftn(a), torch.fft.rfftn(b)
    n = a.shape[0]
    f0 = torch.fft.fftfreq(n, d=1.0 / n, device=a.device)
    f2 = torch.fft.rfftfreq(n, d=1.0 / n, device=a.device)
    r = torch.sqrt(f0[:, None, None] ** 2 + f0[None, :, None] ** 2 + f2[None, None, :] ** 2).to(torch.int64)
    r = r.reshape(-1)
    keep = r <= rmax
    r = r[keep]
    fa, fb = fa.reshape(-1)[keep], fb.reshape(-1)[keep]
    num = torch.bincount(r, weights=(fa * fb.conj()).real.double(), minlength=rmax + 1)
    da = torch.bincount(r, weights=(fa.abs() ** 2).double(), minlength=rmax + 1)
    db = torch.bincount(r, weights=(fb.abs() ** 2).double(), minlength=rmax + 1)
    return (num / torch.sqrt(da * db).clamp_min(1e-300)).cpu().numpy()


def gpu_run(args, tang, timgs, ps):
    import torch
    from cryoim_amd import lib
    d = derived()
    ctx = lib.Context(0)
    n = len(ps.images)
    common = dict(apply_ctf=True, mask_on=True, apix=APIX, imgmask=d["mask"], searchstep=STEP)
    p_plain = lib.make_params(BOX, d["minR"], d["maxR"], search_method=lib.SEARCH_PLAIN, **common)
    p_chosen = lib.make_params(BOX, d["minR"], d["maxR"], search_method=lib.SEARCH_TABLES, search_topk=args.topk,
                               **common)
    p_insert = lib.make_params(BOX, 0, INSERT_MAXR, apply_ctf=True, mask_on=True, norm_on=True, apix=APIX,
                               imgmask=d["mask"])
    ctx.search_templates(p_chosen, timgs, np.c_[tang, np.zeros((len(tang), 2))])
    # second context (its own stream) for the pipelined chosen path: the insert of
    # batch k-1 runs there while the search of batch k runs on ctx
    ctx_i = lib.Context(0)
    images = torch.from_numpy(ps.images).to("cuda:0")
    if args.calib_copy:
        # 1 GiB copy after 1 GiB of unrelated writes, so the source is read from HBM
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        flush = torch.zeros(256 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst, flush
    torch.cuda.synchronize()
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 3:5] = ps.shifts
    P[:, 6:8] = ps.df
    P[:, 8] = ps.astig
    P[:, 9:11] = ps.shifts[0]          # reconstruct.f90:92 masks around the first particle
    P[:, 11] = 1.0

    def step(psearch):
        out = ctx.search_global(psearch, images, P)
        Q = P.copy()
        Q[:, 0:3] = out[:, 0:3]
        ctx.rec_insert(p_insert, images, Q)
        return out

    stream = torch.cuda.ExternalStream(ctx.stream)

    def timed(psearch, layout, steps, warmup):
        ctx.rec_begin(BOX, INSERT_MAXR, layout=layout, rcut=INSERT_MAXR, wg_init=0.01)
        for _ in range(warmup):
            step(psearch)
        ctx.sync()
        ctx.timing(True)
        ctx.counter_reset()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            out = step(psearch)
        e1.record(stream)
        ctx.sync()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kt = {lab: ctx.timing_get(lab) for lab in ("cl_search", "cl_polar", "cl_table", "cl_coarse_table",
                                                   "cl_gather", "cl_rescore",
                                                   "cl_fallback", "insert")}
        cnt = {k: ctx.counter(k) for k in ("cl_search_samples", "insert_samples", "cl_fallback_particles",
                                           "cl_table_pairs", "cl_gather_points", "cl_rescore_samples",
                                           "cl_rescore_pass", *RANK_KEYS)}
        ctx.timing(False)
        return out, e0.elapsed_time(e1), wall, kt, cnt

    def timed_pipelined(psearch, layout, steps, warmup):
        """steps batches through search and insert, the search of batch k on ctx
        overlapping the insert of batch k-1 on ctx_i (two host threads; the
        library calls release the GIL).  The timed region holds the first
        search alone and the last insert alone, so it is exactly steps batches."""
        from concurrent.futures import ThreadPoolExecutor

        def insert(out):
            Q = P.copy()
            Q[:, 0:3] = out[:, 0:3]
            ctx_i.rec_insert(p_insert, images, Q)

        ctx_i.rec_begin(BOX, INSERT_MAXR, layout=layout, rcut=INSERT_MAXR, wg_init=0.01)
        for _ in range(warmup):
            insert(ctx.search_global(psearch, images, P))
        ctx.sync()
        ctx_i.sync()
        for c in (ctx, ctx_i):
            c.timing(True)
            c.counter_reset()
        torch.cuda.synchronize()
        out = None
        with ThreadPoolExecutor(max_workers=2) as pool:
            t0 = time.perf_counter()
            prev = None
            for k in range(steps + 1):
                fs = pool.submit(ctx.search_global, psearch, images, P) if k < steps else None
                fi = pool.submit(insert, prev) if prev is not None else None
                nxt = fs.result() if fs is not None else None
                if fi is not None:
                    fi.result()
                if nxt is not None:
                    out = nxt
                prev = nxt
            ctx.sync()
            ctx_i.sync()
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
        kt = {lab: ctx.timing_get(lab) for lab in ("cl_search", "cl_polar", "cl_table", "cl_coarse_table",
                                                   "cl_gather", "cl_rescore",
                                                   "cl_fallback")}
        kt["insert"] = ctx_i.timing_get("insert")
        cnt = {k: ctx.counter(k) for k in ("cl_search_samples", "cl_fallback_particles", "cl_table_pairs",
                                           "cl_gather_points", "cl_rescore_samples", "cl_rescore_pass", *RANK_KEYS)}
        cnt["insert_samples"] = ctx_i.counter("insert_samples")
        for c in (ctx, ctx_i):
            c.timing(False)
        return out, wall * 1e3, wall, kt, cnt

    if args.chosen_only:
        out_c, ev_c, wall_c, kt_c, cnt_c = timed(p_chosen, COMPENSATED, args.steps, args.warmup)
        ctx.close()
        ctx_i.close()
        return dict(n=n, d=d, ev_c=ev_c, wall_c=wall_c, kt_c=kt_c, cnt_c=cnt_c)
    # chosen path, timed: pipelined (the value) and sequential (reported beside it)
    if args.no_pipeline:
        out_c, ev_c, wall_c, kt_c, cnt_c = timed(p_chosen, COMPENSATED, args.steps, args.warmup)
        ev_seq, kt_seq = ev_c, kt_c
    else:
        out_c, ev_c, wall_c, kt_c, cnt_c = timed_pipelined(p_chosen, COMPENSATED, args.steps, args.warmup)
        _, ev_seq, _, kt_seq, _ = timed(p_chosen, COMPENSATED, args.steps, 0)
    # the same step captured as one HIP graph (cim_step_capture) and replayed step after step
    ctx.rec_begin(BOX, INSERT_MAXR, layout=COMPENSATED, rcut=INSERT_MAXR, wg_init=0.01)
    parts_dev = torch.from_numpy(P.copy()).to("cuda:0")
    ctx.step_capture(p_chosen, p_insert, images, parts_dev)
    for _ in range(args.warmup):
        ctx.step_replay()
    g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    g0.record(stream)
    for _ in range(args.steps):
        ctx.step_replay()
    g1.record(stream)
    torch.cuda.synchronize()
    ev_graph = g0.elapsed_time(g1)
    graph_same = int(np.all(parts_dev.cpu().numpy()[:, [0, 1, 2, 5]] == out_c[:, [0, 1, 2, 5]], axis=1).sum())
    ctx.step_release()
    # plain path on the same particles: one timed step after one untimed
    out_p, ev_p, wall_p, kt_p, cnt_p = timed(p_plain, 0, 1, 1)
    # one map per path from one pass over the batch, and their FSC
    maps = []
    for psearch_out, layout in ((out_c, COMPENSATED), (out_p, 0)):
        ctx.rec_begin(BOX, INSERT_MAXR, layout=layout, rcut=INSERT_MAXR, wg_init=0.01)
        Q = P.copy()
        Q[:, 0:3] = psearch_out[:, 0:3]
        ctx.rec_insert(p_insert, images, Q)
        m = torch.empty((BOX, BOX, BOX), dtype=torch.float32, device="cuda:0")
        ctx.rec_finish(BOX, m)
        ctx.sync()
        maps.append(m)
    fsc = _fsc(maps[0], maps[1], INSERT_MAXR)
    map_rel = float((torch.linalg.vector_norm(maps[0] - maps[1]) / torch.linalg.vector_norm(maps[1])).item())
    ctx.close()
    ctx_i.close()
    dang = np.abs(((out_c[:, 0:3] - out_p[:, 0:3] + 180) % 360) - 180).max(axis=1)
    identical = int(np.all(out_c[:, [0, 1, 2, 5]] == out_p[:, [0, 1, 2, 5]], axis=1).sum())
    within = int((dang <= STEP + 1e-4).sum())
    ang_err = np.abs(((out_p[:, 0:3] - ps.angles + 180) % 360) - 180)
    return dict(n=n, d=d, ev_c=ev_c, wall_c=wall_c, kt_c=kt_c, cnt_c=cnt_c, ev_p=ev_p, wall_p=wall_p, kt_p=kt_p,
                cnt_p=cnt_p, identical=identical, within=within, fsc=fsc, map_rel=map_rel, ang_err=ang_err,
                ev_seq=ev_seq, kt_seq=kt_seq, ev_graph=ev_graph, graph_same=graph_same)


CENTER_BATCH = 224       # one working chunk at box 384
CENTER_METRIC = "particles/s centre finding (icosprocess mode 3, model3d), box 384"


def center_run(ph, ps, steps: int, warmup: int):
    """icosprocess mode 3 on CENTER_BATCH particles resident in HBM: calc_center
    against the phantom map (CTF, mask, low-pass at maxR) and
    calc_center_selfrot180 on the same images."""
    import torch
    from cryoim_amd import lib, synth
    d = derived()
    n = CENTER_BATCH
    idx = np.arange(n) % len(ps.images)
    ctx = lib.Context(0)
    p = lib.make_params(BOX, 0, d["maxR"], apply_ctf=True, mask_on=True, apix=APIX, imgmask=d["mask"])
    ctx.center_model(p, synth.phantom_map(ph))
    images = torch.from_numpy(np.ascontiguousarray(ps.images[idx])).to("cuda:0")
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3] = ps.angles[idx]
    P[:, 3:5] = ps.shifts[idx]
    P[:, 6:8] = ps.df[idx]
    P[:, 11] = 1.0
    stream = torch.cuda.ExternalStream(ctx.stream)
    res = {}
    for name, method in (("model", lib.CENTER_MODEL), ("selfrot180", lib.CENTER_SELFROT180)):
        for _ in range(warmup):
            ctx.center(p, method, images, P)
        ctx.sync()
        ctx.timing(True)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            out, cc = ctx.center(p, method, images, P)
        e1.record(stream)
        ctx.sync()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        kt = {k: ctx.timing_get(k) for k in ("fft2d", "cen_stage", "cen_slice", "cen_elem", "cen_peak")}
        ctx.timing(False)
        err = np.abs(out[:, 3:5] - P[:, 3:5]).max(axis=1)
        res[name] = {"particles_per_s": round(n / (ms / 1e3), 1), "ms_per_step": round(ms, 3),
                     "kernels": {k: {"avg_ms": round(v[0], 4), "launches": v[1]} for k, v in kt.items() if v[1]},
                     "median_shift_error_px": round(float(np.median(err)), 3),
                     "cc_median": round(float(np.median(cc)), 4)}
    ctx.close()
    return res


def refine_run(tang, timgs, ps, steps: int, warmup: int):
    """icosprocess mode 5 (ort_refinement: Ncycle 6 cycles of the 3^5 stencil
    around the starting orientation and shift, deltaAngle and deltaTrans of
    0calc_initialvalues) on the batch, started from the true orientations
    rounded to the search grid.  Host images (cim_refine_local): the wall
    time includes their copy to the device; the kernel time does not."""
    from cryoim_amd import lib
    d = derived()
    n = len(ps.images)
    ctx = lib.Context(0)
    maxres = d["maxR"]
    p = lib.make_params(BOX, d["minR"], d["maxR"], apply_ctf=True, mask_on=True, apix=APIX, imgmask=d["mask"],
                        searchstep=STEP, ncycle=6, delta_angle=360.0 / (maxres * np.sqrt(5.0) * 3.14159),
                        delta_trans=BOX / (maxres * np.sqrt(5.0)))
    ctx.search_templates(p, timgs, np.c_[tang, np.zeros((len(tang), 2))])
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3] = np.round(ps.angles / STEP) * STEP
    P[:, 3:5] = np.round(ps.shifts)
    P[:, 6:8] = ps.df
    P[:, 11] = 1.0
    for _ in range(warmup):
        ctx.refine_local(p, ps.images, P)
    ctx.timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        out = ctx.refine_local(p, ps.images, P)
    wall = (time.perf_counter() - t0) / steps
    kms, launches = ctx.timing_get("refine")
    ctx.timing(False)
    ctx.close()
    return {"particles_per_s": round(n / wall, 1), "ms_per_step": round(wall * 1e3, 3),
            "kernel_avg_ms": round(kms, 3), "kernel_particles_per_s": round(n / (kms / 1e3), 1),
            "median_pr_deg": round(float(np.median(out[:, 5])), 4),
            "median_shift_change_px": round(float(np.median(np.abs(out[:, 3:5] - P[:, 3:5]))), 3)}


_PORT_STATE = {}


def _port_center_job(i):
    """calc_center (model3d) of particle i by the C oracle (FP32)."""
    O, _ = _oracle32()
    ps, fr, fi = _PORT_STATE["ps"], _PORT_STATE["fr"], _PORT_STATE["fi"]
    d = derived()
    t0 = time.perf_counter()
    ctf = O.ctf2d(BOX, ps.df[i, 0], ps.df[i, 1], 0.0, APIX)
    O.calc_center_one(fr, fi, ps.images[i], *ps.shifts[i], ps.angles[i], ctf, d["mask"], d["maxR"])
    return time.perf_counter() - t0


def center_cpu_baseline(ph, ps):
    """icosprocess mode 3 (calc_center with model3d) by the C oracle port on
    the host cores, 4 particles per process; the model's masked 3D transform
    is computed once beforehand, as the reference does once per run.  The
    reference's own rate is quoted from profiles/."""
    from cryoim_amd import synth
    O, _ = _oracle32()
    ncores = min(8, os.cpu_count() or 1)
    t0 = time.perf_counter()
    fr, fi = O.center_model(synth.phantom_map(ph), derived()["mask"])
    t_model = time.perf_counter() - t0
    _PORT_STATE.update(ps=ps, fr=fr, fi=fi)
    K = 4
    secs, wall = _port_pool(_port_center_job, [i % len(ps.images) for i in range(K * ncores)], ncores)
    _PORT_STATE.clear()
    per = float(np.mean(secs))
    return {"value": round(ncores / per, 3), "unit": "particles/s", "cores": ncores, "kind": "port",
            "sample": (f"C oracle (FP32) calc_center in {ncores} processes, {K} particles each (mean {per:.2f} s "
                       f"per particle, {wall:.1f} s wall) after the model transform ({t_model:.1f} s, once)"),
            "reference_here": _quoted_reference("center")}


def center_main(args):
    ph, _, _, ps = make_workload(args.batch, args.templates)
    cpu = None if args.no_cpu_baseline else center_cpu_baseline(ph, ps)      # before any GPU use
    r = center_run(ph, ps, args.steps, args.warmup)
    m = r["model"]
    n = CENTER_BATCH
    fft = m["kernels"]["fft2d"]
    # one fft2d call: row and column passes, each reading and writing n box^2 complex64
    nbytes = 4 * n * BOX * BOX * 8
    achieved = nbytes / (fft["avg_ms"] / 1e3) / 1e9
    line = {"metric": CENTER_METRIC, "value": m["particles_per_s"], "unit": "particles/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (the main workload's particles, tiled to the batch)",
            "config": {"workload": f"calc_center on {n} particles of box {BOX}, maxR {derived()['maxR']}",
                       "box": BOX, "batch": n},
            "roofline": {"bound": "hbm", "kernel": "fft2d (k_fft_lines row + column pass)",
                         "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(achieved / PEAK_HBM_GBS, 3), "traffic": None,
                         "algorithmic_bytes_per_launch": nbytes, "avg_launch_ms": fft["avg_ms"],
                         "launches": fft["launches"]},
            "cpu_baseline": cpu,
            "extra": r}
    print(json.dumps(line))


C1_N, C1_STEP, C1_PORT_STEP, C1_SECOND = 2000, 1.0, 2.0, 8
C1_METRIC = ("particles/s search+insert, box 128 (20-blob phantom, 1 deg AU grid x 360 omega, maxR 24, T=4; "
             "60-fold insert at maxR 62): chosen vs plain FP32 GPU")


def _port_c1_job(i):
    """Particle i of the box-128 workload through the C oracle (FP32):
    ort_globesearch on the C1_PORT_STEP grid, then the 60-fold insert at
    the workload's insert radius."""
    O, pipelines = _oracle32()
    w = _PORT_STATE["w"]
    d, box, apix, ps = w["d"], w["box"], w["apix"], w["ps"]
    t0 = time.perf_counter()
    ta, tb = pipelines.templates(O, w["timgs"], np.zeros((len(w["tang"]), 2)), d["mask"])
    ctf = O.ctf2d(box, ps.df[i, 0], ps.df[i, 1], 0.0, apix)
    a, b, tf = pipelines.template_inputs(O, ta, tb, ctf, d["minR"], d["maxR"])
    _, (pa, pb), fc0 = pipelines.particle_inputs(O, ps.images[i], *ps.shifts[i], d["mask"], d["minR"], d["maxR"])
    t_prep = time.perf_counter() - t0
    t0 = time.perf_counter()
    o4 = O.globesearch(pa, pb, fc0, a, b, tf, w["tang"], d["minR"], d["maxR"], C1_PORT_STEP)
    t_search = time.perf_counter() - t0
    maxR = w["insert_maxR"]
    vol = O.new_volume(box, maxR, compact=True, wg_init=0.0)
    t0 = time.perf_counter()
    fa, fb = O.particle_ft(ps.images[i], *ps.shifts[i], mask_r=d["mask"], norm=True)
    O.insert(vol, fa, fb, ctf, o4[:3], maxR, rcut=maxR - 2, compact=True)
    return t_prep, t_search, time.perf_counter() - t0


def c1_cpu_baseline(w):
    """The C oracle port (FP32) on the host cores, one particle per process:
    transforms, the search's scoring loop on the C1_PORT_STEP grid scaled to
    the C1_STEP grid by the ratio of scored AU-cell x omega points, and the
    insert.  The reference's own rate is quoted from profiles/."""
    ncores = min(8, os.cpu_count() or 1)
    _PORT_STATE.update(w=w)
    res, wall = _port_pool(_port_c1_job, list(range(ncores)), ncores)
    _PORT_STATE.clear()
    scale = _grid_work(C1_STEP) / _grid_work(C1_PORT_STEP)
    per = float(np.mean([tp + ts * scale + ti for tp, ts, ti in res]))
    m = np.mean(res, axis=0)
    return {"value": ncores / per, "unit": "particles/s", "cores": ncores, "kind": "port",
            "sample": (f"C oracle (FP32) in {ncores} processes, one particle each: transforms (mean {m[0]:.2f} s) "
                       f"+ the search's scoring on the {C1_PORT_STEP} deg grid (mean {m[1]:.2f} s, x{scale:.1f} to "
                       f"the {C1_STEP} deg grid by AU-cell x omega count) + 60-fold insert at maxR "
                       f"{w['insert_maxR']} (mean {m[2]:.3f} s); {wall:.1f} s wall"),
            "reference_here": _quoted_reference("c1")}


def c1_main(args):
    """The box-128 workload (the 20-blob phantom at 128^3, C1_N particles at
    uniform AU orientations, +-3 px shifts, defocus U(1, 3) um, SNR 0.1):
    one step is the whole particle set through the global search on the
    1 deg grid and the 60-fold insert at Nyquist-2, chosen path against
    plain path, and the map of each."""
    import torch
    from cryoim_amd import lib
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pipelines
    n = args.particles if args.particles != 5120 else C1_N
    w = pipelines.config1_workload(n)
    cpu = None if args.no_cpu_baseline else c1_cpu_baseline(w)          # host cores, before any GPU use
    box, apix, d, ps = w["box"], w["apix"], w["d"], w["ps"]
    maxR = w["insert_maxR"]
    ctx = lib.Context(0)
    common = dict(apply_ctf=True, mask_on=True, apix=apix, imgmask=d["mask"], searchstep=C1_STEP)
    p_plain = lib.make_params(box, d["minR"], d["maxR"], search_method=lib.SEARCH_PLAIN, **common)
    # at maxR 24 the table ranking is coarser: 8 second windows per particle (measured on the
    # 2,000 particles: 1,987 identical to the plain search without them, 1,995 with them)
    p_chosen = lib.make_params(box, d["minR"], d["maxR"], search_method=lib.SEARCH_TABLES, search_topk=args.topk,
                               search_second=C1_SECOND, **common)
    p_insert = lib.make_params(box, 0, maxR, apply_ctf=True, mask_on=True, norm_on=True, apix=apix,
                               imgmask=d["mask"])
    T = len(w["tang"])
    ctx.search_templates(p_chosen, w["timgs"], np.c_[w["tang"], np.zeros((T, 2))])
    images = torch.from_numpy(ps.images).to("cuda:0")
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 3:5] = ps.shifts
    P[:, 6:8] = ps.df
    P[:, 9:11] = ps.shifts[0]          # reconstruct.f90:92 masks around the first particle
    P[:, 11] = 1.0
    stream = torch.cuda.ExternalStream(ctx.stream)

    def run(psearch, layout, steps, warmup):
        ctx.rec_begin(box, maxR, layout=layout, rcut=maxR, wg_init=0.01)
        out = None
        for k in range(warmup + steps):
            if k == warmup:
                ctx.sync()
                ctx.timing(True)
                ctx.counter_reset()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                torch.cuda.synchronize()
                e0.record(stream)
            out = ctx.search_global(psearch, images, P)
            Q = P.copy()
            Q[:, 0:3] = out[:, 0:3]
            ctx.rec_insert(p_insert, images, Q)
        e1.record(stream)
        ctx.sync()
        torch.cuda.synchronize()
        kt = {lab: ctx.timing_get(lab) for lab in ("cl_search", "cl_polar", "cl_table", "cl_coarse_table",
                                                   "cl_gather", "cl_rescore", "cl_fallback", "insert")}
        cnt = {k: ctx.counter(k) for k in ("cl_fallback_particles", "cl_rescore_samples", "cl_rescore_pass",
                                           "cl_table_pairs", "insert_samples", *RANK_KEYS)}
        ctx.timing(False)
        return out, e0.elapsed_time(e1), kt, cnt

    def one_map(out, layout):
        ctx.rec_begin(box, maxR, layout=layout, rcut=maxR, wg_init=0.01)
        Q = P.copy()
        Q[:, 0:3] = out[:, 0:3]
        ctx.rec_insert(p_insert, images, Q)
        m = torch.empty((box, box, box), dtype=torch.float32, device="cuda:0")
        ctx.rec_finish(box, m)
        ctx.sync()
        return m

    out_c, ms_c, kt_c, cnt_c = run(p_chosen, COMPENSATED, args.steps, args.warmup)
    out_p, ms_p, kt_p, _ = run(p_plain, 0, 1, 1)
    map_c, map_p = one_map(out_c, COMPENSATED), one_map(out_p, 0)
    ctx.close()
    fsc = _fsc(map_c, map_p, maxR)
    rel = float((torch.linalg.vector_norm(map_c - map_p) / torch.linalg.vector_norm(map_p)).item())
    from cryoim_amd import geometry
    eq = np.all(out_c[:, [0, 1, 2, 5]] == out_p[:, [0, 1, 2, 5]], axis=1)
    same = int(eq.sum())
    dang = np.abs(((out_c[:, 0:3] - out_p[:, 0:3] + 180) % 360) - 180).max(axis=1)
    sym = geometry.sym_distance_deg(out_c[:, 0:3], out_p[:, 0:3])
    differing = [{"index": int(i), "true": [round(float(v), 3) for v in ps.angles[i]],
                  "chosen": [round(float(v), 4) for v in out_c[i, [0, 1, 2, 5]]],
                  "plain": [round(float(v), 4) for v in out_p[i, [0, 1, 2, 5]]],
                  "sym_distance_deg": round(float(sym[i]), 3)} for i in np.nonzero(~eq)[0]]
    value = n * args.steps / (ms_c / 1e3)
    plain = n / (ms_p / 1e3)
    kt = {k: {"avg_ms": round(v[0], 3), "launches": v[1]} for k, v in kt_c.items() if v[1]}
    line = {"metric": C1_METRIC, "value": round(value, 1), "unit": "particles/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_c / args.steps, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded 20-blob 60-fold phantom at 128^3, projections with CTF, SNR 0.1)",
            "config": {"workload": f"{n} particles of box {box} per step: global common-lines search (1 deg AU grid x "
                                   f"360 omega, maxR {d['maxR']}, T={T}; chosen: tabulated ranking, top-{args.topk} "
                                   f"+ {C1_SECOND} second windows exact re-score) + 60-fold insert at maxR {maxR} "
                                   f"(chosen: compensated FP32)",
                       "box": box, "batch": n, "templates": T, "search_maxR": d["maxR"], "insert_maxR": maxR,
                       "searchstep": C1_STEP, "apix": apix, "topk": args.topk, "second_windows": C1_SECOND},
            "cpu_baseline": cpu,
            "extra": {"plain_particles_per_s": round(plain, 1), "chosen_over_plain": round(value / plain, 2),
                      "orientation_identical": same, "orientation_within_one_step": int((dang <= C1_STEP + 1e-4).sum()),
                      "particles_compared": n, "fallback_particles": int(cnt_c["cl_fallback_particles"]),
                      "fsc_min_to_nyquist_minus_2": round(float(np.min(fsc[1:])), 6), "map_rel_l2_chosen_vs_plain": rel,
                      "orientation_within_one_step_by_view": int((sym <= C1_STEP + 1e-4).sum()),
                      "exact_minimum_rank": {k[8:]: int(cnt_c[k]) for k in RANK_KEYS},
                      "chosen_kernels": kt, "plain_search_avg_ms": round(kt_p["cl_search"][0], 3),
                      "differing": differing[:40]}}
    print(json.dumps(line))


LBOX_BATCH, LBOX_SUB, LBOX_Z = 256, 96, 140.0
LBOX_METRIC = "particles/s local_box stack mode, box 384: 12 vertex sub-particles of 96^2 per particle"


MISS_METRIC = "chosen vs plain global search on distinct particles, box 384: decisions that differ"


def missrate_main(args):
    """The chosen (tabulated) and the plain search on args.particles distinct
    particles of the bench recipe, generated on the GPU batch by batch (seeds
    1000 + batch); every particle's orientation and PR compared, fallbacks
    and ranks counted, and the maps of both sets of orientations compared."""
    import torch
    from scipy.stats import beta
    from cryoim_amd import geometry, lib, synth
    d = derived()
    B = args.batch
    nbat = (args.particles + B - 1) // B
    s = BOX / 128.0
    ph = synth.make_phantom(BOX, nblobs=20, sigma=(2.0 * s, 4.0 * s), radius=(40.0 * s, 55.0 * s), seed=1)
    tang = synth.random_au_orientations(args.templates, np.random.default_rng(2))
    timgs = synth.project(ph, tang)
    ctx = lib.Context(0)
    common = dict(apply_ctf=True, mask_on=True, apix=APIX, imgmask=d["mask"], searchstep=STEP)
    p_plain = lib.make_params(BOX, d["minR"], d["maxR"], search_method=lib.SEARCH_PLAIN, **common)
    p_chosen = lib.make_params(BOX, d["minR"], d["maxR"], search_method=lib.SEARCH_TABLES, search_topk=args.topk,
                               **common)
    p_insert = lib.make_params(BOX, 0, INSERT_MAXR, apply_ctf=True, mask_on=True, norm_on=True, apix=APIX,
                               imgmask=d["mask"])
    ctx.search_templates(p_chosen, timgs, np.c_[tang, np.zeros((len(tang), 2))])
    maps = [lib.Context(0), lib.Context(0)]
    for m in maps:
        m.rec_begin(BOX, INSERT_MAXR, layout=COMPENSATED, rcut=INSERT_MAXR, wg_init=0.01)
    ctx.counter_reset()
    t_c = t_p = 0.0
    n = ident = within = 0
    misses, err_c, err_p = [], [], []
    fb = 0
    for b in range(nbat):
        m = min(B, args.particles - b * B)
        ps = synth.make_particles_torch(ph, m, seed=1000 + b, shift_max=3.0, df_range=(0.8, 3.5), snr=0.05,
                                        apix=APIX, mask_radius=BOX / 2 - 4)
        torch.cuda.synchronize()
        P = np.zeros((m, lib.PART_STRIDE), np.float32)
        P[:, 3:5] = ps.shifts
        P[:, 6:8] = ps.df
        P[:, 9:11] = ps.shifts[0]
        P[:, 11] = 1.0
        t0 = time.perf_counter()
        oc = ctx.search_global(p_chosen, ps.images, P)
        t_c += time.perf_counter() - t0
        t0 = time.perf_counter()
        op = ctx.search_global(p_plain, ps.images, P)
        t_p += time.perf_counter() - t0
        for mc, o in zip(maps, (oc, op)):
            Q = P.copy()
            Q[:, 0:3] = o[:, 0:3]
            mc.rec_insert(p_insert, ps.images, Q)
        same = np.all(oc[:, [0, 1, 2, 5]] == op[:, [0, 1, 2, 5]], axis=1)
        dang = np.abs(((oc[:, 0:3] - op[:, 0:3] + 180) % 360) - 180).max(axis=1)
        ident += int(same.sum())
        within += int((dang <= STEP + 1e-4).sum())
        for i in np.nonzero(~same)[0]:
            misses.append({"batch": b, "index": int(i), "true": [round(float(v), 3) for v in ps.angles[i]],
                           "chosen": [round(float(v), 4) for v in oc[i, [0, 1, 2, 5]]],
                           "plain": [round(float(v), 4) for v in op[i, [0, 1, 2, 5]]],
                           "sym_distance_deg": round(float(geometry.sym_distance_deg(oc[i:i + 1, 0:3],
                                                                                      op[i:i + 1, 0:3])[0]), 3)})
        # view errors against the true orientations, over the 60 symmetry-equivalent forms
        err_c.append(geometry.sym_distance_deg(oc[:, 0:3], ps.angles))
        err_p.append(geometry.sym_distance_deg(op[:, 0:3], ps.angles))
        n += m
        print(f"missrate batch {b + 1}/{nbat}: {n} particles, {n - ident} differ so far", file=sys.stderr, flush=True)
        del ps
    fb = ctx.counter("cl_fallback_particles")
    ranks = {k[8:]: int(ctx.counter(k)) for k in RANK_KEYS}
    vols = []
    for mc in maps:
        v = torch.empty((BOX, BOX, BOX), dtype=torch.float32, device="cuda:0")
        mc.rec_finish(BOX, v)
        mc.sync()
        vols.append(v)
    fsc = _fsc(vols[0], vols[1], INSERT_MAXR)
    rel = float((torch.linalg.vector_norm(vols[0] - vols[1]) / torch.linalg.vector_norm(vols[1])).item())
    for c in (ctx, *maps):
        c.close()
    k = n - ident
    upper = float(beta.ppf(0.95, k + 1, n - k)) if k < n else 1.0
    err_c, err_p = np.concatenate(err_c), np.concatenate(err_p)
    print(json.dumps({"metric": MISS_METRIC, "particles": n, "batches": nbat, "identical": ident,
                      "within_one_step": within, "differ": k, "differ_rate": k / n,
                      "differ_rate_upper95": upper, "fallback_particles": int(fb), "exact_minimum_rank": ranks,
                      "chosen_particles_per_s": round(n / t_c, 1), "plain_particles_per_s": round(n / t_p, 2),
                      "fsc_min_to_r190": round(float(np.min(fsc[1:])), 6), "map_rel_l2_chosen_vs_plain": rel,
                      "differ_sym_distance_max_deg": max([m_["sym_distance_deg"] for m_ in misses], default=0.0),
                      "view_error_vs_truth_deg": {"chosen_median": float(np.median(err_c)),
                                                  "plain_median": float(np.median(err_p)),
                                                  "chosen_p99": float(np.percentile(err_c, 99)),
                                                  "plain_p99": float(np.percentile(err_p, 99)),
                                                  "chosen_max": float(err_c.max()), "plain_max": float(err_p.max()),
                                                  "note": "smallest rotation to the true view over the 60 "
                                                          "icosahedral equivalents"},
                      "differing": misses[:20],
                      "config": {"workload": f"box {BOX}, {STEP} deg AU grid x 720 omega, maxR {d['maxR']}, "
                                             f"T={args.templates}, top-{args.topk}; particles of the bench recipe "
                                             f"generated on the GPU (seeds 1000 + batch)", "batch": B}}))


def vertex_ops():
    """The 12 operators of local_box's I5 set (5-fold along z) that take the
    z axis to the 12 distinct five-fold vertices: x0 = (0, 0, z) gives one
    sub-particle box per vertex, as the 12-vertex matrix files do."""
    from cryoim_amd import lib
    ops = lib.local_box_sym_ops("i5")
    keep, seen = [], []
    for S in ops:
        v = S[2]                       # S^T z
        if not any(np.abs(v - u).max() < 1e-3 for u in seen):
            seen.append(v)
            keep.append(S)
    assert len(keep) == 12
    return np.array(keep)


def lbox_run(ps, steps: int, warmup: int):
    """local_box stack mode on LBOX_BATCH particles resident in HBM."""
    import torch
    from cryoim_amd import lib
    n = LBOX_BATCH
    idx = np.arange(n) % len(ps.images)
    ops = vertex_ops()
    ctx = lib.Context(0)
    images = torch.from_numpy(np.ascontiguousarray(ps.images[idx])).to("cuda:0")
    boxes = torch.empty(n * len(ops) * LBOX_SUB * LBOX_SUB, dtype=torch.float32, device="cuda:0")
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3] = ps.angles[idx]
    P[:, 3:5] = ps.shifts[idx]
    x0 = (0.0, 0.0, LBOX_Z)
    stream = torch.cuda.ExternalStream(ctx.stream)
    for _ in range(warmup):
        ctx.local_box_stack(images, P, ops, x0, LBOX_SUB, APIX, boxes=boxes, box=BOX)
    ctx.sync()
    ctx.timing(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        _, rec = ctx.local_box_stack(images, P, ops, x0, LBOX_SUB, APIX, boxes=boxes, box=BOX)
    e1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    kt = {k: ctx.timing_get(k) for k in ("lbox_geom", "lbox_rows", "lbox_cols", "lbox_cut")}
    ctx.timing(False)
    b = boxes.view(n * len(ops), -1)
    mean_err = float(b.mean(dim=1).abs().max())
    rms_err = float((b.double().pow(2).mean(dim=1).sqrt() - 1).abs().max())
    ctx.close()
    return {"particles_per_s": round(n / (ms / 1e3), 1), "ms_per_step": round(ms, 3),
            "sub_particles_per_s": round(n * len(ops) / (ms / 1e3), 1),
            "kernels": {k: {"avg_ms": round(v[0], 4), "launches": v[1]} for k, v in kt.items() if v[1]},
            "box_mean_max_abs": mean_err, "box_rms_max_dev": rms_err}


def _port_lbox_job(i):
    """local_box stack mode of particle i (12 vertex boxes) by the C oracle (FP32)."""
    O, _ = _oracle32()
    ps, ops = _PORT_STATE["ps"], _PORT_STATE["ops"]
    t0 = time.perf_counter()
    O.local_box_stack(ps.images[i], ps.angles[i], ps.shifts[i], ops, (0.0, 0.0, LBOX_Z), LBOX_SUB, APIX)
    return time.perf_counter() - t0


def lbox_cpu_baseline(ps):
    """local_box stack mode by the C oracle port on the host cores, 8
    particles per process; the reference's own rate is quoted from profiles/."""
    ncores = min(8, os.cpu_count() or 1)
    _PORT_STATE.update(ps=ps, ops=vertex_ops())
    K = 8
    secs, wall = _port_pool(_port_lbox_job, [i % len(ps.images) for i in range(K * ncores)], ncores)
    _PORT_STATE.clear()
    per = float(np.mean(secs))
    return {"value": round(ncores / per, 3), "unit": "particles/s", "cores": ncores, "kind": "port",
            "sample": (f"C oracle (FP32) local_box stack in {ncores} processes, {K} particles each (mean "
                       f"{per * 1e3:.1f} ms per particle, {wall:.1f} s wall)"),
            "reference_here": _quoted_reference("lbox")}


def lbox_main(args):
    _, _, _, ps = make_workload(args.batch, args.templates)
    cpu = None if args.no_cpu_baseline else lbox_cpu_baseline(ps)      # before any GPU use
    r = lbox_run(ps, args.steps, args.warmup)
    n, nops = LBOX_BATCH, 12
    kt = r["kernels"]
    # algorithmic bytes per launch of each stage: rows read the real images and
    # write complex rows; columns read those and write the real shifted images;
    # the cut reads each shifted image once and writes the boxes
    b2 = n * BOX * BOX
    per = {"lbox_rows": b2 * (4 + 8), "lbox_cols": b2 * (8 + 4),
           "lbox_cut": n * nops * LBOX_SUB * LBOX_SUB * 4 + b2 * 4}
    dom = max((k for k in per if k in kt), key=lambda k: kt[k]["avg_ms"] * kt[k]["launches"])
    achieved = per[dom] / (kt[dom]["avg_ms"] / 1e3) / 1e9
    names = {"lbox_rows": "k_lbox_shift<0> (row transform, ramp, inverse)",
             "lbox_cols": "k_lbox_shift<1> (column transform, ramp, inverse, real part)",
             "lbox_cut": "k_lbox_cut (cut + normalise)"}
    line = {"metric": LBOX_METRIC, "value": r["particles_per_s"], "unit": "particles/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (the main workload's particles, tiled to the batch)",
            "config": {"workload": f"local_box stack mode on {n} particles of box {BOX}: 12 vertex boxes of "
                                   f"{LBOX_SUB}^2 at local_z {LBOX_Z} px", "box": BOX, "batch": n,
                       "sub_FFTsize": LBOX_SUB, "nops": nops},
            "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 1),
                         "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 3),
                         "traffic": None, "algorithmic_bytes_per_launch": per[dom],
                         "avg_launch_ms": kt[dom]["avg_ms"], "launches": kt[dom]["launches"]},
            "cpu_baseline": cpu,
            "extra": r}
    print(json.dumps(line))


LREF_BATCH, LREF_MAXRES_A, LREF_STEP, LREF_NCYCLE, LREF_SYM = 256, 6.0, 0.25, 12, "c5"
LREF_METRIC = ("sub-particles/s local_refine (refine_PR), 96^2 vertex sub-particles of box-384 particles: "
               "C5 candidates, then 12 cycles of 0.25 deg / shift steps")


def _lref_setup(ph, ps, coords):
    """The vertex sub-particles of LREF_BATCH particles (coords: the
    local_box_coords records of the particles at local_z LBOX_Z, (n, 12, 8)),
    started 1.5 deg off their local angles, and the 96^3 model of the phantom
    around the vertex (refine_PR.f90:24-54 takes its transform)."""
    from cryoim_amd import lib, synth
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pipelines
    n = LREF_BATCH
    idx = np.arange(n) % len(ps.images)
    ops = vertex_ops()
    x0 = np.array([0.0, 0.0, LBOX_Z], np.float32)
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3], P[:, 3:5] = ps.angles[idx], ps.shifts[idx]
    rec = np.asarray(coords(P, ops, x0), np.float32).reshape(-1, 8)
    m = len(rec)
    lx = np.repeat(idx, len(ops))
    d = pipelines.lref_derived(LBOX_SUB, APIX, maxRes=LREF_MAXRES_A, imgmask_A=(LBOX_SUB / 2 - 4) * APIX)
    p = lib.make_params(LBOX_SUB, 0, LBOX_SUB // 2 - 2, apply_ctf=True, mask_on=True, conj_out=False, apix=APIX,
                        imgmask=d["mask"], vol=300.0, df_scale=1.0, ncycle=LREF_NCYCLE, delta_angle=LREF_STEP,
                        delta_trans=d["refine_shift"])
    rng = np.random.default_rng(11)
    L = np.zeros((m, lib.PART_STRIDE), np.float32)
    L[:, 0:3] = rec[:, 0:3] + rng.uniform(-1.5, 1.5, (m, 3)).astype(np.float32)
    L[:, 3:5] = rec[:, 5:7]
    L[:, lib.P_DF1] = ps.df[lx, 0] * 1.0e4            # Angstrom, as local_refine reads them
    L[:, lib.P_DF2] = ps.df[lx, 1] * 1.0e4
    L[:, lib.P_BOXX:lib.P_BOXY + 1] = rec[:, 3:5]
    model = synth.phantom_map(synth.Phantom(ph.centres - x0.astype(np.float64), ph.sigmas, ph.amps, LBOX_SUB))
    return dict(idx=idx, lx=lx, L=L, p=p, d=d, model=model.astype(np.float32), m=m)


def lref_run(ps, s, steps: int, warmup: int):
    """cim_lref_refine on the m sub-particles (host records and images, as the
    Fortran program passes them): the refinement kernel's HIP events and the
    whole call's host clock (upload, preparation, refinement, records back)."""
    from cryoim_amd import lib
    ctx = lib.Context(0)
    ctx.lref_model(s["p"], s["model"])
    full = np.ascontiguousarray(ps.images[s["lx"]])
    for _ in range(warmup):
        ctx.lref_refine(s["p"], full, s["L"], sym=LREF_SYM, convF=0.8, maxres=s["d"]["maxres"])
    ctx.sync()
    ctx.timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        out = ctx.lref_refine(s["p"], full, s["L"], sym=LREF_SYM, convF=0.8, maxres=s["d"]["maxres"])
    wall = (time.perf_counter() - t0) / steps
    kt = ctx.timing_get("lref_refine")
    ctx.timing(False)
    ctx.close()
    dang = np.abs((out[:, 0:3] - s["L"][:, 0:3] + 180.0) % 360.0 - 180.0)
    return out, {"kernel_ms": kt[0], "launches": kt[1], "call_ms": wall * 1e3,
                 "median_angle_change_deg": float(np.median(dang)), "median_pr_deg": float(np.median(out[:, 5]))}


def _port_lref_job(k):
    """local_refine of sub-particle k by the C oracle (FP32): its preparation
    (norm, cut, mask, FFT) and the same C5 + 12-cycle refinement."""
    O, pl = _oracle32()
    st = _PORT_STATE
    s, L = st["s"], st["s"]["L"]
    t0 = time.perf_counter()
    bx, by, d1, d2 = st["cols"]
    a20, b20, w = pl.lref_particle(O, st["images"][s["lx"][k]], LBOX_SUB, int(L[k, bx]), int(L[k, by]),
                                   float(L[k, 3]), float(L[k, 4]), s["d"]["mask"], s["d"]["maxres"])
    ctf = O.ctf2d(LBOX_SUB, float(L[k, d1]), float(L[k, d2]), 0.0, APIX, vol=300.0, df_scale=1.0)
    O.lref_refine(st["fr"], st["fi"], a20, b20, ctf, w, st["symops"], LREF_NCYCLE, LREF_STEP,
                  s["d"]["refine_shift"], 0.8, s["d"]["maxres"], L[k, 0:5].astype(np.float64))
    return time.perf_counter() - t0


def lref_cpu_baseline(ph, ps):
    """The port's own local_box_coords, preparation and refinement of a
    sample of the sub-particles on the host cores."""
    from cryoim_amd import lib
    O, pl = _oracle32()
    s = _lref_setup(ph, ps, lambda P, ops, x0: np.stack(
        [O.local_box_coords(P[i, 0:3], P[i, 3:5], ops, x0, BOX, APIX) for i in range(len(P))]))
    fr, fi = pl.lref_model_ft(O, s["model"], s["d"]["mask"])
    ncores = min(8, os.cpu_count() or 1)
    _PORT_STATE.update(s=s, images=ps.images, fr=fr, fi=fi, symops=O.lref_symops(0, 5),
                       cols=(lib.P_BOXX, lib.P_BOXY, lib.P_DF1, lib.P_DF2))
    K = 64
    secs, wall = _port_pool(_port_lref_job, [i % s["m"] for i in range(K * ncores)], ncores)
    _PORT_STATE.clear()
    per = float(np.mean(secs))
    return {"value": round(ncores / per, 2), "unit": "sub-particles/s", "cores": ncores, "kind": "port",
            "sample": (f"C oracle (FP32) local_refine in {ncores} processes, {K} sub-particles each: preparation + "
                       f"C5 + {LREF_NCYCLE} cycles (mean {per * 1e3:.1f} ms per sub-particle, {wall:.1f} s wall)")}


def lref_main(args):
    ph, _, _, ps = make_workload(args.batch, args.templates)
    cpu = None if args.no_cpu_baseline else lref_cpu_baseline(ph, ps)     # before any GPU use
    from cryoim_amd import lib
    ctx = lib.Context(0)
    s = _lref_setup(ph, ps, lambda P, ops, x0: ctx.local_box_coords(P, ops, x0, BOX, APIX))
    ctx.close()
    out, r = lref_run(ps, s, args.steps, args.warmup)
    m, nq = s["m"], int(np.count_nonzero(
        np.hypot(*np.meshgrid(np.arange(-48, 48), np.arange(-48, 48))) <= s["d"]["maxres"]))
    ms = r["kernel_ms"]
    # per sub-particle: 5 symmetry candidates + 7 slices per cycle (the centre slice feeds the 5
    # shift candidates), 8 trilinear taps of 8 bytes per disk pixel; HBM: the model transform
    # once per launch, and per sub-particle its disk values g (8 B) and template factor (4 B)
    slices = 5 + 7 * LREF_NCYCLE
    tap_bytes = m * slices * nq * 8 * 8
    hbm_bytes = LBOX_SUB ** 3 * 8 + m * nq * 12
    achieved = hbm_bytes / (ms / 1e3) / 1e9
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic_lref_box96.json")) as f:
            traffic = json.load(f)["k_lref_refine"]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        traffic = None
    line = {"metric": LREF_METRIC, "value": round(m / (ms / 1e3), 1), "unit": "sub-particles/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (the main workload's phantom and particles; 96^3 model of the phantom around the vertex)",
            "config": {"workload": (f"local_refine of {m} sub-particles ({LREF_BATCH} box-{BOX} particles x 12 "
                                    f"vertices at local_z {LBOX_Z} px), box {LBOX_SUB}, maxres {s['d']['maxres']:.1f} px "
                                    f"({nq} disk pixels), {LREF_SYM} then {LREF_NCYCLE} cycles from 1.5 deg off"),
                       "box": LBOX_SUB, "sub_particles": m, "ncycle": LREF_NCYCLE, "delta_angle": LREF_STEP,
                       "sym": LREF_SYM, "disk_pixels": nq},
            "roofline": {"bound": "hbm", "kernel": "k_lref_refine (central slices of the model transform + phase "
                                                   "residual of 11 candidates per cycle, one workgroup per sub-particle)",
                         "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                         "traffic_source": "profiles/hbm_traffic_lref_box96.json (rocprofv3 FETCH_SIZE and "
                                           "WRITE_SIZE, copy-calibrated)",
                         "algorithmic_bytes_per_launch": hbm_bytes, "avg_launch_ms": round(ms, 4),
                         "launches": r["launches"], "timing": "HIP events on the library's stream",
                         "slice_taps_GBps": round(tap_bytes / (ms / 1e3) / 1e9, 1),
                         "note": ("HBM bytes are the model transform once and each sub-particle's disk values; the "
                                  "slices' 8-tap gathers (slice_taps_GBps) are served by the L2 and the MALL")},
            "cpu_baseline": cpu,
            "extra": {"call_sub_particles_per_s": round(m / (r["call_ms"] / 1e3), 1),
                      "call_ms": round(r["call_ms"], 2),
                      "call_note": "cim_lref_refine from host images and records: upload of the full images, "
                                   "preparation, refinement, records back",
                      "median_angle_change_deg": round(r["median_angle_change_deg"], 3),
                      "median_pr_deg": round(r["median_pr_deg"], 3)}}
    print(json.dumps(line))


def _grid_work(step: float) -> int:
    from cryoim_amd import geometry
    pts = geometry.au_grid_points(step)
    cells = sum(1 if ip == 0 else 2 for _, ip in pts)
    return cells * geometry.search_grid(step)[2]


REF_TIMINGS = os.path.join(ROOT, "profiles", "cpu_reference_box384.json")


def _quoted_reference(stage):
    """The reference's own rate for this stage, measured where its sources
    were built (oracle/time_reference.py), as recorded in profiles/."""
    try:
        with open(REF_TIMINGS) as f:
            r = json.load(f)[stage]
    except (OSError, KeyError, ValueError):
        return None
    return {k: r[k] for k in ("value", "unit", "cores", "sample", "date", "host") if k in r}


def _port_pool(fn, jobs, ncores):
    """Run the oracle port over jobs on ncores forked processes (before any
    GPU use); returns the per-job results and the wall time."""
    import multiprocessing as mp
    with mp.get_context("fork").Pool(ncores) as pool:
        t0 = time.perf_counter()
        out = pool.map(fn, jobs, chunksize=1)
        return out, time.perf_counter() - t0


def _oracle32():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import Oracle
    import pipelines
    return Oracle(32), pipelines


PORT_COARSE = 1.5      # search grid step of the port sample (deg); scaled to STEP by the grid-work ratio


def _port_main_job(job):
    """One particle through the C oracle (FP32): ort_globesearch on the
    PORT_COARSE grid, then the 60-fold insert at INSERT_MAXR (the
    reconstruct_substep form)."""
    timgs, tang, img, shift, df = job
    O, pipelines = _oracle32()
    d = derived()
    t0 = time.perf_counter()
    ta, tb = pipelines.templates(O, timgs, np.zeros((len(tang), 2)), d["mask"])
    ctf = O.ctf2d(BOX, df[0], df[1], 0.0, APIX)
    a, b, tf = pipelines.template_inputs(O, ta, tb, ctf, d["minR"], d["maxR"])
    _, (pa, pb), fc0 = pipelines.particle_inputs(O, img, *shift, d["mask"], d["minR"], d["maxR"])
    t_prep = time.perf_counter() - t0
    t0 = time.perf_counter()
    o4 = O.globesearch(pa, pb, fc0, a, b, tf, tang, d["minR"], d["maxR"], PORT_COARSE)
    t_search = time.perf_counter() - t0
    vol = O.new_volume(BOX, INSERT_MAXR, compact=True, wg_init=0.0)
    t0 = time.perf_counter()
    fa, fb = O.particle_ft(img, *shift, mask_r=d["mask"], norm=True)
    O.insert(vol, fa, fb, ctf, o4[:3], INSERT_MAXR, rcut=INSERT_MAXR - 2, compact=True)
    return t_prep, t_search, time.perf_counter() - t0


def cpu_baseline(args, tang, timgs, ps):
    """The C oracle port (FP32, one process per core) on a bounded sample:
    one particle per core: its template and particle transforms, the
    search's scoring loop on the PORT_COARSE grid (that time alone scaled to
    the STEP grid by the exact ratio of scored AU-cell x omega points) and
    the insert at INSERT_MAXR.  The reference's own rate is quoted from
    profiles/ beside it."""
    ncores = min(8, os.cpu_count() or 1)
    jobs = [(timgs, tang, ps.images[i], tuple(ps.shifts[i]), tuple(ps.df[i])) for i in range(ncores)]
    res, wall = _port_pool(_port_main_job, jobs, ncores)
    scale = _grid_work(STEP) / _grid_work(PORT_COARSE)
    per = float(np.mean([tp + ts * scale + ti for tp, ts, ti in res]))      # seconds per particle on one core
    m = np.mean(res, axis=0)
    return {"value": ncores / per, "unit": "particles/s", "cores": ncores, "kind": "port",
            "sample": (f"C oracle (FP32) in {ncores} processes, one particle each: transforms (mean {m[0]:.2f} s) "
                       f"+ the search's scoring on the {PORT_COARSE} deg grid (mean {m[1]:.2f} s, x{scale:.0f} to "
                       f"the {STEP} deg grid by AU-cell x omega count) + 60-fold insert at maxR {INSERT_MAXR} "
                       f"(mean {m[2]:.2f} s); {wall:.1f} s wall"),
            "reference_here": _quoted_reference("main")}


def _roofline(r, args):
    """Roofline of the kernel with the most time per chosen step, from its
    HIP-event durations over the sequential timed steps (each kernel alone on
    the chip; in the pipelined steps the insert shares it with the next
    batch's search) and the library's own work counters: FP32 operations for
    the exact re-score, BF16 matrix operations for the residual tables, table
    bytes for the gather, sample bytes for the insert."""
    kt, cnt, n = r.get("kt_seq") or r["kt_c"], r["cnt_c"], r["n"]
    tot = {k: v[0] * v[1] for k, v in kt.items() if v[1]}
    dom = max(tot, key=tot.get)
    ms, launches = kt[dom]
    lc = r["kt_c"][dom][1]          # launches the work counters cover
    sec = ms / 1e3
    base = {"avg_launch_ms": round(ms, 3), "launches": launches,
            "timing": "HIP events on the library's stream, sequential timed steps", "traffic": None}
    if dom == "cl_rescore":
        calls = cnt["cl_rescore_samples"] / 2 / lc          # paired calls (two columns each)
        passed = cnt["cl_rescore_pass"] / 2 / lc
        flops = passed * FLOPS_PAIR_CALL_PASS + (calls - passed) * FLOPS_PAIR_CALL_GATED
        achieved = flops / sec / 1e12
        return dom, {"bound": "valu", "kernel": "k_cl_rescore<4,true> (exact phase residual of the top-32 AU "
                                                "points x 16 omega columns and their Friedel pairs, FP32)",
                     "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), **base,
                     "algorithmic_flops_per_launch": int(flops),
                     "flops_rule": f"{FLOPS_PAIR_CALL_PASS} FP32 operations per paired sample call that passes the "
                                   f"particle amplitude gate, {FLOPS_PAIR_CALL_GATED} per call that fails it "
                                   f"(library counters: {int(calls)} calls, {passed / calls:.3f} passing, per launch)",
                     "note": "FP32 vector arithmetic (no matrix instructions): the peak is MI355X's published FP32 "
                             "figure (vector and matrix alike, counting packed FP32); this build compiles without "
                             "packed FP32, so its issue limit is half the peak"}
    if dom == "cl_table":
        flops = cnt["cl_table_pairs"] / lc * TABLE_FEATURES_PER_RADIUS
        achieved = flops / sec / 1e12
        return dom, {"bound": "mfma", "kernel": "k_cl_table_mfma<4> (residual tables, v_mfma_f32_32x32x16_bf16)",
                     "achieved": round(achieved, 1), "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_BF16_TFLOPS, 4), **base, "algorithmic_flops_per_launch": int(flops)}
    if dom == "cl_gather":
        npsi = int(round(360.0 / STEP))
        tb = n * args.templates * (npsi // 2) * npsi * 4 + n * args.templates * (npsi // 4) * (npsi // 2) * 2
        achieved = tb / sec / 1e9
        return dom, {"bound": "hbm", "kernel": "k_cl_gather (coarse E4M3 and fine bfloat16 table sums)",
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), **base, "algorithmic_bytes_per_launch": int(tb),
                     "bytes_rule": "every coarse and fine table entry of the batch read once",
                     "note": "the sums read each table row once per (AU point, line) from L2, not HBM (limits: "
                             "hbm_frac); the scattered 2- and 4-byte reads keep the texture-address units and L2 "
                             "channels busy (ta_busy, l2_req_per_channel_cycle) and the VALU at its issue rate "
                             "(valu_issue_per_simd_cycle against valu_full_rate): those bound it together"}
    # the least HBM traffic of a launch: each particle's staged sample values read once
    # ({a ctf, b ctf, ctf^2, in}, 16 B on the (2R+1)(R+1) half-plane grid) and the accumulator's
    # hi and lo cells within radius R+1 of the origin read and written once (16 B each)
    R = INSERT_MAXR
    g = np.arange(-R - 1, R + 2)
    sphere = int(np.count_nonzero(g[:, None, None] ** 2 + g[None, :, None] ** 2 + g[None, None, :] ** 2
                                  <= (R + 1) ** 2))
    per = n * (2 * R + 1) * (R + 1) * 16 + 2 * 2 * 16 * sphere
    achieved = per / sec / 1e9
    return dom, {"bound": "hbm", "kernel": "k_insert_brickq (60-fold coset insert, compensated FP32)",
                 "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                 "frac": round(achieved / PEAK_HBM_GBS, 4), **base, "algorithmic_bytes_per_launch": int(per),
                 "bytes_rule": (f"sample values read once ({n} particles x (2R+1)(R+1) x 16 B) + the accumulator's "
                                f"hi and lo cells within radius R+1 ({sphere} per map) read and written once, "
                                f"16 B each; R = {R}"),
                 "note": ("the kernel sums each batch in LDS (64-bit fixed-point atomics, 3 per corner of 5 coset "
                          "planes per particle) and touches each accumulator cell once per launch, so HBM does not "
                          "bound it: LDS atomics (and their bank conflicts) and the point enumeration's VALU work "
                          "do (profiles/sq_counters_box384.md); in the pipelined steps it also shares the chip "
                          "with the next batch's search")}


CLOCK_GHZ = 2.4            # MI355X peak engine clock
SIMDS, CUS = 1024, 256


def _limits(keys, roof, tr, n):
    """Utilisation of each unit that can bound the roofline kernel: its
    per-launch SQ counts (profiles/sq_counters_box384.json, rocprofv3 --pmc,
    same batch) over this run's live launch time.  A wave64 VALU instruction
    holds its SIMD for 4 cycles, so valu_busy = 4 ACTIVE_INST_VALU / SIMD
    cycles; MFMA busy cycles are counted per SIMD, LDS instructions and bank
    conflict cycles per CU."""
    path = os.path.join(ROOT, "profiles", "sq_counters_box384.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        sq = json.load(f)
    if not all(k in sq and sq[k].get("batch") == n for k in keys):
        return None
    def tot(c):
        return sum(sq[k]["per_launch"].get(c, 0.0) for k in keys)
    cyc = CLOCK_GHZ * 1e9 * roof["avg_launch_ms"] / 1e3
    out = {"valu_issue_per_simd_cycle": round(tot("SQ_ACTIVE_INST_VALU") / (SIMDS * cyc), 4),
           "valu_full_rate": 0.25,
           "mfma_busy": round(tot("SQ_VALU_MFMA_BUSY_CYCLES") / (SIMDS * cyc), 3),
           "lds_instr_per_cu_cycle": round(tot("SQ_INSTS_LDS") / (CUS * cyc), 4),
           "lds_bank_conflict_frac": round(tot("SQ_LDS_BANK_CONFLICT") / (CUS * cyc), 3)}
    if roof.get("traffic"):
        out["hbm_frac"] = round(roof["traffic"] / (roof["avg_launch_ms"] / 1e3) / (PEAK_HBM_GBS * 1e9), 4)
    # texture-address and L2 rates within the counter run itself (its own clock: GRBM_GUI_ACTIVE
    # sums the 8 XCDs; TA_BUSY_avr averages the CUs' address units; 16 L2 channels per XCD)
    gui = tot("GRBM_GUI_ACTIVE") / len(keys) / 8
    if gui > 0:
        out["ta_busy"] = round(tot("TA_BUSY_avr") / len(keys) / gui, 3)
        out["l2_req_per_channel_cycle"] = round(tot("TCC_REQ_sum") / (128 * gui * len(keys)), 3)
        req = tot("TCC_REQ_sum")
        out["l2_hit"] = round(tot("TCC_HIT_sum") / req, 3) if req else None
    out["rule"] = (f"counts per launch from profiles/sq_counters_box384.json over {CLOCK_GHZ} GHz x the live launch "
                   f"time; {SIMDS} SIMDs, {CUS} CUs; a SIMD issues one full-rate wave64 FP32 instruction per 4 "
                   f"cycles (valu_full_rate)")
    return out


def main():
    args = _parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.stage == "center":
        center_main(args)
        return
    if args.stage == "lbox":
        lbox_main(args)
        return
    if args.stage == "lref":
        lref_main(args)
        return
    if args.stage == "missrate":
        missrate_main(args)
        return
    if args.stage == "c1":
        c1_main(args)
        return
    t_setup = time.perf_counter()
    ph, tang, timgs, ps = make_workload(args.batch, args.templates)
    cpu = None
    if not (args.no_cpu_baseline or args.chosen_only):       # host cores only, before any GPU use
        try:
            cpu = cpu_baseline(args, tang, timgs, ps)
        except Exception as e:
            cpu = {"value": None, "unit": "particles/s", "cores": 0, "kind": "port", "sample": f"failed: {e}"}
    r = gpu_run(args, tang, timgs, ps)
    n = r["n"]
    value = n * args.steps / (r["ev_c"] / 1e3)
    dom, roof = _roofline(r, args)
    if args.chosen_only:
        print(json.dumps({"metric": METRIC + " (chosen path only, step after step)", "value": round(value, 3),
                          "unit": "particles/s", "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": round(r["ev_c"] / args.steps, 3), "roofline": roof,
                          "chosen_kernels": {k: {"avg_ms": round(v[0], 3), "launches": v[1]}
                                             for k, v in r["kt_c"].items() if v[1]},
                          "counters": {k: int(v) for k, v in r["cnt_c"].items()}}))
        return
    cen = center_run(ph, ps, args.steps, args.warmup)
    ref5 = refine_run(tang, timgs, ps, min(args.steps, 3), 1)
    plain = n / (r["ev_p"] / 1e3)
    tj = os.path.join(ROOT, "profiles", "hbm_traffic_box384.json")
    if os.path.exists(tj):   # rocprofv3 FETCH_SIZE + WRITE_SIZE passes, copy-calibrated, per launch
        with open(tj) as f:
            tr = json.load(f)
        keys = {"insert": ["k_insert_brickq"], "cl_gather": ["k_cl_gather_coarse", "k_cl_gather_fine"],
                "cl_table": ["k_cl_table_mfma"]}.get(dom, ["k_cl_rescore"])
        if all(k in tr and tr[k].get("batch") == n for k in keys):
            roof["traffic"] = int(sum(tr[k]["hbm_bytes_per_launch"] for k in keys))
            roof["traffic_source"] = "profiles/hbm_traffic_box384.json (rocprofv3 FETCH_SIZE and WRITE_SIZE, copy-calibrated)"
        roof["limits"] = _limits(keys, roof, tr, n)
    d = r["d"]
    fsc = r["fsc"]
    kt = {k: {"avg_ms": round(v[0], 3), "launches": v[1]} for k, v in r["kt_c"].items() if v[1]}
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ev_c"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 60-fold Gaussian-blob phantom, analytic projections, CTF, SNR 0.05)",
        "config": {"workload": f"chosen path vs plain FP32 path, box {BOX}: global common-lines search (0.5 deg "
                               f"AU grid x 720 omega, maxR {d['maxR']}, T={args.templates} templates; chosen: "
                               f"tabulated ranking, top-{args.topk} exact re-score) + 60-fold insert (maxR "
                               f"{INSERT_MAXR}; chosen: compensated FP32)" +
                               ("" if args.no_pipeline else "; chosen steps pipelined: the search of batch k on one "
                                                            "stream overlaps the insert of batch k-1 on another"),
                   "box": BOX, "batch": n, "templates": args.templates, "search_maxR": d["maxR"],
                   "insert_maxR": INSERT_MAXR, "searchstep": STEP, "apix": APIX, "topk": args.topk,
                   "pipelined": not args.no_pipeline},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": {"plain_particles_per_s": round(plain, 3),
                  "sequential_particles_per_s": round(n * args.steps / (r["ev_seq"] / 1e3), 3),
                  "graph_particles_per_s": round(n * args.steps / (r["ev_graph"] / 1e3), 3),
                  "graph_orientation_identical": r["graph_same"],
                  "chosen_over_plain": round(value / plain, 2),
                  "plain_ms_per_step": round(r["ev_p"], 3),
                  "orientation_identical": r["identical"], "orientation_within_one_step": r["within"],
                  "orientation_misses": n - r["within"], "particles_compared": n,
                  "fsc_min_to_r190": round(float(np.min(fsc[1:])), 6),
                  "fsc_at_r190": round(float(fsc[-1]), 6),
                  "map_rel_l2_chosen_vs_plain": r["map_rel"],
                  "fallback_particles": int(r["cnt_c"]["cl_fallback_particles"]),
                  "exact_minimum_rank": {k[8:]: int(r["cnt_c"][k]) for k in RANK_KEYS},
                  "chosen_kernels": kt,
                  "plain_search_avg_ms": round(r["kt_p"]["cl_search"][0], 3),
                  "search_angle_err_deg_median": float(np.median(r["ang_err"])),
                  "center_mode3": {k: {kk: v[kk] for kk in ("particles_per_s", "ms_per_step")}
                                   for k, v in cen.items()},
                  "refine_mode5": ref5,
                  "wall_s": round(r["wall_c"], 3),
                  "setup_s": round(time.perf_counter() - t_setup, 1)},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
