"""Record search decisions at the measured configurations' scale, written in
this container by the FP64 oracle and by the reference's own icosprocess
(built by oracle/build_ref.sh into oracle/_ref/), into tests/golden/scale.npz.

* config 1 (box 128, 1 deg grid x 360 omega, maxR 24, T = 4 templates):
  N1 particles of tests/pipelines.config1_workload through the FP64 oracle's
  ort_globesearch (one process per core) and through the reference's
  icosprocess mode 4 (8 processes over first/last ranges, the reference's
  own parallel mode), with the reference's wall-clock times for mode 4 and
  mode 1 (insert at maxR 62 and the 3D inverse FFT) on the same particles;
* box 384 (0.5 deg grid x 720 omega, maxR 72, T = 4): N384 particles of
  tests/pipelines.box384_workload through the FP64 oracle.

The particle images are regenerated from their seeds by the tests; the
fixture keeps their float64 sum and norm so a test can tell if the
regenerated set differs.

Usage: python tests/golden/make_scale.py [--n1 400] [--n384 4] [--jobs 8] [--only c1|b384]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import pipelines  # noqa: E402
from cryoim_amd import formats  # noqa: E402

_W = {}


def _init(kind, n):
    _W["w"] = pipelines.config1_workload(n) if kind == "c1" else pipelines.box384_workload(n)
    from oracle import Oracle
    _W["O"] = Oracle(64)


CACHE = os.environ.get("CIM_SCALE_CACHE", os.path.join(tempfile.gettempdir(), "cim_scale_cache"))


def _one(i):
    t0 = time.perf_counter()
    r = pipelines.oracle_search_one(_W["O"], _W["w"], i)
    return i, r, time.perf_counter() - t0


def oracle_decisions(kind, n, jobs):
    """The oracle's decisions for particles 0..n-1; each one is also kept in
    CACHE (per workload and particle count), so an interrupted run resumes."""
    out = np.zeros((n, 4))
    secs = np.zeros(n)
    cdir = os.path.join(CACHE, f"{kind}_{n}")
    os.makedirs(cdir, exist_ok=True)
    todo = []
    for i in range(n):
        f = os.path.join(cdir, f"{i}.npy")
        if os.path.exists(f):
            v = np.load(f)
            out[i], secs[i] = v[:4], v[4]
        else:
            todo.append(i)
    print(f"{kind}: {n - len(todo)} of {n} particles cached in {cdir}", flush=True)
    with cf.ProcessPoolExecutor(max_workers=jobs, initializer=_init, initargs=(kind, n)) as ex:
        for i, r, t in ex.map(_one, todo):
            out[i] = r
            secs[i] = t
            np.save(os.path.join(cdir, f"{i}.npy"), np.r_[r, t])
            if i % 50 == 0:
                print(f"{kind} oracle particle {i}: {r} ({t:.1f} s)", flush=True)
    return out, secs


def _env1():
    env = dict(os.environ)
    env.update(MKL_NUM_THREADS="1", OMP_NUM_THREADS="1", MKL_DYNAMIC="FALSE")
    return env


def reference_config1(w, n, jobs):
    """icosprocess mode 4 and mode 1 (the reference) on the n particles, as
    jobs concurrent processes over contiguous first/last ranges."""
    exe = os.path.join(REF, "icosprocess")
    ps, d, apix, box = w["ps"], w["d"], w["apix"], w["box"]
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        formats.write_mrc(os.path.join(tmp, "stk.mrc"), ps.images[:n])
        formats.write_mrc(os.path.join(tmp, "tmpl.mrc"), w["timgs"])
        formats.write_template_ort(os.path.join(tmp, "tort.txt"), w["tang"])
        parts = [formats.Particle("stk.mrc", i + 1, 0.0, 0.0, 0.0, *ps.shifts[i], 0.0, i + 1, i + 1,
                                  ps.df[i, 0], ps.df[i, 1], 0.0) for i in range(n)]
        formats.write_ortfile(os.path.join(tmp, "ort.txt"), parts)
        bounds = np.linspace(0, n, jobs + 1).astype(int)
        common = [f"apix={apix}", "maxRes=8.0", "minRes=300.0", f"imgmask={d['mask'] * apix}"]

        def run_all(argsets):
            t0 = time.perf_counter()
            procs = [subprocess.Popen([exe, *a], cwd=tmp, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                                      env=_env1()) for a in argsets]
            rcs = [p.wait() for p in procs]
            if any(rcs):
                raise RuntimeError(f"icosprocess exit codes {rcs}")
            return time.perf_counter() - t0

        res["mode4_s"] = run_all([["mode=4", "templateort=tort.txt", "templatestck=tmpl.mrc", "searchstep=1.0",
                                   f"newortfile=m4_{k}.txt", "ortfile=ort.txt", f"first={bounds[k] + 1}",
                                   f"last={bounds[k + 1]}", *common] for k in range(jobs)])
        rows = []
        for k in range(jobs):
            rows += [[p.theta, p.phi, p.omega, p.x, p.y, p.pr]
                     for p in formats.read_ortfile900(os.path.join(tmp, f"m4_{k}.txt"))]
        res["mode4_out"] = np.array(rows, np.float64)
        # mode 1 at the true orientations, insert radius Nyquist-2 (maxRes = apix box / 62), as
        # subrec partial sums per process plus the one merge a full run would do
        parts = [formats.Particle("stk.mrc", i + 1, *ps.angles[i], *ps.shifts[i], 0.0, i + 1, i + 1,
                                  ps.df[i, 0], ps.df[i, 1], 0.0) for i in range(n)]
        formats.write_ortfile(os.path.join(tmp, "ort1.txt"), parts)
        maxres = apix * box / w["insert_maxR"]
        res["mode1_s"] = run_all([["mode=1", "subrec=y", f"result3d=s{k}.tnf", "ortfile=ort1.txt",
                                   f"first={bounds[k] + 1}", f"last={bounds[k + 1]}", f"apix={apix}",
                                   f"maxRes={maxres}", f"imgmask={d['mask'] * apix}"] for k in range(jobs)])
        merge = os.path.join(REF, "merge_icosprocess_subrec")
        t0 = time.perf_counter()
        subprocess.run([merge, *[f"s{k}.tnf" for k in range(jobs)], "map.mrc"], cwd=tmp, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, env=_env1())
        res["merge_s"] = time.perf_counter() - t0
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n1", type=int, default=400)
    ap.add_argument("--n384", type=int, default=4)
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--only", choices=("c1", "b384"), default=None)
    args = ap.parse_args()
    path = os.path.join(HERE, "scale.npz")
    g = dict(np.load(path)) if os.path.exists(path) else {}
    if args.only in (None, "c1"):
        w = pipelines.config1_workload(args.n1)
        im = w["ps"].images.astype(np.float64)
        g.update(c1_n=np.int32(args.n1), c1_img_sum=im.sum(), c1_img_norm=np.linalg.norm(im))
        t0 = time.perf_counter()
        o64, secs = oracle_decisions("c1", args.n1, args.jobs)
        g.update(c1_o64=o64, c1_o64_s=secs, c1_o64_wall=time.perf_counter() - t0, c1_o64_jobs=np.int32(args.jobs))
        if os.path.exists(os.path.join(REF, "icosprocess")):
            r = reference_config1(w, args.n1, args.jobs)
            g.update(c1_ref=r["mode4_out"], c1_ref_mode4_s=r["mode4_s"], c1_ref_mode1_s=r["mode1_s"],
                     c1_ref_merge_s=r["merge_s"], c1_ref_jobs=np.int32(args.jobs))
            print("reference:", {k: v for k, v in r.items() if k.endswith("_s")}, flush=True)
        print("c1 oracle vs reference, steps:", np.bincount(np.ceil(pipelines.angle_steps(
            g["c1_o64"], g.get("c1_ref", g["c1_o64"]), 1.0) - 1e-6).astype(int)))
    if args.only in (None, "b384"):
        w = pipelines.box384_workload(args.n384)
        im = w["ps"].images.astype(np.float64)
        g.update(b384_n=np.int32(args.n384), b384_img_sum=im.sum(), b384_img_norm=np.linalg.norm(im))
        o64, secs = oracle_decisions("b384", args.n384, min(args.jobs, args.n384))
        g.update(b384_o64=o64, b384_o64_s=secs)
        print("b384 oracle:", o64, secs)
    np.savez_compressed(path, **g)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
