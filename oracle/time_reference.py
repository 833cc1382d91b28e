"""Time the reference's own programs (oracle/_ref/, built from the reference
sources by oracle/build_ref.sh) on this container's host cores, on bounded
samples of the box-384 workloads of bench.py, and write the rates to
profiles/cpu_reference_box384.json.  bench.py quotes that file beside the C
oracle port it times itself; the reference never runs where bench.py runs
on a GPU.

Each program runs as N concurrent single-threaded processes (MKL and OpenMP
on one thread each) over disjoint particle ranges: the reference's only
parallel mode.

Usage: python oracle/time_reference.py [--stages main center lbox c1]
"""
from __future__ import annotations

import argparse
import datetime
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
REF = os.path.join(HERE, "_ref")

import bench  # noqa: E402
from bench import APIX, BOX, INSERT_MAXR, LBOX_SUB, LBOX_Z, MAXRES, MINRES, STEP  # noqa: E402

OUT = os.path.join(ROOT, "profiles", "cpu_reference_box384.json")


def center_cpu_baseline(ph, ps):
    """Reference icosprocess mode 3 with model3d on the host cores: the cost
    per particle is the time of K particles minus that of 1, per process
    (the difference removes each process's model read and 3D FFT)."""
    from cryoim_amd import formats, synth
    exe = os.path.join(REF, "icosprocess")
    if not os.path.exists(exe):
        return None
    d = bench.derived()
    ncores = min(8, os.cpu_count() or 1)
    m = len(ps.images)
    K = 41
    with tempfile.TemporaryDirectory() as tmp:
        formats.write_mrc(os.path.join(tmp, "stk.mrc"), ps.images)
        formats.write_mrc(os.path.join(tmp, "model.mrc"), synth.phantom_map(ph))
        parts = [formats.Particle("stk.mrc", i + 1, *ps.angles[i % m], *ps.shifts[i % m], 0.0, i + 1, (i % m) + 1,
                                  ps.df[i % m, 0], ps.df[i % m, 1], 0.0) for i in range(K * ncores)]
        formats.write_ortfile(os.path.join(tmp, "ort.txt"), parts)
        t = {}
        for k in (1, K):
            t0 = time.perf_counter()
            procs = [subprocess.Popen([exe, "mode=3", "model3d=model.mrc", f"newortfile=c{i}.txt",
                                       f"first={i * K + 1}", f"last={i * K + k}", f"apix={APIX}",
                                       f"maxRes={MAXRES}", f"imgmask={d['mask'] * APIX}", "ortfile=ort.txt"],
                                      cwd=tmp, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, env=_ref_env())
                     for i in range(ncores)]
            rcs = [pr.wait() for pr in procs]
            t[k] = time.perf_counter() - t0
            if any(rcs):
                return None
    per = (t[K] - t[1]) / (K - 1)
    return {"value": round(ncores / per, 3), "unit": "particles/s", "cores": ncores, "kind": "reference",
            "sample": (f"icosprocess mode 3 (flang -O2, MKL FFTW) as {ncores} concurrent processes, {K} particles "
                       f"each ({t[K]:.1f} s) minus 1 each ({t[1]:.1f} s: model read and 3D FFT)")}




def lbox_cpu_baseline(ps):
    """Reference local_box stack mode on the host cores: the cost per
    particle is the time of K particles minus that of 1, per process (the
    difference removes each process's stack read and start-up)."""
    from cryoim_amd import formats
    exe = os.path.join(REF, "local_box")
    if not os.path.exists(exe):
        return None
    ncores = min(8, os.cpu_count() or 1)
    m = len(ps.images)
    K = 121
    ops = bench.vertex_ops().astype(np.float32)
    with tempfile.TemporaryDirectory() as tmp:
        formats.write_mrc(os.path.join(tmp, "stk.mrc"), ps.images)
        parts = [formats.Particle("stk.mrc", i + 1, *ps.angles[i % m], *ps.shifts[i % m], 0.0, i + 1, (i % m) + 1,
                                  ps.df[i % m, 0], ps.df[i % m, 1], 0.0) for i in range(K * ncores)]
        formats.write_ortfile(os.path.join(tmp, "ort.txt"), parts)
        with open(os.path.join(tmp, "v12.txt"), "w") as f:
            for S in ops:
                for col in S.T:
                    f.write(" ".join(f"{float(v):.9g}" for v in col) + "\n")
        t = {}
        for k in (1, K):
            t0 = time.perf_counter()
            procs = [subprocess.Popen([exe, "icos_ort=ort.txt", f"local_ort=l{i}.txt", f"local_imgstck=s{i}.mrc",
                                       f"sub_FFTsize={LBOX_SUB}", "sym_matrix=v12.txt", "local_x=0", "local_y=0",
                                       f"local_z={LBOX_Z}", f"apix={APIX}", f"first={i * K + 1}",
                                       f"last={i * K + k}"], cwd=tmp, stdout=subprocess.DEVNULL,
                                      stderr=subprocess.DEVNULL, env=_ref_env())
                     for i in range(ncores)]
            rcs = [pr.wait() for pr in procs]
            t[k] = time.perf_counter() - t0
            if any(rcs):
                return None
    per = (t[K] - t[1]) / (K - 1)
    return {"value": round(ncores / per, 3), "unit": "particles/s", "cores": ncores, "kind": "reference",
            "sample": (f"local_box stack mode (flang -O2, MKL FFTW) as {ncores} concurrent processes, {K} "
                       f"particles each ({t[K]:.1f} s) minus 1 each ({t[1]:.1f} s: stack read, start-up)")}




def _ref_env():
    """Environment of the reference processes: MKL and OpenMP on one thread
    each, so that N concurrent processes use exactly N host cores."""
    env = dict(os.environ)
    env.update(MKL_NUM_THREADS="1", OMP_NUM_THREADS="1", MKL_DYNAMIC="FALSE")
    return env




def _ref_baseline(exe, args, tang, timgs, ps, ncores):
    """icosprocess mode 4 at two coarse search steps, one particle per
    process: the time is fitted as a + b x (AU cells x omega) and b carried
    to the bench's step (a: start-up, template and particle transforms); then
    mode 1 subrec at the insert radius."""
    from cryoim_amd import formats
    d = bench.derived()
    coarse_steps = (4.0, 2.5)
    m = len(ps.images)
    with tempfile.TemporaryDirectory() as tmp:
        formats.write_mrc(os.path.join(tmp, "stk.mrc"), ps.images)
        formats.write_mrc(os.path.join(tmp, "tmpl.mrc"), timgs)
        formats.write_template_ort(os.path.join(tmp, "tort.txt"), tang)
        parts = [formats.Particle("stk.mrc", i + 1, 0.0, 0.0, 0.0, *ps.shifts[i % m], 0.0, i + 1, (i % m) + 1,
                                  ps.df[i % m, 0], ps.df[i % m, 1], 0.0) for i in range(ncores)]
        formats.write_ortfile(os.path.join(tmp, "ort.txt"), parts)
        common = [f"apix={APIX}", f"maxRes={MAXRES}", f"minRes={MINRES}", f"imgmask={d['mask'] * APIX}",
                  "ortfile=ort.txt"]
        t_search = {}
        for coarse in coarse_steps:
            t0 = time.perf_counter()
            procs = [subprocess.Popen([exe, "mode=4", "templateort=tort.txt", "templatestck=tmpl.mrc",
                                       f"searchstep={coarse}", f"newortfile=o{i}.txt", f"first={i + 1}",
                                       f"last={i + 1}", *common], cwd=tmp, stdout=subprocess.DEVNULL,
                                      stderr=subprocess.DEVNULL, env=_ref_env()) for i in range(ncores)]
            rcs = [p.wait() for p in procs]
            t_search[coarse] = time.perf_counter() - t0
            if any(rcs):
                raise RuntimeError(f"icosprocess mode 4 exit codes {rcs}")
        # insert: reconstruct_substep at maxR 190 (maxRes = apix*box/190), 2 particles per process
        maxres_ins = APIX * BOX / INSERT_MAXR
        nins = 2
        parts = [formats.Particle("stk.mrc", i + 1, *ps.angles[i % m], *ps.shifts[i % m], 0.0,
                                  i + 1, (i % m) + 1, ps.df[i % m, 0], ps.df[i % m, 1], 0.0)
                 for i in range(ncores * nins)]
        formats.write_ortfile(os.path.join(tmp, "ort2.txt"), parts)
        t0 = time.perf_counter()
        procs = [subprocess.Popen([exe, "mode=1", "subrec=y", f"result3d=s{i}.tnf", f"first={i * nins + 1}",
                                   f"last={(i + 1) * nins}", f"apix={APIX}", f"maxRes={maxres_ins}",
                                   f"imgmask={d['mask'] * APIX}", "ortfile=ort2.txt"], cwd=tmp,
                                  stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, env=_ref_env())
                 for i in range(ncores)]
        rcs = [p.wait() for p in procs]
        t_ins = time.perf_counter() - t0
        if any(rcs):
            raise RuntimeError(f"icosprocess mode 1 exit codes {rcs}")
    (s1, s2) = coarse_steps
    w1, w2 = bench._grid_work(s1), bench._grid_work(s2)
    b = (t_search[s2] - t_search[s1]) / (w2 - w1)              # seconds per AU cell x omega
    a = t_search[s1] - b * w1
    per_particle_core = a + b * bench._grid_work(STEP) + t_ins / nins      # seconds per particle on one core
    rate = ncores / per_particle_core
    return {"value": rate, "unit": "particles/s", "cores": ncores, "kind": "reference",
            "sample": (f"icosprocess (flang -O2, MKL FFTW) as {ncores} concurrent processes: mode 4 on 1 particle "
                       f"each at searchstep {s1} ({t_search[s1]:.1f} s) and {s2} ({t_search[s2]:.1f} s), fitted as "
                       f"{a:.2f} s + {b * 1e6:.2f} us per AU cell x omega and carried to the {STEP} deg grid "
                       f"({bench._grid_work(STEP)} cells x omega); + mode 1 subrec at maxR {INSERT_MAXR} on {nins} "
                       f"particles each ({t_ins:.1f} s)")}




def _run_ref(exe, argv_of, ncores, tmp):
    """ncores concurrent reference processes (argv_of(i) each); wall seconds."""
    t0 = time.perf_counter()
    procs = [subprocess.Popen([exe, *argv_of(i)], cwd=tmp, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                              env=_ref_env()) for i in range(ncores)]
    rcs = [pr.wait() for pr in procs]
    if any(rcs):
        raise RuntimeError(f"{os.path.basename(exe)} exit codes {rcs}")
    return time.perf_counter() - t0


def c1_baseline(ncores):
    """The box-128 workload of bench.py --stage c1 (tests/pipelines.py
    config1_workload) through the reference icosprocess on the host cores:
    mode 3 (calc_center with model3d), mode 4 on the 1 deg grid and mode 1
    (subrec) at maxR 62.  Per-particle costs are differences of runs with
    K and 1 particles per process, which removes start-up, template and
    model transforms."""
    from cryoim_amd import formats, synth
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pipelines
    exe = os.path.join(REF, "icosprocess")
    w = pipelines.config1_workload(64)
    box, apix, d, ps, tang = w["box"], w["apix"], w["d"], w["ps"], w["tang"]
    m = len(ps.images)
    K4, K3, K1 = 2, 41, 41
    common = [f"apix={apix}", "maxRes=8.0", "minRes=300.0", f"imgmask={d['mask'] * apix}"]
    t = {}
    with tempfile.TemporaryDirectory() as tmp:
        formats.write_mrc(os.path.join(tmp, "stk.mrc"), ps.images)
        formats.write_mrc(os.path.join(tmp, "tmpl.mrc"), w["timgs"])
        formats.write_mrc(os.path.join(tmp, "model.mrc"), synth.phantom_map(w["phantom"]))
        formats.write_template_ort(os.path.join(tmp, "tort.txt"), tang)
        K = max(K4, K3, K1)
        parts = [formats.Particle("stk.mrc", (i % m) + 1, *ps.angles[i % m], *ps.shifts[i % m], 0.0, i + 1,
                                  (i % m) + 1, ps.df[i % m, 0], ps.df[i % m, 1], 0.0) for i in range(K * ncores)]
        formats.write_ortfile(os.path.join(tmp, "ort.txt"), parts)
        for k in (1, K4):
            t[("4", k)] = _run_ref(exe, lambda i, k=k: ["mode=4", "templateort=tort.txt", "templatestck=tmpl.mrc",
                                                        "searchstep=1.0", f"newortfile=o{i}.txt", f"first={i * K + 1}",
                                                        f"last={i * K + k}", "ortfile=ort.txt", *common], ncores, tmp)
        for k in (1, K3):
            t[("3", k)] = _run_ref(exe, lambda i, k=k: ["mode=3", "model3d=model.mrc", f"newortfile=c{i}.txt",
                                                        f"first={i * K + 1}", f"last={i * K + k}", "ortfile=ort.txt",
                                                        *common], ncores, tmp)
        maxres_ins = apix * box / w["insert_maxR"]
        for k in (1, K1):
            t[("1", k)] = _run_ref(exe, lambda i, k=k: ["mode=1", "subrec=y", f"result3d=s{i}.tnf",
                                                        f"first={i * K + 1}", f"last={i * K + k}", f"apix={apix}",
                                                        f"maxRes={maxres_ins}", f"imgmask={d['mask'] * apix}",
                                                        "ortfile=ort.txt"], ncores, tmp)
    s4 = (t[("4", K4)] - t[("4", 1)]) / (K4 - 1)
    s3 = (t[("3", K3)] - t[("3", 1)]) / (K3 - 1)
    s1 = (t[("1", K1)] - t[("1", 1)]) / (K1 - 1)
    stages = {"mode3_particles_per_s": round(ncores / s3, 3), "mode4_particles_per_s": round(ncores / s4, 4),
              "mode1_particles_per_s": round(ncores / s1, 2)}
    return {"value": ncores / (s4 + s1), "unit": "particles/s", "cores": ncores, "kind": "reference",
            "stages": stages,
            "sample": (f"icosprocess (flang -O2, MKL FFTW) as {ncores} concurrent processes on the box-128 workload "
                       f"(T={len(tang)}): mode 4 on the 1 deg grid, {K4} particles each ({t[('4', K4)]:.1f} s) minus "
                       f"1 each ({t[('4', 1)]:.1f} s); mode 1 subrec at maxR {w['insert_maxR']}, {K1} particles each "
                       f"({t[('1', K1)]:.1f} s) minus 1 ({t[('1', 1)]:.1f} s); value = search + insert.  Mode 3 "
                       f"(calc_center, model3d) apart: {K3} each ({t[('3', K3)]:.1f} s) minus 1 ({t[('3', 1)]:.1f} s)")}


def _host():
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu": cpu, "logical_cpus": os.cpu_count(), "machine": platform.machine()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stages", nargs="*", default=["main", "center", "lbox"])
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--templates", type=int, default=4)
    args = ap.parse_args()
    if not os.path.exists(os.path.join(REF, "icosprocess")):
        raise SystemExit("oracle/_ref is not built; run oracle/build_ref.sh first")
    res = json.load(open(OUT)) if os.path.exists(OUT) else {}
    ph, tang, timgs, ps = bench.make_workload(args.batch, args.templates)
    ncores = min(8, os.cpu_count() or 1)
    stamp = datetime.date.today().isoformat()
    if "main" in args.stages:
        r = _ref_baseline(os.path.join(REF, "icosprocess"), args, tang, timgs, ps, ncores)
        res["main"] = dict(r, date=stamp, host=_host())
        print("main", res["main"], flush=True)
    if "center" in args.stages:
        r = center_cpu_baseline(ph, ps)
        res["center"] = dict(r, date=stamp, host=_host())
        print("center", res["center"], flush=True)
    if "c1" in args.stages:
        r = c1_baseline(ncores)
        res["c1"] = dict(r, date=stamp, host=_host())
        print("c1", res["c1"], flush=True)
    if "lbox" in args.stages:
        r = lbox_cpu_baseline(ps)
        res["lbox"] = dict(r, date=stamp, host=_host())
        print("lbox", res["lbox"], flush=True)
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1)
    print(OUT)


if __name__ == "__main__":
    main()
