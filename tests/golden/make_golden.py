"""Record golden vectors from the reference's own Fortran (built by
oracle/build_ref.sh into oracle/_ref/).  Runs only where /root/reference was
available to build it; the GPU box and the CPU test suite only read the
fixtures this script writes (tests/golden/*.npz).

Two kinds of vectors:
* unit vectors: oracle/_ref/icos_units and lref_units call one reference
  routine on seeded inputs (symmetry table, euler_mat, getCTF2d, img2d_mask,
  img2d_norm, img2d_FFT, img2d_FT_centshift, FT_amp_expfit,
  icos_crosslin_angle, return_icos_asymmetry_unit, cal_crosslin_PR_omega,
  cal_crosslin_PR, FFT3d_slice, cal_PR_FT, img2d_FTphase, img3d_softmask,
  img3d_FFT);
* program vectors: the real icosprocess (modes 1, 1+subrec, 4, 5) and
  local_refine executables run on a small synthetic data set written in the
  reference's own MRC / ortfile formats;
* FSC / merge vectors: reconstruct_half + reconstruct_FSC (through
  oracle/_ref/icosprocess_fsc, whose main parses the FSC key) and the real
  merge_icosprocess_subrec on half and partial tnf files of the same set.

Usage: python tests/golden/make_golden.py [--only NAME ...]
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
REFERENCE_ROOT = os.environ.get("CIM_REFERENCE", "/root/reference")   # read for data fixtures only
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from cryoim_amd import formats, synth  # noqa: E402

f4 = np.float32
i4 = np.int32


def _run_unit(exe, case, payload: bytes, nout: int) -> np.ndarray:
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(payload)
        subprocess.run([os.path.join(REF, exe), case, fin, fout], check=True, stdout=subprocess.DEVNULL)
        out = np.fromfile(fout, dtype="<f4")
    assert out.size == nout, (case, out.size, nout)
    return out


def _b(*arrs) -> bytes:
    return b"".join(np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<")).tobytes() for a in arrs)


def unit_vectors(rng) -> dict:
    g = {}
    g["symops"] = _run_unit("icos_units", "symops", b"", 540).reshape(60, 3, 3)
    ang = np.c_[rng.uniform(0, 180, 16), rng.uniform(-180, 180, 16), rng.uniform(0, 360, 16)].astype(f4)
    g["euler_in"] = ang
    g["euler_out"] = _run_unit("icos_units", "euler", _b(np.array([16], i4), ang), 16 * 9).reshape(16, 3, 3)
    # CTF (icosprocess units: um, kV, mm)
    ctfp = np.array([[1.5, 1.7, 30.0, 1.5, 300.0, 2.7, 0.1], [2.8, 2.8, 0.0, 2.0, 200.0, 2.0, 0.07]], f4)
    g["ctf_in"] = ctfp
    g["ctf_out"] = np.stack([_run_unit("icos_units", "ctf", _b(np.array([40], i4), c), 1600).reshape(40, 40)
                             for c in ctfp])
    img = rng.standard_normal((2, 40, 40)).astype(f4) + 0.3
    mp = np.array([[14.0, 1.5, -2.0], [17.0, 0.0, 0.0]], f4)
    g["mask_in"], g["mask_par"] = img, mp
    g["mask_out"] = np.stack([_run_unit("icos_units", "mask", _b(np.array([40], i4), mp[k], img[k]), 1600)
                              .reshape(40, 40) for k in range(2)])
    g["norm_in"] = img[0] * 3 + 1
    g["norm_out"] = _run_unit("icos_units", "norm", _b(np.array([40], i4), g["norm_in"]), 1600).reshape(40, 40)
    a = rng.standard_normal((24, 24)).astype(f4)
    b = rng.standard_normal((24, 24)).astype(f4)
    o = _run_unit("icos_units", "fft2", _b(np.array([24], i4), a, b), 4 * 576).reshape(4, 24, 24)
    g["fft2_in"] = np.stack([a, b])
    g["fft2_out"] = o                      # fwd a, fwd b, then backward of that
    g["shift_in"] = np.stack([a, b])
    g["shift_xy"] = np.array([2.3, -1.7], f4)
    g["shift_out"] = _run_unit("icos_units", "centshift", _b(np.array([24], i4), g["shift_xy"], a, b),
                               2 * 576).reshape(2, 24, 24)
    # FT_amp_expfit on a synthetic decaying transform
    nx, mn, mx = 64, 1, 20
    ky, hx = np.meshgrid(np.arange(nx) - nx // 2, np.arange(nx) - nx // 2, indexing="ij")
    r = np.sqrt(hx ** 2 + ky ** 2)
    amp = np.exp(-0.08 * r) * (1 + 0.3 * rng.standard_normal((nx, nx)))
    ph = rng.uniform(-np.pi, np.pi, (nx, nx))
    ea, eb = (amp * np.cos(ph)).astype(f4), (amp * np.sin(ph)).astype(f4)
    g["expfit_in"] = np.stack([ea, eb])
    g["expfit_r"] = np.array([mn, mx], i4)
    g["expfit_out"] = _run_unit("icos_units", "expfit", _b(np.array([nx, mn, mx], i4), ea, eb), mx - mn + 1)
    pairs = np.c_[rng.uniform(60, 95, 12), rng.uniform(-40, 40, 12), rng.uniform(0, 360, 12),
                  rng.uniform(0, 180, 12), rng.uniform(-180, 180, 12), rng.uniform(0, 360, 12)].astype(f4)
    g["crosslin_in"] = pairs
    g["crosslin_out"] = _run_unit("icos_units", "crosslin", _b(np.array([12], i4), pairs), 12 * 120).reshape(12, 2, 60)
    au = np.c_[rng.uniform(0, 180, 30), rng.uniform(-180, 180, 30), rng.uniform(0, 360, 30)].astype(f4)
    g["au_in"] = au
    g["au_out"] = _run_unit("icos_units", "au", _b(np.array([30], i4), au), 90).reshape(30, 3)
    return g


def pr_vectors(rng) -> dict:
    """cal_crosslin_PR_omega / cal_crosslin_PR on synthetic transforms."""
    from oracle import Oracle
    O = Oracle(32)
    g = {}
    box, mn, mx, T, step = 48, 1, 10, 2, 6.0
    ph = synth.make_phantom(box, radius=(13, 19), sigma=(1.2, 2.0), seed=5)
    tang = np.array([[75.0, 10.0, 30.0], [85.0, -12.0, 200.0]])
    timgs = synth.project(ph, tang)
    ps = synth.make_particles(ph, 1, seed=6, shift_max=1.5, snr=2.0, ctf=True, apix=2.0, df_range=(1.0, 2.0))
    x, y = ps.shifts[0]
    a0, b0 = O.particle_ft(ps.images[0], x, y, mask_r=20.0, shift=False)
    fc0 = O.amp_expfit(a0, b0, mn, mx)
    pa, pb = O.ft_shift(a0, b0, x, y, True)
    ta, tb, tf = [], [], []
    for t in range(T):
        a, b = O.particle_ft(timgs[t], 0.0, 0.0, mask_r=20.0)
        ta.append(a); tb.append(b); tf.append(O.amp_expfit(a, b, mn, mx))
    ta, tb, tf = (np.array(v, f4) for v in (ta, tb, tf))
    ctf = O.ctf2d(box, ps.df[0, 0], ps.df[0, 1], 0.0, 2.0).astype(f4)
    g.update(pr_box=np.array([box, mn, mx, T], i4), pr_step=np.array([step], f4), pr_pa=pa, pr_pb=pb,
             pr_a0=a0, pr_b0=b0, pr_fc0=fc0, pr_ta=ta, pr_tb=tb, pr_tfcut=tf, pr_tang=tang.astype(f4), pr_ctf=ctf)
    cases = [(78.0, 6.0, 0.0, 0.0, 0.0), (84.0, -18.0, 0.0, 0.0, 0.0)]
    outs = []
    for actf in (1, 0):
        for (th, phv, om, xx, yy) in cases:
            payload = _b(np.array([box, mn, mx, T, actf], i4), np.array([step, th, phv, om, xx, yy], f4), pa, pb,
                         fc0, *[v for t in range(T) for v in (ta[t], tb[t])], tang.astype(f4), tf, ctf)
            outs.append(_run_unit("icos_units", "pr_omega", payload, int(360 / step)))
    g["pr_omega_cases"] = np.array(cases, f4)
    g["pr_omega_out"] = np.array(outs).reshape(2, len(cases), -1)
    pcases = [(79.0, 5.0, 40.0, 0.7, -0.4), (72.0, -3.0, 300.0, -1.2, 0.9)]
    pouts = []
    for actf in (1, 0):
        for (th, phv, om, xx, yy) in pcases:
            payload = _b(np.array([box, mn, mx, T, actf], i4), np.array([step, th, phv, om, xx, yy], f4), a0, b0,
                         fc0, *[v for t in range(T) for v in (ta[t], tb[t])], tang.astype(f4), tf, ctf)
            pouts.append(_run_unit("icos_units", "pr", payload, 1)[0])
    g["pr_cases"] = np.array(pcases, f4)
    g["pr_out"] = np.array(pouts, f4).reshape(2, len(pcases))
    return g


def lref_vectors(rng) -> dict:
    g = {}
    nx = 24
    fr = rng.standard_normal((nx, nx, nx)).astype(f4)
    fi = rng.standard_normal((nx, nx, nx)).astype(f4)
    ang = np.array([[33.0, 71.0, 250.0], [90.0, 0.0, 0.0]], f4)
    g["slice_in"] = np.stack([fr, fi])
    g["slice_ang"] = ang
    g["slice_out"] = np.stack([_run_unit("lref_units", "slice", _b(np.array([nx], i4), a, fr, fi), 2 * nx * nx)
                               .reshape(2, nx, nx) for a in ang])
    a1, b1, a2, b2 = (rng.standard_normal((32, 32)).astype(f4) for _ in range(4))
    g["prft_in"] = np.stack([a1, b1, a2, b2])
    g["prft_r"] = np.array([[-1.0, 9.5], [3.0, 14.0]], f4)
    g["prft_out"] = np.array([_run_unit("lref_units", "prft", _b(np.array([32], i4), r, a1, b1, a2, b2), 1)[0]
                              for r in g["prft_r"]], f4)
    g["ftphase_xy"] = np.array([1.3, -2.2], f4)
    g["ftphase_out"] = _run_unit("lref_units", "ftphase", _b(np.array([32], i4), g["ftphase_xy"], a1, b1),
                                 2 * 1024).reshape(2, 32, 32)
    vol = rng.standard_normal((nx, nx, nx)).astype(f4) + 0.5
    g["softmask_in"] = vol
    g["softmask_r"] = np.array([8.5], f4)
    g["softmask_out"] = _run_unit("lref_units", "softmask", _b(np.array([nx], i4), g["softmask_r"], vol),
                                  nx ** 3).reshape(nx, nx, nx)
    v12r, v12i = rng.standard_normal((2, 12, 12, 12)).astype(f4)
    g["fft3_in"] = np.stack([v12r, v12i])
    g["fft3_out"] = _run_unit("lref_units", "fft3", _b(np.array([12], i4), v12r, v12i), 2 * 1728).reshape(2, 12, 12, 12)
    # symmetry generators of the local programs and icos_equalort
    for kind, nn in ((0, 5), (1, 4), (2, 60), (3, 60)):
        g[f"lsym_{kind}_{nn}"] = _run_unit("lref_units", "symops", _b(np.array([kind, nn], i4)),
                                           9 * nn).reshape(nn, 3, 3)
    eq_in = np.array([71.0, 23.0, 140.0], f4)
    g["equalort_in"] = eq_in
    for kind, nn in ((0, 5), (2, 60), (3, 60)):
        g[f"equalort_{kind}_{nn}"] = _run_unit("lref_units", "equalort", _b(np.array([kind, nn], i4), eq_in),
                                               3 * nn).reshape(nn, 3)
    return g


def lref_program_vectors() -> dict:
    """The real local_refine on sub-particles cut from 48-pixel images
    against a 32-voxel model (c1, and c5 with the symmetry stage)."""
    from cryoim_amd import geometry
    g = {}
    box, box0, apix, n = 32, 48, 2.0, 6
    ph = synth.make_phantom(box, nblobs=8, radius=(6, 10), sigma=(1.2, 1.8), seed=17, sym=False)
    model = synth.phantom_map(ph)
    rng = np.random.default_rng(18)
    ang = np.c_[rng.uniform(20, 160, n), rng.uniform(-180, 180, n), rng.uniform(0, 360, n)]
    sh = rng.uniform(-1.0, 1.0, (n, 2))
    proj = synth.project(ph, ang, sh, box=box)
    df = rng.uniform(8000, 20000, n)              # Angstrom, as the local programs take it
    imgs = np.zeros((n, box0, box0), f4)
    cen = rng.integers(18, 30, (n, 2))
    for i in range(n):
        c = synth.ctf2d(box, df[i], df[i], 0.0, apix, vol=200.0, df_scale=1.0)
        p = synth.apply_ctf(proj[i], c) + 0.3 * rng.standard_normal((box, box))
        ys, xs = (np.arange(box) + cen[i, 1] - box // 2) % box0, (np.arange(box) + cen[i, 0] - box // 2) % box0
        imgs[i][np.ix_(ys, xs)] = p
    start = ang + rng.uniform(-2.0, 2.0, (n, 3))
    g.update(lprog_model=model, lprog_images=imgs, lprog_angles=ang.astype(f4), lprog_shifts=sh.astype(f4),
             lprog_df=df.astype(f4), lprog_cen=cen.astype(i4), lprog_start=start.astype(f4),
             lprog_box=np.array([box, box0], i4), lprog_apix=np.array([apix], f4))
    exe = os.path.join(REF, "local_refine")
    with tempfile.TemporaryDirectory() as d:
        formats.write_mrc(os.path.join(d, "model.mrc"), model, is_map=True)
        formats.write_mrc(os.path.join(d, "stk.mrc"), imgs)
        with open(os.path.join(d, "lort.txt"), "w") as f:
            for i in range(n):
                f.write(f"stk.mrc {i + 1} 0 0 0 0 0 0 {i + 1} {i + 1} {df[i]:.3f} {df[i]:.3f} 0 {i + 1} "
                        f"{cen[i, 0]} {cen[i, 1]} {start[i, 0]:.4f} {start[i, 1]:.4f} {start[i, 2]:.4f} "
                        f"{sh[i, 0] * 0.5:.4f} {sh[i, 1] * 0.5:.4f} 0 0\n")
        for tag, extra in (("c1", []), ("c5", ["sym=c5"])):
            r = subprocess.run([exe, "local_ort0=lort.txt", f"local_newort={tag}.txt", "model3d=model.mrc",
                                f"apix={apix}", "maxRes=8", "Ncycle=3", "imgmask=24", *extra], cwd=d,
                               stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stdout)
            rows = []
            with open(os.path.join(d, f"{tag}.txt")) as f:
                for line in f:
                    t = line[64:].split()
                    # id th ph om x y PR imgID imgpid df1 df2 ast limgID hx0 ky0 lth lph lom lx ly lz lPR ...
                    rows.append([float(v) for v in t[15:22]])
            g[f"lprog_{tag}_out"] = np.array(rows, f4)
    _ = geometry
    return g


def lref_defocus_vectors() -> dict:
    """The real local_refine with Zhigh_df=y (refine_PR_Zhigh: the CTF at
    df + local_z0) and with particle_df=y (refine_particle_df: 401 defoci
    local_z0 + 4 (i - 200) A scanned at the starting orientation), sign=+, on
    sub-particles whose true defocus is the ortfile's df plus an offset z."""
    g = {}
    box, box0, apix, n = 32, 48, 2.0, 6
    ph = synth.make_phantom(box, nblobs=8, radius=(6, 10), sigma=(1.2, 1.8), seed=27, sym=False)
    model = synth.phantom_map(ph)
    rng = np.random.default_rng(28)
    ang = np.c_[rng.uniform(20, 160, n), rng.uniform(-180, 180, n), rng.uniform(0, 360, n)]
    sh = rng.uniform(-1.0, 1.0, (n, 2))
    proj = synth.project(ph, ang, sh, box=box)
    df = rng.uniform(8000, 20000, n)              # the true defocus (A)
    z = rng.uniform(-500, 500, n)                  # true height offset: the ortfile holds df - z
    z0 = z + rng.uniform(-300, 300, n)             # the starting local_z0 of the scan
    imgs = np.zeros((n, box0, box0), f4)
    cen = rng.integers(18, 30, (n, 2))
    for i in range(n):
        c = synth.ctf2d(box, df[i], df[i], 0.0, apix, vol=200.0, df_scale=1.0)
        p = synth.apply_ctf(proj[i], c) + 0.2 * rng.standard_normal((box, box))
        ys, xs = (np.arange(box) + cen[i, 1] - box // 2) % box0, (np.arange(box) + cen[i, 0] - box // 2) % box0
        imgs[i][np.ix_(ys, xs)] = p
    start = ang + rng.uniform(-1.5, 1.5, (n, 3))
    dfo = df - z
    g.update(ldf_model=model, ldf_images=imgs, ldf_angles=ang.astype(f4), ldf_shifts=sh.astype(f4),
             ldf_df=dfo.astype(f4), ldf_z=z.astype(f4), ldf_z0=z0.astype(f4), ldf_cen=cen.astype(i4),
             ldf_start=start.astype(f4), ldf_box=np.array([box, box0], i4), ldf_apix=np.array([apix], f4))
    exe = os.path.join(REF, "local_refine")
    with tempfile.TemporaryDirectory() as d:
        formats.write_mrc(os.path.join(d, "model.mrc"), model, is_map=True)
        formats.write_mrc(os.path.join(d, "stk.mrc"), imgs)
        with open(os.path.join(d, "lort.txt"), "w") as f:
            for i in range(n):
                f.write(f"stk.mrc {i + 1} 0 0 0 0 0 0 {i + 1} {i + 1} {dfo[i]:.3f} {dfo[i]:.3f} 0 {i + 1} "
                        f"{cen[i, 0]} {cen[i, 1]} {start[i, 0]:.4f} {start[i, 1]:.4f} {start[i, 2]:.4f} "
                        f"{sh[i, 0] * 0.5:.4f} {sh[i, 1] * 0.5:.4f} {z0[i]:.4f} 0\n")
        for tag, extra in (("zhigh", ["Zhigh_df=y", "sign=+"]), ("pdf", ["particle_df=y", "sign=+"])):
            r = subprocess.run([exe, "local_ort0=lort.txt", f"local_newort={tag}.txt", "model3d=model.mrc",
                                f"apix={apix}", "maxRes=8", "Ncycle=3", "imgmask=24", *extra], cwd=d,
                               stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stdout)
            rows = []
            with open(os.path.join(d, f"{tag}.txt")) as f:
                for line in f:
                    t = line[64:].split()
                    # ... lth lph lom lx ly lz lPR (tokens 15-21)
                    rows.append([float(v) for v in t[15:22]])
            g[f"ldf_{tag}_out"] = np.array(rows, f4)
    return g


def local_reconstruct_vectors() -> dict:
    """The real local_reconstruct (built with the one line flang rejects
    removed) on 40 sub-particles: 32-pixel boxes cut from 48-pixel images at
    (local_hx0, local_ky0), local orientations and shifts, CTF in Angstrom;
    maps for local_sym c5, i2, i3, i5, the substep tnf, and Zhigh_df."""
    g = {}
    box, box0, apix, n = 32, 48, 2.0, 40
    ph = synth.make_phantom(box, nblobs=8, radius=(6, 10), sigma=(1.2, 1.8), seed=27, sym=False)
    rng = np.random.default_rng(28)
    # rounded to the decimals the local ortfile carries, so the stored values are the ones read
    ang = np.round(np.c_[rng.uniform(20, 160, n), rng.uniform(-180, 180, n), rng.uniform(0, 360, n)], 4)
    sh = np.round(rng.uniform(-1.0, 1.0, (n, 2)), 4)
    proj = synth.project(ph, ang, sh, box=box)
    df = np.round(rng.uniform(8000, 20000, n), 3)
    lz = np.round(rng.uniform(-300.0, 300.0, n), 4)
    imgs = np.zeros((n, box0, box0), f4)
    cen = rng.integers(14, 34, (n, 2))
    for i in range(n):
        c = synth.ctf2d(box, df[i], df[i], 0.0, apix, vol=200.0, df_scale=1.0)
        p = synth.apply_ctf(proj[i], c) + 0.3 * rng.standard_normal((box, box))
        ys, xs = (np.arange(box) + cen[i, 1] - box // 2) % box0, (np.arange(box) + cen[i, 0] - box // 2) % box0
        imgs[i][np.ix_(ys, xs)] = p
    g.update(lrec_images=imgs, lrec_angles=ang.astype(f4), lrec_shifts=sh.astype(f4), lrec_df=df.astype(f4),
             lrec_z=lz.astype(f4), lrec_cen=cen.astype(i4), lrec_box=np.array([box, box0], i4),
             lrec_apix=np.array([apix], f4))
    exe = os.path.join(REF, "local_reconstruct")
    with tempfile.TemporaryDirectory() as d:
        formats.write_mrc(os.path.join(d, "stk.mrc"), imgs)
        with open(os.path.join(d, "lort.txt"), "w") as f:
            for i in range(n):
                f.write(f"stk.mrc {i + 1} 0 0 0 0 0 0 {i + 1} {i + 1} {df[i]:.3f} {df[i]:.3f} 0 {i + 1} "
                        f"{cen[i, 0]} {cen[i, 1]} {ang[i, 0]:.4f} {ang[i, 1]:.4f} {ang[i, 2]:.4f} "
                        f"{sh[i, 0]:.4f} {sh[i, 1]:.4f} {lz[i]:.4f} 0\n")
        common = ["local_ort=lort.txt", f"local_FFTsize={box}", f"apix={apix}", "maxRes=6"]
        runs = {"c5": ["local_sym=c5"], "i2": ["local_sym=i2"], "i3": ["local_sym=I3"], "i5": ["local_sym=i5"],
                "zhigh": ["local_sym=c1", "Zhigh_df=y", "sign=-"]}
        for tag, extra in runs.items():
            r = subprocess.run([exe, *common, *extra, f"result3d={tag}.mrc"], cwd=d, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stdout)
            g[f"lrec_{tag}_map"] = np.fromfile(os.path.join(d, f"{tag}.mrc"), dtype="<f4", offset=1024).reshape(
                box, box, box)
        with open(os.path.join(d, "c5.mrc"), "rb") as f:
            g["lrec_hdr"] = np.frombuffer(f.read(1024), dtype="<i4").copy()
        r = subprocess.run([exe, *common, "local_sym=c5", "subrec=y", "imgmask=24", "result3d=c5.tnf"], cwd=d,
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stdout)
        t = formats.read_tnf(os.path.join(d, "c5.tnf"))
        g["lrecsub_fa"], g["lrecsub_fb"], g["lrecsub_wg"] = t["fa"], t["fb"], t["wg"]
        g["lrecsub_hdr"] = np.array([t["fft_size"], t["max_r"], t["fsc"], t["realspaceavg"]], i4)
        g["lrecsub_hdrf"] = np.array([t["apix"], t["imgmask"]], f4)
    return g


def _read_sym_matrix_file(path) -> np.ndarray:
    """The numbers of a local_box symmetry-matrix fixture, 3 per line, read as
    read_sym_matrixa does: column by column (read_sym_matrix.f90:8-12)."""
    vals = [float(v) for v in open(path).read().split()]
    m = np.array(vals, np.float64).reshape(-1, 3, 3)        # [op, column, row]
    return np.transpose(m, (0, 2, 1)).astype(f4)            # [op, row, column]


def _parse_local_ort(path, long):
    """local_box's ortfiles: loc_box_coordinate.f90:45 (long) or loc_box.f90:91."""
    rows = []
    for line in open(path):
        if len(line.rstrip("\n")) < 70:
            continue
        rows.append([float(v) for v in line[64:].split()])
    return np.array(rows, np.float64)


SYM_FILES = ("icos2f_sym_matrix", "icos2f_sym_matrix_53", "icos2f_12vertex_matrix", "icos5f_12vertex_matrix",
             "icos3f_20facecenter", "icos2f_30edgecenter")


def symfile_vectors() -> dict:
    """The symmetry-matrix data files shipped in local_box/SRC, as the operators
    read_sym_matrixa reads from them (column by column), in float64."""
    g = {}
    for name in SYM_FILES:
        vals = [float(v) for v in open(os.path.join(REFERENCE_ROOT, "local_box", "SRC", name + ".txt")).read().split()]
        m = np.array(vals, np.float64).reshape(-1, 3, 3)          # [op, column, row]
        g["sym_" + name] = np.transpose(m, (0, 2, 1)).copy()     # [op, row, column]
    return g


def local_box_vectors() -> dict:
    """The real local_box on the program_vectors particles: local ortfiles
    (coordinate mode) for sym=i5 and for the 12-vertex matrix fixture, and the
    sub-particle stack (stack mode) for the fixture."""
    g = {}
    box, apix, n = 48, 2.0, 30
    ph = synth.make_phantom(box, radius=(13, 19), sigma=(1.2, 2.0), seed=7)
    ps = synth.make_particles(ph, n, seed=8, shift_max=1.5, snr=1.0, ctf=True, apix=apix, df_range=(1.0, 2.5))
    pr = np.zeros(n, f4)
    pr[[3, 11, 17]] = 95.0                          # dropped by PR_threshold = 90
    g.update(lbox_images=ps.images.astype(f4), lbox_angles=np.round(ps.angles, 4).astype(f4),
             lbox_shifts=np.round(ps.shifts, 4).astype(f4), lbox_df=np.round(ps.df, 4).astype(f4), lbox_pr=pr)
    fixture = os.path.join(REFERENCE_ROOT, "local_box", "SRC", "icos5f_12vertex_matrix.txt")
    g["lbox_v12_ops"] = _read_sym_matrix_file(fixture)
    exe = os.path.join(REF, "local_box")
    with tempfile.TemporaryDirectory() as d:
        formats.write_mrc(os.path.join(d, "stk.mrc"), ps.images)
        parts = [formats.Particle("stk.mrc", i + 1, *g["lbox_angles"][i], *g["lbox_shifts"][i], float(pr[i]), i + 1,
                                  i + 1, float(g["lbox_df"][i, 0]), float(g["lbox_df"][i, 1]), 0.0)
                 for i in range(n)]
        formats.write_ortfile(os.path.join(d, "icos.txt"), parts)
        with open(os.path.join(d, "v12.txt"), "w") as f:
            f.write(open(fixture).read())

        def run(args):
            r = subprocess.run([exe, "icos_ort=icos.txt", f"apix={apix}", *args], cwd=d, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stdout)

        run(["local_ort=c_i5.txt", "sym=i5", "local_x=0", "local_y=0", "local_z=18"])
        g["lbox_coords_i5"] = _parse_local_ort(os.path.join(d, "c_i5.txt"), True)
        run(["local_ort=c_v12.txt", "sym_matrix=v12.txt", "local_x=1.5", "local_y=-2", "local_z=17.5"])
        g["lbox_coords_v12"] = _parse_local_ort(os.path.join(d, "c_v12.txt"), True)
        run(["local_ort=s_v12.txt", "local_imgstck=s.mrc", "sub_FFTsize=24", "sym_matrix=v12.txt", "local_x=0",
             "local_y=0", "local_z=14"])
        g["lbox_stack_ort"] = _parse_local_ort(os.path.join(d, "s_v12.txt"), False)
        raw = open(os.path.join(d, "s.mrc"), "rb").read()
        g["lbox_stack_hdr"] = np.frombuffer(raw[:1024], dtype="<i4").copy()
        g["lbox_stack"] = np.frombuffer(raw[1024:], dtype="<f4").reshape(-1, 24, 24).copy()
    return g


def program_vectors() -> dict:
    """Real icosprocess / local_refine runs on a 48-pixel synthetic set."""
    g = {}
    box, apix, n = 48, 2.0, 30
    ph = synth.make_phantom(box, radius=(13, 19), sigma=(1.2, 2.0), seed=7)
    ps = synth.make_particles(ph, n, seed=8, shift_max=1.5, snr=1.0, ctf=True, apix=apix, df_range=(1.0, 2.5))
    tang = np.array([[72.0, 8.0, 20.0], [80.0, -15.0, 140.0], [88.0, 3.0, 300.0]])
    timgs = synth.project(ph, tang)
    g.update(prog_images=ps.images, prog_angles=ps.angles.astype(f4), prog_shifts=ps.shifts.astype(f4),
             prog_df=ps.df.astype(f4), prog_tang=tang.astype(f4), prog_timgs=timgs.astype(f4),
             prog_box=np.array([box], i4), prog_apix=np.array([apix], f4))
    exe = os.path.join(REF, "icosprocess")
    with tempfile.TemporaryDirectory() as d:
        stk = "stack.mrc"
        formats.write_mrc(os.path.join(d, stk), ps.images)
        formats.write_mrc(os.path.join(d, "tmpl.mrc"), timgs)
        formats.write_template_ort(os.path.join(d, "tort.txt"), tang)
        parts = [formats.Particle(stk, i + 1, *ps.angles[i], *ps.shifts[i], 0.0, i + 1, i + 1, ps.df[i, 0],
                                  ps.df[i, 1], 0.0) for i in range(n)]
        formats.write_ortfile(os.path.join(d, "ort.txt"), parts)
        common = ["apix=2.0", "maxRes=6.0", "imgmask=40", "ortfile=ort.txt", f"last={n}"]

        def run(args):
            r = subprocess.run([exe, *args], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stdout)
            return r.stdout

        run(["mode=1", "result3d=map.mrc", *common])
        g["mode1_map"] = formats.read_mrc(os.path.join(d, "map.mrc"))
        run(["mode=1", "subrec=y", "result3d=sub.tnf", *common])
        t = formats.read_tnf(os.path.join(d, "sub.tnf"))
        g["mode1sub_fa"], g["mode1sub_fb"], g["mode1sub_wg"] = t["fa"], t["fb"], t["wg"]
        g["mode1sub_hdr"] = np.array([t["fft_size"], t["max_r"]], i4)
        # search on the first 6 particles, 6 degree grid
        run(["mode=4", "templateort=tort.txt", "templatestck=tmpl.mrc", "searchstep=6.0", "newortfile=m4.txt",
             "apix=2.0", "maxRes=6.0", "minRes=100", "imgmask=40", "ortfile=ort.txt", "last=6"])
        m4 = formats.read_ortfile900(os.path.join(d, "m4.txt"))
        g["mode4_out"] = np.array([[p.theta, p.phi, p.omega, p.x, p.y, p.pr] for p in m4], f4)
        run(["mode=5", "templateort=tort.txt", "templatestck=tmpl.mrc", "Ncycle=3", "newortfile=m5.txt",
             "apix=2.0", "maxRes=6.0", "minRes=100", "imgmask=40", "ortfile=ort.txt", "last=6"])
        m5 = formats.read_ortfile900(os.path.join(d, "m5.txt"))
        g["mode5_out"] = np.array([[p.theta, p.phi, p.omega, p.x, p.y, p.pr] for p in m5], f4)
    return g


def projection_vectors() -> dict:
    """project_template2d through the unit driver on the 48-pixel phantom map,
    icosprocess mode 0 (project2d: the template stack) and mode 4 without a
    templatestck (load_template projects the templates from model3d) on the
    program_vectors data set."""
    g = {}
    box, apix, n = 48, 2.0, 30
    ph = synth.make_phantom(box, radius=(13, 19), sigma=(1.2, 2.0), seed=7)
    vol = synth.phantom_map(ph).astype(f4)
    ang = np.array([[72.0, 8.0, 20.0], [80.0, -15.0, 140.0], [88.0, 3.0, 300.0], [33.3, 71.7, 189.1]], f4)
    g["proj_model"], g["proj_angles"] = vol, ang
    g["proj_out"] = _run_unit("icos_units", "project", _b(np.array([box, len(ang)], i4), ang, vol),
                              len(ang) * box * box).reshape(len(ang), box, box)
    ps = synth.make_particles(ph, n, seed=8, shift_max=1.5, snr=1.0, ctf=True, apix=apix, df_range=(1.0, 2.5))
    tang = np.array([[72.0, 8.0, 20.0], [80.0, -15.0, 140.0], [88.0, 3.0, 300.0]])
    exe = os.path.join(REF, "icosprocess")
    with tempfile.TemporaryDirectory() as d:
        formats.write_mrc(os.path.join(d, "stack.mrc"), ps.images)
        formats.write_mrc(os.path.join(d, "model.mrc"), vol)
        formats.write_template_ort(os.path.join(d, "tort.txt"), tang)
        parts = [formats.Particle("stack.mrc", i + 1, *ps.angles[i], *ps.shifts[i], 0.0, i + 1, i + 1, ps.df[i, 0],
                                  ps.df[i, 1], 0.0) for i in range(n)]
        formats.write_ortfile(os.path.join(d, "ort.txt"), parts)

        def run(args):
            r = subprocess.run([exe, *args], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stdout)
            return r.stdout

        run(["mode=0", "model3d=model.mrc", "templateort=tort.txt", "templatestck=tstk.mrc"])
        g["mode0_stack"] = formats.read_mrc(os.path.join(d, "tstk.mrc"))
        run(["mode=4", "templateort=tort.txt", "model3d=model.mrc", "searchstep=6.0", "newortfile=m4.txt",
             "apix=2.0", "maxRes=6.0", "minRes=100", "imgmask=40", "ortfile=ort.txt", "last=6"])
        m4 = formats.read_ortfile900(os.path.join(d, "m4.txt"))
        g["mode4_model_out"] = np.array([[p.theta, p.phi, p.omega, p.x, p.y, p.pr] for p in m4], f4)
    return g


def _read_fsc_txt(path) -> np.ndarray:
    """FSC.txt rows 'ir FSC sqrt(2FSC/(1+FSC))' (list-directed, NaN allowed)."""
    rows = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) == 3:
                rows.append([float(v) for v in t])
    return np.array(rows, np.float64)


def fsc_program_vectors() -> dict:
    """Half maps, FSC weighting and tnf merges on the program_vectors data set."""
    g = {}
    box, apix, n = 48, 2.0, 30
    ph = synth.make_phantom(box, radius=(13, 19), sigma=(1.2, 2.0), seed=7)
    ps = synth.make_particles(ph, n, seed=8, shift_max=1.5, snr=1.0, ctf=True, apix=apix, df_range=(1.0, 2.5))
    icos, icos_fsc = os.path.join(REF, "icosprocess"), os.path.join(REF, "icosprocess_fsc")
    merge = os.path.join(REF, "merge_icosprocess_subrec")
    with tempfile.TemporaryDirectory() as d:
        stk = "stack.mrc"
        formats.write_mrc(os.path.join(d, stk), ps.images)
        parts = [formats.Particle(stk, i + 1, *ps.angles[i], *ps.shifts[i], 0.0, i + 1, i + 1, ps.df[i, 0],
                                  ps.df[i, 1], 0.0) for i in range(n)]
        formats.write_ortfile(os.path.join(d, "ort.txt"), parts)
        common = ["apix=2.0", "maxRes=6.0", "imgmask=40", "ortfile=ort.txt"]

        def run(exe, args):
            r = subprocess.run([exe, *args], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stdout)
            return r.stdout

        def tnf(name, key):
            t = formats.read_tnf(os.path.join(d, name))
            g[key + "_fa"], g[key + "_fb"], g[key + "_wg"] = t["fa"], t["fb"], t["wg"]
            g[key + "_hdr"] = np.array([t["fft_size"], t["max_r"], t["fsc"], t["realspaceavg"]], i4)
            g[key + "_hdrf"] = np.array([t["apix"], t["imgmask"]], f4)

        def raw_map(name, m):
            return np.fromfile(os.path.join(d, name), dtype="<f4", offset=1024, count=m ** 3).reshape(m, m, m)

        # mode 1 with FSC: reconstruct_half over particles 1,3,.. and 2,4,.., then reconstruct_FSC
        run(icos_fsc, ["mode=1", "FSC=y", "result3d=fsc.mrc", "first=1", f"last={n}", *common])
        tnf("oddtnf", "half_odd")
        tnf("evetnf", "half_eve")
        g["fsc_txt"] = _read_fsc_txt(os.path.join(d, "FSC.txt"))
        g["fsc_map"] = raw_map("fsc.mrc", box)
        # subrec + FSC over two particle ranges, merged with FSC weighting (nFSC = 1)
        run(icos_fsc, ["mode=1", "FSC=y", "subrec=y", "result3d=p1", "first=1", "last=15", *common])
        run(icos_fsc, ["mode=1", "FSC=y", "subrec=y", "result3d=p2", "first=16", f"last={n}", *common])
        for p in ("p1", "p2"):
            tnf(p + "oddtnf", p + "_odd")
            tnf(p + "evetnf", p + "_eve")
        run(merge, ["p1", "p2", "mfsc.mrc"])
        g["mfsc_txt"] = _read_fsc_txt(os.path.join(d, "FSC.txt"))
        g["mfsc_map"] = raw_map("mfsc.mrc", box - 8)
        # plain partial reconstructions merged (nFSC = 0)
        run(icos, ["mode=1", "subrec=y", "result3d=s1.tnf", "first=1", "last=15", *common])
        run(icos, ["mode=1", "subrec=y", "result3d=s2.tnf", "first=16", f"last={n}", *common])
        tnf("s1.tnf", "s1")
        tnf("s2.tnf", "s2")
        run(merge, ["s1.tnf", "s2.tnf", "m0.mrc"])
        g["merge_map"] = raw_map("m0.mrc", box - 8)
        with open(os.path.join(d, "m0.mrc"), "rb") as f:
            g["merge_hdr"] = np.frombuffer(f.read(1024), dtype="<i4").copy()
    return g


def icosavg_program_vectors() -> dict:
    """Real-space icosahedral averaging (realspaceavg=y) on the program_vectors
    data set: icosprocess's img3d_icosavg after mode 1 and after
    reconstruct_FSC, and merge_icosprocess_subrec's img3d_icosavg in the plain
    and the FSC-weighted merge (tnf headers with nicosavg = 1)."""
    g = {}
    box, apix, n = 48, 2.0, 30
    ph = synth.make_phantom(box, radius=(13, 19), sigma=(1.2, 2.0), seed=7)
    ps = synth.make_particles(ph, n, seed=8, shift_max=1.5, snr=1.0, ctf=True, apix=apix, df_range=(1.0, 2.5))
    icos, icos_fsc = os.path.join(REF, "icosprocess"), os.path.join(REF, "icosprocess_fsc")
    merge = os.path.join(REF, "merge_icosprocess_subrec")
    with tempfile.TemporaryDirectory() as d:
        stk = "stack.mrc"
        formats.write_mrc(os.path.join(d, stk), ps.images)
        parts = [formats.Particle(stk, i + 1, *ps.angles[i], *ps.shifts[i], 0.0, i + 1, i + 1, ps.df[i, 0],
                                  ps.df[i, 1], 0.0) for i in range(n)]
        formats.write_ortfile(os.path.join(d, "ort.txt"), parts)
        common = ["apix=2.0", "maxRes=6.0", "imgmask=40", "ortfile=ort.txt", "realspaceavg=y"]

        def run(exe, args):
            r = subprocess.run([exe, *args], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stdout)
            return r.stdout

        def raw_map(name, m):
            return np.fromfile(os.path.join(d, name), dtype="<f4", offset=1024, count=m ** 3).reshape(m, m, m)

        run(icos, ["mode=1", "result3d=plain.mrc", f"last={n}", *common[:-1]])
        g["avg_input_map"] = raw_map("plain.mrc", box)
        run(icos, ["mode=1", "result3d=avg.mrc", f"last={n}", *common])
        g["avg_map"] = raw_map("avg.mrc", box)
        run(icos_fsc, ["mode=1", "FSC=y", "result3d=fsc.mrc", "first=1", f"last={n}", *common])
        g["fscavg_map"] = raw_map("fsc.mrc", box)
        run(icos, ["mode=1", "subrec=y", "result3d=s1.tnf", "first=1", "last=15", *common])
        run(icos, ["mode=1", "subrec=y", "result3d=s2.tnf", "first=16", f"last={n}", *common])
        g["avg_tnf_hdr"] = np.array([formats.read_tnf(os.path.join(d, "s1.tnf"))["realspaceavg"]], i4)
        run(merge, ["s1.tnf", "s2.tnf", "m0.mrc"])
        g["mavg_map"] = raw_map("m0.mrc", box - 8)
        run(icos_fsc, ["mode=1", "FSC=y", "subrec=y", "result3d=p1", "first=1", "last=15", *common])
        run(icos_fsc, ["mode=1", "FSC=y", "subrec=y", "result3d=p2", "first=16", f"last={n}", *common])
        run(merge, ["p1", "p2", "mfsc.mrc"])
        g["mfscavg_map"] = raw_map("mfsc.mrc", box - 8)
    return g


def corrctf_program_vectors() -> dict:
    """reconstruct_corrCTF and reconstruct_corr_CTF_substep (mode 1 with
    corrCTF=step,cycles: a defocus series of 2 cycles x 0.2 um each side) on
    the program_vectors data set, by the key-patched icosprocess_corrctf."""
    g = {}
    box, apix, n = 48, 2.0, 30
    ph = synth.make_phantom(box, radius=(13, 19), sigma=(1.2, 2.0), seed=7)
    ps = synth.make_particles(ph, n, seed=8, shift_max=1.5, snr=1.0, ctf=True, apix=apix, df_range=(1.0, 2.5))
    exe = os.path.join(REF, "icosprocess_corrctf")
    g["corr_step_cycles"] = np.array([0.2, 2.0], f4)
    with tempfile.TemporaryDirectory() as d:
        stk = "stack.mrc"
        formats.write_mrc(os.path.join(d, stk), ps.images)
        parts = [formats.Particle(stk, i + 1, *ps.angles[i], *ps.shifts[i], 0.0, i + 1, i + 1, ps.df[i, 0],
                                  ps.df[i, 1], 0.0) for i in range(n)]
        formats.write_ortfile(os.path.join(d, "ort.txt"), parts)
        common = ["apix=2.0", "maxRes=6.0", "imgmask=40", "ortfile=ort.txt", f"last={n}", "corrCTF=0.2,2"]
        for args in (["mode=1", "result3d=c.mrc"], ["mode=1", "subrec=y", "result3d=c.tnf"]):
            r = subprocess.run([exe, *args, *common], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                               text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stdout)
            if "subrec=y" not in args and "reconstruction:" not in r.stdout:   # reconstruct_corrCTF.f90:15
                raise RuntimeError("corrCTF path not taken:\n" + r.stdout[-2000:])
        g["corr_map"] = np.fromfile(os.path.join(d, "c.mrc"), dtype="<f4", offset=1024).reshape(box, box, box)
        t = formats.read_tnf(os.path.join(d, "c.tnf"))
        g["corrsub_fa"], g["corrsub_fb"], g["corrsub_wg"] = t["fa"], t["fb"], t["wg"]
        g["corrsub_hdr"] = np.array([t["fft_size"], t["max_r"], t["fsc"], t["realspaceavg"]], i4)
    return g


def symc_program_vectors() -> dict:
    """icosprocess mode 1 with a cyclic sym (reconstruct_C1 for c5a and c7,
    reconstruct_C1_substep, img3d_icosavg over the group) on the
    program_vectors data.  (Mode 6 has no c7 operators; the reference program
    ends in a segmentation fault there.)"""
    g = {}
    box, apix, n = 48, 2.0, 30
    ph = synth.make_phantom(box, radius=(13, 19), sigma=(1.2, 2.0), seed=7)
    ps = synth.make_particles(ph, n, seed=8, shift_max=1.5, snr=1.0, ctf=True, apix=apix, df_range=(1.0, 2.5))
    exe = os.path.join(REF, "icosprocess")
    with tempfile.TemporaryDirectory() as d:
        stk = "stack.mrc"
        formats.write_mrc(os.path.join(d, stk), ps.images)
        parts = [formats.Particle(stk, i + 1, *ps.angles[i], *ps.shifts[i], 0.0, i + 1, i + 1, ps.df[i, 0],
                                  ps.df[i, 1], 0.0) for i in range(n)]
        formats.write_ortfile(os.path.join(d, "ort.txt"), parts)
        common = ["apix=2.0", "maxRes=6.0", "imgmask=40", "ortfile=ort.txt", f"last={n}", "sym=c5a"]

        def run(args):
            r = subprocess.run([exe, *args], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stdout)
            return r.stdout

        def raw_map(name, m):
            return np.fromfile(os.path.join(d, name), dtype="<f4", offset=1024, count=m ** 3).reshape(m, m, m)

        run(["mode=1", "result3d=c.mrc", *common])
        g["c5a_map"] = raw_map("c.mrc", box)
        run(["mode=1", "result3d=ca.mrc", "realspaceavg=y", *common])
        g["c5a_avg_map"] = raw_map("ca.mrc", box)
        out = run(["mode=1", "subrec=y", "result3d=c.tnf", "maxcentshift=1.2", *common])
        t = formats.read_tnf(os.path.join(d, "c.tnf"))
        g["c5asub_fa"], g["c5asub_fb"], g["c5asub_wg"] = t["fa"], t["fb"], t["wg"]
        g["c5asub_hdr"] = np.array([t["fft_size"], t["max_r"], t["fsc"], t["realspaceavg"]], i4)
        g["c5asub_maxcentshift"] = np.array([1.2], f4)
        run(["mode=1", "result3d=c7.mrc", *common[:-1], "sym=c7"])
        g["c7_map"] = raw_map("c7.mrc", box)
    return g


def local_program_vectors() -> dict:
    """icosprocess mode 6 (reconstruct_local, reconstruct_local_substep): loc_box windows
    of the program_vectors images around seeded (x0, y0), local symmetry c5a."""
    g = {}
    box, apix, n, lb = 48, 2.0, 30, 32
    ph = synth.make_phantom(box, radius=(13, 19), sigma=(1.2, 2.0), seed=7)
    ps = synth.make_particles(ph, n, seed=8, shift_max=1.5, snr=1.0, ctf=True, apix=apix, df_range=(1.0, 2.5))
    rng = np.random.default_rng(66)
    x0y0 = rng.uniform(-6.0, 6.0, (n, 2)).astype(f4)
    g["loc_x0y0"] = x0y0
    g["loc_box"] = np.array([lb], i4)
    exe = os.path.join(REF, "icosprocess")
    with tempfile.TemporaryDirectory() as d:
        stk = "stack.mrc"
        formats.write_mrc(os.path.join(d, stk), ps.images)
        parts = [formats.Particle(stk, i + 1, *ps.angles[i], *ps.shifts[i], 0.0, i + 1, i + 1, ps.df[i, 0],
                                  ps.df[i, 1], 0.0, x0=float(x0y0[i, 0]), y0=float(x0y0[i, 1])) for i in range(n)]
        formats.write_ortfile(os.path.join(d, "ort6.txt"), parts, with_local_box=True)
        common = ["mode=6", f"loc_box={lb}", "sym=c5a", "apix=2.0", "maxRes=6.0", "imgmask=24", "ortfile=ort6.txt",
                  f"last={n}"]
        for args in (["result3d=loc.mrc"], ["subrec=y", "result3d=loc.tnf"]):
            r = subprocess.run([exe, *common, *args], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                               text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stdout)
        raw = open(os.path.join(d, "loc.mrc"), "rb").read()
        g["loc_hdr"] = np.frombuffer(raw[:1024], dtype="<i4").copy()
        g["loc_map"] = np.frombuffer(raw[1024:], dtype="<f4").reshape(lb, lb, lb).copy()
        t = formats.read_tnf(os.path.join(d, "loc.tnf"))
        g["locsub_fa"], g["locsub_fb"], g["locsub_wg"] = t["fa"], t["fb"], t["wg"]
        g["locsub_hdr"] = np.array([t["fft_size"], t["max_r"], t["fsc"], t["realspaceavg"]], i4)
        g["locsub_hdrf"] = np.array([t["apix"], t["imgmask"]], f4)
    return g


def _parse_center_lines(out: str, first: int, last: int) -> np.ndarray:
    """The 'print*, i, centx, centy, cc' lines of calc_center.f90:86 and
    calc_center_selfrot180.f90:74 (list-directed; NaN allowed)."""
    rows = {}
    for line in out.splitlines():
        t = line.split()
        if len(t) == 4:
            try:
                i = int(t[0])
                rows[i] = [float(v) for v in t[1:]]
            except ValueError:
                continue
    return np.array([rows[i] for i in range(first, last + 1)], np.float64)


def center_program_vectors() -> dict:
    """icosprocess mode 3 (centre finding) with a model (calc_center) and without
    (calc_center_selfrot180) on a 48-pixel set with shifts up to 2.5 px."""
    g = {}
    box, apix, n = 48, 2.0, 12
    ph = synth.make_phantom(box, radius=(13, 19), sigma=(1.2, 2.0), seed=7)
    ps = synth.make_particles(ph, n, seed=9, shift_max=2.5, snr=2.0, ctf=True, apix=apix, df_range=(1.0, 2.5))
    rng = np.random.default_rng(99)
    guess = (ps.shifts + rng.uniform(-0.5, 0.5, ps.shifts.shape)).astype(f4)
    model = synth.phantom_map(ph).astype(f4)
    g.update(cen_model=model, cen_images=ps.images.astype(f4), cen_angles=ps.angles.astype(f4),
             cen_shifts=ps.shifts.astype(f4), cen_guess=guess, cen_df=ps.df.astype(f4))
    exe = os.path.join(REF, "icosprocess")
    with tempfile.TemporaryDirectory() as d:
        stk = "stack.mrc"
        formats.write_mrc(os.path.join(d, stk), ps.images)
        formats.write_mrc(os.path.join(d, "model.mrc"), model)
        parts = [formats.Particle(stk, i + 1, *ps.angles[i], *guess[i], 0.0, i + 1, i + 1, ps.df[i, 0],
                                  ps.df[i, 1], 0.0) for i in range(n)]
        formats.write_ortfile(os.path.join(d, "ort.txt"), parts)
        common = ["mode=3", "apix=2.0", "maxRes=6.0", "imgmask=20", "ortfile=ort.txt", f"last={n}"]
        for key, extra in (("cen_model_out", ["model3d=model.mrc", "newortfile=c1.txt"]),
                           ("cen_self_out", ["newortfile=c2.txt"])):
            r = subprocess.run([exe, *common, *extra], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                               text=True)
            g[key] = _parse_center_lines(r.stdout, 1, n)
            g[key + "_rc"] = np.array([r.returncode], i4)
            ort = os.path.join(d, extra[-1].split("=")[1])
            g[key + "_ortfile"] = np.array([os.path.exists(ort)], i4)
            if os.path.exists(ort):
                rows = formats.read_ortfile900(ort)
                g[key + "_ort"] = np.array([[p.x, p.y, p.pr] for p in rows], f4)
    return g


def main():
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", default=None)
    args = ap.parse_args()
    rng = np.random.default_rng(20261203)
    if not os.path.exists(os.path.join(REF, "icos_units")):
        raise SystemExit("oracle/_ref is not built; run oracle/build_ref.sh first")
    jobs = [("units", lambda: unit_vectors(rng)), ("pr", lambda: pr_vectors(rng)),
            ("lref_units", lambda: lref_vectors(rng)), ("programs", program_vectors),
            ("lref_program", lref_program_vectors), ("fsc", fsc_program_vectors),
            ("local", local_program_vectors), ("center", center_program_vectors),
            ("icosavg", icosavg_program_vectors), ("corrctf", corrctf_program_vectors),
            ("symc", symc_program_vectors), ("lrec", local_reconstruct_vectors),
            ("lbox", local_box_vectors), ("symfiles", symfile_vectors), ("proj", projection_vectors),
            ("lref_df", lref_defocus_vectors)]
    for name, fn in jobs:
        # the seeded rng is shared by the first three: run them in order even when only one is wanted
        if args.only is not None and name not in args.only and name not in ("units", "pr", "lref_units"):
            continue
        vec = fn()
        if args.only is None or name in args.only:
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **vec)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
