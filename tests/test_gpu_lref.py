"""HIP local_refine (central slices + full-plane phase residual) against the
reference's own local_refine outputs (golden, c1 and c5) and the FP64 oracle."""
import os

import numpy as np
import pytest

import pipelines
from cryoim_amd import lib

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def lprog():
    return np.load(os.path.join(G, "lref_program.npz"))


def _records(lprog):
    n = len(lprog["lprog_images"])
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3] = np.round(lprog["lprog_start"], 4)
    P[:, 3:5] = np.round(lprog["lprog_shifts"] * 0.5, 4)
    P[:, lib.P_DF1] = lprog["lprog_df"]
    P[:, lib.P_DF2] = lprog["lprog_df"]
    P[:, lib.P_BOXX:lib.P_BOXY + 1] = lprog["lprog_cen"]
    return P


@pytest.mark.parametrize("sym", ["c1", "c5"])
def test_local_refine_matches_reference_program(ctx, lprog, sym):
    box, box0 = (int(v) for v in lprog["lprog_box"])
    apix = float(lprog["lprog_apix"][0])
    d = pipelines.lref_derived(box, apix, maxRes=8.0, imgmask_A=24.0)
    p = lib.make_params(box, 0, box // 2 - 2, apply_ctf=True, mask_on=True, conj_out=False, apix=apix,
                        imgmask=d["mask"], vol=200.0, df_scale=1.0, ncycle=3, delta_angle=d["refine_angle"],
                        delta_trans=d["refine_shift"])
    ctx.lref_model(p, lprog["lprog_model"])
    out = ctx.lref_refine(p, lprog["lprog_images"], _records(lprog), sym=sym, convF=0.8, maxres=d["maxres"])
    ref = lprog[f"lprog_{sym}_out"]
    same = 0
    for i in range(len(ref)):
        da = (out[i, 0:3] - ref[i, 0:3] + 180) % 360 - 180
        if np.abs(da).max() < 2e-3 and np.abs(out[i, 3:5] - ref[i, 3:5]).max() < 2e-3:
            same += 1
        assert abs(out[i, 5] - ref[i, 6]) < 0.05, (i, out[i], ref[i])
    print(f"{sym}: {same}/{len(ref)} sub-particles on the reference's exact path")
    assert same >= len(ref) - 1


def test_local_refine_fp64_oracle_box64(ctx, O64):
    from cryoim_amd import synth
    box, box0, apix, n = 64, 80, 1.5, 4
    ph = synth.make_phantom(box, nblobs=12, radius=(10, 20), sigma=(1.5, 2.5), seed=31, sym=False)
    model = synth.phantom_map(ph)
    rng = np.random.default_rng(32)
    ang = np.c_[rng.uniform(30, 150, n), rng.uniform(-180, 180, n), rng.uniform(0, 360, n)]
    proj = synth.project(ph, ang, box=box)
    imgs = np.zeros((n, box0, box0), np.float32)
    cen = rng.integers(30, 50, (n, 2))
    df = rng.uniform(9000, 18000, n)
    for i in range(n):
        c = synth.ctf2d(box, df[i], df[i], 0.0, apix, vol=200.0, df_scale=1.0)
        pimg = synth.apply_ctf(proj[i], c) + 0.2 * rng.standard_normal((box, box))
        ys = (np.arange(box) + cen[i, 1] - box // 2) % box0
        xs = (np.arange(box) + cen[i, 0] - box // 2) % box0
        imgs[i][np.ix_(ys, xs)] = pimg
    d = pipelines.lref_derived(box, apix, maxRes=9.0, imgmask_A=40.0)
    start = np.round(ang + rng.uniform(-2, 2, (n, 3)), 3)
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3] = start
    P[:, lib.P_DF1] = df
    P[:, lib.P_DF2] = df
    P[:, lib.P_BOXX:lib.P_BOXY + 1] = cen
    p = lib.make_params(box, 0, box // 2 - 2, apply_ctf=True, mask_on=True, conj_out=False, apix=apix,
                        imgmask=d["mask"], vol=200.0, df_scale=1.0, ncycle=4, delta_angle=d["refine_angle"],
                        delta_trans=d["refine_shift"])
    ctx.lref_model(p, model)
    out = ctx.lref_refine(p, imgs, P, sym="c1", convF=0.8, maxres=d["maxres"])
    fr, fi = pipelines.lref_model_ft(O64, model, d["mask"])
    for i in range(n):
        a20, b20, w = pipelines.lref_particle(O64, imgs[i], box, cen[i, 0], cen[i, 1], 0.0, 0.0, d["mask"],
                                              d["maxres"])
        ctf = O64.ctf2d(box, df[i], df[i], 0.0, apix, vol=200.0, df_scale=1.0)
        io5, pr = O64.lref_refine(fr, fi, a20, b20, ctf, w, None, 4, d["refine_angle"], d["refine_shift"], 0.8,
                                  d["maxres"], np.array([*start[i], 0.0, 0.0]))
        da = (out[i, 0:3] - io5[:3] + 180) % 360 - 180
        assert np.abs(da).max() < 2e-3 and np.abs(out[i, 3:5] - io5[3:5]).max() < 2e-3, (i, out[i, :6], io5, pr)
        assert abs(out[i, 5] - pr) < 1e-2


@pytest.fixture(scope="module")
def ldf():
    return np.load(os.path.join(G, "lref_df.npz"))


def _ldf_setup(ldf):
    box, box0 = (int(v) for v in ldf["ldf_box"])
    apix = float(ldf["ldf_apix"][0])
    d = pipelines.lref_derived(box, apix, maxRes=8.0, imgmask_A=24.0)
    p = lib.make_params(box, 0, box // 2 - 2, apply_ctf=True, mask_on=True, conj_out=False, apix=apix,
                        imgmask=d["mask"], vol=200.0, df_scale=1.0, ncycle=3, delta_angle=d["refine_angle"],
                        delta_trans=d["refine_shift"])
    n = len(ldf["ldf_images"])
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3] = np.round(ldf["ldf_start"], 4)
    P[:, 3:5] = np.round(ldf["ldf_shifts"] * 0.5, 4)
    P[:, lib.P_DF1] = np.round(ldf["ldf_df"], 3)
    P[:, lib.P_DF2] = np.round(ldf["ldf_df"], 3)
    P[:, lib.P_BOXX:lib.P_BOXY + 1] = ldf["ldf_cen"]
    P[:, lib.P_Z] = np.round(ldf["ldf_z0"], 4)
    return p, d, P


def test_particle_df_matches_reference_program(ctx, ldf):
    """refine_particle_df: the 401-defocus scan at the starting orientation
    (cim_lref_defocus) against the real local_refine particle_df=y sign=+."""
    p, d, P = _ldf_setup(ldf)
    ctx.lref_model(p, ldf["ldf_model"])
    out = ctx.lref_defocus(p, ldf["ldf_images"], P, sign=1.0, maxres=d["maxres"])
    ref = ldf["ldf_pdf_out"]
    same = 0
    for i in range(len(ref)):
        print(f"particle {i}: z {out[i, lib.P_Z]:.2f} PR {out[i, lib.P_PR]:.4f} vs reference z {ref[i, 5]:.2f} "
              f"PR {ref[i, 6]:.4f}")
        assert abs(out[i, lib.P_PR] - ref[i, 6]) < 2e-3
        same += abs(out[i, lib.P_Z] - ref[i, 5]) < 1e-2
        # the orientation and shift are left as they came in
        assert np.array_equal(out[i, 0:5], P[i, 0:5])
    assert same >= len(ref) - 1


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_particle_df_scan_grid_and_sign(ctx, ldf, sign):
    """The scanned heights lie on z0 + 4 (i - 200) A, i = 0..400, and sign
    -1 scans df - z: with the z0 column negated the same defoci are visited,
    so the chosen z is the negated one (within the residual's ties)."""
    p, d, P = _ldf_setup(ldf)
    ctx.lref_model(p, ldf["ldf_model"])
    Q = P.copy()
    Q[:, lib.P_Z] *= sign
    out = ctx.lref_defocus(p, ldf["ldf_images"], Q, sign=sign, maxres=d["maxres"])
    step = (out[:, lib.P_Z] - Q[:, lib.P_Z]) / 4.0
    assert np.allclose(step, np.round(step), atol=1e-3) and np.all(np.abs(step) <= 200)
    ref = ctx.lref_defocus(p, ldf["ldf_images"], P, sign=1.0, maxres=d["maxres"])
    assert np.allclose(sign * out[:, lib.P_Z], ref[:, lib.P_Z], atol=1e-2)
    assert np.allclose(out[:, lib.P_PR], ref[:, lib.P_PR], atol=1e-5)


def test_zhigh_matches_reference_program(ctx, ldf):
    """refine_PR_Zhigh is refine_PR under the CTF at df + local_z0 (sign=+)."""
    p, d, P = _ldf_setup(ldf)
    ctx.lref_model(p, ldf["ldf_model"])
    Q = P.copy()
    Q[:, lib.P_DF1] += Q[:, lib.P_Z]
    Q[:, lib.P_DF2] += Q[:, lib.P_Z]
    out = ctx.lref_refine(p, ldf["ldf_images"], Q, sym="c1", convF=0.8, maxres=d["maxres"])
    ref = ldf["ldf_zhigh_out"]
    same = 0
    for i in range(len(ref)):
        da = (out[i, 0:3] - ref[i, 0:3] + 180) % 360 - 180
        same += np.abs(da).max() < 2e-3 and np.abs(out[i, 3:5] - ref[i, 3:5]).max() < 2e-3
        assert abs(out[i, 5] - ref[i, 6]) < 0.05, (i, out[i], ref[i])
    assert same >= len(ref) - 1


DRV = os.path.join(os.path.dirname(G), "..", "cryoim_amd", "bin", "local_refine_hip")


def _driver_rows(path):
    rows = []
    with open(path) as f:
        for line in f:
            rows.append([float(v) for v in line[64:].split()[15:]])
    return np.array(rows)


def _driver_dir(g, pre, z0):
    import tempfile
    from cryoim_amd import formats
    if not os.path.exists(DRV):
        pytest.skip("Fortran driver not built (amdflang missing)")
    d = tempfile.mkdtemp(prefix="cimlrf_")
    formats.write_mrc(os.path.join(d, "model.mrc"), g[f"{pre}_model"], is_map=True)
    formats.write_mrc(os.path.join(d, "stk.mrc"), g[f"{pre}_images"])
    df, cen, st, sh = g[f"{pre}_df"], g[f"{pre}_cen"], g[f"{pre}_start"], g[f"{pre}_shifts"]
    with open(os.path.join(d, "lort.txt"), "w") as f:
        for i in range(len(df)):
            f.write(f"stk.mrc {i + 1} 0 0 0 0 0 0 {i + 1} {i + 1} {df[i]:.3f} {df[i]:.3f} 0 {i + 1} "
                    f"{cen[i, 0]} {cen[i, 1]} {st[i, 0]:.4f} {st[i, 1]:.4f} {st[i, 2]:.4f} "
                    f"{sh[i, 0] * 0.5:.4f} {sh[i, 1] * 0.5:.4f} {z0[i]:.4f} 0\n")
    return d


def _run_driver(d, tag, *extra):
    import subprocess
    r = subprocess.run([DRV, "local_ort0=lort.txt", f"local_newort={tag}.txt", "model3d=model.mrc", "apix=2.0",
                        "maxRes=8", "Ncycle=3", "imgmask=24", *extra], cwd=d, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    return _driver_rows(os.path.join(d, f"{tag}.txt"))


def _same_path(out, ref, tol_pr):
    same = 0
    for i in range(len(ref)):
        da = (out[i, 0:3] - ref[i, 0:3] + 180) % 360 - 180
        same += np.abs(da).max() < 2e-3 and np.abs(out[i, 3:5] - ref[i, 3:5]).max() < 2e-3
        assert abs(out[i, 6] - ref[i, 6]) < tol_pr, (i, out[i], ref[i])
    return same


@pytest.mark.parametrize("sym", ["c1", "c5"])
def test_local_refine_driver(lprog, sym):
    """cryoim_amd/bin/local_refine_hip with the reference program's command line."""
    d = _driver_dir(lprog, "lprog", np.zeros(len(lprog["lprog_df"])))
    out = _run_driver(d, sym, *([] if sym == "c1" else [f"sym={sym}"]))
    assert out.shape[1] == 12                 # lth..lPR and the five deltas
    ref = lprog[f"lprog_{sym}_out"]
    assert _same_path(out, ref, 0.05) >= len(ref) - 1
    assert np.allclose(out[:, 7:12], out[:, 0:5] - np.c_[np.round(lprog["lprog_start"], 4),
                                                         np.round(lprog["lprog_shifts"] * 0.5, 4)], atol=2e-3)


@pytest.mark.parametrize("tag,extra", [("zhigh", ["Zhigh_df=y", "sign=+"]), ("pdf", ["particle_df=y", "sign=+"])])
def test_local_refine_driver_defocus(ldf, tag, extra):
    d = _driver_dir(ldf, "ldf", ldf["ldf_z0"])
    out = _run_driver(d, tag, *extra)
    assert out.shape[1] == 14                 # lth..lPR, dth dph dom dx dy dz dPR
    ref = ldf[f"ldf_{tag}_out"]
    if tag == "zhigh":
        assert _same_path(out, ref, 0.05) >= len(ref) - 1
        assert np.allclose(out[:, 5], ref[:, 5], atol=1e-3)
    else:
        assert np.sum(np.abs(out[:, 5] - ref[:, 5]) < 1e-2) >= len(ref) - 1
        assert np.allclose(out[:, 6], ref[:, 6], atol=2e-3)
        assert np.allclose(out[:, 0:5], ref[:, 0:5], atol=1e-3)
        assert np.allclose(out[:, 12], out[:, 5] - np.round(ldf["ldf_z0"], 4), atol=1e-2)
