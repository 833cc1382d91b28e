"""local_reconstruct through the C ABI and the Fortran drop-in
(cryoim_amd/bin/local_reconstruct_hip) against the reference program's own
outputs (tests/golden/lrec.npz), and the mask-then-norm substep against the
FP64 oracle at box 64."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pipelines
from conftest import rel_l2
from cryoim_amd import formats, lib, synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cryoim_amd", "bin", "local_reconstruct_hip")


@pytest.fixture(scope="module")
def lr():
    return np.load(pipelines.golden_path("lrec.npz"))


@pytest.fixture(scope="module")
def lrdir(lr):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (amdflang missing)")
    d = tempfile.mkdtemp(prefix="cimlr_")
    formats.write_mrc(os.path.join(d, "stk.mrc"), lr["lrec_images"])
    ang, sh, df, z, cen = lr["lrec_angles"], lr["lrec_shifts"], lr["lrec_df"], lr["lrec_z"], lr["lrec_cen"]
    with open(os.path.join(d, "lort.txt"), "w") as f:
        for i in range(len(ang)):
            f.write(f"stk.mrc {i + 1} 0 0 0 0 0 0 {i + 1} {i + 1} {df[i]:.3f} {df[i]:.3f} 0 {i + 1} "
                    f"{cen[i, 0]} {cen[i, 1]} {ang[i, 0]:.4f} {ang[i, 1]:.4f} {ang[i, 2]:.4f} "
                    f"{sh[i, 0]:.4f} {sh[i, 1]:.4f} {z[i]:.4f} 0\n")
    return d


def _run(d, *args):
    r = subprocess.run([EXE, "local_ort=lort.txt", "local_FFTsize=32", "apix=2.0", "maxRes=6", *args], cwd=d,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    return r.stdout


@pytest.mark.parametrize("tag,extra", [("c5", ["local_sym=c5"]), ("i2", ["local_sym=i2"]), ("i3", ["local_sym=I3"]),
                                       ("i5", ["local_sym=i5"]),
                                       ("zhigh", ["local_sym=c1", "Zhigh_df=y", "sign=-"])])
def test_local_reconstruct_driver_map(lrdir, lr, tag, extra):
    _run(lrdir, *extra, f"result3d={tag}.mrc")
    raw = open(os.path.join(lrdir, f"{tag}.mrc"), "rb").read()
    assert np.array_equal(np.frombuffer(raw[:1024], "<i4"), lr["lrec_hdr"])
    err = rel_l2(np.frombuffer(raw[1024:], "<f4").reshape(32, 32, 32), lr[f"lrec_{tag}_map"])
    print(f"local_reconstruct {tag} map vs reference: rel-L2 {err:.2e}")
    assert err < 1e-5


def test_local_reconstruct_driver_substep(lrdir, lr):
    _run(lrdir, "local_sym=c5", "subrec=y", "imgmask=24", "result3d=c5.tnf")
    t = formats.read_tnf(os.path.join(lrdir, "c5.tnf"))
    assert [t["fft_size"], t["max_r"], t["fsc"], t["realspaceavg"]] == list(lr["lrecsub_hdr"])
    assert np.allclose([t["apix"], t["imgmask"]], lr["lrecsub_hdrf"])
    for k in ("fa", "fb", "wg"):
        err = rel_l2(t[k], lr[f"lrecsub_{k}"])
        print(f"local_reconstruct substep {k}: rel-L2 {err:.2e}")
        assert err < 1e-5


def test_mask_then_norm_substep_vs_fp64_oracle_box64(ctx, O64):
    """norm_on = 2 (mask, then normalise) with I5 operators on 64-pixel
    sub-particle boxes cut from 96-pixel images."""
    box, box0, apix, n, maxR = 64, 96, 1.5, 12, 20
    ph = synth.make_phantom(box, nblobs=10, radius=(10, 20), sigma=(1.5, 2.5), seed=51, sym=False)
    rng = np.random.default_rng(52)
    ang = np.round(np.c_[rng.uniform(20, 160, n), rng.uniform(-180, 180, n), rng.uniform(0, 360, n)], 4)
    sh = np.round(rng.uniform(-1.5, 1.5, (n, 2)), 4)
    df = np.round(rng.uniform(9000, 18000, n), 3)
    cen = rng.integers(30, 66, (n, 2))
    proj = synth.project(ph, ang, sh, box=box)
    imgs = (0.2 * rng.standard_normal((n, box0, box0))).astype(np.float32)
    for i in range(n):
        ys, xs = (np.arange(box) + cen[i, 1] - box // 2) % box0, (np.arange(box) + cen[i, 0] - box // 2) % box0
        imgs[i][np.ix_(ys, xs)] += proj[i]
    mask = 26.0
    vol = pipelines.local_reconstruct_volume(O64, imgs, cen, ang, sh, df, apix, box, maxR, "i5", compact=True,
                                             mask_r=mask)
    boxes = np.stack([O64.local_box(imgs[i], int(cen[i, 0]), int(cen[i, 1]), box) for i in range(n)]).astype(
        np.float32)
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3], P[:, 3:5], P[:, 6], P[:, 7] = ang, sh, df, df
    P[:, 9:11], P[:, 11] = sh, 1.0
    p = lib.make_params(box, 0, maxR, apply_ctf=True, mask_on=True, norm_on=2, apix=apix, imgmask=mask, vol=200.0,
                        df_scale=1.0)
    ctx.rec_begin(box, maxR, layout=1, rcut=maxR - 1, wg_init=0.1, ops=lib.local_reconstruct_sym_ops("i5"))
    ctx.rec_insert(p, boxes, P)
    for o, v, k in zip(ctx.rec_get_tnf(maxR), vol, ("fa", "fb", "wg")):
        err = rel_l2(o, v)
        print(f"box 64 substep {k}: rel-L2 {err:.2e} vs FP64 oracle")
        assert err < 1e-5
