"""local_box through the C ABI and the Fortran drop-in
(cryoim_amd/bin/local_box_hip) against the reference program's own outputs
(tests/golden/lbox.npz: coordinate mode with sym=i5 and with the 12-vertex
matrix file, stack mode with that file), and the stack path against the
FP64 oracle at box 384 with the 60 I2 operators."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import pipelines
from conftest import rel_l2
from cryoim_amd import formats, lib

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cryoim_amd", "bin", "local_box_hip")
X0_I5, X0_V12, X0_STACK = (0.0, 0.0, 18.0), (1.5, -2.0, 17.5), (0.0, 0.0, 14.0)


@pytest.fixture(scope="module")
def g():
    return np.load(pipelines.golden_path("lbox.npz"))


def good(g):
    pr, sh = g["lbox_pr"], g["lbox_shifts"]
    return [i for i in range(len(pr)) if pr[i] <= 90.0 and abs(sh[i, 0]) <= 100 and abs(sh[i, 1]) <= 100]


def records(g, idx):
    P = np.zeros((len(idx), lib.PART_STRIDE), np.float32)
    P[:, 0:3] = g["lbox_angles"][idx]
    P[:, 3:5] = g["lbox_shifts"][idx]
    return P


def angdiff(a, b):
    d = np.abs(np.asarray(a, np.float64) - b) % 360.0
    return np.minimum(d, 360.0 - d)


@pytest.mark.parametrize("case", ["i5", "v12"])
def test_coordinates_vs_reference(ctx, g, case):
    ops, x0 = (lib.local_box_sym_ops("i5"), X0_I5) if case == "i5" else (g["lbox_v12_ops"], X0_V12)
    idx = good(g)
    rec = ctx.local_box_coords(records(g, idx), ops, x0, 48, 2.0).reshape(-1, lib.LBOX_REC)
    ref = g[f"lbox_coords_{case}"]
    assert np.array_equal(rec[:, 3:5], ref[:, 13:15])
    da, df = angdiff(rec[:, 0:3], ref[:, 15:18]).max(), np.abs(rec[:, 5:8] - ref[:, 18:21]).max()
    print(f"local_box coordinates {case}: max angle diff {da:.1e} deg, fraction/height {df:.1e}")
    assert da < 1.5e-4 and df < 1.2e-4


def test_stack_vs_reference(ctx, g):
    idx = good(g)
    boxes, rec = ctx.local_box_stack(g["lbox_images"][idx], records(g, idx), g["lbox_v12_ops"], X0_STACK, 24, 2.0)
    rec = rec.reshape(-1, lib.LBOX_REC)
    ref = g["lbox_stack_ort"]
    assert angdiff(rec[:, 0:3], ref[:, 1:4]).max() < 1.5e-4
    assert np.abs(rec[:, 5:7] - ref[:, 4:6]).max() < 1.2e-4 and np.abs(rec[:, 7] - ref[:, 12]).max() < 1.2e-4
    err = rel_l2(boxes.reshape(-1, 24, 24), g["lbox_stack"])
    print(f"local_box stack vs reference: rel-L2 {err:.2e}")
    assert err < 1e-5


def test_stack_box384_vs_fp64_oracle(ctx, O64):
    """Box 384, 96-pixel boxes around a point 60 pixels out, 60 I2 operators;
    every box against the FP64 oracle, and each box's mean 0 and rms 1."""
    n, box, sub, x0 = 3, 384, 96, (12.0, -20.0, 55.0)
    rng = np.random.default_rng(11)
    imgs = rng.normal(size=(n, box, box)).astype(np.float32)
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3] = np.c_[rng.uniform(10, 170, n), rng.uniform(-180, 180, n), rng.uniform(-180, 180, n)]
    P[:, 3:5] = rng.uniform(-4, 4, (n, 2))
    ops = lib.local_box_sym_ops("i2")
    boxes, rec = ctx.local_box_stack(imgs, P, ops, x0, sub, 1.3)
    assert np.abs(boxes.mean(axis=(-1, -2))).max() < 1e-5
    assert np.abs(np.sqrt((boxes.astype(np.float64) ** 2).mean(axis=(-1, -2))) - 1).max() < 1e-5
    worst = 0.0
    for i in range(n):
        ob, om = O64.local_box_stack(imgs[i], P[i, 0:3], P[i, 3:5], ops, x0, sub, 1.3)
        assert np.abs(rec[i, :, 5:7] - om[:, 3:5]).max() < 2e-3
        worst = max(worst, rel_l2(boxes[i], ob))
        assert angdiff(rec[i, :, 0:3], om[:, 0:3]).max() < 2e-3
        assert np.abs(rec[i, :, 7] - om[:, 5]).max() < 2e-3
    print(f"box 384 local_box stack vs FP64 oracle: worst rel-L2 {worst:.2e}")
    assert worst < 2e-5


def test_stack_device_variant_matches_host(ctx, g):
    from devmem import DevArray
    idx = good(g)[:9]
    ops = lib.local_box_sym_ops("c5")
    host, hrec = ctx.local_box_stack(g["lbox_images"][idx], records(g, idx), ops, (3.0, 1.0, 9.0), 32, 1.0)
    dimg = DevArray.from_numpy(g["lbox_images"][idx])
    dbox = DevArray(len(idx) * 5 * 32 * 32)
    try:
        _, drec = ctx.local_box_stack(dimg, records(g, idx), ops, (3.0, 1.0, 9.0), 32, 1.0, boxes=dbox, box=48)
        assert np.array_equal(dbox.numpy(host.shape), host)
        assert np.array_equal(drec, hrec)
    finally:
        dimg.free()
        dbox.free()


def test_stack_refuses_boxes_beyond_one_wrap(ctx, g):
    idx = good(g)[:2]
    with pytest.raises(lib.CimError):
        ctx.local_box_stack(g["lbox_images"][idx], records(g, idx), np.eye(3)[None], (0, 0, 0), 64, 1.0)
    with pytest.raises(lib.CimError):
        ctx.local_box_stack(g["lbox_images"][idx], records(g, idx), np.eye(3)[None], (0, 0, 47.0), 24, 1.0)


# ---- the Fortran drop-in -----------------------------------------------------

@pytest.fixture(scope="module")
def lbdir(g):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (amdflang missing)")
    d = tempfile.mkdtemp(prefix="cimlb_")
    formats.write_mrc(os.path.join(d, "stk.mrc"), g["lbox_images"])
    parts = [formats.Particle("stk.mrc", i + 1, *g["lbox_angles"][i], *g["lbox_shifts"][i], float(g["lbox_pr"][i]),
                              i + 1, i + 1, float(g["lbox_df"][i, 0]), float(g["lbox_df"][i, 1]), 0.0)
             for i in range(len(g["lbox_pr"]))]
    formats.write_ortfile(os.path.join(d, "icos.txt"), parts)
    with open(os.path.join(d, "v12.txt"), "w") as f:
        for m in g["lbox_v12_ops"]:
            for col in m.T:
                f.write(" ".join(f"{float(v):.9g}" for v in col) + "\n")
            f.write("\n")
    yield d
    shutil.rmtree(d, ignore_errors=True)


def _run(d, *args):
    r = subprocess.run([EXE, "icos_ort=icos.txt", "apix=2.0", *args], cwd=d, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    return r.stdout


def _rows(path):
    return np.array([[float(v) for v in line[64:].split()] for line in open(path) if len(line.rstrip("\n")) >= 70])


@pytest.mark.parametrize("case,args", [
    ("i5", ["sym=i5", "local_x=0", "local_y=0", "local_z=18"]),
    ("v12", ["sym_matrix=v12.txt", "local_x=1.5", "local_y=-2", "local_z=17.5"])])
def test_driver_coordinates(lbdir, g, case, args):
    _run(lbdir, f"local_ort=c_{case}.txt", *args)
    got, ref = _rows(os.path.join(lbdir, f"c_{case}.txt")), g[f"lbox_coords_{case}"]
    assert got.shape == ref.shape
    assert np.array_equal(got[:, :15], ref[:, :15]) and np.array_equal(got[:, 21], ref[:, 21])
    assert angdiff(got[:, 15:18], ref[:, 15:18]).max() < 2.01e-4
    assert np.abs(got[:, 18:21] - ref[:, 18:21]).max() < 2.01e-4


def test_driver_stack(lbdir, g):
    _run(lbdir, "local_ort=s.txt", "local_imgstck=s.mrc", "sub_FFTsize=24", "sym_matrix=v12.txt", "local_x=0",
         "local_y=0", "local_z=14", "Nsym=5")
    raw = open(os.path.join(lbdir, "s.mrc"), "rb").read()
    assert np.array_equal(np.frombuffer(raw[:1024], "<i4"), g["lbox_stack_hdr"])
    err = rel_l2(np.frombuffer(raw[1024:], "<f4").reshape(-1, 24, 24), g["lbox_stack"])
    print(f"local_box_hip stack vs reference: rel-L2 {err:.2e}")
    assert err < 1e-5
    got, ref = _rows(os.path.join(lbdir, "s.txt")), g["lbox_stack_ort"]
    assert got.shape == ref.shape
    assert np.array_equal(got[:, [0, 6, 7, 8, 9, 10, 11]], ref[:, [0, 6, 7, 8, 9, 10, 11]])
    assert angdiff(got[:, 1:4], ref[:, 1:4]).max() < 2.01e-4
    assert np.abs(got[:, [4, 5, 12]] - ref[:, [4, 5, 12]]).max() < 2.01e-4


def test_driver_stack_from_icos_imgstck_and_first(lbdir, g):
    """icos_imgstck: images in order from particle first (loc_box.f90:14-20,
    35-36); first/last select the records."""
    _run(lbdir, "local_ort=s2.txt", "local_imgstck=s2.mrc", "icos_imgstck=stk.mrc", "sub_FFTsize=24",
         "sym_matrix=v12.txt", "local_z=14", "first=5", "last=20")
    raw = open(os.path.join(lbdir, "s2.mrc"), "rb").read()
    sel = [i for i in good(g) if 4 <= i < 20]
    boxes = np.frombuffer(raw[1024:], "<f4").reshape(-1, 12, 24, 24)
    ref = g["lbox_stack"].reshape(-1, 12, 24, 24)[[good(g).index(i) for i in sel]]
    assert boxes.shape == ref.shape and rel_l2(boxes, ref) < 1e-5
