"""Template bank projected from a model (icosprocess/SRC/load_template.f90:32-33,
project_template2d.f90:3-80; mode 0's project2d.f90) on the GPU, against the
reference's own projections and search results (tests/golden/proj.npz) and
the FP64 oracle."""
import numpy as np
import pytest

from conftest import rel_l2
from cryoim_amd import lib, synth
import pipelines

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pj():
    return np.load(pipelines.golden_path("proj.npz"))


def test_projection_matches_reference(ctx, O64, pj):
    got = ctx.project_templates(pj["proj_model"], pj["proj_angles"])
    for k, ang in enumerate(pj["proj_angles"]):
        e_ref = rel_l2(got[k], pj["proj_out"][k])
        e_64 = rel_l2(got[k], O64.project_template2d(pj["proj_model"], ang))
        print(f"projection at {ang}: rel-L2 {e_ref:.2e} vs reference, {e_64:.2e} vs FP64 oracle")
        assert e_ref < 1e-6 and e_64 < 1e-5


def test_projection_box128_vs_fp64(ctx, O64):
    box = 128
    ph = synth.make_phantom(box, radius=(40, 55), sigma=(2.0, 4.0), seed=101)
    vol = synth.phantom_map(ph)
    ang = np.array([[71.7, 4.3, 346.9], [88.0, -20.0, 100.0]])
    got = ctx.project_templates(vol, ang)
    for k in range(len(ang)):
        e = rel_l2(got[k], O64.project_template2d(vol, ang[k]))
        print(f"box 128 projection {ang[k]}: rel-L2 {e:.2e} vs FP64 oracle")
        assert e < 1e-5


def test_search_with_model_templates_matches_reference(ctx, pj, prog):
    """mode 4 without a templatestck: the templates are projected from model3d
    (the reference's mode-4 output on the 48-pixel set, first 6 particles, 6 deg)."""
    box, apix = 48, 2.0
    d = pipelines.derived(box, apix, maxRes=6.0, minRes=100.0, imgmask_A=40)
    p = lib.make_params(box, d["minR"], d["maxR"], apply_ctf=True, mask_on=True, apix=apix, imgmask=d["mask"],
                        searchstep=6.0)
    tang = prog["prog_tang"]
    ctx.search_templates_model(p, pj["proj_model"], np.c_[tang, np.zeros((len(tang), 2))])
    P = np.zeros((6, lib.PART_STRIDE), np.float32)
    P[:, 3:5] = prog["prog_shifts"][:6]
    P[:, 6:8] = prog["prog_df"][:6]
    P[:, 11] = 1
    out = ctx.search_global(p, prog["prog_images"][:6], P)
    ref = pj["mode4_model_out"]
    for o, r in zip(out, ref):
        assert np.abs(o[[0, 1, 2]] - r[[0, 1, 2]]).max() < 1e-3, (o, r)
        assert abs(o[5] - r[5]) < 2e-3, (o[5], r[5])
