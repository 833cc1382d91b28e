"""The oracle's refine_PR_Zhigh and refine_particle_df (local_refine/SRC/
refine_PR_Zhigh.f90, refine_particle_df.f90) against the real local_refine
run with Zhigh_df=y and particle_df=y, sign=+ (tests/golden/lref_df.npz).
Zhigh is refine_PR under the CTF at df + local_z0; particle_df scans 401
defoci z0 + 4 (i - 200) A at the starting orientation and keeps the first
minimum.  Also what the arithmetic fixes: the scan's residual at its own
minimum equals the plain residual of the chosen defocus."""
import numpy as np
import pytest

import pipelines


@pytest.fixture(scope="module")
def ldf():
    return np.load(pipelines.golden_path("lref_df.npz"))


def _inputs(O, g, i):
    box, box0 = (int(v) for v in g["ldf_box"])
    apix = float(g["ldf_apix"][0])
    d = pipelines.lref_derived(box, apix, maxRes=8.0, imgmask_A=24.0)
    fr, fi = pipelines.lref_model_ft(O, g["ldf_model"], d["mask"])
    x0, y0 = (float(np.float32(round(float(v), 4))) for v in g["ldf_shifts"][i] * 0.5)
    a20, b20, w = pipelines.lref_particle(O, g["ldf_images"][i], box, int(g["ldf_cen"][i, 0]),
                                          int(g["ldf_cen"][i, 1]), x0, y0, d["mask"], d["maxres"])
    start = np.array([round(float(v), 4) for v in g["ldf_start"][i]])
    df = float(np.float32(round(float(g["ldf_df"][i]), 3)))
    z0 = float(np.float32(round(float(g["ldf_z0"][i]), 4)))
    return box, apix, d, fr, fi, a20, b20, w, start, x0, y0, df, z0


@pytest.mark.parametrize("bits", [32, 64])
def test_particle_df_matches_reference(bits, O32, O64, ldf):
    O = O32 if bits == 32 else O64
    ref = ldf["ldf_pdf_out"]
    same = 0
    for i in range(len(ref)):
        box, apix, d, fr, fi, a20, b20, w, start, x0, y0, df, z0 = _inputs(O, ldf, i)
        z, pr, prs = O.lref_defocus(fr, fi, a20, b20, w, start, x0, y0, z0, df, df, 0.0, apix, 1.0, d["maxres"])
        near = np.sort(prs)[1] - prs.min()
        print(f"FP{bits} particle {i}: z {z:.2f} PR {pr:.4f} vs reference z {ref[i, 5]:.2f} PR {ref[i, 6]:.4f} "
              f"(next-best residual +{near:.2e})")
        assert abs(pr - ref[i, 6]) < 2e-3
        same += abs(z - ref[i, 5]) < 1e-2
    assert same >= len(ref) - (0 if bits == 32 else 1)


@pytest.mark.parametrize("bits", [32, 64])
def test_zhigh_matches_reference(bits, O32, O64, ldf):
    O = O32 if bits == 32 else O64
    ref = ldf["ldf_zhigh_out"]
    same = 0
    for i in range(len(ref)):
        box, apix, d, fr, fi, a20, b20, w, start, x0, y0, df, z0 = _inputs(O, ldf, i)
        ctf = O.ctf2d(box, df + z0, df + z0, 0.0, apix, vol=200.0, df_scale=1.0)
        io5, pr = O.lref_refine(fr, fi, a20, b20, ctf, w, None, 3, d["refine_angle"], d["refine_shift"], 0.8,
                                d["maxres"], np.array([*start, x0, y0]))
        da = (io5[:3] - ref[i, :3] + 180) % 360 - 180
        same += np.abs(da).max() < 2e-3 and np.abs(io5[3:5] - ref[i, 3:5]).max() < 2e-3
        assert abs(pr - ref[i, 6]) < 0.05
    assert same >= len(ref) - (0 if bits == 32 else 1)
    # Zhigh keeps the input height as local_z
    assert np.allclose(ref[:, 5], np.round(ldf["ldf_z0"], 4), atol=1e-3)
