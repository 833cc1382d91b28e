"""Search and map parity at the measured configurations' scale, against the
FP64 oracle and the reference's own decisions (tests/golden/scale.npz,
written by tests/golden/make_scale.py in a container holding the reference).

* box 128, 1 deg grid x 360 omega, maxR 24, T = 4 (the config-1 recipe of
  tests/pipelines.config1_workload): the plain and the tabulated GPU search
  of every particle against the FP64 oracle's ort_globesearch
  (ort_globesearch.f90:34-134) and the reference's mode 4; then the 60-fold
  insert at Nyquist-2 (maxR 62) and the map (reconstruct.f90:53-247) at the
  oracle's orientations against the FP64 oracle's map: rel-L2 and FSC.
* box 384, 0.5 deg grid x 720 omega, maxR 72 (tests/pipelines.box384_workload):
  both GPU searches against the FP64 oracle's decisions.

The gate: at least 99.5% of orientations within one grid step of the
oracle's, map rel-L2 <= 1e-4 and FSC >= 0.999 in every shell to Nyquist-2."""
import numpy as np
import pytest

from conftest import rel_l2
from cryoim_amd import lib
import pipelines

pytestmark = pytest.mark.gpu

COMPENSATED = 16


@pytest.fixture(scope="module")
def scale():
    return np.load(pipelines.golden_path("scale.npz"))


def _records(ps):
    n = len(ps.images)
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 3:5] = ps.shifts
    P[:, 6:8] = ps.df
    P[:, 11] = 1.0
    return P


def _check_images(w, g, key):
    im = w["ps"].images.astype(np.float64)
    assert abs(im.sum() - float(g[key + "_img_sum"])) <= 1e-5 * float(g[key + "_img_norm"])
    assert abs(np.linalg.norm(im) / float(g[key + "_img_norm"]) - 1.0) < 1e-9


def _searches(ctx, w, batch_env=None):
    d = w["d"]
    common = dict(apply_ctf=True, mask_on=True, apix=w["apix"], imgmask=d["mask"], searchstep=w["step"])
    out = {}
    for name, method in (("plain", lib.SEARCH_PLAIN), ("tables", lib.SEARCH_TABLES)):
        p = lib.make_params(w["box"], d["minR"], d["maxR"], search_method=method, **common)
        T = len(w["tang"])
        ctx.search_templates(p, w["timgs"], np.c_[w["tang"], np.zeros((T, 2))])
        out[name] = ctx.search_global(p, w["ps"].images, _records(w["ps"]))
    return out


def _report(name, got, ref, step, what):
    s = pipelines.angle_steps(got[:, :3], ref[:, :3], step)
    within = float(np.mean(s <= 1.0 + 1e-4))
    same = int(np.sum(s < 1e-4))
    dpr = np.abs(got[:, 5] - ref[:, 3 if ref.shape[1] == 4 else 5])[s < 1e-4]
    print(f"{name} vs {what}: {same}/{len(s)} identical orientations, {100 * within:.2f}% within one step, "
          f"PR |diff| on identical median {np.median(dpr) if len(dpr) else 0:.2e} max {dpr.max() if len(dpr) else 0:.2e}"
          f" deg; worst {s.max():.1f} steps", flush=True)
    return within


def test_config1_search_and_map(ctx, O64, scale):
    n = int(scale["c1_n"])
    w = pipelines.config1_workload(n)
    _check_images(w, scale, "c1")
    print(f"config 1: {n} particles, box 128, 1 deg, maxR {w['d']['maxR']}", flush=True)
    out = _searches(ctx, w)
    o64 = scale["c1_o64"]
    for name in ("plain", "tables"):
        assert _report(name, out[name], o64, 1.0, "FP64 oracle") >= 0.995
        if "c1_ref" in scale:
            _report(name, out[name], scale["c1_ref"], 1.0, "reference mode 4")
    if "c1_ref" in scale:
        _report("FP64 oracle", np.c_[o64[:, :3], np.zeros((n, 2)), o64[:, 3]], scale["c1_ref"], 1.0,
                "reference mode 4")
    # the map at the oracle's orientations: 60-fold insert at maxR 62 (the coset insert, compensated)
    maxR = w["insert_maxR"]
    ps, d = w["ps"], w["d"]
    ang = o64[:, :3].astype(np.float32)
    p = lib.make_params(w["box"], 0, maxR, apply_ctf=True, mask_on=True, norm_on=True, apix=w["apix"],
                        imgmask=d["mask"])
    P = _records(ps)
    P[:, 0:3] = ang
    P[:, 9:11] = ps.shifts[0]          # reconstruct.f90:92 masks around the first particle's shift
    ctx.rec_begin(w["box"], maxR, layout=COMPENSATED, rcut=maxR, wg_init=0.01)
    ctx.rec_insert(p, ps.images, P)
    gmap = ctx.rec_finish(w["box"])
    print("FP64 oracle map ...", flush=True)
    vol = pipelines.mode1_volume(O64, ps.images, ang, ps.shifts, ps.df, w["apix"], maxR, d["mask"])
    omap = O64.fold_and_map(vol, maxR)
    err = rel_l2(gmap, omap)
    fsc = pipelines.shell_fsc(gmap, omap, maxR)
    print(f"config 1 map: rel-L2 {err:.2e} vs FP64 oracle; FSC min {fsc[1:].min():.7f} over shells 1..{maxR}")
    assert err <= 1e-4
    assert fsc[1:].min() >= 0.999


def test_box384_search_decisions(ctx, scale):
    n = int(scale["b384_n"])
    w = pipelines.box384_workload(n)
    _check_images(w, scale, "b384")
    print(f"box 384: {n} particles, 0.5 deg, maxR {w['d']['maxR']}", flush=True)
    out = _searches(ctx, w)
    o64 = scale["b384_o64"]
    for name in ("plain", "tables"):
        assert _report(name, out[name], o64, 0.5, "FP64 oracle") == 1.0
