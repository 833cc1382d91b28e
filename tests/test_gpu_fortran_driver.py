"""The Fortran drop-in (cryoim_amd/bin/icosprocess_hip, ISO_C_BINDING over the
C ABI) run with icosprocess's own command lines and files, against the
outputs the reference icosprocess wrote for the same inputs (golden)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import rel_l2
from cryoim_amd import formats

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cryoim_amd", "bin", "icosprocess_hip")
G = os.path.join(ROOT, "tests", "golden", "programs.npz")


@pytest.fixture(scope="module")
def workdir():
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (amdflang missing)")
    g = np.load(G)
    d = tempfile.mkdtemp(prefix="cimf_")
    n = len(g["prog_images"])
    formats.write_mrc(os.path.join(d, "stack.mrc"), g["prog_images"])
    formats.write_mrc(os.path.join(d, "tmpl.mrc"), g["prog_timgs"])
    formats.write_template_ort(os.path.join(d, "tort.txt"), g["prog_tang"])
    parts = [formats.Particle("stack.mrc", i + 1, *g["prog_angles"][i], *g["prog_shifts"][i], 0.0, i + 1, i + 1,
                              g["prog_df"][i, 0], g["prog_df"][i, 1], 0.0) for i in range(n)]
    formats.write_ortfile(os.path.join(d, "ort.txt"), parts)
    return d, g


def _run(d, *args):
    r = subprocess.run([EXE, *args], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    return r.stdout


def test_mode1_map(workdir):
    d, g = workdir
    _run(d, "mode=1", "result3d=map.mrc", "apix=2.0", "maxRes=6.0", "imgmask=40", "ortfile=ort.txt", "last=30")
    m = formats.read_mrc(os.path.join(d, "map.mrc"))
    err = rel_l2(m, g["mode1_map"])
    print(f"mode 1 map vs reference: rel-L2 {err:.2e}")
    assert err < 1e-4


def test_mode1_realspaceavg_map(workdir):
    """reconstruct.f90:253-255: img3d_icosavg on the map before it is written."""
    d, g = workdir
    a = np.load(os.path.join(ROOT, "tests", "golden", "icosavg.npz"))
    _run(d, "mode=1", "result3d=avg.mrc", "apix=2.0", "maxRes=6.0", "imgmask=40", "ortfile=ort.txt", "last=30",
         "realspaceavg=y")
    m = formats.read_mrc(os.path.join(d, "avg.mrc"))
    err = rel_l2(m, a["avg_map"])
    print(f"mode 1 realspaceavg map vs reference: rel-L2 {err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("sym,key", [("c5a", "c5a_map"), ("c7", "c7_map")])
def test_mode1_cyclic_sym_map(workdir, sym, key):
    """sym != icos in mode 1: reconstruct_C1 with the cyclic group."""
    d, g = workdir
    sc = np.load(os.path.join(ROOT, "tests", "golden", "symc.npz"))
    _run(d, "mode=1", f"sym={sym}", f"result3d={sym}.mrc", "apix=2.0", "maxRes=6.0", "imgmask=40",
         "ortfile=ort.txt", "last=30")
    err = rel_l2(formats.read_mrc(os.path.join(d, f"{sym}.mrc")), sc[key])
    print(f"mode 1 sym={sym} map vs reference: rel-L2 {err:.2e}")
    assert err < 1e-4


def test_mode1_cyclic_sym_average_and_substep(workdir):
    """reconstruct_C1 with realspaceavg=y (average over c5a), and
    reconstruct_C1_substep (identity only, maxcentshift filter)."""
    d, g = workdir
    sc = np.load(os.path.join(ROOT, "tests", "golden", "symc.npz"))
    common = ["mode=1", "sym=c5a", "apix=2.0", "maxRes=6.0", "imgmask=40", "ortfile=ort.txt", "last=30"]
    _run(d, *common, "realspaceavg=y", "result3d=ca.mrc")
    assert rel_l2(formats.read_mrc(os.path.join(d, "ca.mrc")), sc["c5a_avg_map"]) < 1e-4
    _run(d, *common, "subrec=y", "maxcentshift=1.2", "result3d=c.tnf")
    t = formats.read_tnf(os.path.join(d, "c.tnf"))
    assert [t["fft_size"], t["max_r"], t["fsc"], t["realspaceavg"]] == list(sc["c5asub_hdr"])
    for k in ("fa", "fb", "wg"):
        assert rel_l2(t[k], sc[f"c5asub_{k}"]) < 1e-4


def test_mode1_subrec_tnf(workdir):
    d, g = workdir
    _run(d, "mode=1", "subrec=y", "result3d=sub.tnf", "apix=2.0", "maxRes=6.0", "imgmask=40", "ortfile=ort.txt",
         "last=30")
    t = formats.read_tnf(os.path.join(d, "sub.tnf"))
    assert [t["fft_size"], t["max_r"]] == list(g["mode1sub_hdr"])
    for k in ("fa", "fb", "wg"):
        assert rel_l2(t[k], g[f"mode1sub_{k}"]) < 1e-4


@pytest.mark.parametrize("mode", [4, 5])
def test_search_modes(workdir, mode):
    d, g = workdir
    extra = ["searchstep=6.0"] if mode == 4 else ["Ncycle=3"]
    _run(d, f"mode={mode}", "templateort=tort.txt", "templatestck=tmpl.mrc", f"newortfile=m{mode}.txt",
         "apix=2.0", "maxRes=6.0", "minRes=100", "imgmask=40", "ortfile=ort.txt", "last=6", *extra)
    out = formats.read_ortfile900(os.path.join(d, f"m{mode}.txt"))
    ref = g[f"mode{mode}_out"]
    assert len(out) == len(ref)
    for o, r in zip(out, ref):
        got = np.array([o.theta, o.phi, o.omega, o.x, o.y])
        assert np.abs(got - r[:5]).max() < 2e-3, (got, r)     # format 900 keeps 4 decimals
        assert abs(o.pr - r[5]) < 2e-3


def test_mode0_template_stack(workdir):
    """mode 0 (project2d.f90): the template stack projected from model3d."""
    d, g = workdir
    pj = np.load(os.path.join(ROOT, "tests", "golden", "proj.npz"))
    formats.write_mrc(os.path.join(d, "model.mrc"), pj["proj_model"])
    _run(d, "mode=0", "model3d=model.mrc", "templateort=tort.txt", "templatestck=tstk.mrc")
    st = formats.read_mrc(os.path.join(d, "tstk.mrc"))
    assert st.shape == pj["mode0_stack"].shape
    err = rel_l2(st, pj["mode0_stack"])
    print(f"mode 0 template stack vs reference: rel-L2 {err:.2e}")
    assert err < 1e-6


def test_mode4_with_model_templates(workdir):
    """mode 4 without templatestck: load_template projects the templates from model3d."""
    d, g = workdir
    pj = np.load(os.path.join(ROOT, "tests", "golden", "proj.npz"))
    formats.write_mrc(os.path.join(d, "model.mrc"), pj["proj_model"])
    _run(d, "mode=4", "templateort=tort.txt", "model3d=model.mrc", "newortfile=m4model.txt", "searchstep=6.0",
         "apix=2.0", "maxRes=6.0", "minRes=100", "imgmask=40", "ortfile=ort.txt", "last=6")
    out = formats.read_ortfile900(os.path.join(d, "m4model.txt"))
    ref = pj["mode4_model_out"]
    assert len(out) == len(ref)
    for o, r in zip(out, ref):
        got = np.array([o.theta, o.phi, o.omega, o.x, o.y])
        assert np.abs(got - r[:5]).max() < 2e-3, (got, r)
        assert abs(o.pr - r[5]) < 2e-3


@pytest.fixture(scope="module")
def local_ort(workdir):
    d, g = workdir
    loc = np.load(os.path.join(ROOT, "tests", "golden", "local.npz"))
    n = len(g["prog_images"])
    parts = [formats.Particle("stack.mrc", i + 1, *g["prog_angles"][i], *g["prog_shifts"][i], 0.0, i + 1, i + 1,
                              g["prog_df"][i, 0], g["prog_df"][i, 1], 0.0, x0=float(loc["loc_x0y0"][i, 0]),
                              y0=float(loc["loc_x0y0"][i, 1])) for i in range(n)]
    formats.write_ortfile(os.path.join(d, "ort6.txt"), parts, with_local_box=True)
    return d, loc


LOCAL_ARGS = ["mode=6", "loc_box=32", "sym=c5a", "apix=2.0", "maxRes=6.0", "imgmask=24", "ortfile=ort6.txt",
              "last=30"]


def test_mode6_local_map(local_ort):
    """reconstruction_local.f90: loc_box^3 map with the c5a local operators."""
    d, loc = local_ort
    _run(d, *LOCAL_ARGS, "result3d=loc.mrc")
    raw = open(os.path.join(d, "loc.mrc"), "rb").read()
    assert np.array_equal(np.frombuffer(raw[:1024], "<i4"), loc["loc_hdr"])
    m = np.frombuffer(raw[1024:], "<f4").reshape(loc["loc_map"].shape)
    err = rel_l2(m, loc["loc_map"])
    print(f"mode 6 map vs reference: rel-L2 {err:.2e}")
    assert err < 1e-4


def test_mode6_local_substep_tnf(local_ort):
    """reconstruction_local_substep.f90: compact tnf with its (loc_box, maxR, apix, imgmask) header."""
    d, loc = local_ort
    _run(d, *LOCAL_ARGS, "subrec=y", "result3d=loc.tnf")
    t = formats.read_tnf(os.path.join(d, "loc.tnf"))
    assert [t["fft_size"], t["max_r"], t["fsc"], t["realspaceavg"]] == list(loc["locsub_hdr"])
    assert np.allclose([t["apix"], t["imgmask"]], loc["locsub_hdrf"])
    for k in ("fa", "fb", "wg"):
        assert rel_l2(t[k], loc[f"locsub_{k}"]) < 1e-4


def test_mode6_rejects_unknown_symmetry(local_ort):
    d, _ = local_ort
    r = subprocess.run([EXE, *[a if not a.startswith("sym=") else "sym=c7" for a in LOCAL_ARGS], "result3d=x.mrc"],
                       cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode != 0 and "local symmetry" in r.stdout


@pytest.fixture(scope="module")
def center_dir():
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (amdflang missing)")
    c = np.load(os.path.join(ROOT, "tests", "golden", "center.npz"))
    d = tempfile.mkdtemp(prefix="cimc_")
    n = len(c["cen_images"])
    formats.write_mrc(os.path.join(d, "stack.mrc"), c["cen_images"])
    formats.write_mrc(os.path.join(d, "model.mrc"), c["cen_model"])
    parts = [formats.Particle("stack.mrc", i + 1, *c["cen_angles"][i], *c["cen_guess"][i], 0.0, i + 1, i + 1,
                              c["cen_df"][i, 0], c["cen_df"][i, 1], 0.0) for i in range(n)]
    formats.write_ortfile(os.path.join(d, "ort.txt"), parts)
    return d, c


@pytest.mark.parametrize("with_model", [True, False])
def test_mode3_center(center_dir, with_model):
    """calc_center (model3d=) / calc_center_selfrot180, then output_ort."""
    d, c = center_dir
    key = "cen_model_out" if with_model else "cen_self_out"
    extra = ["model3d=model.mrc"] if with_model else []
    out = _run(d, "mode=3", "apix=2.0", "maxRes=6.0", "imgmask=20", "ortfile=ort.txt", "newortfile=c.txt",
               f"last={len(c['cen_images'])}", *extra)
    rows = formats.read_ortfile900(os.path.join(d, "c.txt"))
    got = np.array([[p.x, p.y, p.pr] for p in rows])
    ref = c[key + "_ort"]
    assert np.abs(got[:, :2] - ref[:, :2]).max() < 2e-3, (got, ref)
    assert np.abs(got[:, 2] - ref[:, 2]).max() < 2e-2          # PR = acos(cc): d/dcc ~ 2.3 rad near cc 0.9
    printed = np.array([[float(v) for v in line.split()[1:]] for line in out.splitlines() if len(line.split()) == 4])
    assert np.abs(printed - c[key]).max() < 2e-4
