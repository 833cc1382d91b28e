"""The search + insert step captured as one HIP graph (cim_step_capture /
cim_step_replay) against the same step run call by call: the searched
orientations and PR of every replay equal the eager ones bit for bit, with
eager work and synchronisations between the replays; the map after three
replays equals the map of three eager inserts; a replay after the images
are replaced in place gives the eager result for the new images."""
import numpy as np
import pytest

import pipelines
from conftest import rel_l2
from cryoim_amd import lib

pytestmark = pytest.mark.gpu
COMPENSATED = 16


def _setup(w):
    d = w["d"]
    common = dict(apply_ctf=True, mask_on=True, apix=w["apix"], imgmask=d["mask"], searchstep=w["step"])
    ps_ = lib.make_params(w["box"], d["minR"], d["maxR"], search_method=lib.SEARCH_TABLES, **common)
    pi_ = lib.make_params(w["box"], 0, w["insert_maxR"], apply_ctf=True, mask_on=True, norm_on=True, apix=w["apix"],
                          imgmask=d["mask"])
    return ps_, pi_


def _records(ps):
    n = len(ps.images)
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 3:5] = ps.shifts
    P[:, 6:8] = ps.df
    P[:, 9:11] = ps.shifts[0]
    P[:, 11] = 1.0
    return P


def test_step_graph_replays_equal_eager():
    import torch
    n = 64
    w = pipelines.config1_workload(n)
    w2 = pipelines.config1_workload(2 * n)
    ps_, pi_ = _setup(w)
    T = len(w["tang"])
    tort = np.c_[w["tang"], np.zeros((T, 2))]
    images = torch.from_numpy(np.ascontiguousarray(w["ps"].images)).cuda()
    other = torch.from_numpy(np.ascontiguousarray(w2["ps"].images[n:])).cuda()
    P0 = _records(w["ps"])
    P1 = _records(w2["ps"])[n:]
    P1[:, 9:11] = P0[0, 9:11]
    box, maxR = w["box"], w["insert_maxR"]

    # eager reference: the search, then three inserts at its orientations
    ce = lib.Context(0)
    ce.search_templates(ps_, w["timgs"], tort)
    out_e = ce.search_global(ps_, images, P0)
    out_e2 = ce.search_global(ps_, other, P1)
    Q = P0.copy()
    Q[:, 0:3] = out_e[:, 0:3]
    ce.rec_begin(box, maxR, layout=COMPENSATED, rcut=maxR, wg_init=0.01)
    for _ in range(3):
        ce.rec_insert(pi_, images, Q)
    map_e = ce.rec_finish(box)
    ce.close()

    cg = lib.Context(0)
    cg.search_templates(ps_, w["timgs"], tort)
    cg.search_global(ps_, images, P0)          # sets up the search's constant tables for n
    cg.rec_begin(box, maxR, layout=COMPENSATED, rcut=maxR, wg_init=0.01)
    parts = torch.from_numpy(P0.copy()).cuda()
    cg.step_capture(ps_, pi_, images, parts)
    for rep in range(3):
        parts.copy_(torch.from_numpy(P0))
        cg.step_replay()
        got = parts.cpu().numpy()
        same = np.array_equal(got[:, [0, 1, 2, 5]], out_e[:, [0, 1, 2, 5]])
        print(f"replay {rep}: {int(np.all(got[:, [0, 1, 2, 5]] == out_e[:, [0, 1, 2, 5]], axis=1).sum())}/{n} "
              f"identical to the eager search", flush=True)
        assert same
        assert np.array_equal(got[:, 3:5], P0[:, 3:5])
        # eager work and synchronisation between the replays
        x = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        del x
    map_g = cg.rec_finish(box)
    err = rel_l2(map_g, map_e)
    print(f"map after 3 replays vs 3 eager inserts: rel-L2 {err:.2e}, bit-identical {np.array_equal(map_g, map_e)}")
    assert err < 1e-6
    # new images in the same buffers: the replay searches them
    images.copy_(other)
    parts.copy_(torch.from_numpy(P1))
    cg.step_replay()
    got = parts.cpu().numpy()
    assert np.array_equal(got[:, [0, 1, 2, 5]], out_e2[:, [0, 1, 2, 5]])
    cg.step_release()
    cg.close()


def test_step_capture_needs_set_up_search():
    import torch
    n = 16
    w = pipelines.config1_workload(n)
    ps_, pi_ = _setup(w)
    T = len(w["tang"])
    c = lib.Context(0)
    c.search_templates(ps_, w["timgs"], np.c_[w["tang"], np.zeros((T, 2))])
    c.rec_begin(w["box"], w["insert_maxR"], layout=COMPENSATED, rcut=w["insert_maxR"], wg_init=0.01)
    images = torch.from_numpy(np.ascontiguousarray(w["ps"].images)).cuda()
    parts = torch.from_numpy(_records(w["ps"])).cuda()
    with pytest.raises(RuntimeError, match="host records"):
        c.step_capture(ps_, pi_, images, parts)
    with pytest.raises(RuntimeError, match="cim_step_capture"):
        c.step_replay()
    c.close()
