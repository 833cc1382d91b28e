"""Host arithmetic of the learned proposals (cryoim_amd/proposals.py): the
targets are the 60 symmetry-equivalent views, the AU angles of a rotation
are the angles it was made from (for AU views), the loss gradient matches
finite differences, the grid order is the search grid's, and the cone
candidates hold the grid point nearest to a hypothesis."""
import numpy as np

from cryoim_amd import geometry, proposals, synth


def _au_angles(n, seed):
    return synth.random_au_orientations(n, np.random.default_rng(seed))


def test_au_euler_recovers_au_angles():
    a = _au_angles(64, 1)
    R = geometry.euler_mat(a[:, 0], a[:, 1], a[:, 2])
    b = proposals.au_euler(R)
    assert np.all(geometry.sym_distance_deg(a, b) < 1e-4)
    # an equivalent view g R maps back into the AU too
    g = geometry.icos_ops()[17]
    c = proposals.au_euler(np.einsum("ij,njk->nik", g, R))
    assert np.all(geometry.sym_distance_deg(a, c) < 1e-4)


def test_targets_are_equivalent_views():
    a = _au_angles(5, 2)
    t = proposals.targets(a)
    assert t.shape == (5, 60, 6)
    R = geometry.euler_mat(a[:, 0], a[:, 1], a[:, 2])
    # operator 0 of the table is the identity: the first target is R itself
    assert np.allclose(geometry.icos_ops()[0], np.eye(3))
    assert np.allclose(t[:, 0, 0:3], R[:, :, 0]) and np.allclose(t[:, 0, 3:6], R[:, :, 1])
    # every target completes to a rotation that shows the same view
    for k in (5, 33, 59):
        b1, b2 = t[:, k, 0:3], t[:, k, 3:6]
        M = np.stack([b1, b2, np.cross(b1, b2)], axis=-1)
        assert np.all(geometry.sym_distance_deg(a, proposals.au_euler(M)) < 1e-4)


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    out = rng.standard_normal((6, proposals.OUT))
    tgt = proposals.targets(_au_angles(6, 4))
    loss, g, _ = proposals.loss_and_grad(out, tgt)
    h = 1e-6
    for i, j in [(0, 0), (1, 4), (2, 6), (3, 13), (4, 20), (5, 27)]:
        o = out.copy()
        o[i, j] += h
        lp = proposals.loss_and_grad(o, tgt)[0]
        o[i, j] -= 2 * h
        lm = proposals.loss_and_grad(o, tgt)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i, j]) <= 1e-5 * max(1.0, abs(fd)), (i, j, fd, g[i, j])


def test_grid_order_and_cones():
    step = 1.0
    pts = proposals.grid_points(step)
    half = geometry.au_grid_points(step)
    assert len(pts) == sum(1 if ip == 0 else 2 for _, ip in half)
    a = _au_angles(8, 5)
    ang = np.repeat(a[:, None, :], proposals.HYP, axis=1)
    conf = np.tile([0.9, 0.1, 0.1, 0.1], (8, 1))
    cand, omega = proposals.cone_candidates(ang, conf, step, cone=3.0, min_conf=0.5)
    d = proposals._dirs(pts[:, 0], pts[:, 1])
    for i in range(8):
        nearest = int(np.argmax(d @ proposals._dirs(a[i, 0], a[i, 1])))
        assert nearest in set(cand[i].tolist())
        assert omega[i, 0] == int(round(a[i, 2] / step)) % 360
