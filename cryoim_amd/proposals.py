"""Learned proposals for the global search (icosprocess mode 4): a network
(include/cryoim.h, cim_cnn_*) maps a 128 x 128 particle image to
CNN_HYP hypotheses of its orientation, each with a confidence; the exact
score (cim_search_candidates_dev, the plain search's arithmetic) then runs
only over the AU grid points within a cone around each confident
hypothesis, over an omega window around its in-plane angle.  Particles whose
best confidence is below a threshold take the tabulated search instead.

The reference has no counterpart: its search scores every grid orientation
(ort_globesearch.f90:68-106).  The network's forward and backward passes run
in the library's HIP kernels; the loss, its gradient with respect to the
network's outputs and the Adam update are host arithmetic on small arrays.

Targets: the rotation's first two columns.  The model is invariant under
the 60 icosahedral operators g and a particle shows R^T P of each model
point P, so R and g R are the same view (g R turns the view direction, the
third column, by g): the loss of
a particle is the smallest squared distance between any hypothesis and any
of the 60 equivalent targets (the hypotheses compete: only the winner is
pulled towards its target), plus a logistic loss that teaches each
hypothesis's confidence whether it won.
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

from . import geometry, lib, synth

IMG = 128
HYP, OUT = lib.CNN_HYP, lib.CNN_OUT
CONF_WEIGHT = 0.1


def param_layout():
    """[(w_offset, w_shape, b_offset, nb)] per layer then the head, and the total
    (the layout of cim_cnn_forward's parameters)."""
    out, off, cin = [], 0, 1
    for _ in range(6):
        shp = (3, 3, cin, 64)
        out.append((off, shp, off + int(np.prod(shp)), 64))
        off += int(np.prod(shp)) + 64
        cin = 64
    out.append((off, (256, OUT), off + 256 * OUT, OUT))
    return out, off + 256 * OUT + OUT


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases."""
    lay, total = param_layout()
    rng = np.random.default_rng(seed)
    p = np.zeros(total, np.float32)
    for wo, shp, _, _ in lay:
        n = int(np.prod(shp))
        p[wo:wo + n] = (rng.standard_normal(n) * np.sqrt(2.0 / np.prod(shp[:-1]))).astype(np.float32)
    return p


def targets(angles: np.ndarray) -> np.ndarray:
    """(n, 60, 6): the first two columns (column-major) of g R for the 60
    icosahedral operators g, R = euler_mat of each particle's angles."""
    a = np.asarray(angles, np.float64).reshape(-1, 3)
    R = geometry.euler_mat(a[:, 0], a[:, 1], a[:, 2])
    Rg = np.einsum("gij,njk->ngik", geometry.icos_ops(), R)
    return np.concatenate([Rg[..., :, 0], Rg[..., :, 1]], axis=-1)


def loss_and_grad(out: np.ndarray, tgt: np.ndarray):
    """Mean loss over the batch and its gradient with respect to the network
    outputs (n, OUT)."""
    n = len(out)
    o = out.reshape(n, HYP, 7).astype(np.float64)
    rot, conf = o[..., :6], o[..., 6]
    d2 = ((rot[:, :, None, :] - tgt[:, None, :, :]) ** 2).sum(-1)           # (n, HYP, 60)
    flat = d2.reshape(n, -1).argmin(axis=1)
    h, g = flat // tgt.shape[1], flat % tgt.shape[1]
    idx = np.arange(n)
    grad = np.zeros((n, HYP, 7))
    grad[idx, h, :6] = 2.0 * (rot[idx, h] - tgt[idx, g]) / n
    label = np.zeros((n, HYP))
    label[idx, h] = 1.0
    p = 1.0 / (1.0 + np.exp(-conf))
    eps = 1e-12
    bce = -(label * np.log(p + eps) + (1 - label) * np.log(1 - p + eps)).sum(axis=1)
    grad[..., 6] = CONF_WEIGHT * (p - label) / n
    loss = float(np.mean(d2[idx, h, g]) + CONF_WEIGHT * np.mean(bce))
    return loss, grad.reshape(n, OUT).astype(np.float32), float(np.mean(d2[idx, h, g]))


def rotations(out: np.ndarray):
    """Hypotheses as rotations (n, HYP, 3, 3) by Gram-Schmidt of the two
    predicted columns, and their confidences (n, HYP) in (0, 1)."""
    o = np.asarray(out, np.float64).reshape(-1, HYP, 7)
    a1, a2 = o[..., 0:3], o[..., 3:6]
    b1 = a1 / np.maximum(np.linalg.norm(a1, axis=-1, keepdims=True), 1e-12)
    c2 = a2 - (b1 * a2).sum(-1, keepdims=True) * b1
    b2 = c2 / np.maximum(np.linalg.norm(c2, axis=-1, keepdims=True), 1e-12)
    b3 = np.cross(b1, b2)
    return np.stack([b1, b2, b3], axis=-1), 1.0 / (1.0 + np.exp(-o[..., 6]))


def au_euler(R: np.ndarray) -> np.ndarray:
    """(theta, phi, omega) in degrees of the equivalent g R whose view
    direction (third column) lies in the asymmetric unit, for R (..., 3, 3)."""
    shp = R.shape[:-2]
    Rf = R.reshape(-1, 3, 3)
    ops = geometry.icos_ops()
    out = np.zeros((len(Rf), 3))
    found = np.zeros(len(Rf), bool)
    for g in ops:
        M = g @ Rf
        v = M[:, :, 2]
        ok = geometry.in_au(v[:, 0], v[:, 1], v[:, 2]) & ~found
        if ok.any():
            th = np.degrees(np.arccos(np.clip(M[ok, 2, 2], -1, 1)))
            ph = np.degrees(np.arctan2(M[ok, 1, 2], M[ok, 0, 2]))
            om = np.degrees(np.arctan2(M[ok, 2, 1], -M[ok, 2, 0])) % 360.0
            out[ok] = np.c_[th, ph, om]
            found |= ok
    return out.reshape(*shp, 3)


def grid_points(step: float) -> np.ndarray:
    """(npts, 2) theta, phi (degrees) of the search grid's AU points in the
    library's order: theta index outer; per phi index ip >= 0 the point at
    +ip, then at -ip when ip > 0."""
    pts = []
    for it, ip in geometry.au_grid_points(step):
        th = 69.0 + step * it
        pts.append((th, step * ip))
        if ip > 0:
            pts.append((th, -step * ip))
    return np.array(pts)


def _dirs(theta, phi):
    t, p = np.radians(theta), np.radians(phi)
    return np.stack([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)], axis=-1)


def cone_candidates(angles: np.ndarray, conf: np.ndarray, step: float, cone: float = 3.0, min_conf: float = 0.0,
                    kmax: int = 128):
    """Candidates of each particle from its hypotheses (angles (n, HYP, 3) in
    the AU, conf (n, HYP)): the grid points whose view lies within cone
    degrees of a hypothesis with confidence >= min_conf (most confident
    first), each with the omega column of that hypothesis.  Returns
    (cand, omega) of shape (n, K) with K <= kmax, rows padded by repeating
    their first entry."""
    pts = grid_points(step)
    pd = _dirs(pts[:, 0], pts[:, 1])
    No = int(round(360.0 / step))
    cos_c = np.cos(np.radians(cone))
    rows = []
    for i in range(len(angles)):
        seen, c, o = set(), [], []
        for h in np.argsort(-conf[i]):
            if conf[i, h] < min_conf:
                continue
            d = _dirs(angles[i, h, 0], angles[i, h, 1])
            near = np.nonzero(pd @ d >= cos_c)[0]
            near = near[np.argsort(-(pd[near] @ d))]
            k = int(round(angles[i, h, 2] / step)) % No
            for q in near:
                if q not in seen and len(c) < kmax:
                    seen.add(q)
                    c.append(int(q))
                    o.append(k)
        if not c:   # no confident hypothesis: the nearest point to the first one
            h = int(np.argmax(conf[i]))
            d = _dirs(angles[i, h, 0], angles[i, h, 1])
            c, o = [int(np.argmax(pd @ d))], [int(round(angles[i, h, 2] / step)) % No]
        rows.append((c, o))
    K = max(len(c) for c, _ in rows)
    cand = np.array([c + [c[0]] * (K - len(c)) for c, _ in rows], np.int32)
    omega = np.array([o + [o[0]] * (K - len(o)) for _, o in rows], np.int32)
    return cand, omega


def normalise(images):
    """Per-image zero mean and unit variance (torch tensors on the device)."""
    m = images.mean(dim=(1, 2), keepdim=True)
    s = images.std(dim=(1, 2), keepdim=True).clamp_min(1e-6)
    return ((images - m) / s).contiguous()


class Adam:
    def __init__(self, n, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.m, self.v, self.t = np.zeros(n), np.zeros(n), 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, p: np.ndarray, g: np.ndarray):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        p -= (self.lr * mh / (np.sqrt(vh) + self.eps)).astype(np.float32)


@dataclasses.dataclass
class Recipe:
    """Renders of the config-1 map: box 128, 1.5 A/px, defocus U(1, 3) um,
    shifts +-3 px, SNR 0.1 (tests/pipelines.config1_workload's recipe)."""
    box: int = 128
    apix: float = 1.5
    snr: float = 0.1
    df_range: tuple = (1.0, 3.0)
    shift_max: float = 3.0


def render(ph, n, seed, recipe: Recipe = Recipe()):
    return synth.make_particles_torch(ph, n, seed=seed, shift_max=recipe.shift_max, df_range=recipe.df_range,
                                      snr=recipe.snr, apix=recipe.apix, mask_radius=recipe.box / 2 - 4)


def train(ctx: "lib.Context", ph, steps: int, batch: int = 64, lr: float = 1e-3, seed: int = 0, params=None,
          recipe: Recipe = Recipe(), log=None, lr_decay_at=None):
    """Adam on fresh renders every step (seeds seed * 1_000_000 + step);
    returns the parameters and the per-step (loss, rotation loss)."""
    import torch
    p = init_params(seed) if params is None else np.array(params, np.float32)
    opt = Adam(len(p), lr=lr)
    hist = []
    out = torch.empty((batch, OUT), dtype=torch.float32, device="cuda:0")
    for k in range(steps):
        if lr_decay_at and k in lr_decay_at:
            opt.lr *= 0.3
        ps = render(ph, batch, seed * 1_000_000 + k, recipe)
        x = normalise(ps.images)
        ctx.cnn_forward(x, p, out)
        loss, dout, lrot = loss_and_grad(out.cpu().numpy(), targets(ps.angles))
        g = ctx.cnn_backward(torch.from_numpy(dout).to("cuda:0"))
        opt.step(p, g.astype(np.float64))
        hist.append((loss, lrot))
        if log and (k % 50 == 0 or k == steps - 1):
            log(f"step {k}: loss {loss:.4f} rotation {lrot:.4f}")
    return p, np.array(hist)


def propose(ctx: "lib.Context", params, images, chunk: int = 512):
    """Hypotheses of device images (n, 128, 128): AU angles (n, HYP, 3) and
    confidences (n, HYP)."""
    import torch
    n = images.shape[0]
    outs = []
    for i0 in range(0, n, chunk):
        x = normalise(images[i0:i0 + chunk])
        out = torch.empty((x.shape[0], OUT), dtype=torch.float32, device=x.device)
        ctx.cnn_forward(x, params, out)
        outs.append(out.cpu().numpy())
    R, conf = rotations(np.concatenate(outs))
    return au_euler(R), conf


def config1_phantom():
    """The config-1 map: box 128, 20 blobs (sigma U(2, 4) px, radius U(40, 55)
    px) expanded by the 60 operators, seed 101 (tests/pipelines.config1_workload)."""
    return synth.make_phantom(128, nblobs=20, sigma=(2.0, 4.0), radius=(40.0, 55.0), seed=101)


DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
WEIGHTS_C1 = os.path.join(DATA, "proposal_c1.npz")


def load_weights(path: str = WEIGHTS_C1) -> np.ndarray:
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32)


def main(argv=None):
    """python -m cryoim_amd.proposals --out cryoim_amd/data/proposal_c1.npz
    [--steps 5000 --batch 128 --lr 1e-3 --seed 1]: train on renders of the
    config-1 map and write the parameters."""
    import argparse
    import time
    import torch
    torch.cuda.init()       # torch's HIP runtime reaches the GPU before the library's
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=WEIGHTS_C1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    ctx = lib.Context(0)
    t0 = time.perf_counter()
    p, hist = train(ctx, config1_phantom(), a.steps, batch=a.batch, lr=a.lr, seed=a.seed,
                    log=lambda s: print(s, flush=True), lr_decay_at=(int(a.steps * 0.6), int(a.steps * 0.85)))
    secs = time.perf_counter() - t0
    np.savez(a.out, params=p, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed, train_seconds=secs,
             final_rotation_loss=float(hist[-100:, 1].mean()))
    print(f"{a.steps} steps of {a.batch} renders in {secs:.1f} s; rotation loss {hist[-100:, 1].mean():.4f} -> {a.out}")
    ctx.close()


if __name__ == "__main__":
    main()
