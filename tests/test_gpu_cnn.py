"""The proposal network's HIP forward and backward passes (cim_cnn_forward /
cim_cnn_backward, bfloat16 MFMA products with float32 sums) against the
FP64 restatement (oracle/cnn_oracle.py, itself pinned by finite
differences); exact scoring of given candidates (cim_search_candidates_dev)
against the plain search; and a short training run that lowers the loss."""
import os
import sys

import numpy as np
import pytest

import pipelines
from conftest import rel_l2
from cryoim_amd import geometry, lib, proposals

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))


def test_cnn_forward_backward_vs_fp64_oracle(ctx):
    import torch
    import cnn_oracle
    rng = np.random.default_rng(11)
    B = 3
    imgs = rng.standard_normal((B, 128, 128)).astype(np.float32)
    imgs[:, 40:90, 30:100] += 2.0
    p = proposals.init_params(1)
    assert len(p) == cnn_oracle.layout()[1]
    x = torch.from_numpy(imgs).cuda()
    out = torch.empty((B, lib.CNN_OUT), dtype=torch.float32, device="cuda")
    ctx.cnn_forward(x, p, out)
    ref, cache = cnn_oracle.forward(imgs, p)
    refq, cacheq = cnn_oracle.forward(imgs, p, rounded=True)
    got = out.cpu().numpy()
    err, errq = rel_l2(got, ref), rel_l2(got, refq)
    print(f"forward: rel-L2 {err:.2e} vs FP64, {errq:.2e} vs FP64 with bfloat16 operands")
    # the kernels' products take bfloat16 operands: against the oracle that rounds the same operands
    # only the float32 sums differ at first (the first layer's activations agree to 1e-8), but rounding
    # to bfloat16 is discontinuous: a last-bit difference of a float32 sum now and then flips an
    # operand of the next layer by one bfloat16 step, and the flips compound over the six layers
    # (measured: 3e-4 after the sixth); against exact FP64 the rounding itself shows
    assert errq < 3e-3
    assert err < 3e-2
    dout = rng.standard_normal((B, lib.CNN_OUT)).astype(np.float32)
    g = ctx.cnn_backward(torch.from_numpy(dout).cuda())
    gref = cnn_oracle.backward(p, cache, dout.astype(np.float64))
    grefq = cnn_oracle.backward(p, cacheq, dout.astype(np.float64))
    lay, _ = cnn_oracle.layout()
    for l, (wo, shp, bo, nb) in enumerate(lay):
        n = int(np.prod(shp))
        ew, eb = rel_l2(g[wo:wo + n], grefq[wo:wo + n]), rel_l2(g[bo:bo + nb], grefq[bo:bo + nb])
        e64 = rel_l2(g[wo:wo + n], gref[wo:wo + n])
        print(f"layer {l}: dW rel-L2 {ew:.2e}, db {eb:.2e} vs bfloat16-operand FP64; dW {e64:.2e} vs FP64")
        assert ew < 3e-3 and eb < 3e-3


def test_candidates_holding_the_plain_minimum_give_the_plain_result(ctx):
    n = 12
    w = pipelines.config1_workload(n)
    d = w["d"]
    common = dict(apply_ctf=True, mask_on=True, apix=w["apix"], imgmask=d["mask"], searchstep=w["step"])
    pp = lib.make_params(w["box"], d["minR"], d["maxR"], search_method=lib.SEARCH_PLAIN, **common)
    T = len(w["tang"])
    ctx.search_templates(pp, w["timgs"], np.c_[w["tang"], np.zeros((T, 2))])
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 3:5] = w["ps"].shifts
    P[:, 6:8] = w["ps"].df
    P[:, 11] = 1.0
    import torch
    images = torch.from_numpy(np.ascontiguousarray(w["ps"].images)).cuda()
    plain = ctx.search_global(pp, images, P)
    pts = proposals.grid_points(w["step"])
    rng = np.random.default_rng(5)
    cand = rng.integers(0, len(pts), (n, 9)).astype(np.int32)
    omega = rng.integers(0, 360, (n, 9)).astype(np.int32)
    for i in range(n):
        q = int(np.nonzero(np.isclose(pts[:, 0], plain[i, 0]) & np.isclose(pts[:, 1], plain[i, 1]))[0][0])
        cand[i, 4] = q
        omega[i, 4] = int(round(plain[i, 2] / w["step"])) % 360
    got = ctx.search_candidates(pp, images, P, cand, omega)
    assert np.array_equal(got[:, [0, 1, 2, 5]], plain[:, [0, 1, 2, 5]])


def test_training_lowers_the_loss(ctx):
    w = pipelines.config1_workload(4)
    _, hist = proposals.train(ctx, w["phantom"], steps=60, batch=32, lr=1e-3, seed=3)
    first, last = hist[:10, 1].mean(), hist[-10:, 1].mean()
    print(f"rotation loss: first 10 steps {first:.3f}, last 10 {last:.3f}")
    assert np.all(np.isfinite(hist))
    assert last < 0.9 * first
