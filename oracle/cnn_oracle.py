"""FP64 restatement of the proposal network (cryoim_amd/csrc/cnn.hip) for the
tests: forward and backward of six 3x3 convolutions (64 channels, stride 2,
padding 1, ReLU) on 128 x 128 single-channel images, then the linear head of
28 outputs.  Test infrastructure only (the checker of the GPU kernels).

The network has no counterpart in the reference (its global search scores
every orientation of the grid, ort_globesearch.f90:68-106), so it is pinned
by what the mathematics fixes: the convolution against a direct sum over
taps, and the backward pass against finite differences of the forward pass
(tests/test_oracle_cnn.py).

Parameter layout (as cim_cnn_nparams / cim_cnn_forward): per layer
W[ky][kx][cin][cout] then the 64 biases; then the head W[256][28] (features
in (y, x, channel) order of the 2 x 2 x 64 output) and its 28 biases.
"""
from __future__ import annotations

import numpy as np

IMG, CH, LAYERS, OUT = 128, 64, 6, 28
FEAT = 2 * 2 * CH


def layout():
    """[(w_offset, w_shape, b_offset, nb)] per layer, the head last, and the total."""
    out, off, cin, = [], 0, 1
    for _ in range(LAYERS):
        shp = (3, 3, cin, CH)
        out.append((off, shp, off + int(np.prod(shp)), CH))
        off += int(np.prod(shp)) + CH
        cin = CH
    out.append((off, (FEAT, OUT), off + FEAT * OUT, OUT))
    off += FEAT * OUT + OUT
    return out, off


def bf16(x):
    """x rounded to bfloat16 (through float32, to nearest, ties to even), as
    float64: the operands of the kernels' matrix products."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _q(x, on):
    return bf16(x) if on else x


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases (float32)."""
    lay, total = layout()
    rng = np.random.default_rng(seed)
    p = np.zeros(total, np.float32)
    for wo, shp, bo, nb in lay:
        fan_in = int(np.prod(shp[:-1]))
        p[wo:wo + int(np.prod(shp))] = (rng.standard_normal(int(np.prod(shp))) * np.sqrt(2.0 / fan_in)).astype(
            np.float32)
    return p


def _im2col(x):
    """x (B, H, H, C) -> cols (B, H/2, H/2, 9, C): 3x3 patches, stride 2, padding 1."""
    B, H, _, C = x.shape
    Ho = H // 2
    xp = np.zeros((B, H + 2, H + 2, C), x.dtype)
    xp[:, 1:H + 1, 1:H + 1] = x
    cols = np.empty((B, Ho, Ho, 9, C), x.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky * 3 + kx] = xp[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Ho:2]
    return cols


def _col2im(dcols, H):
    """Adjoint of _im2col."""
    B, Ho, _, _, C = dcols.shape
    dxp = np.zeros((B, H + 2, H + 2, C), dcols.dtype)
    for ky in range(3):
        for kx in range(3):
            dxp[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Ho:2] += dcols[:, :, :, ky * 3 + kx]
    return dxp[:, 1:H + 1, 1:H + 1]


def forward(images, params, rounded: bool = False):
    """images (B, 128, 128), params (float) -> outputs (B, 28) float64 and the
    cache.  rounded: every matrix product's operands rounded to bfloat16 (the
    kernels' products; sums stay float64)."""
    lay, _ = layout()
    p = np.asarray(params, np.float64)
    x = np.asarray(images, np.float64)[..., None]
    cache = []
    for l in range(LAYERS):
        wo, shp, bo, nb = lay[l]
        W = p[wo:wo + int(np.prod(shp))].reshape(9 * shp[2], CH)
        cols = _im2col(x)
        y = np.maximum(_q(cols.reshape(*cols.shape[:3], -1), rounded) @ _q(W, rounded) + p[bo:bo + nb], 0.0)
        cache.append((x, y))
        x = y
    wo, shp, bo, nb = lay[LAYERS]
    feat = x.reshape(len(x), -1)
    out = _q(feat, rounded) @ _q(p[wo:wo + FEAT * OUT].reshape(FEAT, OUT), rounded) + p[bo:bo + nb]
    return out, (cache, feat, rounded)


def backward(params, cache, dout):
    """Gradient (float64, the layout of params) of sum(out * dout); rounded
    as the forward pass was."""
    lay, total = layout()
    p = np.asarray(params, np.float64)
    g = np.zeros(total)
    conv, feat, rd = cache
    q = (lambda a: bf16(a)) if rd else (lambda a: a)
    wo, shp, bo, nb = lay[LAYERS]
    g[wo:wo + FEAT * OUT] = (q(feat).T @ q(dout)).ravel()
    g[bo:bo + nb] = dout.sum(axis=0)
    dy = (q(dout) @ q(p[wo:wo + FEAT * OUT].reshape(FEAT, OUT)).T).reshape(conv[-1][1].shape)
    for l in range(LAYERS - 1, -1, -1):
        wo, shp, bo, nb = lay[l]
        x, y = conv[l]
        dy = dy * (y > 0)
        cols = _im2col(x)
        B, Ho = y.shape[0], y.shape[1]
        K = 9 * shp[2]
        g[wo:wo + K * CH] = (q(cols.reshape(-1, K)).T @ q(dy.reshape(-1, CH))).ravel()
        g[bo:bo + nb] = dy.reshape(-1, CH).sum(axis=0)
        if l == 0:
            break
        W = p[wo:wo + K * CH].reshape(K, CH)
        dcols = (q(dy.reshape(-1, CH)) @ q(W).T).reshape(B, Ho, Ho, 9, shp[2])
        dy = _col2im(dcols, x.shape[1])
    return g


def conv_direct(x, W, b):
    """One layer by its definition (test check of _im2col): y[b, oy, ox, co] =
    relu(b[co] + sum_{ky, kx, ci} x[b, 2 oy + ky - 1, 2 ox + kx - 1, ci] W[ky, kx, ci, co])."""
    B, H, _, C = x.shape
    Ho = H // 2
    y = np.zeros((B, Ho, Ho, W.shape[-1])) + b
    for ky in range(3):
        for kx in range(3):
            for oy in range(Ho):
                iy = 2 * oy + ky - 1
                if iy < 0 or iy >= H:
                    continue
                for ox in range(Ho):
                    ix = 2 * ox + kx - 1
                    if ix < 0 or ix >= H:
                        continue
                    y[:, oy, ox] += x[:, iy, ix] @ W[ky, kx]
    return np.maximum(y, 0.0)
