"""The FP64 restatement of the proposal network (oracle/cnn_oracle.py),
pinned by the mathematics: each layer's patch product equals the
convolution's direct sum over taps, and the backward pass equals central
finite differences of the forward pass for parameters of every layer.
Also: the library's parameter count is the oracle's."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import cnn_oracle as C  # noqa: E402


def test_patch_product_is_the_direct_convolution():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 10, 10, 3))
    W = rng.standard_normal((3, 3, 3, 5))
    b = rng.standard_normal(5)
    cols = C._im2col(x)
    y = np.maximum(cols.reshape(2, 5, 5, -1) @ W.reshape(27, 5) + b, 0.0)
    assert np.allclose(y, C.conv_direct(x, W, b), atol=1e-12)


def test_col2im_is_the_adjoint_of_im2col():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 8, 8, 3))
    d = rng.standard_normal((2, 4, 4, 9, 3))
    assert np.isclose(np.sum(C._im2col(x) * d), np.sum(x * C._col2im(d, 8)), rtol=1e-12)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = C.init_params(7).astype(np.float64)
    img = rng.standard_normal((2, 128, 128))
    dout = rng.standard_normal((2, C.OUT))
    out, cache = C.forward(img, p)
    g = C.backward(p, cache, dout)
    lay, total = C.layout()
    assert total == len(p)
    # a small step keeps the ReLU kinks (pre-activations within h of zero) out of the differences
    h = 1e-7
    worst = 0.0
    for wo, shp, bo, nb in lay:
        n = int(np.prod(shp))
        for idx in list(rng.choice(n, 3, replace=False) + wo) + [bo + int(rng.integers(nb))]:
            q = p.copy()
            q[idx] += h
            fp = np.sum(C.forward(img, q)[0] * dout)
            q[idx] -= 2 * h
            fm = np.sum(C.forward(img, q)[0] * dout)
            fd = (fp - fm) / (2 * h)
            err = abs(fd - g[idx]) / max(1e-6, abs(fd), abs(g[idx]))
            worst = max(worst, err)
            assert err < 1e-4, (idx, fd, g[idx])
    print(f"worst relative gradient error vs central differences: {worst:.2e}")


def test_library_parameter_count():
    from cryoim_amd import lib
    L = lib.load()
    assert L.cim_cnn_nparams() == C.layout()[1]


def test_bf16_rounding_is_to_nearest_even():
    # bfloat16 keeps 8 significand bits: the spacing above 1 is 2^-7
    x = np.array([1.0, 1 + 2 ** -9, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7 + 2 ** -8, -2.5, 0.0])
    want = np.array([1.0, 1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -6, -2.5, 0.0])
    assert np.array_equal(C.bf16(x), want)
