"""HIP FFT against the float64 DFT (numpy.fft) in the reference's conventions:
forward scaled 1/N, backward unscaled, optional b = -Im (FFTW.f90:179-219,
img2d_FFT.f90:18-19)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [6, 8, 12, 20, 24, 48, 64, 96, 128, 384, 512])
@pytest.mark.parametrize("conj", [False, True])
def test_fft2d_forward(ctx, n, conj):
    rng = np.random.default_rng(n)
    batch = 3
    re = rng.standard_normal((batch, n, n)).astype(np.float32)
    im = rng.standard_normal((batch, n, n)).astype(np.float32)
    gr, gi = ctx.fft2d(re, im, 1, conj)
    ref = np.fft.fft2(re.astype(np.float64) + 1j * im.astype(np.float64)) / (n * n)
    if conj:
        ref = np.conj(ref)
    assert rel_l2(gr + 1j * gi, ref) < 2e-6


@pytest.mark.parametrize("n", [12, 64, 384])
def test_fft2d_roundtrip(ctx, n):
    rng = np.random.default_rng(7)
    re = rng.standard_normal((2, n, n)).astype(np.float32)
    im = np.zeros_like(re)
    fr, fi = ctx.fft2d(re, im, 1, False)
    br, bi = ctx.fft2d(fr, fi, -1, False)
    assert rel_l2(br, re) < 2e-6
    assert np.abs(bi).max() < 1e-5


@pytest.mark.parametrize("n", [8, 12, 24, 32, 64])
def test_fft3d_backward(ctx, n):
    rng = np.random.default_rng(n + 1)
    re = rng.standard_normal((n, n, n)).astype(np.float32)
    im = rng.standard_normal((n, n, n)).astype(np.float32)
    gr, gi = ctx.fft3d(re, im, -1, False)
    ref = np.fft.ifftn(re.astype(np.float64) + 1j * im) * n ** 3
    assert rel_l2(gr + 1j * gi, ref) < 2e-6
    fr, fi = ctx.fft3d(re, im, 1, True)
    ref = np.conj(np.fft.fftn(re.astype(np.float64) + 1j * im) / n ** 3)
    assert rel_l2(fr + 1j * fi, ref) < 2e-6


@pytest.mark.parametrize("n", [256, 384, 512])
def test_fft3d_backward_full_size(ctx, n):
    """The inverse 3D transform of every map at the measured boxes against
    scipy's complex128 inverse FFT (float64 arithmetic), scaled to the
    unscaled backward convention."""
    import scipy.fft
    rng = np.random.default_rng(n + 1)
    re = rng.standard_normal((n, n, n), dtype=np.float32)
    im = rng.standard_normal((n, n, n), dtype=np.float32)
    print(f"fft3d backward n={n}", flush=True)
    gr, gi = ctx.fft3d(re, im, -1, False)
    z = re.astype(np.complex128)
    z.imag = im
    del re, im
    ref = scipy.fft.ifftn(z, overwrite_x=True, workers=8)
    ref *= float(n) ** 3
    num = np.sqrt(np.sum((gr - ref.real) ** 2, dtype=np.float64) + np.sum((gi - ref.imag) ** 2, dtype=np.float64))
    err = float(num / np.linalg.norm(ref.ravel()))
    print(f"n={n}: rel-L2 {err:.2e} vs float64")
    assert err < 2e-6
