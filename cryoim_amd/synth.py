"""Seeded synthetic data in the reference's conventions.

Phantom: Gaussian blobs expanded by the 60 icosahedral operators.  Particle
images are analytic projections of the blobs (a projected 3D Gaussian is a
2D Gaussian), placed with the reference's conventions:

* a model point P appears at image position R^T P (first two components)
  relative to the box centre nx/2, where R = euler_mat(theta, phi, omega)
  (the real-space projector icosprocess/SRC/project_template2d.f90:33-73
  samples the model at R (x', y', z'));
* the particle centre sits at (nx/2 + x, nx/2 + y), which is what
  img2d_FT_centshift removes (icosprocess/SRC/img2d_FT_centshift.f90:11);
* the CTF of getCTF2d (icosprocess/SRC/getCTF2d.f90:2-29) multiplies the
  image transform.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from . import geometry


@dataclasses.dataclass
class Phantom:
    centres: np.ndarray   # (nb, 3) model-frame positions in pixels (from the box centre)
    sigmas: np.ndarray    # (nb,)
    amps: np.ndarray      # (nb,)
    box: int


def make_phantom(box: int, nblobs: int = 20, sigma=(2.0, 4.0), radius=(40.0, 55.0),
                 seed: int = 0, sym: bool = True) -> Phantom:
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(nblobs, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(radius[0], radius[1], size=nblobs)
    s = rng.uniform(sigma[0], sigma[1], size=nblobs)
    a = rng.uniform(0.5, 1.0, size=nblobs)
    p = d * r[:, None]
    if sym:
        ops = geometry.icos_ops()
        p = np.einsum("kij,bj->kbi", ops, p).reshape(-1, 3)
        s = np.tile(s, 60)
        a = np.tile(a, 60)
    return Phantom(p, s, a, box)


def phantom_map(ph: Phantom, box: Optional[int] = None) -> np.ndarray:
    """Voxel map (nz, ny, nx), origin at index box/2 on each axis."""
    n = box or ph.box
    c = n // 2
    ax = np.arange(n, dtype=np.float64) - c
    inv = 1.0 / (2 * ph.sigmas ** 2)
    gx = np.exp(-(ax[None, :] - ph.centres[:, 0:1]) ** 2 * inv[:, None])   # (nb, n)
    gy = np.exp(-(ax[None, :] - ph.centres[:, 1:2]) ** 2 * inv[:, None])
    gz = np.exp(-(ax[None, :] - ph.centres[:, 2:3]) ** 2 * inv[:, None]) * ph.amps[:, None]
    # the map is a sum of nb separable blobs: vol[z, y*n + x] = gz^T (gy (x) gx),
    # one matrix product per block of blobs
    vol = np.zeros((n, n * n))
    step = max(1, (1 << 27) // (n * n))
    for b0 in range(0, len(inv), step):
        b1 = min(len(inv), b0 + step)
        gyx = (gy[b0:b1, :, None] * gx[b0:b1, None, :]).reshape(b1 - b0, n * n)
        vol += gz[b0:b1].T @ gyx
    return vol.reshape(n, n, n).astype(np.float32)


def project(ph: Phantom, angles: np.ndarray, shifts: Optional[np.ndarray] = None,
            box: Optional[int] = None) -> np.ndarray:
    """Analytic projections (n, box, box) for angles (n, 3) in degrees."""
    n = box or ph.box
    c = n // 2
    angles = np.asarray(angles, dtype=np.float64).reshape(-1, 3)
    R = geometry.euler_mat(angles[:, 0], angles[:, 1], angles[:, 2])
    uv = np.einsum("nji,bj->nbi", R, ph.centres)[..., :2]      # R^T P
    if shifts is not None:
        uv = uv + np.asarray(shifts, dtype=np.float64)[:, None, :]
    ax = np.arange(n, dtype=np.float64) - c
    w = ph.amps * ph.sigmas * np.sqrt(2 * np.pi)
    out = np.empty((len(angles), n, n), dtype=np.float32)
    inv = 1.0 / (2 * ph.sigmas ** 2)
    for i in range(len(angles)):
        gx = np.exp(-(ax[None, :] - uv[i, :, 0:1]) ** 2 * inv[:, None])    # (nb, n)
        gy = np.exp(-(ax[None, :] - uv[i, :, 1:2]) ** 2 * inv[:, None])
        out[i] = (gy * w[:, None]).T @ gx
    return out


def ctf2d(nx: int, df1: float, df2: float, astig: float, apix: float, vol: float = 300.0,
          cs: float = 2.7, ampwg: float = 0.1, df_scale: float = 1.0e4) -> np.ndarray:
    """getCTF2d on the centred grid (float64 restatement of getCTF2d.f90:10-27)."""
    ky, hx = np.meshgrid(np.arange(nx) - nx // 2, np.arange(nx) - nx // 2, indexing="ij")
    lam = 0.387818 / np.sqrt(vol * (1.0 + 0.978459e-3 * vol))
    phi = np.arctan2(ky, hx)
    phi = np.where(phi < 0, phi + 2 * np.pi, phi) - np.deg2rad(astig)
    df = df1 * df_scale * np.cos(phi) ** 2 + df2 * df_scale * np.sin(phi) ** 2
    res = np.sqrt(hx ** 2 + ky ** 2) / nx / apix
    phase = 2 * np.pi * (-cs * 1e7 * lam ** 3 * res ** 4 / 4 + df * lam * res ** 2 / 2)
    return -(np.sqrt(1 - ampwg ** 2) * np.sin(phase) + ampwg * np.cos(phase))


def apply_ctf(img: np.ndarray, ctf_centred: np.ndarray) -> np.ndarray:
    f = np.fft.fft2(img)
    return np.real(np.fft.ifft2(f * np.fft.ifftshift(ctf_centred))).astype(np.float32)


def random_au_orientations(n: int, rng: np.random.Generator) -> np.ndarray:
    """Uniform directions restricted to the AU by rejection, omega ~ U[0, 360)."""
    out = []
    while len(out) < n:
        v = rng.normal(size=(4 * n + 64, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        ok = geometry.in_au(v[:, 0], v[:, 1], v[:, 2])
        for x, y, z in v[ok]:
            th = np.rad2deg(np.arccos(z))
            ph = np.rad2deg(np.arctan2(y, x))
            out.append((th, ph, rng.uniform(0, 360)))
            if len(out) == n:
                break
    return np.array(out)


@dataclasses.dataclass
class ParticleSet:
    images: np.ndarray     # (n, box, box) f32
    angles: np.ndarray     # (n, 3) true theta, phi, omega
    shifts: np.ndarray     # (n, 2) true x, y
    df: np.ndarray         # (n, 2) df1, df2 (um for icosprocess)
    astig: np.ndarray      # (n,)


def make_particles(ph: Phantom, n: int, seed: int = 1, shift_max: float = 3.0,
                   df_range=(1.0, 3.0), snr: Optional[float] = 0.1, apix: float = 1.5,
                   vol: float = 300.0, cs: float = 2.7, ampwg: float = 0.1,
                   ctf: bool = True, mask_radius: Optional[float] = None) -> ParticleSet:
    rng = np.random.default_rng(seed)
    ang = random_au_orientations(n, rng)
    sh = rng.uniform(-shift_max, shift_max, size=(n, 2))
    df = rng.uniform(df_range[0], df_range[1], size=n)
    dfs = np.stack([df, df], axis=1)
    astig = np.zeros(n)
    imgs = project(ph, ang, sh)
    nx = ph.box
    if ctf:
        for i in range(n):
            imgs[i] = apply_ctf(imgs[i], ctf2d(nx, dfs[i, 0], dfs[i, 1], 0.0, apix, vol, cs, ampwg))
    if snr is not None:
        r = mask_radius if mask_radius is not None else nx / 2 - 2
        yy, xx = np.meshgrid(np.arange(nx) - nx // 2, np.arange(nx) - nx // 2, indexing="ij")
        inside = (xx ** 2 + yy ** 2) <= r * r
        for i in range(n):
            var = float(np.var(imgs[i][inside]))
            imgs[i] = imgs[i] + rng.normal(scale=np.sqrt(var / snr), size=(nx, nx)).astype(np.float32)
    return ParticleSet(imgs.astype(np.float32), ang, sh, dfs, astig)


def make_particles_torch(ph: Phantom, n: int, seed: int = 1, shift_max: float = 3.0, df_range=(1.0, 3.0),
                         snr: Optional[float] = 0.1, apix: float = 1.5, vol: float = 300.0, cs: float = 2.7,
                         ampwg: float = 0.1, mask_radius: Optional[float] = None, device: str = "cuda:0"):
    """The particles of make_particles (same recipe: AU orientations, shifts,
    defoci from a numpy generator seeded with seed; projections, CTF and
    noise) generated on a torch device, for workloads too large to render on
    the host.  The noise comes from a torch generator seeded with seed, so
    images differ from make_particles' in their noise only.  Returns a
    ParticleSet whose images are a float32 (n, box, box) device tensor."""
    import torch
    rng = np.random.default_rng(seed)
    ang = random_au_orientations(n, rng)
    sh = rng.uniform(-shift_max, shift_max, size=(n, 2))
    df = rng.uniform(df_range[0], df_range[1], size=n)
    nx = ph.box
    c = nx // 2
    R = geometry.euler_mat(ang[:, 0], ang[:, 1], ang[:, 2])
    uv = np.einsum("nji,bj->nbi", R, ph.centres)[..., :2] + sh[:, None, :]
    dev = torch.device(device)
    ax = torch.arange(nx, dtype=torch.float64, device=dev) - c
    inv = torch.as_tensor(1.0 / (2 * ph.sigmas ** 2), device=dev)
    w = torch.as_tensor(ph.amps * ph.sigmas * np.sqrt(2 * np.pi), device=dev)
    uvt = torch.as_tensor(uv, device=dev)
    # separable Gaussians: image = Gy^T diag(w) Gx, one product per particle
    gx = torch.exp(-(ax[None, None, :] - uvt[:, :, 0:1]) ** 2 * inv[None, :, None]).float()
    gy = (torch.exp(-(ax[None, None, :] - uvt[:, :, 1:2]) ** 2 * inv[None, :, None]) * w[None, :, None]).float()
    imgs = torch.bmm(gy.transpose(1, 2).contiguous(), gx.contiguous())
    del gx, gy
    # CTF of getCTF2d per particle, on the uncentred FFT grid
    f = torch.fft.fftfreq(nx, d=1.0, device=dev, dtype=torch.float64) * nx
    ky, hx = torch.meshgrid(f, f, indexing="ij")
    lam = 0.387818 / np.sqrt(vol * (1.0 + 0.978459e-3 * vol))
    res2 = (hx ** 2 + ky ** 2) / (nx * apix) ** 2
    dft = torch.as_tensor(df * 1.0e4, device=dev)[:, None, None]
    phase = 2 * np.pi * (-cs * 1e7 * lam ** 3 * res2[None] ** 2 / 4 + dft * lam * res2[None] / 2)
    ctf = (-(np.sqrt(1 - ampwg ** 2) * torch.sin(phase) + ampwg * torch.cos(phase))).float()
    imgs = torch.fft.ifft2(torch.fft.fft2(imgs) * ctf).real.contiguous()
    del ctf, phase
    if snr is not None:
        r = mask_radius if mask_radius is not None else nx / 2 - 2
        yy, xx = torch.meshgrid(ax, ax, indexing="ij")
        inside = (xx ** 2 + yy ** 2) <= r * r
        var = imgs[:, inside].var(dim=1, unbiased=False)
        g = torch.Generator(device=dev)
        g.manual_seed(int(seed))
        imgs += torch.randn(imgs.shape, generator=g, device=dev) * torch.sqrt(var / snr)[:, None, None]
    return ParticleSet(imgs.float().contiguous(), ang, sh, np.stack([df, df], axis=1), np.zeros(n))
