"""Residual tables of the tabulated search (search.hip k_cl_table,
k_cl_table_mfma) against float64 restatements of the same sums, through the
kernel-level entry point of include/cryoim_debug.h.

For every pair of rays (template ray jT, particle ray jP) the tables hold
N = sum_r w |fold(phT - phP)| and D = sum_r w, w = max(hT + hP, 0) with
h = |F|/2, or w = 0 when either sample is gated (the phase residual of
cal_crosslin_PR_omega.f90:48-69 before the line sums).  The VALU kernel
(nh = 0) evaluates |fold(d)| exactly; the MFMA kernel evaluates the series
pi/2 - (4/pi) sum_{k odd <= 2 nh - 1} cos(k d) / k^2 with bfloat16 features
and float32 sums.  Both store N and D as bfloat16."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

GATED = -1.0e30


def _bf16(x):
    """float32 -> nearest bfloat16 (ties to even), as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _rays(rng, n, nr, npsi, gated=0.2):
    ph = rng.uniform(-np.pi, np.pi, (n, nr, npsi))
    h = rng.uniform(0.05, 2.0, (n, nr, npsi))
    h[rng.uniform(size=h.shape) < gated] = GATED
    return np.stack([ph, h], -1).astype(np.float32)


def _reference(P, Tm, T, nh, per_particle):
    """float64 N, D of every (particle, template) item, with the features the
    kernel forms (bfloat16 for nh > 0)."""
    npart, nr, npsi = P.shape[:3]
    nrow = npsi // 2
    Ns, Ds = [], []
    for p in range(npart):
        for t in range(T):
            tp = Tm[p * T + t if per_particle else t][:, :nrow]
            pp = P[p]
            gT = (tp[..., 1] > 0.5 * GATED).astype(np.float64)
            gP = (pp[..., 1] > 0.5 * GATED).astype(np.float64)
            aT, aP = gT * tp[..., 1], gP * pp[..., 1]
            if nh == 0:
                w = gT[:, :, None] * gP[:, None, :] * (aT[:, :, None] + aP[:, None, :])
                d = np.abs(tp[..., 0].astype(np.float64)[:, :, None] - pp[..., 0][:, None, :])
                d = np.minimum(d, 2 * np.pi - d)
                Ns.append(np.einsum("rij,rij->ij", w, d))
                Ds.append(w.sum(0))
                continue
            b = _bf16
            D = np.einsum("ri,rj->ij", b(aT), b(gP)) + np.einsum("ri,rj->ij", b(gT), b(aP))
            H = np.zeros_like(D)
            for kk in range(nh):
                k = 2 * kk + 1
                c = np.float32(-4.0 / np.pi / k ** 2)
                cT, sT = np.cos(k * tp[..., 0].astype(np.float64)), np.sin(k * tp[..., 0].astype(np.float64))
                cP, sP = np.cos(k * pp[..., 0].astype(np.float64)), np.sin(k * pp[..., 0].astype(np.float64))
                H += (np.einsum("ri,rj->ij", b(aT * cT), b(c * gP * cP)) + np.einsum("ri,rj->ij", b(aT * sT), b(c * gP * sP))
                      + np.einsum("ri,rj->ij", b(gT * cT), b(c * aP * cP)) + np.einsum("ri,rj->ij", b(gT * sT), b(c * aP * sP)))
            Ns.append(np.pi / 2 * D + H)
            Ds.append(D)
    return np.array(Ns), np.array(Ds)


@pytest.mark.parametrize("nh", [0, 1, 2, 4, 8])
@pytest.mark.parametrize("npsi,nr", [(72, 9), (720, 73), (200, 21)])
def test_tables_match_float64(ctx, nh, npsi, nr):
    rng = np.random.default_rng(npsi + nr + nh)
    npart, T = 2, 3
    P = _rays(rng, npart, nr, npsi)
    Tm = _rays(rng, npart * T, nr, npsi)
    N, D = ctx.debug_cl_table(nh, P, Tm, T, True)
    Nr, Dr = _reference(P, Tm, T, nh, True)
    eN, eD = rel_l2(N, Nr), rel_l2(D, Dr)
    mD = float(np.max(np.abs(D - Dr) / np.maximum(np.abs(Dr), 1e-3)))
    print(f"nh {nh} npsi {npsi} nr {nr}: N rel-L2 {eN:.2e}, D rel-L2 {eD:.2e}, D max rel {mD:.2e}")
    # bfloat16 storage of N and D: half an ulp is 2^-9 of the value
    assert eN < 3e-3 and eD < 3e-3 and mD < 8e-3


def test_series_tables_approach_the_exact_residual(ctx):
    """With more harmonics the MFMA tables converge on the exact residual
    (the series of |d| converges as 1/k^2)."""
    rng = np.random.default_rng(5)
    P = _rays(rng, 1, 25, 360)
    Tm = _rays(rng, 2, 25, 360)
    N0, D0 = ctx.debug_cl_table(0, P, Tm, 2, True)
    errs = []
    for nh in (1, 2, 4, 8):
        N, D = ctx.debug_cl_table(nh, P, Tm, 2, True)
        errs.append(rel_l2(N / np.maximum(D, 1e-3), N0 / np.maximum(D0, 1e-3)))
    print("mean residual vs exact, nh 1/2/4/8:", " ".join(f"{e:.3e}" for e in errs))
    # (nh 8 sits at the bfloat16 storage noise, 2^-9 of each entry)
    assert errs[0] > errs[1] > errs[2]
    assert max(errs[2], errs[3]) < 0.03
