"""Host-side rotation geometry in float64.

* ``euler_mat``: ZYZ matrix R = Rz(phi) Ry(theta) Rz(omega) mapping the image
  frame to the model frame (icosprocess/SRC/euler_mat.f90:3-21).
* ``icos_ops``: the 60 icosahedral rotations in the reference's order
  (icosprocess/SRC/icos2f_sym_mat.f90:1-248), generated by closure exactly
  as cryoim_amd/csrc/cim_symops.h does.
* ``in_au`` / ``return_au``: the asymmetric unit of
  icosprocess/SRC/ort_globesearch.f90:17-18,80-82 and
  icosprocess/SRC/return_icos_asymmetry_unit.f90:1-33.
"""
from __future__ import annotations

import numpy as np

ICOS60_ORDER = (
    0, 17, 24, 8, 10, 16, 49, 30, 7, 34, 57, 47, 13, 36, 58, 48, 14, 12, 46, 33,
    9, 1, 3, 6, 11, 41, 38, 35, 19, 28, 22, 39, 43, 20, 37, 42, 31, 44, 40, 45,
    32, 18, 5, 15, 4, 23, 50, 51, 25, 2, 21, 55, 53, 56, 26, 54, 52, 29, 59, 27)


def euler_mat(theta, phi, omega) -> np.ndarray:
    """Rotation(s) for angles in degrees; broadcasts, returns (..., 3, 3)."""
    th, ph, om = (np.deg2rad(np.asarray(v, dtype=np.float64)) for v in (theta, phi, omega))
    c1, c2, c3 = np.cos(ph), np.cos(th), np.cos(om)
    s1, s2, s3 = np.sin(ph), np.sin(th), np.sin(om)
    R = np.empty(np.broadcast(th, ph, om).shape + (3, 3))
    R[..., 0, 0] = c1 * c2 * c3 - s1 * s3
    R[..., 0, 1] = -c1 * c2 * s3 - c3 * s1
    R[..., 0, 2] = c1 * s2
    R[..., 1, 0] = c2 * c3 * s1 + c1 * s3
    R[..., 1, 1] = -c2 * s1 * s3 + c1 * c3
    R[..., 1, 2] = s1 * s2
    R[..., 2, 0] = -c3 * s2
    R[..., 2, 1] = s2 * s3
    R[..., 2, 2] = c2
    return R


def _axis_rotation(axis, ang):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)


def icos_ops_canonical() -> np.ndarray:
    """Breadth-first closure of the 5-fold about (1,0,phi) and the 2-fold about z."""
    gphi = (1 + 5 ** 0.5) / 2
    gens = [_axis_rotation((1, 0, gphi), 2 * np.pi / 5), _axis_rotation((0, 0, 1), np.pi)]
    out = [np.eye(3)]
    head = 0
    while head < len(out) and len(out) < 60:
        x = out[head]
        head += 1
        for g in gens:
            y = g @ x
            if not any(np.abs(y - z).max() < 1e-9 for z in out):
                out.append(y)
            if len(out) == 60:
                break
    return np.array(out)


_ICOS = None


def icos_ops() -> np.ndarray:
    """(60, 3, 3) float64 in the reference's operator order."""
    global _ICOS
    if _ICOS is None:
        _ICOS = icos_ops_canonical()[list(ICOS60_ORDER)]
    return _ICOS.copy()


def cn_ops(n: int) -> np.ndarray:
    """Cn about z (local_reconstruct/SRC/C_matrix.f90:1-17 generates these)."""
    out = []
    for k in range(n):
        a = 2 * np.pi * k / n
        out.append(np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1.0]]))
    return np.array(out)


_D2R = np.arctan2(1.0, 1.0) / 45.0
_A1 = -np.sin(31.717 * _D2R) * np.cos(69.09 * _D2R)
_B1 = -np.cos(31.717 * _D2R) * np.cos(69.09 * _D2R)
_C1 = np.sin(31.717 * _D2R) * np.sin(69.09 * _D2R)
_B2 = -_B1


def in_au(x, y, z):
    """Direction (x, y, z) inside the icosahedral asymmetric unit."""
    z1 = -(_A1 * x + _B1 * y) / _C1
    z2 = -(_A1 * x + _B2 * y) / _C1
    return (z >= 0) & (z <= z1) & (z <= z2)


def return_au(theta, phi, omega):
    """Fold an orientation into the AU, first operator that lands there wins
    (falls through to the last operator, return_icos_asymmetry_unit.f90:17-29)."""
    E = euler_mat(theta, phi, omega)
    ops = icos_ops()
    th = ph = om = 0.0
    for S in ops:
        M = S @ E
        th = np.arccos(np.clip(M[2, 2], -1, 1))
        ph = np.arctan2(M[1, 2], M[0, 2])
        om = np.arctan2(M[2, 1], -M[2, 0])
        if in_au(np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)):
            break
    return np.rad2deg(th), np.rad2deg(ph), np.rad2deg(om)


def search_grid(step: float):
    """(Ntheta, Nphi, Nomega) of ort_globesearch.f90:19-21 (float32 division as in Fortran)."""
    s = np.float32(step)
    return int(np.float32(21.0) / s), int(np.float32(32.0) / s), int(np.float32(360.0) / s)


def au_grid_points(step: float):
    """(itheta, iphi) pairs inside the AU in the order ort_globesearch visits them."""
    nt, nph, _ = search_grid(step)
    s = np.float32(step)
    d2r = np.float32(np.arctan2(np.float32(1), np.float32(1)) / np.float32(45))
    pts = []
    for it in range(nt + 1):
        th = (np.float32(69.0) + s * np.float32(it)) * d2r
        z, z0 = np.cos(th), np.sin(th)
        for ip in range(nph + 1):
            ph = s * np.float32(ip) * d2r
            if in_au(z0 * np.cos(ph), z0 * np.sin(ph), z):
                pts.append((it, ip))
    return pts


def sym_distance_deg(a, b, ops: np.ndarray | None = None) -> np.ndarray:
    """Smallest rotation angle (degrees) between orientations a and b (n, 3)
    Euler angles, over the symmetry-equivalent forms of b: the model is
    invariant under the operators, so R and R g (model frame) are the same
    view for every operator g (60 icosahedral operators by default)."""
    ops = icos_ops() if ops is None else ops
    a = np.asarray(a, dtype=np.float64).reshape(-1, 3)
    b = np.asarray(b, dtype=np.float64).reshape(-1, 3)
    Ra = euler_mat(a[:, 0], a[:, 1], a[:, 2])
    Rb = euler_mat(b[:, 0], b[:, 1], b[:, 2])
    best = np.full(len(a), 180.0)
    for g in ops:
        for M in (Rb @ g, g @ Rb):
            # trace of Ra^T M gives the angle of the relative rotation
            tr = np.einsum("nij,nij->n", Ra, M)
            ang = np.rad2deg(np.arccos(np.clip((tr - 1.0) / 2.0, -1.0, 1.0)))
            best = np.minimum(best, ang)
    return best
