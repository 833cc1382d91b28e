"""cryoIM file formats, kept byte-compatible with the reference programs.

* MRC: 1024-byte header laid out as ``type mrchead``
  (icosprocess/SRC/0common_icosproc.f90:61-113), mode 2 little-endian f32
  data; stacks hold one image per section
  (icosprocess/SRC/reconstruct.f90:63-76).  Maps are written with the header
  fields set by ``maphead`` (icosprocess/SRC/mrchead.f90:2-58).
* icosprocess ortfile: list-directed input
  ``stck id theta phi omega x y PR imgID imgparticleID df1 df2 astig [x0 y0]``
  (icosprocess/SRC/load_original_ortcent.f90:12-22) and output format 900
  (icosprocess/SRC/ort_globesearch.f90:128-139).
* template ortfile ``id theta phi omega x y`` (icosprocess/SRC/load_template.f90:18-21).
* local ortfile (local_box output / local_refine input, format 901 on output:
  local_refine/SRC/refine_PR.f90:235-256, load_original_ort.f90:12-16).
* tnf partial volume: ``FFTsize maxR apix imgmask fsc realspaceavg`` then
  ``fa fb wg`` of (2maxR+1)^2 (maxR+1) f32 each
  (icosprocess/SRC/reconstruct_substep.f90:148-155).
"""
from __future__ import annotations

import dataclasses
import struct
from typing import List, Optional, Sequence

import numpy as np

MRC_HEADER_BYTES = 1024


def _mrc_header(nx: int, ny: int, nz: int, apix: float = 1.0, is_map: bool = False) -> bytes:
    """Build a 1024-byte MRC header.  For maps the fields follow maphead."""
    h = bytearray(MRC_HEADER_BYTES)
    if is_map:
        start = (-nx // 2, -ny // 2, -nz // 2)
        m = (nx - 1, ny - 1, nz - 1)
        cell = (m[0] * apix, m[1] * apix, m[1] * apix)
        origin = (-1.0 * nx, -1.0 * ny, -1.0 * nz)
    else:
        start = (0, 0, 0)
        m = (nx, ny, nz)
        cell = (nx * apix, ny * apix, nz * apix)
        origin = (0.0, 0.0, 0.0)
    struct.pack_into("<10i", h, 0, nx, ny, nz, 2, start[0], start[1], start[2], m[0], m[1], m[2])
    struct.pack_into("<6f", h, 40, cell[0], cell[1], cell[2], 90.0, 90.0, 90.0)
    struct.pack_into("<3i", h, 64, 1, 2, 3)
    struct.pack_into("<3f", h, 76, 0.0, 0.0, 0.0)
    struct.pack_into("<2i", h, 88, 0, 0)
    struct.pack_into("<3f", h, 196, origin[0], origin[1], origin[2])
    h[208:212] = b"MAP "
    h[212:216] = b"DA\x00\x00"
    return bytes(h)


def write_mrc(path: str, data: np.ndarray, apix: float = 1.0, is_map: bool = False) -> None:
    """Write a 2D image, a stack (n, ny, nx) or a map (nz, ny, nx) as mode 2."""
    a = np.ascontiguousarray(data, dtype="<f4")
    if a.ndim == 2:
        a = a[None]
    nz, ny, nx = a.shape
    with open(path, "wb") as f:
        f.write(_mrc_header(nx, ny, nz, apix, is_map))
        f.write(a.tobytes())


def read_mrc_header(path: str) -> dict:
    with open(path, "rb") as f:
        h = f.read(MRC_HEADER_BYTES)
    nx, ny, nz, mode = struct.unpack_from("<4i", h, 0)
    return {"nx": nx, "ny": ny, "nz": nz, "mode": mode}


def read_mrc(path: str, mmap: bool = False) -> np.ndarray:
    """Read a mode-2 MRC file as (nz, ny, nx) f32 (the reference ignores nsymbt)."""
    hd = read_mrc_header(path)
    if hd["mode"] != 2:
        raise ValueError(f"{path}: MRC mode {hd['mode']} unsupported (only mode 2, as the reference)")
    shape = (hd["nz"], hd["ny"], hd["nx"])
    if mmap:
        return np.memmap(path, dtype="<f4", mode="r", offset=MRC_HEADER_BYTES, shape=shape)
    return np.fromfile(path, dtype="<f4", offset=MRC_HEADER_BYTES).reshape(shape)


@dataclasses.dataclass
class Particle:
    """One ortfile record (icosprocess/SRC/0common_icosproc.f90:28-35)."""
    stckfile: str
    particle_id: int
    theta: float
    phi: float
    omega: float
    x: float
    y: float
    pr: float
    img_id: int
    img_particle_id: int
    df1: float
    df2: float
    astig: float
    x0: float = 0.0      # mode 6 only
    y0: float = 0.0
    dtheta: float = 0.0
    dphi: float = 0.0
    domega: float = 0.0
    dx: float = 0.0
    dy: float = 0.0
    dpr: float = 0.0


def read_ortfile(path: str, with_local_box: bool = False) -> List[Particle]:
    """List-directed read of an icosprocess ortfile (load_original_ortcent.f90:12-22)."""
    out = []
    with open(path) as f:
        for line in f:
            t = line.replace(",", " ").split()
            if len(t) < 13:
                continue
            p = Particle(t[0], int(t[1]), float(t[2]), float(t[3]), float(t[4]), float(t[5]),
                         float(t[6]), float(t[7]), int(t[8]), int(t[9]), float(t[10]),
                         float(t[11]), float(t[12]))
            if with_local_box:
                p.x0, p.y0 = float(t[13]), float(t[14])
            out.append(p)
    return out


def write_ortfile(path: str, parts: Sequence[Particle], with_local_box: bool = False) -> None:
    """Write an input-style ortfile (one record per line, list-directed readable)."""
    with open(path, "w") as f:
        for p in parts:
            s = (f"{p.stckfile} {p.particle_id} {p.theta:.6f} {p.phi:.6f} {p.omega:.6f} {p.x:.6f} "
                 f"{p.y:.6f} {p.pr:.6f} {p.img_id} {p.img_particle_id} {p.df1:.6f} {p.df2:.6f} "
                 f"{p.astig:.6f}")
            if with_local_box:
                s += f" {p.x0:.6f} {p.y0:.6f}"
            f.write(s + "\n")


def _f10(v: float) -> str:
    s = f"{v:10.4f}"
    return s if len(s) == 10 else "*" * 10


def format900(p: Particle) -> str:
    """Output record, format 900 (a54,i8,6f10.4,i12,i6,9f10.4)."""
    s = p.stckfile[:54].ljust(54) + f"{p.particle_id:8d}"
    for v in (p.theta, p.phi, p.omega, p.x, p.y, p.pr):
        s += _f10(v)
    s += f"{p.img_id:12d}{p.img_particle_id:6d}"
    for v in (p.df1, p.df2, p.astig, p.dtheta, p.dphi, p.domega, p.dx, p.dy, p.dpr):
        s += _f10(v)
    return s


def write_ortfile900(path: str, parts: Sequence[Particle]) -> None:
    with open(path, "w") as f:
        for p in parts:
            f.write(format900(p) + "\n")


def read_ortfile900(path: str) -> List[Particle]:
    """Read format-900 output written by ort_globesearch / ort_refinement."""
    out = []
    with open(path) as f:
        for line in f:
            if len(line.rstrip("\n")) < 54 + 8:
                continue
            stck = line[:54].strip()
            rest = line[54:]
            fields = [rest[0:8]]
            pos = 8
            for _ in range(6):
                fields.append(rest[pos:pos + 10]); pos += 10
            fields.append(rest[pos:pos + 12]); pos += 12
            fields.append(rest[pos:pos + 6]); pos += 6
            for _ in range(9):
                fields.append(rest[pos:pos + 10]); pos += 10
            v = [float(x) if "*" not in x else float("nan") for x in fields]
            p = Particle(stck, int(v[0]), v[1], v[2], v[3], v[4], v[5], v[6], int(v[7]), int(v[8]),
                         v[9], v[10], v[11], dtheta=v[12], dphi=v[13], domega=v[14], dx=v[15],
                         dy=v[16], dpr=v[17])
            out.append(p)
    return out


def write_template_ort(path: str, angles: np.ndarray, shifts: Optional[np.ndarray] = None) -> None:
    """Template ortfile ``id theta phi omega x y`` (load_template.f90:18-21)."""
    with open(path, "w") as f:
        for i, (t, p, o) in enumerate(np.asarray(angles, float)):
            x, y = (0.0, 0.0) if shifts is None else shifts[i]
            f.write(f"{i + 1} {t:.6f} {p:.6f} {o:.6f} {x:.6f} {y:.6f}\n")


def read_template_ort(path: str) -> np.ndarray:
    rows = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) >= 6:
                rows.append([float(v) for v in t[1:6]])
    return np.asarray(rows, dtype=np.float64).reshape(-1, 5)


# --- tnf partial volumes (reconstruct_substep.f90:148-155) ---------------

def write_tnf(path: str, fft_size: int, max_r: int, apix: float, imgmask: float, fa: np.ndarray,
              fb: np.ndarray, wg: np.ndarray, fsc: int = 0, realspaceavg: int = 0) -> None:
    n = (2 * max_r + 1) ** 2 * (max_r + 1)
    for a in (fa, fb, wg):
        if a.size != n:
            raise ValueError("tnf arrays must hold (2maxR+1)^2 (maxR+1) values")
    with open(path, "wb") as f:
        f.write(struct.pack("<2i2f2i", fft_size, max_r, apix, imgmask, fsc, realspaceavg))
        for a in (fa, fb, wg):
            f.write(np.ascontiguousarray(a, dtype="<f4").tobytes())


def read_tnf(path: str) -> dict:
    with open(path, "rb") as f:
        hd = f.read(24)
        fft_size, max_r, apix, imgmask, fsc, rsa = struct.unpack("<2i2f2i", hd)
        n = (2 * max_r + 1) ** 2 * (max_r + 1)
        data = np.fromfile(f, dtype="<f4", count=3 * n)
    if data.size != 3 * n:
        raise IOError(f"{path}: truncated tnf ({data.size} of {3 * n} values)")
    shape = (max_r + 1, 2 * max_r + 1, 2 * max_r + 1)
    return {"fft_size": fft_size, "max_r": max_r, "apix": apix, "imgmask": imgmask, "fsc": fsc,
            "realspaceavg": rsa, "fa": data[:n].reshape(shape), "fb": data[n:2 * n].reshape(shape),
            "wg": data[2 * n:].reshape(shape)}
