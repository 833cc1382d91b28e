"""ctypes binding of libcryoim_hip.so (include/cryoim.h).

The HIP library is the only compute path: if it is missing or fails to
load, every call raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcryoim_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cryoim.h")

PART_STRIDE = 12
P_Z = 11          # local_refine records: local_z0 in, chosen z out (cim_lref_defocus)
CNN_HYP, CNN_OUT = 4, 28   # proposal network: hypotheses, outputs (7 per hypothesis)
SEARCH_PLAIN, SEARCH_TABLES = 0, 1
CENTER_MODEL, CENTER_SELFROT180 = 0, 1
ICOSAVG_FULL, ICOSAVG_MERGE = 0, 1
LBOX_REC = 8
P_THETA, P_PHI, P_OMEGA, P_X, P_Y, P_PR, P_DF1, P_DF2, P_ASTIG, P_MASKX, P_MASKY, P_USE = range(12)
P_BOXX, P_BOXY = P_MASKX, P_MASKY


def _ops(fn, sym: str) -> np.ndarray:
    n = ctypes.c_int()
    if fn(sym.encode(), None, ctypes.byref(n)) != 0:
        raise CimError(f"unknown symmetry {sym!r}")
    out = np.zeros(9 * n.value, np.float64)
    fn(sym.encode(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(n))
    return out.reshape(-1, 3, 3)


def sym_ops(sym: str) -> np.ndarray:
    """Symmetry operators of the local programs (cim_sym_ops), (n, 3, 3)."""
    return _ops(load().cim_sym_ops, sym)


def local_reconstruct_sym_ops(sym: str) -> np.ndarray:
    """local_reconstruct's operators (cim_local_reconstruct_sym_ops), (n, 3, 3)."""
    return _ops(load().cim_local_reconstruct_sym_ops, sym)


def local_box_sym_ops(sym: str) -> np.ndarray:
    """local_box's operators (cim_local_box_sym_ops), (n, 3, 3)."""
    return _ops(load().cim_local_box_sym_ops, sym)


def icosprocess_sym_ops(sym: str) -> np.ndarray:
    """Local symmetry operators of icosprocess mode 6 (cim_icosprocess_sym_ops), (n, 3, 3)."""
    return _ops(load().cim_icosprocess_sym_ops, sym)


class CimParams(ctypes.Structure):
    _fields_ = [
        ("box", ctypes.c_int), ("minR", ctypes.c_int), ("maxR", ctypes.c_int),
        ("apply_ctf", ctypes.c_int), ("mask_on", ctypes.c_int), ("norm_on", ctypes.c_int),
        ("ncycle", ctypes.c_int), ("conj_out", ctypes.c_int), ("search_method", ctypes.c_int),
        ("search_topk", ctypes.c_int), ("corr_cycles", ctypes.c_int), ("search_second", ctypes.c_int),
        ("reserved_i", ctypes.c_int * 4),
        ("apix", ctypes.c_float), ("imgmask", ctypes.c_float), ("vol", ctypes.c_float),
        ("cs", ctypes.c_float), ("ampwg", ctypes.c_float), ("df_scale", ctypes.c_float),
        ("searchstep", ctypes.c_float), ("delta_angle", ctypes.c_float),
        ("delta_trans", ctypes.c_float), ("corr_step", ctypes.c_float), ("reserved_f", ctypes.c_float * 6),
    ]


def make_params(box, minR=0, maxR=None, apply_ctf=True, mask_on=True, norm_on=False, ncycle=6,
                conj_out=True, apix=1.0, imgmask=None, vol=300.0, cs=2.7, ampwg=0.1, df_scale=1.0e4,
                searchstep=2.0, delta_angle=0.0, delta_trans=0.0, search_method=0, search_topk=0,
                corr_step=0.0, corr_cycles=0, search_second=0) -> CimParams:
    p = CimParams()
    p.box = int(box)
    p.minR = int(minR)
    p.maxR = int(maxR if maxR is not None else box // 2 - 2)
    p.apply_ctf = int(bool(apply_ctf))
    p.mask_on = int(bool(mask_on))
    p.norm_on = int(norm_on)          # 0, 1 (before the mask) or 2 (after it); True/False as 1/0
    p.ncycle = int(ncycle)
    p.conj_out = int(bool(conj_out))
    p.apix = float(apix)
    p.imgmask = float(imgmask if imgmask is not None else box / 2 - 2)
    p.vol, p.cs, p.ampwg, p.df_scale = float(vol), float(cs), float(ampwg), float(df_scale)
    p.searchstep = float(searchstep)
    p.delta_angle, p.delta_trans = float(delta_angle), float(delta_trans)
    p.search_method = int(search_method)
    p.search_topk = int(search_topk)
    p.search_second = int(search_second)
    p.corr_step, p.corr_cycles = float(corr_step), int(corr_cycles)
    return p


class CimError(RuntimeError):
    pass


_LIB = None


def load() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise CimError(f"{LIB_PATH} is not built (run __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    vp, ip, fp, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
    P = ctypes.POINTER(CimParams)
    sig = {
        "cim_init": (ip, [ctypes.POINTER(vp), ip]),
        "cim_finalize": (ip, [vp]),
        "cim_last_error": (ctypes.c_char_p, [vp]),
        "cim_get_stream": (vp, [vp]),
        "cim_sync": (ip, [vp]),
        "cim_timing": (ip, [vp, ip]),
        "cim_timing_get": (ip, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_longlong)]),
        "cim_counter_get": (ip, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_longlong)]),
        "cim_counter_reset": (ip, [vp]),
        "cim_fft2d": (ip, [vp, fp, fp, ip, ip, ip, ip]),
        "cim_fft3d": (ip, [vp, fp, fp, ip, ip, ip]),
        "cim_fft2d_dev": (ip, [vp, vp, ip, ip, ip, ip]),
        "cim_fft3d_dev": (ip, [vp, vp, ip, ip, ip]),
        "cim_rec_begin": (ip, [vp, ip, ip, ip, ctypes.c_float, ctypes.c_float, ip, dp]),
        "cim_rec_insert": (ip, [vp, P, ip, fp, fp]),
        "cim_rec_insert_dev": (ip, [vp, P, ip, vp, fp]),
        "cim_rec_insert_ft": (ip, [vp, ip, fp, fp, fp, fp]),
        "cim_rec_get_tnf": (ip, [vp, fp, fp, fp]),
        "cim_rec_add_tnf": (ip, [vp, fp, fp, fp]),
        "cim_rec_finish": (ip, [vp, fp]),
        "cim_rec_finish_dev": (ip, [vp, vp]),
        "cim_merge_map": (ip, [vp, ip, ctypes.c_float, fp]),
        "cim_fsc_map": (ip, [vp, ip, ip, ctypes.c_float, ip, ip, fp, fp, fp, fp, fp, fp, fp, fp]),
        "cim_icosavg": (ip, [vp, fp, ctypes.c_longlong, ip, ctypes.c_float, ip, ip, dp]),
        "cim_icosavg_dev": (ip, [vp, vp, ctypes.c_longlong, ip, ctypes.c_float, ip, ip, dp]),
        "cim_search_templates": (ip, [vp, P, ip, fp, fp]),
        "cim_search_global": (ip, [vp, P, ip, fp, fp]),
        "cim_search_global_dev": (ip, [vp, P, ip, vp, fp]),
        "cim_step_capture": (ip, [vp, P, P, ip, vp, vp]),
        "cim_step_replay": (ip, [vp]),
        "cim_step_release": (ip, [vp]),
        "cim_search_pr_array": (ip, [vp, P, fp, fp, fp]),
        "cim_refine_local": (ip, [vp, P, ip, fp, fp]),
        "cim_sym_ops": (ip, [ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_int)]),
        "cim_icosprocess_sym_ops": (ip, [ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_int)]),
        "cim_local_reconstruct_sym_ops": (ip, [ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_int)]),
        "cim_local_box_sym_ops": (ip, [ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_int)]),
        "cim_local_box_coords": (ip, [vp, ip, fp, ip, dp, fp, ip, ctypes.c_float, fp]),
        "cim_local_box_stack": (ip, [vp, ip, fp, fp, ip, ip, dp, fp, ip, ctypes.c_float, fp, fp]),
        "cim_local_box_stack_dev": (ip, [vp, ip, vp, fp, ip, ip, dp, fp, ip, ctypes.c_float, vp, fp]),
        "cim_local_box_stack_noalloc": (ip, [vp, ip, vp, fp, ip, ip, dp, fp, ip, ctypes.c_float, vp, fp]),
        "cim_search_templates_model": (ip, [vp, P, ip, fp, fp]),
        "cim_project_templates": (ip, [vp, ip, ip, fp, fp, fp]),
        "cim_debug_cl_table": (ip, [vp, ip, ip, ip, ip, ip, ip, fp, fp, ctypes.POINTER(ctypes.c_uint32)]),
        "cim_lref_model": (ip, [vp, P, fp]),
        "cim_center_model": (ip, [vp, P, fp]),
        "cim_center": (ip, [vp, P, ip, ip, fp, fp, fp]),
        "cim_center_dev": (ip, [vp, P, ip, ip, vp, fp, fp]),
        "cim_lref_refine": (ip, [vp, P, ip, fp, ip, fp, ctypes.c_char_p, ctypes.c_float, ctypes.c_float]),
        "cim_lref_defocus": (ip, [vp, P, ip, fp, ip, fp, ctypes.c_float, ctypes.c_float]),
        "cim_search_candidates_dev": (ip, [vp, P, ip, vp, fp, ip, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int)]),
        "cim_cnn_nparams": (ip, []),
        "cim_debug_cnn_act": (ip, [vp, ip, fp]),
        "cim_cnn_forward": (ip, [vp, ip, vp, fp, vp]),
        "cim_cnn_backward": (ip, [vp, vp, fp]),
    }
    for name, (res, args) in sig.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
    _LIB = lib
    return lib


def header_functions(path: str = HEADER):
    """Names of the functions include/cryoim.h declares."""
    import re
    txt = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:int|void\s*\*|const char\s*\*)\s*(cim_\w+)\s*\(", txt, re.M)))


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class Context:
    """One cim_ctx on one device; all calls are synchronous.  A process that
    also uses torch on the GPU initialises torch's device first
    (torch.cuda.init()): torch carries its own copy of the HIP runtime, and
    the second runtime to reach the GPU finds none."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.h = ctypes.c_void_p()
        rc = self.lib.cim_init(ctypes.byref(self.h), int(device))
        if rc != 0:
            raise CimError(f"cim_init(device={device}) failed with {rc} (no HIP device?)")

    def close(self):
        if self.h:
            self.lib.cim_finalize(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.cim_last_error(self.h)
            raise CimError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    @property
    def stream(self) -> int:
        return self.lib.cim_get_stream(self.h) or 0

    def sync(self):
        self._check(self.lib.cim_sync(self.h), "cim_sync")

    def timing(self, on: bool):
        self._check(self.lib.cim_timing(self.h, int(on)), "cim_timing")

    def timing_get(self, name: str):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        self._check(self.lib.cim_timing_get(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)),
                    "cim_timing_get")
        return ms.value, n.value

    def counter(self, name: str) -> int:
        v = ctypes.c_longlong()
        self._check(self.lib.cim_counter_get(self.h, name.encode(), ctypes.byref(v)), "cim_counter_get")
        return v.value

    def counter_reset(self):
        self._check(self.lib.cim_counter_reset(self.h), "cim_counter_reset")

    # -- FFT -------------------------------------------------------------
    def fft2d(self, re: np.ndarray, im: np.ndarray, direction: int = 1, conj_out: bool = False):
        re, im = _f32(re).copy(), _f32(im).copy()
        n = re.shape[-1]
        batch = re.size // (n * n)
        self._check(self.lib.cim_fft2d(self.h, _fptr(re), _fptr(im), n, batch, direction, int(conj_out)),
                    "cim_fft2d")
        return re, im

    def fft3d(self, re: np.ndarray, im: np.ndarray, direction: int = 1, conj_out: bool = False):
        re, im = _f32(re).copy(), _f32(im).copy()
        n = re.shape[-1]
        self._check(self.lib.cim_fft3d(self.h, _fptr(re), _fptr(im), n, direction, int(conj_out)), "cim_fft3d")
        return re, im

    # -- reconstruction ---------------------------------------------------
    def rec_begin(self, box: int, maxR: int, layout: int = 0, rcut: Optional[float] = None,
                  wg_init: float = 0.01, ops: Optional[np.ndarray] = None):
        if ops is None:
            nops, optr = 60, None
        else:
            o = np.ascontiguousarray(ops, dtype=np.float64).reshape(-1, 9)
            self._ops_keep = o
            nops, optr = o.shape[0], o.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rc = self.lib.cim_rec_begin(self.h, box, maxR, layout, float(maxR if rcut is None else rcut),
                                    float(wg_init), nops, optr)
        self._check(rc, "cim_rec_begin")

    def rec_insert(self, params: CimParams, images, parts: np.ndarray):
        parts = _f32(parts).reshape(-1, PART_STRIDE)
        n = parts.shape[0]
        if hasattr(images, "data_ptr"):       # device tensor
            rc = self.lib.cim_rec_insert_dev(self.h, ctypes.byref(params), n, ctypes.c_void_p(images.data_ptr()),
                                             _fptr(parts))
        else:
            im = _f32(images)
            rc = self.lib.cim_rec_insert(self.h, ctypes.byref(params), n, _fptr(im), _fptr(parts))
        self._check(rc, "cim_rec_insert")

    def rec_insert_ft(self, ft_re, ft_im, euler, ctf=None):
        re, im, eu = _f32(ft_re), _f32(ft_im), _f32(euler).reshape(-1, 3)
        c = None if ctf is None else _f32(ctf)
        rc = self.lib.cim_rec_insert_ft(self.h, eu.shape[0], _fptr(re), _fptr(im),
                                        None if c is None else _fptr(c), _fptr(eu))
        self._check(rc, "cim_rec_insert_ft")

    def rec_get_tnf(self, maxR: int):
        n = (2 * maxR + 1) ** 2 * (maxR + 1)
        fa, fb, wg = (np.empty(n, np.float32) for _ in range(3))
        self._check(self.lib.cim_rec_get_tnf(self.h, _fptr(fa), _fptr(fb), _fptr(wg)), "cim_rec_get_tnf")
        shape = (maxR + 1, 2 * maxR + 1, 2 * maxR + 1)
        return fa.reshape(shape), fb.reshape(shape), wg.reshape(shape)

    def rec_add_tnf(self, fa, fb, wg):
        fa, fb, wg = _f32(fa), _f32(fb), _f32(wg)
        self._check(self.lib.cim_rec_add_tnf(self.h, _fptr(fa), _fptr(fb), _fptr(wg)), "cim_rec_add_tnf")

    def rec_finish(self, box: int, out=None):
        if out is not None and hasattr(out, "data_ptr"):
            self._check(self.lib.cim_rec_finish_dev(self.h, ctypes.c_void_p(out.data_ptr())), "cim_rec_finish_dev")
            return out
        m = np.empty((box, box, box), np.float32)
        self._check(self.lib.cim_rec_finish(self.h, _fptr(m)), "cim_rec_finish")
        return m

    def merge_map(self, box: int, icosavg: bool = False, imgmask: float = 0.0) -> np.ndarray:
        """merge_reconstruct.f90 map of the accumulated tnf sums, (box-8)^3;
        icosavg: the header's nicosavg = 1 (merge's real-space average, radius imgmask)."""
        m = np.empty((box - 8,) * 3, np.float32)
        self._check(self.lib.cim_merge_map(self.h, int(bool(icosavg)), float(imgmask), _fptr(m)), "cim_merge_map")
        return m

    def fsc_map(self, box: int, maxR: int, odd, even, imgmask: float = 0.0, merge: bool = False,
                icosavg: bool = False):
        """Half-map FSC + FSC-weighted map: (fsc[maxR+1], map).  merge=False is
        reconstruct_FSC.f90 (box^3), merge=True merge_reconstruct_FSC.f90 ((box-8)^3);
        icosavg: realspaceavg=y / nicosavg = 1."""
        arrs = [_f32(a).reshape(-1) for a in (*odd, *even)]
        n = (2 * maxR + 1) ** 2 * (maxR + 1)
        if any(a.size != n for a in arrs):
            raise CimError(f"tnf arrays must hold (2maxR+1)^2 (maxR+1) = {n} values")
        f = np.empty(maxR + 1, np.float32)
        m = np.empty((box - 8,) * 3 if merge else (box,) * 3, np.float32)
        rc = self.lib.cim_fsc_map(self.h, box, maxR, float(imgmask), 1 if merge else 0, int(bool(icosavg)),
                                  *[_fptr(a) for a in arrs], _fptr(f), _fptr(m))
        self._check(rc, "cim_fsc_map")
        return f, m

    def icosavg(self, vol, r0: float, variant: int = 0, n: Optional[int] = None, ops=None):
        """Real-space average (cim_icosavg / cim_icosavg_dev) of the first n^3
        values of vol (n = its box) over ops (default: the 60 icosahedral
        operators): a new array for a host array, in place for a device tensor
        (the values past n^3 are read, not written)."""
        if ops is None:
            nops, optr = 0, None
        else:
            o = np.ascontiguousarray(ops, dtype=np.float64).reshape(-1, 9)
            nops, optr = o.shape[0], o.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        if hasattr(vol, "data_ptr"):
            nn = vol.shape[-1] if n is None else n
            self._check(self.lib.cim_icosavg_dev(self.h, ctypes.c_void_p(vol.data_ptr()), vol.numel(), int(nn),
                                                 float(r0), int(variant), nops, optr), "cim_icosavg_dev")
            return vol
        v = _f32(vol).copy()
        nn = v.shape[-1] if n is None else n
        self._check(self.lib.cim_icosavg(self.h, _fptr(v), v.size, int(nn), float(r0), int(variant), nops, optr),
                    "cim_icosavg")
        return v

    # -- local_box ----------------------------------------------------------
    def local_box_coords(self, parts, ops, x0, box: int, apix: float) -> np.ndarray:
        """cim_local_box_coords: (n, nops, LBOX_REC) records."""
        parts = _f32(parts).reshape(-1, PART_STRIDE)
        o = np.ascontiguousarray(ops, dtype=np.float64).reshape(-1, 9)
        x = _f32(x0).reshape(3)
        rec = np.zeros((parts.shape[0], o.shape[0], LBOX_REC), np.float32)
        rc = self.lib.cim_local_box_coords(self.h, parts.shape[0], _fptr(parts), o.shape[0],
                                           o.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _fptr(x), int(box),
                                           float(apix), _fptr(rec))
        self._check(rc, "cim_local_box_coords")
        return rec

    def local_box_stack(self, images, parts, ops, x0, sub: int, apix: float, boxes=None, box: Optional[int] = None):
        """cim_local_box_stack (host images) or cim_local_box_stack_dev (device
        arrays with data_ptr(): images of n*box^2 floats, boxes of
        n*nops*sub^2 floats).  box defaults to the images' last dimension.
        Returns (boxes (n, nops, sub, sub), records (n, nops, LBOX_REC))."""
        parts = _f32(parts).reshape(-1, PART_STRIDE)
        n = parts.shape[0]
        o = np.ascontiguousarray(ops, dtype=np.float64).reshape(-1, 9)
        optr = o.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        x = _f32(x0).reshape(3)
        rec = np.zeros((n, o.shape[0], LBOX_REC), np.float32)
        box = int(images.shape[-1] if box is None else box)
        if hasattr(images, "data_ptr"):
            if boxes is None or boxes.numel() != n * o.shape[0] * sub * sub:
                raise CimError("device images need a device boxes tensor of n*nops*sub^2 floats")
            rc = self.lib.cim_local_box_stack_dev(self.h, n, ctypes.c_void_p(images.data_ptr()), _fptr(parts), box,
                                                  o.shape[0], optr, _fptr(x), int(sub), float(apix),
                                                  ctypes.c_void_p(boxes.data_ptr()), _fptr(rec))
            self._check(rc, "cim_local_box_stack_dev")
            return boxes, rec
        im = _f32(images).reshape(n, box, box)
        out = np.zeros((n, o.shape[0], sub, sub), np.float32)
        rc = self.lib.cim_local_box_stack(self.h, n, _fptr(im), _fptr(parts), box, o.shape[0], optr, _fptr(x),
                                          int(sub), float(apix), _fptr(out), _fptr(rec))
        self._check(rc, "cim_local_box_stack")
        return out, rec

    def local_box_stack_noalloc(self, images, parts, ops, x0, sub: int, apix: float, boxes, box: int):
        """cim_local_box_stack_noalloc: as the device form of local_box_stack,
        with every working buffer in the library's static scratch array (no
        device allocation).  images, boxes: device arrays with data_ptr()."""
        parts = _f32(parts).reshape(-1, PART_STRIDE)
        n = parts.shape[0]
        o = np.ascontiguousarray(ops, dtype=np.float64).reshape(-1, 9)
        optr = o.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        x = _f32(x0).reshape(3)
        rec = np.zeros((n, o.shape[0], LBOX_REC), np.float32)
        if boxes.numel() != n * o.shape[0] * sub * sub:
            raise CimError("boxes must hold n*nops*sub^2 floats")
        rc = self.lib.cim_local_box_stack_noalloc(self.h, n, ctypes.c_void_p(images.data_ptr()), _fptr(parts),
                                                  int(box), o.shape[0], optr, _fptr(x), int(sub), float(apix),
                                                  ctypes.c_void_p(boxes.data_ptr()), _fptr(rec))
        self._check(rc, "cim_local_box_stack_noalloc")
        return boxes, rec

    def debug_cl_table(self, nh: int, ppol, tpol, T: int, tpol_per_particle: bool = True):
        """cim_debug_cl_table (include/cryoim_debug.h): residual tables from polar
        rays ppol (np, nr, npsi, 2) and tpol; returns (N, D) as float32 arrays
        (np*T, npsi/2, npsi) decoded from the bfloat16 pairs."""
        pp = _f32(ppol)
        tp = _f32(tpol)
        np_, nr, npsi = pp.shape[:3]
        out = np.zeros((np_ * T, npsi // 2, npsi), np.uint32)
        rc = self.lib.cim_debug_cl_table(self.h, int(nh), np_, int(T), int(bool(tpol_per_particle)), nr, npsi,
                                         _fptr(pp), _fptr(tp), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        self._check(rc, "cim_debug_cl_table")
        return (out << 16).view(np.float32), (out & 0xFFFF0000).view(np.float32)

    # -- search -------------------------------------------------------------
    def search_templates_model(self, params: CimParams, model, tort):
        """Template bank projected from the box^3 model (load_template.f90:32-33)."""
        m = _f32(model)
        to = _f32(tort).reshape(-1, 5)
        self._check(self.lib.cim_search_templates_model(self.h, ctypes.byref(params), to.shape[0], _fptr(m),
                                                        _fptr(to)), "cim_search_templates_model")

    def project_templates(self, model, angles):
        """Real-space projections (T, box, box) of the model (project_template2d.f90)."""
        m = _f32(model)
        ang = _f32(angles).reshape(-1, 3)
        box = m.shape[-1]
        out = np.zeros((ang.shape[0], box, box), np.float32)
        self._check(self.lib.cim_project_templates(self.h, box, ang.shape[0], _fptr(m), _fptr(ang), _fptr(out)),
                    "cim_project_templates")
        return out

    def search_templates(self, params: CimParams, images, tort):
        im = _f32(images)
        to = _f32(tort).reshape(-1, 5)
        self._check(self.lib.cim_search_templates(self.h, ctypes.byref(params), to.shape[0], _fptr(im), _fptr(to)),
                    "cim_search_templates")

    def search_global(self, params: CimParams, images, parts: np.ndarray) -> np.ndarray:
        parts = _f32(parts).reshape(-1, PART_STRIDE).copy()
        n = parts.shape[0]
        if hasattr(images, "data_ptr"):
            rc = self.lib.cim_search_global_dev(self.h, ctypes.byref(params), n, ctypes.c_void_p(images.data_ptr()),
                                                _fptr(parts))
        else:
            im = _f32(images)
            rc = self.lib.cim_search_global(self.h, ctypes.byref(params), n, _fptr(im), _fptr(parts))
        self._check(rc, "cim_search_global")
        return parts

    # -- learned proposals -------------------------------------------------------
    def search_candidates(self, params: CimParams, images, parts: np.ndarray, cand: np.ndarray,
                          omega: np.ndarray) -> np.ndarray:
        """cim_search_candidates_dev: exact scores of the (n, K) candidate AU
        points, each over an omega window centred on omega[i, k]."""
        parts = _f32(parts).reshape(-1, PART_STRIDE).copy()
        n = parts.shape[0]
        c = np.ascontiguousarray(cand, np.int32).reshape(n, -1)
        o = np.ascontiguousarray(omega, np.int32).reshape(n, -1)
        ip_ = ctypes.POINTER(ctypes.c_int)
        self._check(self.lib.cim_search_candidates_dev(self.h, ctypes.byref(params), n,
                                                       ctypes.c_void_p(images.data_ptr()), _fptr(parts), c.shape[1],
                                                       c.ctypes.data_as(ip_), o.ctypes.data_as(ip_)),
                    "cim_search_candidates_dev")
        return parts

    def cnn_forward(self, images, params: np.ndarray, out):
        """cim_cnn_forward: images (B, 128, 128) float32 and out (B, CNN_OUT)
        float32, both device tensors; params host float32."""
        prm = np.ascontiguousarray(params, np.float32)
        self._check(self.lib.cim_cnn_forward(self.h, int(images.shape[0]), ctypes.c_void_p(images.data_ptr()),
                                             _fptr(prm), ctypes.c_void_p(out.data_ptr())), "cim_cnn_forward")

    def debug_cnn_act(self, layer: int, batch: int) -> np.ndarray:
        """cim_debug_cnn_act (include/cryoim_debug.h): the kept activations of layer."""
        h = 128 >> layer
        out = np.empty((batch, h, h, 1 if layer == 0 else 64), np.float32)
        self._check(self.lib.cim_debug_cnn_act(self.h, int(layer), _fptr(out)), "cim_debug_cnn_act")
        return out

    def cnn_backward(self, dout) -> np.ndarray:
        """cim_cnn_backward: dout (B, CNN_OUT) device tensor; returns the
        parameter gradient (float32, host)."""
        g = np.empty(self.lib.cim_cnn_nparams(), np.float32)
        self._check(self.lib.cim_cnn_backward(self.h, ctypes.c_void_p(dout.data_ptr()), _fptr(g)), "cim_cnn_backward")
        return g

    def step_capture(self, psearch: CimParams, pinsert: CimParams, images, parts_dev):
        """cim_step_capture: search + insert of the device records parts_dev
        (torch float32 (n, PART_STRIDE) on the GPU) and device images,
        captured as one HIP graph."""
        n = parts_dev.shape[0]
        self._check(self.lib.cim_step_capture(self.h, ctypes.byref(psearch), ctypes.byref(pinsert), n,
                                              ctypes.c_void_p(images.data_ptr()),
                                              ctypes.c_void_p(parts_dev.data_ptr())), "cim_step_capture")

    def step_replay(self):
        self._check(self.lib.cim_step_replay(self.h), "cim_step_replay")

    def step_release(self):
        self._check(self.lib.cim_step_release(self.h), "cim_step_release")

    def search_pr_array(self, params: CimParams, image, part, shape) -> np.ndarray:
        im = _f32(image)
        pa = _f32(part).reshape(PART_STRIDE)
        out = np.empty(int(np.prod(shape)), np.float32)
        self._check(self.lib.cim_search_pr_array(self.h, ctypes.byref(params), _fptr(im), _fptr(pa), _fptr(out)),
                    "cim_search_pr_array")
        return out.reshape(shape)

    # -- local_refine ---------------------------------------------------------
    def lref_model(self, params: CimParams, vol):
        v = _f32(vol)
        self._check(self.lib.cim_lref_model(self.h, ctypes.byref(params), _fptr(v)), "cim_lref_model")

    def lref_refine(self, params: CimParams, images, parts, sym: str = "c1", convF: float = 0.8,
                    maxres: float = 10.0) -> np.ndarray:
        im = _f32(images)
        parts = _f32(parts).reshape(-1, PART_STRIDE).copy()
        box0 = im.shape[-1]
        self._check(self.lib.cim_lref_refine(self.h, ctypes.byref(params), parts.shape[0], _fptr(im), box0,
                                             _fptr(parts), sym.encode(), float(convF), float(maxres)),
                    "cim_lref_refine")
        return parts

    def lref_defocus(self, params: CimParams, images, parts, sign: float = 1.0, maxres: float = 10.0) -> np.ndarray:
        """refine_particle_df (cim_lref_defocus): records with the chosen z in
        slot P_Z and its residual in P_PR."""
        im = _f32(images)
        parts = _f32(parts).reshape(-1, PART_STRIDE).copy()
        box0 = im.shape[-1]
        self._check(self.lib.cim_lref_defocus(self.h, ctypes.byref(params), parts.shape[0], _fptr(im), box0,
                                              _fptr(parts), float(sign), float(maxres)), "cim_lref_defocus")
        return parts

    def center_model(self, params: CimParams, model):
        """calc_center's model transform (cim_center_model), kept on the device."""
        v = _f32(model)
        self._check(self.lib.cim_center_model(self.h, ctypes.byref(params), _fptr(v)), "cim_center_model")

    def center(self, params: CimParams, method: int, images, parts: np.ndarray):
        """icosprocess mode 3 (cim_center / cim_center_dev): returns the updated
        records (x, y, PR) and the unclipped fitted peak values."""
        parts = _f32(parts).reshape(-1, PART_STRIDE).copy()
        n = parts.shape[0]
        cc = np.zeros(n, np.float32)
        if hasattr(images, "data_ptr"):
            rc = self.lib.cim_center_dev(self.h, ctypes.byref(params), int(method), n,
                                         ctypes.c_void_p(images.data_ptr()), _fptr(parts), _fptr(cc))
        else:
            im = _f32(images)
            rc = self.lib.cim_center(self.h, ctypes.byref(params), int(method), n, _fptr(im), _fptr(parts), _fptr(cc))
        self._check(rc, "cim_center")
        return parts, cc

    def refine_local(self, params: CimParams, images, parts: np.ndarray) -> np.ndarray:
        parts = _f32(parts).reshape(-1, PART_STRIDE).copy()
        im = _f32(images)
        self._check(self.lib.cim_refine_local(self.h, ctypes.byref(params), parts.shape[0], _fptr(im), _fptr(parts)),
                    "cim_refine_local")
        return parts
