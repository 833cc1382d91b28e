"""ctypes front end of the C oracle (oracle/cim_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Never part of the product
path.  ``Oracle(64)`` is the FP64 oracle, ``Oracle(32)`` the FP32-faithful
restatement.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class Oracle:
    def __init__(self, bits: int = 64):
        path = os.path.join(HERE, f"liboracle{bits}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        self.L = ctypes.CDLL(path)
        self.dt = np.float64 if bits == 64 else np.float32
        self.ct = ctypes.c_double if bits == 64 else ctypes.c_float
        assert self.L.o_real_size() == np.dtype(self.dt).itemsize
        ret_real = ["o_crosslin_pr", "o_refine_particle", "o_pr_ft"]
        for f in ret_real:
            getattr(self.L, f).restype = self.ct

    # helpers
    def _a(self, x):
        return np.ascontiguousarray(x, dtype=self.dt)

    def _p(self, a):
        return a.ctypes.data_as(ctypes.c_void_p)

    def _r(self, v):
        return self.ct(float(v))

    # ---- geometry --------------------------------------------------------
    def euler_mat(self, th, ph, om):
        R = np.zeros(9, self.dt)
        self.L.o_euler_mat(self._r(th), self._r(ph), self._r(om), self._p(R))
        return R.reshape(3, 3)

    def icos_ops(self):
        o = np.zeros(540, self.dt)
        self.L.o_icos_ops(self._p(o))
        return o.reshape(60, 3, 3)

    def crosslin_angle(self, a1, a2, ops=None):
        ops = self.icos_ops() if ops is None else self._a(ops)
        ops = self._a(ops).reshape(-1, 9)
        n = ops.shape[0]
        al1, al2 = np.zeros(n, self.dt), np.zeros(n, self.dt)
        self.L.o_crosslin_angle(*[self._r(v) for v in (*a1, *a2)], n, self._p(ops), self._p(al1), self._p(al2))
        return al1, al2

    def return_au(self, th, ph, om):
        v = [self.ct(float(th)), self.ct(float(ph)), self.ct(float(om))]
        self.L.o_return_au(ctypes.byref(v[0]), ctypes.byref(v[1]), ctypes.byref(v[2]))
        return v[0].value, v[1].value, v[2].value

    # ---- images ------------------------------------------------------------
    def ctf2d(self, nx, df1, df2, astig, apix, vol=300.0, cs=2.7, ampwg=0.1, df_scale=1.0e4):
        c = np.zeros(nx * nx, self.dt)
        self.L.o_getctf2d(nx, *[self._r(v) for v in (df1, df2, astig, apix, vol, cs, ampwg, df_scale)], self._p(c))
        return c.reshape(nx, nx)

    def mask(self, img, mask, ox, oy):
        a = self._a(img).copy()
        self.L.o_img2d_mask(self._p(a), a.shape[-1], self._r(mask), self._r(ox), self._r(oy))
        return a

    def norm(self, img):
        a = self._a(img).copy()
        self.L.o_img2d_norm(self._p(a), a.shape[-1])
        return a

    def checker(self, img):
        a = self._a(img).copy()
        self.L.o_checker2d(self._p(a), a.shape[-1])
        return a

    def fft2(self, a, b, direction=1, conj_out=True):
        a, b = self._a(a).copy(), self._a(b).copy()
        self.L.o_img2d_fft(self._p(a), self._p(b), a.shape[-1], direction, int(conj_out))
        return a, b

    def fft3(self, a, b, direction=1, conj_out=False):
        a, b = self._a(a).copy(), self._a(b).copy()
        self.L.o_img3d_fft(self._p(a), self._p(b), a.shape[-1], direction, int(conj_out))
        return a, b

    def dft2_brute(self, re, im, sign):
        re, im = np.ascontiguousarray(re, np.float64), np.ascontiguousarray(im, np.float64)
        o1, o2 = np.zeros_like(re), np.zeros_like(im)
        self.L.o_dft2d_brute(self._p(re), self._p(im), re.shape[-1], sign, self._p(o1), self._p(o2))
        return o1, o2

    def ft_shift(self, a, b, x0, y0, recentre=True):
        a, b = self._a(a).copy(), self._a(b).copy()
        self.L.o_ft_shift(self._p(a), self._p(b), a.shape[-1], self._r(x0), self._r(y0), int(recentre))
        return a, b

    def amp_expfit(self, a, b, minR, maxR):
        a, b = self._a(a), self._a(b)
        f = np.zeros(maxR - minR + 1, self.dt)
        self.L.o_amp_expfit(self._p(a), self._p(b), a.shape[-1], int(minR), int(maxR), self._p(f))
        return f

    # ---- pipelines -------------------------------------------------------------
    def particle_ft(self, img, x, y, mask_r=None, mask_xy=None, norm=False, conj_out=True, shift=True,
                    mask_first=False):
        """norm -> mask -> checkerboard -> FFT -> centre shift (reconstruct.f90:79-96);
        mask_first: mask, then norm (local_reconstruct/SRC/reconstruct_substep.f90:93-94)."""
        a = self._a(img).copy()
        if norm and not mask_first:
            a = self.norm(a)
        if mask_r is not None:
            mx, my = (x, y) if mask_xy is None else mask_xy
            a = self.mask(a, mask_r, mx, my)
        if norm and mask_first:
            a = self.norm(a)
        a = self.checker(a)
        fa, fb = self.fft2(a, np.zeros_like(a), 1, conj_out)
        if shift:
            fa, fb = self.ft_shift(fa, fb, x, y, True)
        return fa, fb

    def insert(self, vol, a, b, ctf, angles, maxR, rcut=None, compact=False, ops=None):
        fa, fb, wg = vol
        nx = a.shape[-1]
        ops = self.icos_ops() if ops is None else self._a(ops)
        ops = self._a(ops).reshape(-1, 9)
        R = self._a(self.euler_mat(*angles)).reshape(9)
        c = None if ctf is None else self._a(ctf)
        self.L.o_insert(self._p(fa), self._p(fb), self._p(wg), int(compact), nx, int(maxR),
                        self._r(maxR if rcut is None else rcut), self._p(self._a(a)), self._p(self._a(b)),
                        None if c is None else self._p(c), self._p(R), ops.shape[0], self._p(ops))

    def insert_series(self, vol, a, b, ctf3d, ctf_last, angles, maxR, rcut=None, compact=False, ops=None):
        """reconstruct_corrCTF's insert: ctf3d (ncyc, nx, nx) for fa, fb; ctf_last^2 per cycle for wg."""
        fa, fb, wg = vol
        nx = a.shape[-1]
        ops = self._a(self.icos_ops() if ops is None else ops).reshape(-1, 9)
        R = self._a(self.euler_mat(*angles)).reshape(9)
        c3 = self._a(ctf3d)
        self.L.o_insert_series(self._p(fa), self._p(fb), self._p(wg), int(compact), nx, int(maxR),
                               self._r(maxR if rcut is None else rcut), self._p(self._a(a)), self._p(self._a(b)),
                               self._p(c3), c3.shape[0], self._p(self._a(ctf_last)), self._p(R), ops.shape[0],
                               self._p(ops))

    def new_volume(self, nx, maxR, compact=False, wg_init=0.01):
        if compact:
            shape = (maxR + 1, 2 * maxR + 1, 2 * maxR + 1)
        else:
            shape = (nx, nx, nx)
        return (np.zeros(shape, self.dt), np.zeros(shape, self.dt), np.full(shape, wg_init, self.dt))

    def fold_and_map(self, vol, maxR, int_radius=False, imag_sign=-1):
        """Fold + normalise + inverse FFT of a full-box volume (reconstruct.f90:151-247;
        int_radius/imag_sign select the reconstruct_FSC / merge variants)."""
        fa, fb, wg = (self._a(v).copy() for v in vol)
        nx = fa.shape[-1]
        self.L.o_fold_normalise(self._p(fa), self._p(fb), self._p(wg), nx, int(maxR), int(bool(int_radius)))
        m = np.zeros_like(fa)
        self.L.o_volume_to_map(self._p(fa), self._p(fb), nx, int(imag_sign), self._p(m))
        return m

    def embed_tnf(self, fa, fb, wg, nx, maxR, wg_fill=0.0):
        """Place compact tnf arrays ((maxR+1, 2maxR+1, 2maxR+1), lz slowest) at the
        centre of an nx^3 volume (lz >= 0 half); weights elsewhere wg_fill."""
        c = nx // 2
        out = [np.zeros((nx, nx, nx), self.dt) for _ in range(2)] + [np.full((nx, nx, nx), wg_fill, self.dt)]
        sl = (slice(c, c + maxR + 1), slice(c - maxR, c + maxR + 1), slice(c - maxR, c + maxR + 1))
        for o, v in zip(out, (fa, fb, wg)):
            o[sl] = self._a(v).reshape(maxR + 1, 2 * maxR + 1, 2 * maxR + 1)
        return tuple(out)

    LOCAL_SYM = {"c1": (1, 0.0), "c2": (2, 180.0), "c3": (3, 120.0), "c4": (4, 90.0), "c5a": (5, 72.0),
                 "c6": (6, 60.0), "c7": (7, -1.0), "c8": (8, 45.0), "c12": (12, 30.0), "c15": (15, 24.0)}

    def icosprocess_local_ops(self, sym):
        """icosprocess's cyclic groups (icos2f_sym_mat.f90:250-540), (N, 3, 3)."""
        N, step = self.LOCAL_SYM[sym]
        o = np.zeros(9 * N, self.dt)
        self.L.o_icosprocess_local_ops(N, ctypes.c_float(step), self._p(o))
        return o.reshape(N, 3, 3)

    def tnf_fsc(self, odd, even, maxR):
        """reconstruct_FSC.f90:30-97: (raw FSC per shell, merged fa, fb, wg)."""
        ins = [self._a(v).reshape(-1).copy() for v in (*odd, *even)]
        n = ins[0].size
        outs = [np.zeros(n, self.dt) for _ in range(3)]
        f = np.zeros(maxR + 1, self.dt)
        self.L.o_tnf_fsc(int(maxR), *[self._p(v) for v in ins], self._p(f), *[self._p(o) for o in outs])
        return (f, *outs)

    def fsc_map(self, odd, even, nx, maxR, imgmask=None, merge=False, icosavg=False):
        """Map of reconstruct_FSC.f90 (merge=False, nx^3) or of
        merge_reconstruct_FSC.f90 (merge=True: masked, then the first (nx-8)^3
        values of the flat array, as its writer at :204-206 emits them).
        icosavg: realspaceavg / nicosavg = 1 -- icosprocess's img3d_icosavg on
        the map (reconstruct_FSC.f90:185), or merge's on the first (nx-8)^3
        values taken as an (nx-8)^3 cube (merge_reconstruct_FSC.f90:202, after
        maphead has set FFTsize = nx-8)."""
        f, fa, fb, wg = self.tnf_fsc(odd, even, maxR)
        vol = self.embed_tnf(fa, fb, wg, nx, maxR)
        if not merge:
            m = self.fold_and_map(vol, maxR, int_radius=True, imag_sign=+1)
            if icosavg:
                m = self.icosavg_full(m, imgmask)
            return f, m
        m = self.fold_and_map(vol, maxR, int_radius=True, imag_sign=-1)
        self.L.o_sphere_zero(self._p(m), nx, self._r(imgmask))
        M = nx - 8
        if icosavg:
            m = self.icosavg_merge(m, imgmask, n=M)
        return f, m.reshape(-1)[: M ** 3].reshape(M, M, M).copy()

    def merge_map(self, tnfs, nx, maxR, imgmask=None, icosavg=False):
        """merge_reconstruct.f90: sum of tnf triples onto weights 0.01, integer-radius
        fold, inverse FFT, [merge's img3d_icosavg when nicosavg = 1, :142], then
        the (nx-8)^3 crop of :143-149."""
        T = 2 * maxR + 1
        fa, fb = np.zeros((maxR + 1, T, T), self.dt), np.zeros((maxR + 1, T, T), self.dt)
        wg = np.full((maxR + 1, T, T), 0.01, self.dt)
        for t in tnfs:
            fa += self._a(t[0]).reshape(fa.shape)
            fb += self._a(t[1]).reshape(fa.shape)
            wg += self._a(t[2]).reshape(fa.shape)
        vol = self.embed_tnf(fa, fb, wg, nx, maxR, wg_fill=0.01)
        m = self.fold_and_map(vol, maxR, int_radius=True, imag_sign=-1)
        if icosavg:
            m = self.icosavg_merge(m, imgmask)
        out = np.zeros((nx - 8) ** 3, self.dt)
        self.L.o_merge_crop(self._p(m), nx, self._p(out))
        return out.reshape(nx - 8, nx - 8, nx - 8)

    # ---- common lines ------------------------------------------------------------
    def search_grid(self, step):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.L.o_search_grid(self._r(step), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def pr_omega(self, pa, pb, fcut0, ta, tb, tfcut, teul, minR, maxR, theta, phi, step):
        nx = pa.shape[-1]
        _, _, No = self.search_grid(step)
        ops = self.icos_ops().reshape(-1)
        out = np.zeros(No, self.dt)
        T = len(teul)
        self.L.o_crosslin_pr_omega(self._p(self._a(pa)), self._p(self._a(pb)), self._p(self._a(fcut0)), T,
                                   self._p(self._a(ta)), self._p(self._a(tb)), self._p(self._a(tfcut)),
                                   self._p(self._a(teul)), nx, int(minR), int(maxR), 60, self._p(self._a(ops)),
                                   self._r(theta), self._r(phi), No, self._r(step), self._p(out))
        return out

    def globesearch(self, pa, pb, fcut0, ta, tb, tfcut, teul, minR, maxR, step, want_array=False):
        nx = pa.shape[-1]
        Nt, Np, No = self.search_grid(step)
        ops = self._a(self.icos_ops().reshape(-1))
        out4 = np.zeros(4, self.dt)
        arr = np.zeros((Nt + 1) * (2 * Np + 1) * No, self.dt) if want_array else None
        self.L.o_globesearch(self._p(self._a(pa)), self._p(self._a(pb)), self._p(self._a(fcut0)), len(teul),
                             self._p(self._a(ta)), self._p(self._a(tb)), self._p(self._a(tfcut)),
                             self._p(self._a(teul)), nx, int(minR), int(maxR), 60, self._p(ops), self._r(step),
                             self._p(out4), None if arr is None else self._p(arr))
        if want_array:
            return out4, arr.reshape(Nt + 1, 2 * Np + 1, No)
        return out4

    def crosslin_pr(self, pa, pb, fcut0, ta, tb, tfcut, teul, minR, maxR, th, ph, om, x, y):
        nx = pa.shape[-1]
        ops = self._a(self.icos_ops().reshape(-1))
        return self.L.o_crosslin_pr(self._p(self._a(pa)), self._p(self._a(pb)), self._p(self._a(fcut0)), len(teul),
                                    self._p(self._a(ta)), self._p(self._a(tb)), self._p(self._a(tfcut)),
                                    self._p(self._a(teul)), nx, int(minR), int(maxR), 60, self._p(ops),
                                    *[self._r(v) for v in (th, ph, om, x, y)])

    def refine_particle(self, pa, pb, fcut0, ta, tb, tfcut, teul, minR, maxR, ncycle, dang, dtran, start5):
        nx = pa.shape[-1]
        ops = self._a(self.icos_ops().reshape(-1))
        io5 = self._a(start5).copy()
        pr = self.L.o_refine_particle(self._p(self._a(pa)), self._p(self._a(pb)), self._p(self._a(fcut0)),
                                      len(teul), self._p(self._a(ta)), self._p(self._a(tb)),
                                      self._p(self._a(tfcut)), self._p(self._a(teul)), nx, int(minR), int(maxR), 60,
                                      self._p(ops), ncycle, self._r(dang), self._r(dtran), self._p(io5))
        return io5, pr

    # ---- local refine ----------------------------------------------------------------
    def fft3d_slice(self, fr, fi, th, ph, om):
        nx = fr.shape[-1]
        a, b = np.zeros((nx, nx), self.dt), np.zeros((nx, nx), self.dt)
        self.L.o_fft3d_slice(self._p(self._a(fr)), self._p(self._a(fi)), nx, self._r(th), self._r(ph),
                             self._r(om), self._p(a), self._p(b))
        return a, b

    def pr_ft(self, a1, b1, a2, b2, minR, maxR):
        nx = a1.shape[-1]
        return self.L.o_pr_ft(*[self._p(self._a(v)) for v in (a1, b1, a2, b2)], nx, self._r(minR), self._r(maxR))

    def softmask3d(self, vol, mask):
        v = self._a(vol).copy()
        self.L.o_img3d_softmask(self._p(v), v.shape[-1], self._r(mask))
        return v

    # ---- local_refine ----------------------------------------------------------------
    def local_box(self, img, hx0, ky0, n):
        a = self._a(img)
        out = np.zeros((n, n), self.dt)
        self.L.o_local_box(self._p(a), a.shape[-1], int(hx0), int(ky0), self._p(out), int(n))
        return out

    def lref_symops(self, kind, n):
        r = np.zeros(9 * max(n, 60), self.dt)
        m = self.L.o_lref_symops(int(kind), int(n), self._p(r))
        return r[:9 * m].reshape(m, 3, 3)

    def icos_equalort(self, th, ph, om, ops):
        ops = self._a(ops).reshape(-1, 9)
        out = np.zeros(3 * len(ops), self.dt)
        self.L.o_icos_equalort(self._r(th), self._r(ph), self._r(om), len(ops), self._p(ops), self._p(out))
        return out.reshape(-1, 3)

    def lref_refine(self, fr, fi, a20, b20, ctf, wlp, ops, ncycle, refine_angle, refine_shift, convF, maxres,
                    start5):
        nx = a20.shape[-1]
        io5 = self._a(start5).copy()
        if ops is None:
            nsym, optr = 0, None
        else:
            o = self._a(ops).reshape(-1, 9)
            nsym, optr = len(o), self._p(o)
        self.L.o_lref_refine.restype = self.ct
        pr = self.L.o_lref_refine(self._p(self._a(fr)), self._p(self._a(fi)), nx, self._p(self._a(a20)),
                                  self._p(self._a(b20)), self._p(self._a(ctf)), self._p(self._a(wlp)), nsym, optr,
                                  int(ncycle), self._r(refine_angle), self._r(refine_shift), self._r(convF),
                                  self._r(maxres), self._p(io5))
        return io5, pr

    def lref_defocus(self, fr, fi, a20, b20, wlp, angles, x0, y0, z0, df1, df2, astig, apix, sign, maxres,
                     vol=200.0, cs=2.7, ampwg=0.1):
        """refine_particle_df for one sub-particle: (z, PR, the 401 residuals)."""
        nx = a20.shape[-1]
        out = np.zeros(2, self.dt)
        prs = np.zeros(401, self.dt)
        self.L.o_lref_defocus(self._p(self._a(fr)), self._p(self._a(fi)), nx, self._p(self._a(a20)),
                              self._p(self._a(b20)), self._p(self._a(wlp)),
                              *[self._r(v) for v in (*angles, x0, y0, z0, df1, df2, astig, apix, vol, cs, ampwg,
                                                     sign, maxres)], self._p(out), self._p(prs))
        return float(out[0]), float(out[1]), prs

    # ---- centre finding (icosprocess mode 3) -------------------------------
    def img3d_mask(self, vol, mask, origin=(0.0, 0.0, 0.0)):
        v = self._a(vol).copy()
        self.L.o_img3d_mask(self._p(v), v.shape[-1], self._r(mask), *[self._r(o) for o in origin])
        return v

    def center_model(self, vol, mask=None):
        """calc_center.f90:19-41: (fr, fi) of the (masked) model, fi = -Im."""
        v = self._a(vol).copy()
        fr, fi = np.zeros_like(v), np.zeros_like(v)
        self.L.o_center_model(self._p(v), v.shape[-1], self._r(-1.0 if mask is None else mask), self._p(fr),
                              self._p(fi))
        return fr, fi

    def lowpass(self, img, lp):
        a = self._a(img).copy()
        self.L.o_img2d_lowpass(self._p(a), a.shape[-1], int(lp))
        return a

    def apply_ctf(self, img, ctf):
        a = self._a(img).copy()
        self.L.o_img2d_applyctf(self._p(a), a.shape[-1], self._p(self._a(ctf)))
        return a

    def project_template2d(self, vol, angles):
        """project_template2d.f90:3-80: the real-space projection (nx, nx) of
        the nx^3 model at orientation angles (theta, phi, omega)."""
        v = self._a(vol)
        n = v.shape[-1]
        R = self._a(self.euler_mat(*angles)).reshape(9)
        out = np.zeros((n, n), self.dt)
        self.L.o_project_template2d(self._p(v), n, self._p(R), self._p(out))
        return out

    def project2d_ft(self, fr, fi, th, ph, om):
        fr, fi = self._a(fr), self._a(fi)
        n = fr.shape[-1]
        out = np.zeros((n, n), self.dt)
        self.L.o_project2d_ft(self._p(fr), self._p(fi), n, self._r(th), self._r(ph), self._r(om), self._p(out))
        return out

    def cross_correlation(self, a1, a2):
        a, b = self._a(a1).copy(), self._a(a2).copy()
        self.L.o_cross_correlation(self._p(a), self._p(b), a.shape[-1])
        return a

    def gauss_regression(self, z):
        """z[j+1, i+1] = z(i, j) of guass_regression.f90; returns (dx, dy, cc)."""
        zz = self._a(z).reshape(9)
        dx, dy, cc = self.ct(), self.ct(), self.ct()
        self.L.o_gauss_regression(self._p(zz), ctypes.byref(dx), ctypes.byref(dy), ctypes.byref(cc))
        return dx.value, dy.value, cc.value

    def center_peak(self, ccmap, halve=False):
        c = self._a(ccmap)
        out = np.zeros(3, self.dt)
        self.L.o_center_peak(self._p(c), c.shape[-1], int(halve), self._p(out))
        return out

    def calc_center_one(self, fr, fi, img, x, y, angles, ctf, mask, maxR):
        """calc_center.f90:44-76: {x, y, cc} of one particle (cc not clipped)."""
        a = self._a(img).copy()
        out = np.zeros(3, self.dt)
        cp = None if ctf is None else self._p(self._a(ctf))
        self.L.o_calc_center_one(self._p(self._a(fr)), self._p(self._a(fi)), self._p(a), a.shape[-1], self._r(x),
                                 self._r(y), *[self._r(v) for v in angles], cp,
                                 self._r(-1.0 if mask is None else mask), int(maxR), self._p(out))
        return out

    def center_selfrot180_one(self, img, mask, maxR):
        """calc_center_selfrot180.f90:33-79: {x, y, cc} of one particle."""
        a = self._a(img).copy()
        out = np.zeros(3, self.dt)
        self.L.o_center_selfrot180_one(self._p(a), a.shape[-1], self._r(-1.0 if mask is None else mask), int(maxR),
                                       self._p(out))
        return out

    # ---- real-space icosahedral averaging --------------------------------
    def icosavg_full(self, vol, r0, ops=None, n=None):
        """icosprocess img3d_icosavg on the first n^3 values of vol (n = its box)."""
        v = self._a(vol).copy()
        flat = v.reshape(-1)
        n = v.shape[-1] if n is None else n
        o = self._a(self.icos_ops() if ops is None else ops).reshape(-1, 9)
        self.L.o_img3d_icosavg_full(self._p(flat), flat.size, int(n), self._r(r0), self._p(o), len(o))
        return v

    def icosavg_merge(self, vol, r0, n=None):
        """merge_icosprocess_subrec img3d_icosavg on the first n^3 values of vol."""
        v = self._a(vol).copy()
        flat = v.reshape(-1)
        n = v.shape[-1] if n is None else n
        o = self._a(self.icos_ops()).reshape(-1, 9)
        self.L.o_img3d_icosavg_merge(self._p(flat), flat.size, int(n), self._r(r0), self._p(o))
        return v

    # ---- local_reconstruct ---------------------------------------------------
    @staticmethod
    def local_reconstruct_sym(sym):
        """(kind, N) of local_reconstruct/SRC/0calc_initialvalues.f90:33-62:
        I2 / I3 / I5 (60 operators) or C_matrix with N from c1..c15 (the
        spellings it accepts; a sym it does not know keeps Nsym = 1)."""
        s = sym.strip()
        for k in (2, 3, 5):
            if s in (f"i{k}", f"I{k}"):
                return k, 60
        for n in range(1, 16):
            names = (f"c{n}", f"C{n}") if n < 10 else (f"c{n}", f"C_{n}")
            if s in names:
                return 0, n
        return 0, 1

    def local_reconstruct_ops(self, sym):
        kind, n = self.local_reconstruct_sym(sym)
        r = np.zeros(9 * 60, self.dt)
        m = self.L.o_local_reconstruct_ops(kind, n, self._p(r))
        return r[:9 * m].reshape(m, 3, 3)

    # ---- local_box -------------------------------------------------------------------
    @staticmethod
    def local_box_sym(sym):
        """(kind, N) of local_box/SRC/0main.f90:80-122: I2 / I3 / I5 (60
        operators) or C_matrix with N from c1..c15 (either case); a sym it
        does not know keeps the default Nsym = 60 and goes to C_matrix."""
        s = sym.strip()
        for k in (2, 3, 5):
            if s in (f"i{k}", f"I{k}"):
                return k, 60
        for n in range(1, 16):
            if s in (f"c{n}", f"C{n}"):
                return 0, n
        return 0, 60

    def local_box_ops(self, sym):
        kind, n = self.local_box_sym(sym)
        r = np.zeros(9 * 60, self.dt)
        m = self.L.o_local_reconstruct_ops(kind, n, self._p(r))
        return r[:9 * m].reshape(m, 3, 3)

    def img2d_realshift(self, img, x0, y0):
        a = self._a(img).copy()
        self.L.o_img2d_realshift(self._p(a), a.shape[-1], self._r(x0), self._r(y0))
        return a

    def local_box_coords(self, ang, xy, ops, x0, nx, apix):
        """Per operator {lth, lph, lom, hx0, ky0, frac x, frac y, z apix}."""
        ops = self._a(ops).reshape(-1, 9)
        out = np.zeros((len(ops), 8), self.dt)
        self.L.o_local_box_coords(*[self._r(v) for v in (*ang, *xy)], len(ops), self._p(ops),
                                  self._p(self._a(x0)), int(nx), self._r(apix), self._p(out))
        return out

    def local_box_stack(self, img, ang, xy, ops, x0, sub, apix):
        """(boxes [nops, sub, sub], meta [nops, {lth, lph, lom, frac x, frac y, z apix}])."""
        a = self._a(img)
        ops = self._a(ops).reshape(-1, 9)
        boxes = np.zeros((len(ops), sub, sub), self.dt)
        meta = np.zeros((len(ops), 6), self.dt)
        self.L.o_local_box_stack(self._p(a), a.shape[-1], *[self._r(v) for v in (*ang, *xy)], len(ops),
                                 self._p(ops), self._p(self._a(x0)), int(sub), self._r(apix), self._p(boxes),
                                 self._p(meta))
        return boxes, meta
