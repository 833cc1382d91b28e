/*
 * cim_oracle.c -- CPU restatement of cryoIM's icosahedral search and
 * direct-Fourier-inversion arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library
 * is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * It is compiled twice by oracle/Makefile:
 *   liboracle64.so  REAL=double  (the FP64 oracle every error is stated against)
 *   liboracle32.so  REAL=float   (FP32-faithful: same operation order as the
 *                                 Fortran, which computes in default REAL*4)
 * FFTs are always evaluated in double precision (own mixed-radix FFT,
 * pinned by a brute-force DFT in the tests); the FFT provider of the
 * reference (Intel MKL's FFTW3 wrapper) is not available, so FFT rounding
 * parity is "unpinned" and only the convention (sign, 1/N scaling,
 * conjugated output) is restated.
 *
 * Layout: images are nx*nx row-major with the x index fastest, i.e. the
 * same memory order as the reference's flat column-major a(i),
 * i = ky*nx + hx + 1.  Complex data are split re/im arrays (a, b), exactly
 * as in the reference.  Rotation matrices are row-major R[3*i+j] =
 * RotMat(i+1, j+1).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL double
#endif

#define EXPORT __attribute__((visibility("default")))

typedef REAL real;

static real r_pi(void) { return (real)4.0 * (real)atan2((real)1.0, (real)1.0); }
static real r_deg2rad(void) { return (real)atan2((real)1.0, (real)1.0) / (real)45.0; }

EXPORT int o_real_size(void) { return (int)sizeof(real); }

/* ------------------------------------------------------------------ */
/* Rotations                                                           */
/* ------------------------------------------------------------------ */

/* euler_mat: icosprocess/SRC/euler_mat.f90:3-21 (ZYZ, degrees). */
EXPORT void o_euler_mat(real theta0, real phi0, real omega0, real* R)
{
    real d2r = r_deg2rad();
    real th = theta0 * d2r, ph = phi0 * d2r, om = omega0 * d2r;
    real c1 = (real)cos(ph), c2 = (real)cos(th), c3 = (real)cos(om);
    real s1 = (real)sin(ph), s2 = (real)sin(th), s3 = (real)sin(om);
    R[0] = c1 * c2 * c3 - s1 * s3;  R[1] = -c1 * c2 * s3 - c3 * s1; R[2] = c1 * s2;
    R[3] = c2 * c3 * s1 + c1 * s3;  R[4] = -c2 * s1 * s3 + c1 * c3; R[5] = s1 * s2;
    R[6] = -c3 * s2;                R[7] = s2 * s3;                 R[8] = c2;
}

/* 60 icosahedral rotations, 2-fold axes on x, y, z and 5-fold axes along
 * (+-1, 0, +-phi) and its cyclic permutations -- the orientation of
 * icos2f_sym_mat (icosprocess/SRC/icos2f_sym_mat.f90:1-248).  The group is
 * generated here by closure of two generators and then put into the
 * reference's element order with an order table; see cim_symops.h. */
#include "../cryoim_amd/csrc/cim_symops.h"

EXPORT void o_icos_ops(real* ops)
{
    double m[540];
    cim_icos60_ref_order(m);
    for (int i = 0; i < 540; ++i) ops[i] = (real)m[i];
}

static void matvec(const real* M, const real* x, real* y)
{
    for (int i = 0; i < 3; ++i) y[i] = M[3 * i] * x[0] + M[3 * i + 1] * x[1] + M[3 * i + 2] * x[2];
}
static void mattvec(const real* M, const real* x, real* y)
{
    for (int i = 0; i < 3; ++i) y[i] = M[i] * x[0] + M[3 + i] * x[1] + M[6 + i] * x[2];
}
static void matmul3(const real* A, const real* B, real* C)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

/* icos_crosslin_angle: icosprocess/SRC/icos_crosslin_angle.f90:2-37.
 * nops = 60 (icosahedral) or the number of rows of ops.  alpha = 1000 marks
 * a skipped line (|z1 x S z2| < 0.5, :25-29). */
EXPORT void o_crosslin_angle(real th1, real ph1, real om1, real th2, real ph2, real om2,
                             int nops, const real* ops, real* alpha1, real* alpha2)
{
    real r2d = (real)45.0 / (real)atan2((real)1.0, (real)1.0);
    real E1[9], E2[9], z0[3] = {0, 0, 1}, z1[3], z2[3];
    o_euler_mat(th1, ph1, om1, E1);
    o_euler_mat(th2, ph2, om2, E2);
    matvec(E1, z0, z1);
    matvec(E2, z0, z2);
    for (int i = 0; i < nops; ++i) {
        const real* S = ops + 9 * i;
        real z2a[3], v[3], v10[3], t[3], v20[3];
        matvec(S, z2, z2a);
        v[0] = z1[1] * z2a[2] - z1[2] * z2a[1];
        v[1] = z1[2] * z2a[0] - z1[0] * z2a[2];
        v[2] = z1[0] * z2a[1] - z1[1] * z2a[0];
        real vv = (real)sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (vv < (real)0.5) { alpha1[i] = 1000; alpha2[i] = 1000; continue; }
        v[0] /= vv; v[1] /= vv; v[2] /= vv;
        mattvec(E1, v, v10);
        mattvec(S, v, t);
        mattvec(E2, t, v20);
        alpha1[i] = (real)atan2(v10[1], v10[0]) * r2d;
        alpha2[i] = (real)atan2(v20[1], v20[0]) * r2d;
    }
}

/* Asymmetric-unit test used by ort_globesearch.f90:17-18,80-82 and
 * return_icos_asymmetry_unit.f90:9-12,22-27.  Evaluated in float32 in both
 * oracle builds: it is a discrete decision the reference takes in REAL*4,
 * and at theta = 90 deg cos() rounds to -4.4e-8 < 0, so REAL*4 excludes
 * the whole equator row of the search grid where float64 would keep it. */
static int in_au(real xr, real yr, real zr)
{
    float x = (float)xr, y = (float)yr, z = (float)zr;
    float d2r = atan2f(1.0f, 1.0f) / 45.0f;
    float s31 = sinf(31.717f * d2r), c31 = cosf(31.717f * d2r);
    float s69 = sinf(69.09f * d2r), c69 = cosf(69.09f * d2r);
    float A1 = -s31 * c69, B1 = -c31 * c69, C1 = s31 * s69;
    float A2 = -s31 * c69, B2 = c31 * c69, C2 = s31 * s69;
    float z1 = -(A1 * x + B1 * y) / C1;
    float z2 = -(A2 * x + B2 * y) / C2;
    return (z >= 0) && (z <= z1) && (z <= z2);
}

/* return_icos_asymmetry_unit: icosprocess/SRC/return_icos_asymmetry_unit.f90:1-33.
 * Falls through to the last operator when none lands in the AU (:17-29). */
EXPORT void o_return_au(real* theta, real* phi, real* omega)
{
    real ops[540], E[9], M[9];
    real d2r = r_deg2rad();
    o_icos_ops(ops);
    o_euler_mat(*theta, *phi, *omega, E);
    real th = 0, ph = 0, om = 0;
    for (int i = 0; i < 60; ++i) {
        matmul3(ops + 9 * i, E, M);
        th = (real)acos(M[8]);
        ph = (real)atan2(M[5], M[2]);
        om = (real)atan2(M[7], -M[6]);
        real x = (real)sin(th) * (real)cos(ph), y = (real)sin(th) * (real)sin(ph), z = (real)cos(th);
        if (in_au(x, y, z)) break;
    }
    *theta = th / d2r;
    *phi = ph / d2r;
    *omega = om / d2r;
}

/* ------------------------------------------------------------------ */
/* FFT (double precision, own mixed radix)                             */
/* ------------------------------------------------------------------ */

/* out[k] = sum_j in[j*stride] exp(sign*2*pi*i*j*k/n); recursive
 * decimation in time over the smallest prime factor. */
static void dft_rec(const double* inr, const double* ini, int n, int stride, int sign,
                    double* outr, double* outi, double* tmpr, double* tmpi)
{
    if (n == 1) { outr[0] = inr[0]; outi[0] = ini[0]; return; }
    int p = 0;
    for (int f = 2; f * f <= n; ++f) if (n % f == 0) { p = f; break; }
    if (p == 0) p = n;                       /* n prime: direct DFT */
    int m = n / p;
    if (m == 1) {
        for (int k = 0; k < n; ++k) {
            double sr = 0, si = 0;
            for (int j = 0; j < n; ++j) {
                double ang = sign * 2.0 * M_PI * (double)((long long)j * k % n) / n;
                double c = cos(ang), s = sin(ang);
                sr += inr[j * stride] * c - ini[j * stride] * s;
                si += inr[j * stride] * s + ini[j * stride] * c;
            }
            outr[k] = sr; outi[k] = si;
        }
        return;
    }
    /* sub-transforms of the p decimated sequences, stored in tmp */
    for (int q = 0; q < p; ++q)
        dft_rec(inr + q * stride, ini + q * stride, m, stride * p, sign,
                tmpr + q * m, tmpi + q * m, outr, outi);
    for (int k = 0; k < n; ++k) {
        double sr = 0, si = 0;
        for (int q = 0; q < p; ++q) {
            double ang = sign * 2.0 * M_PI * (double)((long long)q * k % n) / n;
            double c = cos(ang), s = sin(ang);
            double xr = tmpr[q * m + (k % m)], xi = tmpi[q * m + (k % m)];
            sr += xr * c - xi * s;
            si += xr * s + xi * c;
        }
        outr[k] = sr; outi[k] = si;
    }
}

/* In-place 1D transform of n points at stride s, count lines at line stride ls. */
static void fft_lines(double* re, double* im, int n, long s, long count, long ls, int sign)
{
    double* buf = (double*)malloc(sizeof(double) * 6 * n);
    double *ir = buf, *ii = buf + n, *orr = buf + 2 * n, *oi = buf + 3 * n, *tr = buf + 4 * n, *ti = buf + 5 * n;
    for (long l = 0; l < count; ++l) {
        double* r = re + l * ls;
        double* i = im + l * ls;
        for (int j = 0; j < n; ++j) { ir[j] = r[j * s]; ii[j] = i[j * s]; }
        dft_rec(ir, ii, n, 1, sign, orr, oi, tr, ti);
        for (int j = 0; j < n; ++j) { r[j * s] = orr[j]; i[j * s] = oi[j]; }
    }
    free(buf);
}

/* Generic d-dimensional cube transform (d = 1, 2, 3) of edge n; sign -1 is
 * FFTW_FORWARD, +1 FFTW_BACKWARD.  Unscaled. */
EXPORT void o_dft_cube(double* re, double* im, int n, int d, int sign)
{
    long n2 = (long)n * n;
    if (d == 1) { fft_lines(re, im, n, 1, 1, n, sign); return; }
    if (d == 2) {
        fft_lines(re, im, n, 1, n, n, sign);          /* rows (x)    */
        fft_lines(re, im, n, n, n, 1, sign);          /* columns (y) */
        return;
    }
    fft_lines(re, im, n, 1, n2, n, sign);             /* x */
    for (int z = 0; z < n; ++z)
        fft_lines(re + z * n2, im + z * n2, n, n, n, 1, sign);   /* y */
    fft_lines(re, im, n, n2, n2, 1, sign);            /* z */
}

/* Brute-force DFT of a 2D image, for pinning the FFT in tests. */
EXPORT void o_dft2d_brute(const double* re, const double* im, int n, int sign, double* ore, double* oim)
{
    for (int ky = 0; ky < n; ++ky)
        for (int kx = 0; kx < n; ++kx) {
            double sr = 0, si = 0;
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x) {
                    long ph = ((long)kx * x + (long)ky * y) % n;
                    double ang = sign * 2.0 * M_PI * (double)ph / n;
                    double c = cos(ang), s = sin(ang);
                    sr += re[y * n + x] * c - im[y * n + x] * s;
                    si += re[y * n + x] * s + im[y * n + x] * c;
                }
            ore[ky * n + kx] = sr; oim[ky * n + kx] = si;
        }
}

/* img2d_FFT: icosprocess/SRC/img2d_FFT.f90:4-32 with FFTW.f90:179-193.
 * dir = 1: forward, scaled 1/N; dir = -1: backward, unscaled.
 * conj_out = 1 stores b = -Im (icosprocess, local_reconstruct),
 * conj_out = 0 stores b = +Im (local_refine/SRC/img2d_FFT.f90:19). */
EXPORT void o_img2d_fft(real* a, real* b, int nx, int dir, int conj_out)
{
    long N = (long)nx * nx;
    double* re = (double*)malloc(sizeof(double) * N);
    double* im = (double*)malloc(sizeof(double) * N);
    for (long i = 0; i < N; ++i) { re[i] = a[i]; im[i] = b[i]; }
    o_dft_cube(re, im, nx, 2, dir == 1 ? -1 : 1);
    double sc = dir == 1 ? 1.0 / (double)N : 1.0;
    for (long i = 0; i < N; ++i) {
        a[i] = (real)(re[i] * sc);
        b[i] = (real)(conj_out ? -im[i] * sc : im[i] * sc);
    }
    free(re); free(im);
}

/* img3d_FFT / FFTcmplx3d_f/B: local_refine/SRC/img3d_FFT.f90:2-31,
 * icosprocess/SRC/FFTW.f90:205-219. Same conventions as o_img2d_fft. */
EXPORT void o_img3d_fft(real* a, real* b, int nx, int dir, int conj_out)
{
    long N = (long)nx * nx * nx;
    double* re = (double*)malloc(sizeof(double) * N);
    double* im = (double*)malloc(sizeof(double) * N);
    for (long i = 0; i < N; ++i) { re[i] = a[i]; im[i] = b[i]; }
    o_dft_cube(re, im, nx, 3, dir == 1 ? -1 : 1);
    double sc = dir == 1 ? 1.0 / (double)N : 1.0;
    for (long i = 0; i < N; ++i) {
        a[i] = (real)(re[i] * sc);
        b[i] = (real)(conj_out ? -im[i] * sc : im[i] * sc);
    }
    free(re); free(im);
}

/* ------------------------------------------------------------------ */
/* Per-image operations                                                */
/* ------------------------------------------------------------------ */

/* getCTF2d: icosprocess/SRC/getCTF2d.f90:2-29.  df_scale = 1e4 for the
 * icosprocess variant (defocus in um, :14-15), 1 for the local programs
 * (local_refine/SRC/getCTF2d.f90:14-15, defocus in A). */
EXPORT void o_getctf2d(int nx, real df10, real df20, real astig0, real apix, real vol,
                       real Cs0, real ampwg, real df_scale, real* ctf)
{
    real PI = r_pi();
    real d2r = PI / (real)180.0;
    real lam = (real)0.387818 / (real)sqrt((real)1.0 * vol * ((real)1.0 + (real)0.978459E-3 * vol));
    real Cs = Cs0 * (real)1.0e7;
    real df1 = df10 * df_scale, df2 = df20 * df_scale;
    real ast = astig0 * d2r;
    real sq = (real)sqrt((real)1.0 - ampwg * ampwg);
    for (long i = 0; i < (long)nx * nx; ++i) {
        int ky = (int)(i / nx) - nx / 2;
        int hx = (int)(i % nx) - nx / 2;
        real ph = (real)atan2((real)ky, (real)hx);
        if (ph < 0) ph = ph + (real)2.0 * PI;
        ph = ph - ast;
        real c = (real)cos(ph), s = (real)sin(ph);
        real df = df1 * c * c + df2 * s * s;
        real res = (real)sqrt((real)(hx * hx + ky * ky)) / (real)nx / apix;
        real phase = (real)2.0 * PI * (-Cs * lam * lam * lam * res * res * res * res / (real)4 +
                                       df * lam * res * res / (real)2);
        ctf[i] = -(sq * (real)sin(phase) + ampwg * (real)cos(phase));
    }
}

/* img2d_mask: icosprocess/SRC/img2d_mask.f90:2-49 (ring mean summed in
 * double precision, cosine edge of half width 3 around the mask radius). */
EXPORT void o_img2d_mask(real* d, int nx, real mask, real ox, real oy)
{
    int halfw = 6;
    real pi = r_pi();
    real R1 = mask - (real)(halfw / 2);
    if (R1 <= 0) R1 = 0;
    real R2 = mask + (real)(halfw / 2);
    if (R2 > (real)(nx / 2)) R2 = (real)(nx / 2);
    double summ = 0;
    long num = 0;
    real x0 = ox + (real)(nx / 2), y0 = oy + (real)(nx / 2);
    for (long i = 0; i < (long)nx * nx; ++i) {
        int iy = (int)(i / nx), ix = (int)(i % nx);
        real yy = (-y0 + iy) * (-y0 + iy), xx = (-x0 + ix) * (-x0 + ix);
        real ra = (real)sqrt(yy + xx);
        if (ra > R1 && ra <= R2) { summ += d[i]; num++; }
    }
    if (num != 0) summ = summ / (double)(float)num;
    for (long i = 0; i < (long)nx * nx; ++i) {
        int iy = (int)(i / nx), ix = (int)(i % nx);
        real yy = (-y0 + iy) * (-y0 + iy), xx = (-x0 + ix) * (-x0 + ix);
        real ra = (real)sqrt(yy + xx);
        if (ra > R1 && ra <= R2) {
            real edge = ((real)1.0 + (real)cos(pi * (ra - R1) / (real)halfw)) / (real)2.0;
            d[i] = d[i] * edge + ((real)1.0 - edge) * (real)summ;
        }
        if (ra > R2) d[i] = (real)summ;
    }
}

/* img2d_norm: icosprocess/SRC/img2d_norm.f90:1-22. */
EXPORT void o_img2d_norm(real* d, int nx)
{
    long N = (long)nx * nx;
    real avg = 0, dev = 0;
    for (long i = 0; i < N; ++i) avg += d[i];
    avg = avg / (real)N;
    for (long i = 0; i < N; ++i) d[i] -= avg;
    for (long i = 0; i < N; ++i) dev += d[i] * d[i];
    dev = (real)sqrt(dev / (real)N);
    for (long i = 0; i < N; ++i) d[i] /= dev;
}

/* Multiply by (-1)^(x+y): the shift2d arrays of ort_globesearch.f90:25-29. */
EXPORT void o_checker2d(real* d, int nx)
{
    for (long i = 0; i < (long)nx * nx; ++i)
        if (((i / nx) + (i % nx)) & 1) d[i] = -d[i];
}

/* img2d_FT_centshift (recentre = 1; icosprocess/SRC/img2d_FT_centshift.f90:2-18,
 * phase uses x0 + nx/2) and img2d_FTphase (recentre = 0;
 * local_refine/SRC/img2d_FTphase.f90:1-22, phase uses x0). */
EXPORT void o_ft_shift(real* a, real* b, int nx, real x0, real y0, int recentre)
{
    real pi = r_pi();
    real xs = recentre ? x0 + (real)(nx / 2) : x0;
    real ys = recentre ? y0 + (real)(nx / 2) : y0;
    for (long i = 0; i < (long)nx * nx; ++i) {
        int ky = (int)(i / nx) - nx / 2, hx = (int)(i % nx) - nx / 2;
        real ph = (real)2.0 * pi * (xs * hx + ys * ky) / (real)nx;
        real c = (real)cos(ph), s = (real)sin(ph);
        real fr = a[i] * c + b[i] * s;
        real fi = -a[i] * s + b[i] * c;
        a[i] = fr; b[i] = fi;
    }
}

/* FT_amp_expfit: icosprocess/SRC/FT_amp_expfit.f90:2-47.  Fcut has
 * maxR-minR+1 entries for rings minR..maxR. */
EXPORT void o_amp_expfit(const real* a, const real* b, int nx, int minR, int maxR, real* Fcut)
{
    int nr = maxR - minR + 1;
    real* Favg = (real*)calloc(nr, sizeof(real));
    real* resLN = (real*)calloc(nr, sizeof(real));
    int* Num = (int*)calloc(nr, sizeof(int));
    for (int ky = 0; ky <= maxR; ++ky)
        for (int hx = -maxR; hx <= maxR; ++hx) {
            real rad = (real)sqrt((real)(hx * hx + ky * ky));
            int ir = (int)(rad + (real)0.5);
            if (ir >= minR && ir <= maxR) {
                long i = (long)(ky + nx / 2) * nx + (hx + nx / 2);
                real amp = (real)sqrt(a[i] * a[i] + b[i] * b[i]);
                Favg[ir - minR] += amp;
                Num[ir - minR] += 1;
                resLN[ir - minR] += rad;
            }
        }
    for (int k = 0; k < nr; ++k) {
        if (Num[k] > 0 && Favg[k] > 0) {
            resLN[k] = resLN[k] / (real)Num[k];
            Favg[k] = (real)log(Favg[k] / (real)Num[k]);
        } else {
            Favg[k] = 0;
        }
    }
    /* 2-parameter normal equations for log|F| = x1*r + x2 (:36-42) */
    real a11 = 0, a12 = 0, a22 = 0, b1 = 0, b2 = 0;
    for (int k = 0; k < nr; ++k) {
        a11 += resLN[k] * resLN[k];
        a12 += resLN[k];
        a22 += 1;
        b1 += resLN[k] * Favg[k];
        b2 += Favg[k];
    }
    real x2 = (b2 - b1 * a12 / a11) / (a22 - a12 * a12 / a11);
    real x1 = (b1 - a12 * x2) / a11;
    for (int k = 0; k < nr; ++k) Fcut[k] = (real)exp(resLN[k] * x1 + x2) * (real)0.8;
    free(Favg); free(resLN); free(Num);
}

/* ------------------------------------------------------------------ */
/* Common-lines phase residual (icosprocess search)                    */
/* ------------------------------------------------------------------ */

/* Bilinear sample of a centred FT at (x, y): img2d_fill_line.f90:17-31. */
static void bilin(const real* a, const real* b, int nx, real x, real y, real* re, real* im)
{
    int lx = (int)x, ly = (int)y;
    real ax = x - (real)lx, ay = y - (real)ly;
    long n1 = (long)nx * ly + lx, n2 = (long)nx * (ly + 1) + lx;
    *re = ((real)1.0 - ax) * ((real)1.0 - ay) * a[n1] + ax * ((real)1.0 - ay) * a[n1 + 1] +
          ((real)1.0 - ax) * ay * a[n2] + ax * ay * a[n2 + 1];
    *im = ((real)1.0 - ax) * ((real)1.0 - ay) * b[n1] + ax * ((real)1.0 - ay) * b[n1 + 1] +
          ((real)1.0 - ax) * ay * b[n2] + ax * ay * b[n2 + 1];
}

/* Phase difference folded into [0, pi]: cal_crosslin_PR_omega.f90:57-60. */
static real phase_diff(real re1, real im1, real re2, real im2)
{
    real pi = r_pi(), tpi = (real)2.0 * pi;
    real p1 = (real)atan2(im1, re1), p2 = (real)atan2(im2, re2);
    real diff = (real)fabs(p2 - p1);
    diff = (real)fmod(diff, tpi);
    if (diff > pi) diff = tpi - diff;
    return diff;
}

/* cal_crosslin_PR_omega for one trial direction (theta, phi, omega=0):
 * icosprocess/SRC/cal_crosslin_PR_omega.f90:3-82.
 *   pa,pb     particle FT (masked, checkerboarded, FFT, centre-shifted)
 *   Fcut0     particle Fcut (rings minR..maxR)
 *   ta,tb     T template FTs, nx*nx each; tctf_fcut T Fcut rows used as-is
 *             (the caller has applied the CTF and refitted Fcut when
 *             applyCTF is on -- identical to the per-call refit at :34-38)
 *   teul      T*3 template angles
 *   PR        out, Nomega entries (999 where the gate leaves nothing) */
EXPORT void o_crosslin_pr_omega(const real* pa, const real* pb, const real* Fcut0,
                                int T, const real* ta, const real* tb, const real* tFcut,
                                const real* teul, int nx, int minR, int maxR,
                                int nops, const real* ops, real theta, real phi,
                                int Nomega, real step, real* PR)
{
    real pi = r_pi();
    real r2d = (real)180.0 / pi;
    real d2r = r_deg2rad();
    int nr = maxR - minR + 1;
    real* PR2 = (real*)calloc(Nomega, sizeof(real));
    real* lin2 = (real*)malloc(sizeof(real) * 2 * nr);
    real* coss = (real*)malloc(sizeof(real) * Nomega);
    real* sinn = (real*)malloc(sizeof(real) * Nomega);
    real al1[64 * 6], al2[64 * 6];
    for (int k = 0; k < Nomega; ++k) PR[k] = 0;
    long N2 = (long)nx * nx;
    real half = (real)(nx / 2);
    for (int t = 0; t < T; ++t) {
        const real* a2 = ta + t * N2;
        const real* b2 = tb + t * N2;
        const real* Fc = tFcut + t * nr;
        o_crosslin_angle(theta, phi, (real)0.0, teul[3 * t], teul[3 * t + 1], teul[3 * t + 2],
                         nops, ops, al1, al2);
        for (int j = 0; j < nops; ++j) {
            if (!(al1[j] < (real)999.0 && al2[j] < (real)999.0)) continue;
            /* template line (img2d_fill_line.f90:13-31) */
            for (int FR = minR; FR <= maxR; ++FR) {
                real x = (real)FR * (real)cos(al2[j] * d2r) + half;
                real y = (real)FR * (real)sin(al2[j] * d2r) + half;
                bilin(a2, b2, nx, x, y, &lin2[2 * (FR - minR)], &lin2[2 * (FR - minR) + 1]);
            }
            /* particle line for every omega (img2d_fill_lin_omegearray.f90:13-37) */
            for (int k = 0; k < Nomega; ++k) {
                real om = al1[j] - step * (real)k;
                coss[k] = (real)cos(om * d2r);
                sinn[k] = (real)sin(om * d2r);
            }
            for (int FR = minR; FR <= maxR; ++FR) {
                real t_re = lin2[2 * (FR - minR)], t_im = lin2[2 * (FR - minR) + 1];
                real amp2 = (real)sqrt(t_re * t_re + t_im * t_im);
                if (!(amp2 >= Fc[FR - minR])) continue;
                for (int k = 0; k < Nomega; ++k) {
                    real x = (real)FR * coss[k] + half;
                    real y = (real)FR * sinn[k] + half;
                    real p_re, p_im;
                    bilin(pa, pb, nx, x, y, &p_re, &p_im);
                    real amp1 = (real)sqrt(p_re * p_re + p_im * p_im);
                    if (!(amp1 >= Fcut0[FR - minR])) continue;
                    real diff = phase_diff(p_re, p_im, t_re, t_im);
                    real w = (real)0.5 * (amp1 + amp2);
                    PR[k] += diff * w;
                    PR2[k] += w;
                }
            }
        }
    }
    for (int k = 0; k < Nomega; ++k)
        PR[k] = PR2[k] > (real)1.0e-10 ? PR[k] / PR2[k] * r2d : (real)999.0;
    free(PR2); free(lin2); free(coss); free(sinn);
}

/* Grid sizes of ort_globesearch.f90:19-21. */
EXPORT void o_search_grid(real step, int* Ntheta, int* Nphi, int* Nomega)
{
    *Ntheta = (int)((real)21.0 / step);
    *Nphi = (int)((real)32.0 / step);
    *Nomega = (int)((real)360.0 / step);
}

/* AU membership of grid point (itheta, iphi): ort_globesearch.f90:68-82,
 * in float32 as the reference (see in_au). */
EXPORT int o_grid_in_au(real step, int itheta, int iphi)
{
    float d2r = atan2f(1.0f, 1.0f) / 45.0f;
    float theta = 69.0f + (float)step * (float)itheta;
    float tmp = theta * d2r;
    float z = cosf(tmp), z0 = sinf(tmp);
    float phi = (float)step * (float)iphi;
    tmp = phi * d2r;
    return in_au(z0 * cosf(tmp), z0 * sinf(tmp), z);
}

/* ort_globesearch for one particle: icosprocess/SRC/ort_globesearch.f90:68-106.
 * PR_array (Nomega x (2Nphi+1) x (Ntheta+1), omega fastest) is optional
 * output (may be NULL).  Returns theta, phi, omega, PR of the first
 * minimum in column-major order (Fortran minloc). */
EXPORT void o_globesearch(const real* pa, const real* pb, const real* Fcut0,
                          int T, const real* ta, const real* tb, const real* tFcut, const real* teul,
                          int nx, int minR, int maxR, int nops, const real* ops, real step,
                          real* out4, real* PR_array_out)
{
    int Nt, Np, No;
    o_search_grid(step, &Nt, &Np, &No);
    long plane = (long)No * (2 * Np + 1);
    long total = plane * (Nt + 1);
    real* arr = (real*)malloc(sizeof(real) * total);
    real* pro = (real*)malloc(sizeof(real) * No);
    for (long i = 0; i < total; ++i) arr[i] = (real)1000.0;
    for (int it = 0; it <= Nt; ++it) {
        real theta = (real)69.0 + step * (real)it;
        for (int ip = 0; ip <= Np; ++ip) {
            if (!o_grid_in_au(step, it, ip)) continue;
            real phi = step * (real)ip;
            o_crosslin_pr_omega(pa, pb, Fcut0, T, ta, tb, tFcut, teul, nx, minR, maxR, nops, ops,
                                theta, phi, No, step, pro);
            memcpy(arr + it * plane + (long)(ip + Np) * No, pro, sizeof(real) * No);
            o_crosslin_pr_omega(pa, pb, Fcut0, T, ta, tb, tFcut, teul, nx, minR, maxR, nops, ops,
                                theta, -phi, No, step, pro);
            memcpy(arr + it * plane + (long)(-ip + Np) * No, pro, sizeof(real) * No);
        }
    }
    long best = 0;
    for (long i = 1; i < total; ++i) if (arr[i] < arr[best]) best = i;
    int it = (int)(best / plane);
    int ip = (int)((best % plane) / No) - Np;
    int io = (int)(best % No);
    out4[0] = (real)69.0 + step * (real)it;
    out4[1] = step * (real)ip;
    out4[2] = step * (real)io;
    out4[3] = arr[best];
    if (PR_array_out) memcpy(PR_array_out, arr, sizeof(real) * total);
    free(arr); free(pro);
}

/* cal_crosslin_PR: icosprocess/SRC/cal_crosslin_PR.f90:2-71.  pa,pb is the
 * particle FT without centre shift; the trial shift (x, y) is applied here
 * (:22).  Returns 1000 when the gate leaves nothing (:64-68). */
EXPORT real o_crosslin_pr(const real* pa, const real* pb, const real* Fcut0,
                          int T, const real* ta, const real* tb, const real* tFcut, const real* teul,
                          int nx, int minR, int maxR, int nops, const real* ops,
                          real theta, real phi, real omega, real x, real y)
{
    real pi = r_pi();
    real d2r = r_deg2rad();
    long N2 = (long)nx * nx;
    int nr = maxR - minR + 1;
    real* a1 = (real*)malloc(sizeof(real) * N2);
    real* b1 = (real*)malloc(sizeof(real) * N2);
    memcpy(a1, pa, sizeof(real) * N2);
    memcpy(b1, pb, sizeof(real) * N2);
    o_ft_shift(a1, b1, nx, x, y, 1);
    real al1[64 * 6], al2[64 * 6];
    real PR1 = 0, PR2 = 0;
    real half = (real)(nx / 2);
    for (int t = 0; t < T; ++t) {
        const real* a2 = ta + t * N2;
        const real* b2 = tb + t * N2;
        const real* Fc = tFcut + t * nr;
        o_crosslin_angle(theta, phi, omega, teul[3 * t], teul[3 * t + 1], teul[3 * t + 2],
                         nops, ops, al1, al2);
        for (int j = 0; j < nops; ++j) {
            if (!(al1[j] < (real)999.0 && al2[j] < (real)999.0)) continue;
            for (int FR = minR; FR <= maxR; ++FR) {
                real p_re, p_im, t_re, t_im;
                bilin(a1, b1, nx, (real)FR * (real)cos(al1[j] * d2r) + half,
                      (real)FR * (real)sin(al1[j] * d2r) + half, &p_re, &p_im);
                bilin(a2, b2, nx, (real)FR * (real)cos(al2[j] * d2r) + half,
                      (real)FR * (real)sin(al2[j] * d2r) + half, &t_re, &t_im);
                real amp1 = (real)sqrt(p_re * p_re + p_im * p_im);
                real amp2 = (real)sqrt(t_re * t_re + t_im * t_im);
                if (amp1 >= Fcut0[FR - minR] && amp2 >= Fc[FR - minR]) {
                    real diff = phase_diff(p_re, p_im, t_re, t_im);
                    real w = (real)0.5 * (amp1 + amp2);
                    PR1 += diff * w;
                    PR2 += w;
                }
            }
        }
    }
    free(a1); free(b1);
    return PR2 > (real)1.0e-10 ? PR1 / PR2 * (real)180.0 / pi : (real)1000.0;
}

/* ort_cent_11 (icosprocess variant, with AU folding):
 * icosprocess/SRC/ort_cent_11.f90:1-94.  oc is 11 x 5 (theta, phi, omega, x, y). */
EXPORT void o_ort_cent_11(real th0, real ph0, real om0, real x0, real y0, real dang, real dtran, real* oc)
{
    o_return_au(&th0, &ph0, &om0);
    for (int jj = 0; jj <= 10; ++jj) {
        real* o = oc + 5 * jj;
        o[0] = th0; o[1] = ph0; o[2] = om0; o[3] = x0; o[4] = y0;
        switch (jj) {
            case 1: o[0] = th0 - dang; break;
            case 2: o[0] = th0 + dang; break;
            case 3: o[1] = ph0 - dang; break;
            case 4: o[1] = ph0 + dang; break;
            case 5: o[2] = om0 - dang; break;
            case 6: o[2] = om0 + dang; break;
            case 7: o[3] = x0 - dtran; break;
            case 8: o[3] = x0 + dtran; break;
            case 9: o[4] = y0 - dtran; break;
            case 10: o[4] = y0 + dtran; break;
        }
        o_return_au(&o[0], &o[1], &o[2]);
    }
}

/* ort_refinement for one particle: icosprocess/SRC/ort_refinement.f90:69-95.
 * io5 = (theta, phi, omega, x, y) in/out; returns PR0. */
EXPORT real o_refine_particle(const real* pa, const real* pb, const real* Fcut0,
                              int T, const real* ta, const real* tb, const real* tFcut, const real* teul,
                              int nx, int minR, int maxR, int nops, const real* ops,
                              int Ncycle, real deltaAngle, real deltaTrans, real* io5)
{
    real oc[55], PR[11];
    real PR0 = 0;
    for (int n = 1; n <= Ncycle; ++n) {
        real f = 1;
        for (int k = 1; k < n; ++k) f *= (real)0.8;
        real dang = deltaAngle * f, dtran = deltaTrans * f;
        o_ort_cent_11(io5[0], io5[1], io5[2], io5[3], io5[4], dang, dtran, oc);
        for (int jj = 0; jj <= 10; ++jj)
            PR[jj] = o_crosslin_pr(pa, pb, Fcut0, T, ta, tb, tFcut, teul, nx, minR, maxR, nops, ops,
                                   oc[5 * jj], oc[5 * jj + 1], oc[5 * jj + 2], oc[5 * jj + 3], oc[5 * jj + 4]);
        int best = 0;
        for (int jj = 1; jj <= 10; ++jj) if (PR[jj] < PR[best]) best = jj;
        PR0 = PR[best];
        for (int q = 0; q < 5; ++q) io5[q] = oc[5 * best + q];
    }
    return PR0;
}

/* ------------------------------------------------------------------ */
/* Direct Fourier inversion                                            */
/* ------------------------------------------------------------------ */

/* Insert of one particle: icosprocess/SRC/reconstruct.f90:103-147 (layout
 * full: volume nx^3, offset nx/2, radius cut rcut) and
 * reconstruct_substep.f90:101-145 (layout compact: (2R+1)^2 (R+1), offset
 * R = maxR).  a,b: centred particle FT after centre shift; ctf may be NULL
 * (ctf = 1).  R: particle rotation (euler_mat). */
/* The eight corners and trilinear weights of one sample x00 = R (hx, ky, 0)
 * under operator S (reconstruct.f90:116-138): z < 0 flips the point and
 * returns signz = -1; offsets off on x, y (and z for the full box). */
static int place_corners(const real* S, const real* x00, int compact, real off, long D, long D2, long id[8],
                         real w[8])
{
    real x0[3];
    matvec(S, x00, x0);
    int signz = 1;
    if (x0[2] < 0) { x0[0] = -x0[0]; x0[1] = -x0[1]; x0[2] = -x0[2]; signz = -1; }
    x0[0] += off; x0[1] += off;
    if (!compact) x0[2] += off;
    int lx[3] = {(int)x0[0], (int)x0[1], (int)x0[2]};
    real ax[3] = {x0[0] - lx[0], x0[1] - lx[1], x0[2] - lx[2]};
    w[0] = ((real)1 - ax[0]) * ((real)1 - ax[1]) * ((real)1 - ax[2]);
    w[1] = ax[0] * ((real)1 - ax[1]) * ((real)1 - ax[2]);
    w[2] = ((real)1 - ax[0]) * ax[1] * ((real)1 - ax[2]);
    w[3] = ax[0] * ax[1] * ((real)1 - ax[2]);
    w[4] = ((real)1 - ax[0]) * ((real)1 - ax[1]) * ax[2];
    w[5] = ax[0] * ((real)1 - ax[1]) * ax[2];
    w[6] = ((real)1 - ax[0]) * ax[1] * ax[2];
    w[7] = ax[0] * ax[1] * ax[2];
    long c1 = lx[2] * D2 + lx[1] * D + lx[0];
    long t[8] = {c1, c1 + 1, c1 + D, c1 + D + 1, c1 + D2, c1 + D2 + 1, c1 + D2 + D, c1 + D2 + D + 1};
    for (int k = 0; k < 8; ++k) id[k] = t[k];
    return signz;
}

EXPORT void o_insert(real* fa, real* fb, real* wg, int compact, int nx, int maxR, real rcut,
                     const real* a, const real* b, const real* ctf, const real* R,
                     int nops, const real* ops)
{
    long D = compact ? 2 * maxR + 1 : nx;
    long D2 = D * D;
    /* The compact array is indexed flat (reconstruct_half.f90:127-130): a
     * corner at x = 2maxR+1 wraps into the next row, one at y = 2maxR+1 into
     * the next plane; one past the last plane falls outside the array. */
    long size = compact ? D2 * (maxR + 1) : D2 * nx;
    real off = compact ? (real)maxR : (real)(nx / 2);
    for (int ky = -maxR; ky <= maxR; ++ky)
        for (int hx = 0; hx <= maxR; ++hx) {
            real dr = (real)sqrt((real)(hx * hx + ky * ky));
            if (!(dr <= rcut)) continue;
            long i2 = (long)(ky + nx / 2) * nx + (hx + nx / 2);
            real x[3] = {(real)hx, (real)ky, 0}, x00[3];
            matvec(R, x, x00);
            real c = ctf ? ctf[i2] : (real)1.0;
            for (int s = 0; s < nops; ++s) {
                long id[8];
                real w[8];
                int signz = place_corners(ops + 9 * s, x00, compact, off, D, D2, id, w);
                for (int j = 0; j < 8; ++j) {
                    if (id[j] < 0 || id[j] >= size) continue;
                    fa[id[j]] += a[i2] * w[j] * c;
                    fb[id[j]] += b[i2] * w[j] * c * (real)signz;
                    wg[id[j]] += w[j] * c * c;
                }
            }
        }
}

/* reconstruct_corrCTF.f90:106-150 / reconstruct_corrCTF_substep.f90:103-147:
 * the insert with a defocus series.  ctf3d holds ncyc CTF images (defocus
 * df + k step, k = -C..C); fa, fb take a w ctf3d_k for every k, in that order,
 * and wg takes w ctf2d^2 ncyc times, ctf2d being the last CTF the series
 * computed (k = +C), not ctf3d_k. */
EXPORT void o_insert_series(real* fa, real* fb, real* wg, int compact, int nx, int maxR, real rcut,
                            const real* a, const real* b, const real* ctf3d, int ncyc, const real* ctf2d,
                            const real* R, int nops, const real* ops)
{
    long D = compact ? 2 * maxR + 1 : nx;
    long D2 = D * D;
    long size = compact ? D2 * (maxR + 1) : D2 * nx;
    long N2 = (long)nx * nx;
    real off = compact ? (real)maxR : (real)(nx / 2);
    for (int ky = -maxR; ky <= maxR; ++ky)
        for (int hx = 0; hx <= maxR; ++hx) {
            real dr = (real)sqrt((real)(hx * hx + ky * ky));
            if (!(dr <= rcut)) continue;
            long i2 = (long)(ky + nx / 2) * nx + (hx + nx / 2);
            real x[3] = {(real)hx, (real)ky, 0}, x00[3];
            matvec(R, x, x00);
            for (int s = 0; s < nops; ++s) {
                long id[8];
                real w[8];
                int signz = place_corners(ops + 9 * s, x00, compact, off, D, D2, id, w);
                for (int j = 0; j < 8; ++j) {
                    if (id[j] < 0 || id[j] >= size) continue;
                    real t1 = a[i2] * w[j], t2 = b[i2] * w[j];
                    for (int k = 0; k < ncyc; ++k) {
                        real c = ctf3d[k * N2 + i2];
                        fa[id[j]] = fa[id[j]] + t1 * c;
                        fb[id[j]] = fb[id[j]] + t2 * c * (real)signz;
                        wg[id[j]] = wg[id[j]] + w[j] * (ctf2d[i2] * ctf2d[i2]);
                    }
                }
            }
        }
}

/* Fold of the lz = 0 plane, division by the weights and (-1)^(h+k+l):
 * icosprocess/SRC/reconstruct.f90:151-237.  Full nx^3 volumes in place. */
/* int_radius = 0: dr = sqrt(real) <= maxR (reconstruct.f90:151-237).
 * int_radius = 1: int(sqrt(real)) <= maxR, the shell test of
 * reconstruct_FSC.f90:101-169 and merge_reconstruct.f90:37-107 (integer dr). */
static int o_in_sphere(int hx, int ky, int lz, int maxR, int int_radius)
{
    real dr = (real)sqrt((real)(hx * hx + ky * ky + lz * lz));
    return int_radius ? ((int)dr <= maxR) : (dr <= (real)maxR);
}

EXPORT void o_fold_normalise(real* fa, real* fb, real* wg, int nx, int maxR, int int_radius)
{
    long n2 = (long)nx * nx;
    int c = nx / 2;
#define IDX(h, k, l) ((long)((l) + c) * n2 + (long)((k) + c) * nx + ((h) + c))
    int lz = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int kylo = pass == 0 ? 1 : 0, kyhi = pass == 0 ? maxR : 0;
        for (int ky = kylo; ky <= kyhi; ++ky) {
            int hxlo = pass == 0 ? -maxR : 1;
            for (int hx = hxlo; hx <= maxR; ++hx) {
                long n1 = IDX(hx, ky, lz), nn2 = IDX(-hx, -ky, -lz);
                if (o_in_sphere(hx, ky, lz, maxR, int_radius)) {
                    wg[n1] = wg[n1] + wg[nn2];
                    fa[n1] = (fa[n1] + fa[nn2]) / wg[n1];
                    fb[n1] = (fb[n1] - fb[nn2]) / wg[n1];
                    if (((hx + ky + lz) % 2) != 0) { fa[n1] = -fa[n1]; fb[n1] = -fb[n1]; }
                    fa[nn2] = fa[n1];
                    fb[nn2] = -fb[n1];
                } else {
                    fa[n1] = 0; fb[n1] = 0; fa[nn2] = 0; fb[nn2] = 0;
                }
            }
        }
    }
    {
        long n1 = IDX(0, 0, 0);
        fa[n1] = fa[n1] / wg[n1];
        fb[n1] = 0;
    }
    for (lz = 1; lz <= maxR; ++lz)
        for (int ky = -maxR; ky <= maxR; ++ky)
            for (int hx = -maxR; hx <= maxR; ++hx) {
                long n1 = IDX(hx, ky, lz), nn2 = IDX(-hx, -ky, -lz);
                if (o_in_sphere(hx, ky, lz, maxR, int_radius)) {
                    fa[n1] = fa[n1] / wg[n1];
                    fb[n1] = fb[n1] / wg[n1];
                    if (((hx + ky + lz) % 2) != 0) { fa[n1] = -fa[n1]; fb[n1] = -fb[n1]; }
                    fa[nn2] = fa[n1];
                    fb[nn2] = -fb[n1];
                } else {
                    fa[n1] = 0; fb[n1] = 0; fa[nn2] = 0; fb[nn2] = 0;
                }
            }
#undef IDX
}

/* cmplx(fa, s*fb) -> backward 3D FFT -> shift3dComplex -> real:
 * s = -1 in reconstruct.f90:240-247 and merge_reconstruct.f90:133-139,
 * s = +1 in reconstruct_FSC.f90:173-179 (FFTW.f90:213-219,304-321).
 * merge_reconstruct_FSC.f90:181-185 transforms cmplx(fa, fb) forward with
 * FFT1d (isn = -1); its real part is that of s = -1 here. */
EXPORT void o_volume_to_map(const real* fa, const real* fb, int nx, int imag_sign, real* map)
{
    long N = (long)nx * nx * nx, n2 = (long)nx * nx;
    double* re = (double*)malloc(sizeof(double) * N);
    double* im = (double*)malloc(sizeof(double) * N);
    for (long i = 0; i < N; ++i) { re[i] = fa[i]; im[i] = imag_sign < 0 ? -(double)fb[i] : (double)fb[i]; }
    o_dft_cube(re, im, nx, 3, +1);
    for (long i = 0; i < N; ++i) {
        long z = i / n2, y = (i % n2) / nx, x = i % nx;
        map[i] = (real)(((x + y + z) & 1) ? -re[i] : re[i]);
    }
    free(re); free(im);
}

/* Half-map FSC and FSC-weighted merge of two tnf volumes (compact layout,
 * lz slowest): reconstruct_FSC.f90:30-97, identical in
 * merge_reconstruct_FSC.f90:38-97.  Both weight sums get +1, the shell FSC
 * uses ir = int(r) over the whole half-cube, and the merged sums are scaled
 * by sqrt(2 FSC / (1 + FSC)) (0 where FSC <= 0).  The weighting loop's
 * else-branch zeroes a(i3), b(i3) at the index left from the last in-sphere
 * voxel, so each in-sphere voxel that is followed (in loop order) by an
 * out-of-sphere one ends with a = b = 0; that is kept.
 * fsc_raw[ir] = FSC(ir) as written to FSC.txt; out: merged fa, fb, wg. */
EXPORT void o_tnf_fsc(int maxR, const real* A1, const real* B1, const real* W1, const real* A2, const real* B2,
                      const real* W2, real* fsc_raw, real* fa, real* fb, real* wg)
{
    const long T = 2 * maxR + 1, n = T * T * (maxR + 1);
    real* w2 = (real*)malloc(sizeof(real) * n);
    real* a2 = (real*)malloc(sizeof(real) * n);
    real* b2 = (real*)malloc(sizeof(real) * n);
    real* S = (real*)malloc(sizeof(real) * (maxR + 1));
    real* S1 = (real*)malloc(sizeof(real) * (maxR + 1));
    real* S2 = (real*)malloc(sizeof(real) * (maxR + 1));
    for (long i = 0; i < n; ++i) {
        wg[i] = W1[i] + (real)1;
        w2[i] = W2[i] + (real)1;
        fa[i] = A1[i] / wg[i];
        fb[i] = B1[i] / wg[i];
        a2[i] = A2[i] / w2[i];
        b2[i] = B2[i] / w2[i];
    }
    for (int r = 0; r <= maxR; ++r) { S[r] = 0; S1[r] = (real)1.0e-10; S2[r] = (real)1.0e-10; }
    long i3 = 0;
    for (int lz = 0; lz <= maxR; ++lz)
        for (int ky = -maxR; ky <= maxR; ++ky)
            for (int hx = -maxR; hx <= maxR; ++hx) {
                int ir = (int)(real)sqrt((real)(hx * hx + ky * ky + lz * lz));
                if (ir > maxR) continue;
                i3 = (long)lz * T * T + (long)(ky + maxR) * T + (hx + maxR);
                S[ir] += fa[i3] * a2[i3] + fb[i3] * b2[i3];
                S1[ir] += fa[i3] * fa[i3] + fb[i3] * fb[i3];
                S2[ir] += a2[i3] * a2[i3] + b2[i3] * b2[i3];
            }
    for (int r = 0; r <= maxR; ++r) {
        real f = S[r] / (real)sqrt(S1[r] * S2[r]);
        fsc_raw[r] = f;
        S[r] = f <= 0 ? (real)0 : (real)sqrt((real)2 * f / ((real)1 + f));
    }
    for (long i = 0; i < n; ++i) {
        fa[i] = fa[i] * wg[i];
        fb[i] = fb[i] * wg[i];
        a2[i] = a2[i] * w2[i];
        b2[i] = b2[i] * w2[i];
        fa[i] = fa[i] + a2[i];
        fb[i] = fb[i] + b2[i];
        wg[i] = wg[i] + w2[i] - (real)2;
    }
    for (int lz = 0; lz <= maxR; ++lz)
        for (int ky = -maxR; ky <= maxR; ++ky)
            for (int hx = -maxR; hx <= maxR; ++hx) {
                int ir = (int)(real)sqrt((real)(hx * hx + ky * ky + lz * lz));
                if (ir <= maxR) {
                    i3 = (long)lz * T * T + (long)(ky + maxR) * T + (hx + maxR);
                    fa[i3] = fa[i3] * S[ir];
                    fb[i3] = fb[i3] * S[ir];
                    wg[i3] = wg[i3] * S[ir];
                } else {
                    fa[i3] = 0;
                    fb[i3] = 0;
                }
            }
    for (long i = 0; i < n; ++i) wg[i] = wg[i] + (real)1;
    free(w2); free(a2); free(b2); free(S); free(S1); free(S2);
}

/* Local symmetry of icosprocess mode 6 (icos2f_sym_mat.f90:250-540,
 * reconstruction_local.f90:30-55): N z rotations by k*step degrees, each entry
 * cos(a*3.1415926/180.0) in REAL*4.  ops: 9*N row-major. */
EXPORT void o_icosprocess_local_ops(int N, float step, real* ops)
{
    for (int k = 0; k < N; ++k) {
        /* step > 0: the literal angles k*step of icos2f_sym_matc2..c15;
         * step <= 0: icos2f_sym_matc7's float(i-1)*360.0/N (:520-523) */
        float a = step > 0 ? (float)k * step : (float)k * 360.0f / (float)N;
        float rad = a * 3.1415926f / 180.0f;
        float c = cosf(rad), sn = sinf(rad);
        real* m = ops + 9 * k;
        for (int i = 0; i < 9; ++i) m[i] = 0;
        m[0] = c; m[1] = -sn; m[3] = sn; m[4] = c; m[8] = 1;
    }
}

/* Output crop of merge_reconstruct.f90:143-149: the (nx-8)^3 centre block,
 * with lz taken as i/(nx-8)^2 + 4 from the 1-based i (not i-1), so the
 * last voxel of each output plane comes from the next plane. */
EXPORT void o_merge_crop(const real* map, int nx, real* out)
{
    const long M = nx - 8, M2 = M * M, n2 = (long)nx * nx;
    for (long i = 1; i <= M2 * M; ++i) {
        long lz = i / M2 + 4, ky = ((i - 1) % M2) / M + 4, hx = ((i - 1) % M2) % M + 4;
        out[i - 1] = map[lz * n2 + ky * nx + hx];
    }
}

/* merge_reconstruct_FSC.f90:188-197: zero outside the sphere of radius
 * imgmask about (nx/2, nx/2, nx/2) (the checkerboard inside is the one
 * o_volume_to_map already applies). */
EXPORT void o_sphere_zero(real* map, int nx, real imgmask)
{
    const long n2 = (long)nx * nx;
    for (long i = 0; i < n2 * nx; ++i) {
        long lz = i / n2, ky = (i % n2) / nx, hx = i % nx;
        real d = (real)sqrt((real)((hx - nx / 2) * (hx - nx / 2) + (ky - nx / 2) * (ky - nx / 2) +
                                   (lz - nx / 2) * (lz - nx / 2)));
        if (!(d <= imgmask)) map[i] = 0;
    }
}

/* ------------------------------------------------------------------ */
/* Central-slice phase residual (local_refine)                         */
/* ------------------------------------------------------------------ */

/* FFT3d_slice: local_refine/SRC/FFT3d_slice.f90:3-78. */
EXPORT void o_fft3d_slice(const real* fr, const real* fi, int nx, real theta, real phi, real omega,
                          real* oa, real* ob)
{
    real R[9];
    o_euler_mat(theta, phi, omega, R);
    long n2 = (long)nx * nx;
    long n = 0;
    for (long i = 0; i < n2; ++i) { oa[i] = 0; ob[i] = 0; }
    for (int iy = -nx / 2; iy <= nx / 2 - 1; ++iy)
        for (int ix = -nx / 2; ix <= nx / 2 - 1; ++ix, ++n) {
            real xp[3] = {(real)ix, (real)iy, 0}, X[3];
            matvec(R, xp, X);
            if ((real)sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]) > (real)(nx / 2 - 2)) continue;
            X[0] += (real)(nx / 2); X[1] += (real)(nx / 2); X[2] += (real)(nx / 2);
            int hx = (int)X[0], ky = (int)X[1], lz = (int)X[2];
            real xb = X[0] - hx, yb = X[1] - ky, zb = X[2] - lz;
            long b0 = (long)lz * n2 + (long)ky * nx + hx;
            long o[8] = {b0, b0 + 1, b0 + nx, b0 + nx + 1, b0 + n2, b0 + n2 + 1, b0 + n2 + nx, b0 + n2 + nx + 1};
            real w[8] = {((real)1 - xb) * ((real)1 - yb) * ((real)1 - zb), xb * ((real)1 - yb) * ((real)1 - zb),
                         ((real)1 - xb) * yb * ((real)1 - zb), xb * yb * ((real)1 - zb),
                         ((real)1 - xb) * ((real)1 - yb) * zb, xb * ((real)1 - yb) * zb,
                         ((real)1 - xb) * yb * zb, xb * yb * zb};
            real sr = 0, si = 0;
            for (int q = 0; q < 8; ++q) { sr += w[q] * fr[o[q]]; si += w[q] * fi[o[q]]; }
            oa[n] = sr; ob[n] = si;
        }
}

/* cal_PR_FT: local_refine/SRC/cal_PR_FT.f90:1-36 (weight |F1|+|F2|). */
EXPORT real o_pr_ft(const real* a1, const real* b1, const real* a2, const real* b2, int nx,
                    real minR, real maxR)
{
    real pi = r_pi();
    real PR1 = 0, PR2 = 0;
    for (long i = 0; i < (long)nx * nx; ++i) {
        int ky = (int)(i / nx) - nx / 2, hx = (int)(i % nx) - nx / 2;
        real dr = (real)sqrt((real)(hx * hx + ky * ky));
        if (dr >= minR && dr <= maxR) {
            real amp1 = (real)sqrt(a1[i] * a1[i] + b1[i] * b1[i]);
            real amp2 = (real)sqrt(a2[i] * a2[i] + b2[i] * b2[i]);
            real diff = phase_diff(a1[i], b1[i], a2[i], b2[i]);
            real w = amp1 + amp2;
            PR1 = diff * w + PR1;
            PR2 = w + PR2;
        }
    }
    return PR2 > (real)1.0e-10 ? PR1 / PR2 * (real)180.0 / pi : (real)1000.0;
}

/* img3d_softmask: local_refine/SRC/img3d_softmask.f90:3-46 (sum in REAL). */
EXPORT void o_img3d_softmask(real* d, int nx, real mask)
{
    int halfw = 6;
    real pi = r_pi();
    real inn = mask - (real)(halfw / 2);
    if (inn <= 0) inn = 0;
    real out = mask + (real)(halfw / 2);
    if (out > (real)(nx / 2)) out = (real)(nx / 2);
    long n2 = (long)nx * nx, N = n2 * nx;
    long num = 0;
    real summ = 0;
    for (long i = 0; i < N; ++i) {
        int iz = (int)(i / n2) - nx / 2;
        if ((real)iz <= out) {
            int iy = (int)((i % n2) / nx) - nx / 2, ix = (int)(i % nx) - nx / 2;
            real r = (real)sqrt((real)(ix * ix + iy * iy + iz * iz));
            if (r > inn && r <= out) { summ = d[i] + summ; num++; }
        }
    }
    if (num != 0) summ = summ / (real)num;
    for (long i = 0; i < N; ++i) {
        int iz = (int)(i / n2) - nx / 2, iy = (int)((i % n2) / nx) - nx / 2, ix = (int)(i % nx) - nx / 2;
        real r = (real)sqrt((real)(ix * ix + iy * iy + iz * iz));
        if (r > inn && r <= out) {
            real edge = ((real)1.0 + (real)cos(pi * (r - inn) / (real)halfw)) / (real)2.0;
            d[i] = d[i] * edge + ((real)1.0 - edge) * summ;
        }
        if (r > out) d[i] = summ;
    }
}

/* ------------------------------------------------------------------ */
/* local_refine driver                                                 */
/* ------------------------------------------------------------------ */

/* Symmetry operators of the local programs, in their order:
 * kind 0 = Cn (local_refine/SRC/icos2f_sym_matc.f90:1-18: rotations by
 *          i*2pi/N, i = 1..N), kind 1 = asym (:20-36, steps pi/6/N),
 * kind 2 = I2 (:39-95), kind 3 = I5 (:97-118, I2 conjugated). */
static void mm3(const real* A, const real* B, real* C) { matmul3(A, B, C); }
static void tr3(const real* A, real* T)
{
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) T[3 * j + i] = A[3 * i + j];
}

EXPORT int o_lref_symops(int kind, int N, real* r)
{
    real pi = r_pi();
    if (kind == 0 || kind == 1) {
        real ph = kind == 0 ? (real)2.0 * pi / (real)N : pi / (real)6.0 / (real)N;
        real R1[9] = {(real)cos(ph), -(real)sin(ph), 0, (real)sin(ph), (real)cos(ph), 0, 0, 0, 1};
        memcpy(r, R1, sizeof(R1));
        for (int i = 1; i < N; ++i) mm3(r + 9 * (i - 1), R1, r + 9 * i);
        return N;
    }
    real d2r = pi / (real)180.0;
    real a5 = (real)72.0 * d2r;
    real v = (real)acos((real)cos((real)60.0 * d2r) / (real)sin((real)36.0 * d2r));
    real rz5[9] = {(real)cos(a5), -(real)sin(a5), 0, (real)sin(a5), (real)cos(a5), 0, 0, 0, 1};
    real ryv[9] = {(real)cos(v), 0, (real)sin(v), 0, 1, 0, -(real)sin(v), 0, (real)cos(v)};
    real t1[9], t2[9], ryvT[9];
    tr3(ryv, ryvT);
    mm3(rz5, ryvT, t1);
    mm3(ryv, t1, t2);
    memcpy(rz5, t2, sizeof(t2));
    real rx2[9] = {1, 0, 0, 0, -1, 0, 0, 0, -1};
    real ry2[9] = {-1, 0, 0, 0, 1, 0, 0, 0, -1};
    real r111[9] = {0, 0, 1, 1, 0, 0, 0, 1, 0};
    memset(r, 0, sizeof(real) * 9 * 60);
    r[0] = r[4] = r[8] = 1;
    for (int i = 1; i < 5; ++i) mm3(r + 9 * (i - 1), rz5, r + 9 * i);
    for (int i = 5; i < 10; ++i) mm3(rx2, r + 9 * (i - 5), r + 9 * i);
    for (int i = 10; i < 20; ++i) mm3(ry2, r + 9 * (i - 10), r + 9 * i);
    for (int i = 20; i < 40; ++i) mm3(r111, r + 9 * (i - 20), r + 9 * i);
    for (int i = 40; i < 60; ++i) mm3(r111, r + 9 * (i - 20), r + 9 * i);
    if (kind == 3) {
        real rot0[9] = {(real)cos(v), 0, (real)sin(v), 0, 1, 0, -(real)sin(v), 0, (real)cos(v)};
        real rT[9];
        tr3(rot0, rT);
        for (int i = 0; i < 60; ++i) {
            mm3(r + 9 * i, rT, t1);
            mm3(rot0, t1, r + 9 * i);
        }
    }
    return 60;
}

/* icos_equalort: local_refine/SRC/icos_equalort.f90:2-27. */
EXPORT void o_icos_equalort(real th0, real ph0, real om0, int nsym, const real* ops, real* ang)
{
    real E[9], M[9];
    real r2d = (real)45.0 / (real)atan2((real)1.0, (real)1.0);
    o_euler_mat(th0, -ph0, (real)180.0 - om0, E);
    for (int i = 0; i < nsym; ++i) {
        mm3(ops + 9 * i, E, M);
        ang[3 * i] = (real)acos(M[8]) * r2d;
        ang[3 * i + 1] = -(real)atan2(M[5], M[2]) * r2d;
        ang[3 * i + 2] = (real)atan2(M[7], M[6]) * r2d;
    }
}

/* img2d_local_box: local_refine/SRC/img2d_local_box.f90:1-19 (wrap-around cut). */
EXPORT void o_local_box(const real* img, int n0, int ix, int iy, real* out, int n)
{
    for (long i = 0; i < (long)n * n; ++i) {
        int ky = (int)(i / n), hx = (int)(i % n);
        hx = hx + ix - n / 2;
        ky = ky + iy - n / 2;
        if (hx < 0) hx += n0;
        if (hx >= n0) hx -= n0;
        if (ky < 0) ky += n0;
        if (ky >= n0) ky -= n0;
        out[i] = img[(long)ky * n0 + hx];
    }
}

/* One sub-particle of refine_PR (local_refine/SRC/refine_PR.f90:126-233):
 *   fr, fi   model transform (nx^3, after softmask/checkerboard/FFT/checkerboard)
 *   a20,b20  particle transform times w_lp (nx^2)
 *   ctf,wlp  nx^2
 *   nsym/ops symmetry stage (nsym = 0 skips it, sym = c1)
 *   io5      theta phi omega x y in/out; returns the last PR. */
EXPORT real o_lref_refine(const real* fr, const real* fi, int nx, const real* a20, const real* b20,
                          const real* ctf, const real* wlp, int nsym, const real* ops, int ncycle,
                          real refine_angle, real refine_shift, real convF, real maxres, real* io5)
{
    long n2 = (long)nx * nx;
    real* a1 = (real*)malloc(sizeof(real) * n2);
    real* b1 = (real*)malloc(sizeof(real) * n2);
    real best = 1000;
    real th0 = io5[0], ph0 = io5[1], om0 = io5[2], x0 = io5[3], y0 = io5[4];
#define TEMPLATE_PR(TH, PH, OM, X, Y, OUT)                                           \
    do {                                                                             \
        o_fft3d_slice(fr, fi, nx, TH, PH, OM, a1, b1);                               \
        o_ft_shift(a1, b1, nx, X, Y, 0);                                             \
        for (long q = 0; q < n2; ++q) {                                              \
            real s = (((q / nx) + (q % nx)) & 1) ? (real)-1 : (real)1;               \
            a1[q] = a1[q] * s * ctf[q] * wlp[q];                                     \
            b1[q] = b1[q] * s * ctf[q] * wlp[q];                                     \
        }                                                                            \
        OUT = o_pr_ft(a1, b1, a20, b20, nx, (real)-1.0, maxres);                     \
    } while (0)
    if (nsym > 0) {
        real* ang = (real*)malloc(sizeof(real) * 3 * nsym);
        o_icos_equalort(th0, ph0, om0, nsym, ops, ang);
        int bi = 0;
        for (int i = 0; i < nsym; ++i) {
            real pr;
            TEMPLATE_PR(ang[3 * i], ang[3 * i + 1], ang[3 * i + 2], x0, y0, pr);
            if (i == 0 || pr < best) { best = pr; bi = i; }
        }
        th0 = ang[3 * bi]; ph0 = ang[3 * bi + 1]; om0 = ang[3 * bi + 2];
        free(ang);
    }
    for (int c = 1; c <= ncycle; ++c) {
        real f = 1;
        for (int k = 1; k < c; ++k) f *= convF;
        real da = refine_angle * f, dt = refine_shift * f;
        /* local ort_cent_11 order: local_refine/SRC/ort_cent_11.f90:6-84 */
        real oc[11][5];
        for (int jj = 0; jj <= 10; ++jj) {
            oc[jj][0] = th0; oc[jj][1] = ph0; oc[jj][2] = om0; oc[jj][3] = x0; oc[jj][4] = y0;
        }
        oc[1][3] -= dt; oc[2][3] += dt; oc[3][4] -= dt; oc[4][4] += dt;
        oc[5][0] -= da; oc[6][0] += da; oc[7][1] -= da; oc[8][1] += da; oc[9][2] -= da; oc[10][2] += da;
        int bi = 0;
        for (int jj = 0; jj <= 10; ++jj) {
            real pr;
            TEMPLATE_PR(oc[jj][0], oc[jj][1], oc[jj][2], oc[jj][3], oc[jj][4], pr);
            if (jj == 0 || pr < best) { best = pr; bi = jj; }
        }
        th0 = oc[bi][0]; ph0 = oc[bi][1]; om0 = oc[bi][2]; x0 = oc[bi][3]; y0 = oc[bi][4];
    }
#undef TEMPLATE_PR
    io5[0] = th0; io5[1] = ph0; io5[2] = om0; io5[3] = x0; io5[4] = y0;
    free(a1); free(b1);
    return best;
}

/* One sub-particle of refine_particle_df (local_refine/SRC/refine_particle_df.f90:100-135):
 * the slice at (th, ph, om), phase-shifted by (x0, y0), times shift2d and wlp,
 * against the particle (a20, b20) under the CTF of the 401 defoci
 * df -/+ z_i, z_i = z0 + (i - 200) 4 (sign +1 / -1), by cal_PR_FT; the first
 * minimum.  out2 = {z, PR}; prs (may be NULL) receives all 401 residuals. */
EXPORT void o_lref_defocus(const real* fr, const real* fi, int nx, const real* a20, const real* b20, const real* wlp,
                           real th, real ph, real om, real x0, real y0, real z0, real df1, real df2, real astig,
                           real apix, real vol, real cs, real ampwg, real sign, real maxres, real* out2, real* prs)
{
    long n2 = (long)nx * nx;
    real* a1 = (real*)malloc(sizeof(real) * n2);
    real* b1 = (real*)malloc(sizeof(real) * n2);
    real* a10 = (real*)malloc(sizeof(real) * n2);
    real* b10 = (real*)malloc(sizeof(real) * n2);
    real* ctf = (real*)malloc(sizeof(real) * n2);
    o_fft3d_slice(fr, fi, nx, th, ph, om, a1, b1);
    o_ft_shift(a1, b1, nx, x0, y0, 0);
    for (long q = 0; q < n2; ++q) {
        real s = (((q / nx) + (q % nx)) & 1) ? (real)-1 : (real)1;
        a1[q] = a1[q] * s * wlp[q];
        b1[q] = b1[q] * s * wlp[q];
    }
    real best = 0, bz = z0;
    for (int i = 0; i <= 400; ++i) {
        real zi = z0 + (real)(i - 200) * (real)4.0;
        real d1 = sign > 0 ? df1 + zi : df1 - zi, d2 = sign > 0 ? df2 + zi : df2 - zi;
        o_getctf2d(nx, d1, d2, astig, apix, vol, cs, ampwg, (real)1.0, ctf);
        for (long q = 0; q < n2; ++q) {
            a10[q] = a1[q] * ctf[q];
            b10[q] = b1[q] * ctf[q];
        }
        real pr = o_pr_ft(a10, b10, a20, b20, nx, (real)-1.0, maxres);
        if (prs) prs[i] = pr;
        if (i == 0 || pr < best) { best = pr; bz = zi; }
    }
    out2[0] = bz;
    out2[1] = best;
    free(a1); free(b1); free(a10); free(b10); free(ctf);
}

/* ------------------------------------------------------------------ */
/* Centre finding (icosprocess mode 3)                                  */
/* ------------------------------------------------------------------ */

/* img3d_mask: icosprocess/SRC/img3d_mask.f90:2-49 (shell mean in double,
 * cosine edge of half width 3 around the mask radius, outside = mean). */
EXPORT void o_img3d_mask(real* d, int nx, real mask, real ox, real oy, real oz)
{
    int halfw = 6;
    real pi = r_pi();
    real R1 = mask - (real)(halfw / 2);
    if (R1 <= 0) R1 = 0;
    real R2 = mask + (real)(halfw / 2);
    if (R2 > (real)(nx / 2)) R2 = (real)(nx / 2);
    double summ = 0;
    long num = 0;
    real x0 = ox + (real)(nx / 2), y0 = oy + (real)(nx / 2), z0 = oz + (real)(nx / 2);
    long n2 = (long)nx * nx, N = n2 * nx;
    for (long i = 0; i < N; ++i) {
        int iz = (int)(i / n2), iy = (int)((i % n2) / nx), ix = (int)(i % nx);
        real yy = (-y0 + iy) * (-y0 + iy), xx = (-x0 + ix) * (-x0 + ix), zz = (-z0 + iz) * (-z0 + iz);
        real ra = (real)sqrt(xx + yy + zz);
        if (ra > R1 && ra <= R2) { summ += d[i]; num++; }
    }
    if (num != 0) summ = summ / (double)(float)num;
    for (long i = 0; i < N; ++i) {
        int iz = (int)(i / n2), iy = (int)((i % n2) / nx), ix = (int)(i % nx);
        real yy = (-y0 + iy) * (-y0 + iy), xx = (-x0 + ix) * (-x0 + ix), zz = (-z0 + iz) * (-z0 + iz);
        real ra = (real)sqrt(xx + yy + zz);
        if (ra > R1 && ra <= R2) {
            real edge = ((real)1.0 + (real)cos(pi * (ra - R1) / (real)halfw)) / (real)2.0;
            d[i] = d[i] * edge + ((real)1.0 - edge) * (real)summ;
        }
        if (ra > R2) d[i] = (real)summ;
    }
}

/* The model transform of calc_center.f90:19-41: optional img3d_mask at the
 * origin, checkerboard, forward 3D FFT (scaled 1/N^3), checkerboard;
 * fr = Re, fi = -Im.  vol is overwritten by the masked model. */
EXPORT void o_center_model(real* vol, int nx, real mask, real* fr, real* fi)
{
    long n2 = (long)nx * nx, N = n2 * nx;
    if (mask >= 0) o_img3d_mask(vol, nx, mask, 0, 0, 0);
    for (long i = 0; i < N; ++i) {
        long s = i / n2 + (i % n2) / nx + i % nx;
        fr[i] = (s & 1) ? -vol[i] : vol[i];
        fi[i] = 0;
    }
    o_img3d_fft(fr, fi, nx, 1, 0);
    for (long i = 0; i < N; ++i) {
        long s = i / n2 + (i % n2) / nx + i % nx;
        if (s & 1) { fr[i] = -fr[i]; fi[i] = -fi[i]; }
        fi[i] = -fi[i];
    }
}

/* img2d_lowpass: icosprocess/SRC/img2d_lowpass.f90:2-37.  Gaussian
 * exp(-r^2 / (2 lp^2)) about the centred origin between a checkerboarded
 * forward and backward img2d_FFT; the imaginary part is dropped. */
EXPORT void o_img2d_lowpass(real* a, int nx, int lp)
{
    long N = (long)nx * nx;
    real* b = (real*)calloc(N, sizeof(real));
    real lowpass = (real)2.0 * (real)(lp * lp);
    o_checker2d(a, nx);
    o_img2d_fft(a, b, nx, 1, 1);
    for (long i = 0; i < N; ++i) {
        int ky = (int)(i / nx), hx = (int)(i % nx);
        real r2 = (real)((hx - nx / 2) * (hx - nx / 2) + (ky - nx / 2) * (ky - nx / 2));
        real H = (real)exp(-r2 / lowpass);
        a[i] = a[i] * H / (real)N;
        b[i] = b[i] * H / (real)N;
    }
    o_img2d_fft(a, b, nx, -1, 1);
    o_checker2d(a, nx);
    free(b);
}

/* img2d_applyCTF: icosprocess/SRC/img2d_applyCTF.f90:1-27. */
EXPORT void o_img2d_applyctf(real* a, int nx, const real* ctf)
{
    long N = (long)nx * nx;
    real* b = (real*)calloc(N, sizeof(real));
    o_checker2d(a, nx);
    o_img2d_fft(a, b, nx, 1, 1);
    for (long i = 0; i < N; ++i) { a[i] *= ctf[i]; b[i] *= ctf[i]; }
    o_img2d_fft(a, b, nx, -1, 1);
    o_checker2d(a, nx);
    free(b);
}

/* project_template2d.f90:3-80: the template of one orientation projected in
 * real space from an nx^3 model: for every voxel (ix, iy, iz) strictly inside
 * radius nx/2 - 1 of the centre, X = R (ix, iy, iz) - nx/2 + nx/2 (R of
 * euler_mat, row-major), the trilinear value of the model at X is added to
 * pixel (ix, iy) -- the voxel's own column, iz ascending -- as
 * slice + t1 + ... + t8 with ti = ((wx wy) wz) v.  out: nx^2. */
EXPORT void o_project_template2d(const real* vol, int nx, const real* R, real* out)
{
    const int h = nx / 2;
    const long n2 = (long)nx * nx;
    for (long i = 0; i < n2; ++i) out[i] = 0;
    for (int iz = 0; iz < nx; ++iz)
        for (int iy = 0; iy < nx; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const real xp[3] = {(real)ix - (real)h, (real)iy - (real)h, (real)iz - (real)h};
                const real d = (real)sqrt(xp[0] * xp[0] + xp[1] * xp[1] + xp[2] * xp[2]);
                if (d >= (real)(h - 1)) continue;
                real X[3];
                for (int a = 0; a < 3; ++a) X[a] = R[3 * a] * xp[0] + R[3 * a + 1] * xp[1] + R[3 * a + 2] * xp[2];
                for (int a = 0; a < 3; ++a) X[a] += (real)h;
                const int jx = (int)X[0], jy = (int)X[1], jz = (int)X[2];
                const real xb = X[0] - (real)jx, yb = X[1] - (real)jy, zb = X[2] - (real)jz;
                const long b = (long)jz * n2 + (long)jy * nx + jx;
                const real* v = vol;
                real s = out[(long)iy * nx + ix];
                s = s + ((real)1 - xb) * ((real)1 - yb) * ((real)1 - zb) * v[b];
                s = s + xb * ((real)1 - yb) * ((real)1 - zb) * v[b + 1];
                s = s + ((real)1 - xb) * yb * ((real)1 - zb) * v[b + nx];
                s = s + xb * yb * ((real)1 - zb) * v[b + nx + 1];
                s = s + ((real)1 - xb) * ((real)1 - yb) * zb * v[b + n2];
                s = s + xb * ((real)1 - yb) * zb * v[b + n2 + 1];
                s = s + ((real)1 - xb) * yb * zb * v[b + n2 + nx];
                s = s + xb * yb * zb * v[b + n2 + nx + 1];
                out[(long)iy * nx + ix] = s;
            }
}

/* project2d_FT3dTo2d: icosprocess/SRC/project2d_FT3dTo2d.f90:2-94.  Central
 * slice of (fr, fi) by trilinear interpolation inside radius nx/2-2, then a
 * checkerboarded backward img2d_FFT; the real part is the projection. */
EXPORT void o_project2d_ft(const real* fr, const real* fi, int nx, real th, real ph, real om, real* out)
{
    long N = (long)nx * nx, n2 = N;
    real R[9];
    real* oi = (real*)calloc(N, sizeof(real));
    o_euler_mat(th, ph, om, R);
    long n = 0;
    for (int iy = -nx / 2; iy < nx / 2; ++iy)
        for (int ix = -nx / 2; ix < nx / 2; ++ix, ++n) {
            out[n] = 0;
            real X[3] = {(real)ix, (real)iy, 0}, x[3];
            matvec(R, X, x);
            if ((real)sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) > (real)(nx / 2 - 2)) continue;
            x[0] += (real)(nx / 2); x[1] += (real)(nx / 2); x[2] += (real)(nx / 2);
            int hx = (int)x[0], ky = (int)x[1], lz = (int)x[2];
            real xb = x[0] - (real)hx, yb = x[1] - (real)ky, zb = x[2] - (real)lz;
            long c = (long)lz * n2 + (long)ky * nx + hx;
            const real* src[2] = {fr, fi};
            real v[2];
            for (int k = 0; k < 2; ++k) {
                const real* f = src[k];
                v[k] = ((real)1. - xb) * ((real)1. - yb) * ((real)1. - zb) * f[c] +
                       xb * ((real)1. - yb) * ((real)1. - zb) * f[c + 1] +
                       ((real)1. - xb) * yb * ((real)1. - zb) * f[c + nx] +
                       xb * yb * ((real)1. - zb) * f[c + nx + 1] +
                       ((real)1. - xb) * ((real)1. - yb) * zb * f[c + n2] +
                       xb * ((real)1. - yb) * zb * f[c + n2 + 1] +
                       ((real)1. - xb) * yb * zb * f[c + n2 + nx] +
                       xb * yb * zb * f[c + n2 + nx + 1];
            }
            out[n] = v[0];
            oi[n] = v[1];
        }
    o_checker2d(out, nx);
    o_checker2d(oi, nx);
    o_img2d_fft(out, oi, nx, -1, 1);
    o_checker2d(out, nx);
    free(oi);
}

/* cross_correlation: icosprocess/SRC/cross_correlation.f90:1-28.  a1 is
 * overwritten by the correlation map (peak at the centre for zero shift). */
EXPORT void o_cross_correlation(real* a1, real* a2, int nx)
{
    long N = (long)nx * nx;
    real* b1 = (real*)calloc(N, sizeof(real));
    real* b2 = (real*)calloc(N, sizeof(real));
    o_img2d_fft(a1, b1, nx, 1, 1);
    o_img2d_fft(a2, b2, nx, 1, 1);
    for (long i = 0; i < N; ++i) {
        int ky = (int)(i / nx), hx = (int)(i % nx);
        real x = a1[i] * a2[i] + b1[i] * b2[i];
        real y = b1[i] * a2[i] - a1[i] * b2[i];
        a1[i] = ((hx + ky) & 1) ? -x : x;
        b1[i] = ((hx + ky) & 1) ? -y : y;
    }
    o_img2d_fft(a1, b1, nx, -1, 1);
    free(b1); free(b2);
}

/* guass_regression: icosprocess/SRC/guass_regression.f90:1-56.  Least-
 * squares quadric on log z over the 3x3 neighbourhood z[(j+1)*3 + i+1] =
 * z(i, j); returns the sub-pixel offset and the fitted peak value. */
EXPORT void o_gauss_regression(const real* z, real* dx, real* dy, real* cc)
{
    real c10 = 0, c01 = 0, c11 = 0, c20 = 0, c02 = 0, c00 = 0;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) c10 += (real)i * (real)log(z[(j + 1) * 3 + i + 1]);
    c10 /= (real)6.0;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) c01 += (real)j * (real)log(z[(j + 1) * 3 + i + 1]);
    c01 /= (real)6.0;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) c11 += (real)(i * j) * (real)log(z[(j + 1) * 3 + i + 1]);
    c11 /= (real)4.0;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) c20 += (real)(3 * i * i - 2) * (real)log(z[(j + 1) * 3 + i + 1]);
    c20 /= (real)6.0;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) c02 += (real)(3 * j * j - 2) * (real)log(z[(j + 1) * 3 + i + 1]);
    c02 /= (real)6.0;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) c00 += (real)(5 - 3 * i * i - 3 * j * j) * (real)log(z[(j + 1) * 3 + i + 1]);
    c00 /= (real)9.0;
    real den = (real)4.0 * c20 * c02 - c11 * c11;
    *dx = (c11 * c01 - (real)2.0 * c10 * c02) / den;
    *dy = (c11 * c10 - (real)2.0 * c01 * c20) / den;
    *cc = (real)exp(c00 - c20 * *dx * *dx - c11 * *dx * *dy - c02 * *dy * *dy);
}

/* The peak step shared by calc_center.f90:70-91 and
 * calc_center_selfrot180.f90:58-79: first maximum in memory order, the 3x3
 * values around it at flat indices (iy0+j)*nx + ix0+i (a neighbour past a
 * row end reads the adjacent row; one outside the image reads 0), the
 * quadric fit, the shift (halved for the 180-degree self-correlation),
 * zeroed at |shift| >= nx/16.  out = {x, y, cc}; cc is not clipped. */
EXPORT void o_center_peak(const real* ccmap, int nx, int halve, real* out)
{
    long N = (long)nx * nx, imax = 0;
    for (long i = 1; i < N; ++i)
        if (ccmap[i] > ccmap[imax]) imax = i;
    int iy0 = (int)(imax / nx), ix0 = (int)(imax % nx);
    real z[9];
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) {
            long k = (long)(iy0 + j) * nx + ix0 + i;
            z[(j + 1) * 3 + i + 1] = (k >= 0 && k < N) ? ccmap[k] : (real)0;
        }
    real dx, dy, cc;
    o_gauss_regression(z, &dx, &dy, &cc);
    real cx = (real)(ix0 - nx / 2) + dx, cy = (real)(iy0 - nx / 2) + dy;
    if (halve) { cx = cx / (real)2.0; cy = cy / (real)2.0; }
    if (fabs(cx) >= (real)nx / (real)16.0) cx = 0;
    if (fabs(cy) >= (real)nx / (real)16.0) cy = 0;
    out[0] = cx; out[1] = cy; out[2] = cc;
}

/* calc_center.f90:44-76 for one particle: img is the particle image
 * (overwritten), (fr, fi) the model transform, ctf the particle's CTF (NULL:
 * applyCTF off), mask < 0: imgmask off. */
EXPORT void o_calc_center_one(const real* fr, const real* fi, real* img, int nx, real x, real y, real th, real ph,
                              real om, const real* ctf, real mask, int maxR, real* out)
{
    long N = (long)nx * nx;
    real* a2 = (real*)malloc(sizeof(real) * N);
    if (mask >= 0) o_img2d_mask(img, nx, mask, x, y);
    o_img2d_lowpass(img, nx, maxR);
    o_project2d_ft(fr, fi, nx, th, ph, om, a2);
    if (ctf) o_img2d_applyctf(a2, nx, ctf);
    if (mask >= 0) o_img2d_mask(a2, nx, mask, 0, 0);
    o_img2d_lowpass(a2, nx, maxR);
    o_img2d_norm(img, nx);
    o_img2d_norm(a2, nx);
    o_cross_correlation(img, a2, nx);
    o_center_peak(img, nx, 0, out);
    free(a2);
}

/* calc_center_selfrot180.f90:33-79 for one particle (img overwritten). */
EXPORT void o_center_selfrot180_one(real* img, int nx, real mask, int maxR, real* out)
{
    long N = (long)nx * nx;
    real* a2 = (real*)calloc(N, sizeof(real));
    o_img2d_lowpass(img, nx, maxR);
    if (mask >= 0) o_img2d_mask(img, nx, mask, 0, 0);
    o_img2d_norm(img, nx);
    for (long i = 0; i < N; ++i) {
        int ky = (int)(i / nx), hx = (int)(i % nx);
        if (hx == 0 || ky == 0) img[i] = 0;
        else a2[(long)(nx - ky) * nx + (nx - hx)] = img[i];
    }
    o_cross_correlation(img, a2, nx);
    o_center_peak(img, nx, 1, out);
    free(a2);
}

/* ------------------------------------------------------------------ */
/* Real-space icosahedral averaging                                     */
/* ------------------------------------------------------------------ */

/* Adds the eight trilinear taps of one operator to acc, term by term in
 * the order of img3d_icosavg.f90:75-83 (acc + t1 + ... + t8, left to
 * right): x = S x0 + nx/2, taps by flat index n1 = iz nx^2 + iy nx + ix
 * (int() truncation), read from the first len values of img (0 outside). */
static real icosavg_taps(real acc, const real* img, long len, int nx, const real* S, int hx, int ky, int lz)
{
    real x0[3] = {(real)hx, (real)ky, (real)lz}, x[3];
    matvec(S, x0, x);
    for (int k = 0; k < 3; ++k) x[k] += (real)(nx / 2);
    int ix = (int)x[0], iy = (int)x[1], iz = (int)x[2];
    real ax = x[0] - (real)ix, ay = x[1] - (real)iy, az = x[2] - (real)iz;
    long n2 = (long)nx * nx;
    long n1 = (long)iz * n2 + (long)iy * nx + ix;
    long idx[8] = {n1, n1 + 1, n1 + nx, n1 + nx + 1, n1 + n2, n1 + 1 + n2, n1 + nx + n2, n1 + nx + 1 + n2};
    real v[8];
    for (int k = 0; k < 8; ++k) v[k] = (idx[k] >= 0 && idx[k] < len) ? img[idx[k]] : (real)0;
    acc = acc + ((real)1.0 - ax) * ((real)1.0 - ay) * ((real)1.0 - az) * v[0];
    acc = acc + ax * ((real)1.0 - ay) * ((real)1.0 - az) * v[1];
    acc = acc + ((real)1.0 - ax) * ay * ((real)1.0 - az) * v[2];
    acc = acc + ax * ay * ((real)1.0 - az) * v[3];
    acc = acc + ((real)1.0 - ax) * ((real)1.0 - ay) * az * v[4];
    acc = acc + ax * ((real)1.0 - ay) * az * v[5];
    acc = acc + ((real)1.0 - ax) * ay * az * v[6];
    acc = acc + ax * ay * az * v[7];
    return acc;
}

/* icosprocess/SRC/img3d_icosavg.f90:2-126: every voxel within r0 of the
 * centre is the sum over the nops operators of the interpolated map at
 * S x, divided by nops; voxels beyond r0 become 0.  The first nx^3 values of
 * img are replaced; taps may read anywhere in its first len values. */
EXPORT void o_img3d_icosavg_full(real* img, long len, int nx, real r0, const real* ops, int nops)
{
    long n2 = (long)nx * nx, N = n2 * nx;
    real* acc = (real*)calloc(N, sizeof(real));
    for (int lz = -nx / 2; lz < nx / 2; ++lz)
        for (int ky = -nx / 2; ky < nx / 2; ++ky)
            for (int hx = -nx / 2; hx < nx / 2; ++hx) {
                real r = (real)sqrt((real)(hx * hx + ky * ky + lz * lz));
                if (r > r0) continue;
                long i3 = (long)(lz + nx / 2) * n2 + (long)(ky + nx / 2) * nx + (hx + nx / 2);
                for (int s = 0; s < nops; ++s) acc[i3] = icosavg_taps(acc[i3], img, len, nx, ops + 9 * s, hx, ky, lz);
            }
    for (long i = 0; i < N; ++i) img[i] = acc[i] / (real)nops;
    free(acc);
}

/* merge_icosprocess_subrec/SRC/img3d_icosavg.f90:1-91: the 60-operator sum
 * over the quarter lz, ky in [0, nx/2-1] within r0, then the copies of its
 * second loop in loop order -- to (-hx, -ky, lz), and for lz > 0 also to
 * (-hx, ky, -lz) and (hx, -ky, -lz) -- by flat index, so an x index of nx
 * lands on x = 0 of the next row; copies outside the nx^3 array are
 * dropped.  Divided by 60.  Reads as o_img3d_icosavg_full. */
EXPORT void o_img3d_icosavg_merge(real* img, long len, int nx, real r0, const real* ops)
{
    long n2 = (long)nx * nx, N = n2 * nx;
    real* acc = (real*)calloc(N, sizeof(real));
    for (int lz = 0; lz < nx / 2; ++lz)
        for (int ky = 0; ky < nx / 2; ++ky)
            for (int hx = -nx / 2; hx < nx / 2; ++hx) {
                real r = (real)sqrt((real)(hx * hx + ky * ky + lz * lz));
                if (r > r0) continue;
                long i3 = (long)(lz + nx / 2) * n2 + (long)(ky + nx / 2) * nx + (hx + nx / 2);
                for (int s = 0; s < 60; ++s) acc[i3] = icosavg_taps(acc[i3], img, len, nx, ops + 9 * s, hx, ky, lz);
            }
    for (int lz = 0; lz < nx / 2; ++lz)
        for (int ky = 0; ky < nx / 2; ++ky)
            for (int hx = -nx / 2; hx < nx / 2; ++hx) {
                real r = (real)sqrt((real)(hx * hx + ky * ky + lz * lz));
                if (r > r0) continue;
                long i3 = (long)(lz + nx / 2) * n2 + (long)(ky + nx / 2) * nx + (hx + nx / 2);
                long t[3];
                int nt;
                if (lz == 0) {
                    t[0] = (long)(lz + nx / 2) * n2 + (long)(-ky + nx / 2) * nx + (-hx + nx / 2);
                    nt = 1;
                } else {
                    t[0] = (long)(lz + nx / 2) * n2 + (long)(-ky + nx / 2) * nx + (-hx + nx / 2);
                    t[1] = (long)(-lz + nx / 2) * n2 + (long)(ky + nx / 2) * nx + (-hx + nx / 2);
                    t[2] = (long)(-lz + nx / 2) * n2 + (long)(-ky + nx / 2) * nx + (hx + nx / 2);
                    nt = 3;
                }
                for (int k = 0; k < nt; ++k)
                    if (t[k] >= 0 && t[k] < N) acc[t[k]] = acc[i3];
            }
    for (long i = 0; i < N; ++i) img[i] = acc[i] / (real)60;
    free(acc);
}

/* ------------------------------------------------------------------ */
/* local_reconstruct's symmetry operators                               */
/* ------------------------------------------------------------------ */

static void mm3t(const real* A, const real* B, real* C)   /* C = A B^T */
{
    real T[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T[3 * j + i] = B[3 * i + j];
    matmul3(A, T, C);
}

/* C_matrix.f90:1-17: r1 = rotation about z by 2 pi / N, r_i = r_(i-1) r1
 * (so the identity comes last, up to rounding). */
static void o_c_matrix(int N, real* r)
{
    real pi = r_pi();
    real phi = (real)2.0 * pi / (real)N;
    real* m = r;
    for (int i = 0; i < 9; ++i) m[i] = 0;
    m[0] = (real)cos(phi); m[1] = -(real)sin(phi);
    m[3] = (real)sin(phi); m[4] = (real)cos(phi);
    m[8] = (real)1.0;
    for (int i = 1; i < N; ++i) matmul3(r + 9 * (i - 1), r, r + 9 * i);
}

/* icos_matrix_i2.f90:1-56: the 60 operators built from a 5-fold tilted by
 * v onto the 2-fold frame (ryv rz5 ryv^T), the 2-folds about x and y and the
 * 3-fold permutation about (1,1,1), in that order of products. */
static void o_icos_matrix_i2(real* r)
{
    real d2r = r_pi() / (real)180.0;
    real a5 = (real)72.0 * d2r;
    real v = (real)acos((real)cos((real)60.0 * d2r) / (real)sin((real)36.0 * d2r));
    real rz5[9] = {(real)cos(a5), -(real)sin(a5), 0, (real)sin(a5), (real)cos(a5), 0, 0, 0, 1};
    real ryv[9] = {(real)cos(v), 0, (real)sin(v), 0, 1, 0, -(real)sin(v), 0, (real)cos(v)};
    real t[9], z5[9];
    mm3t(rz5, ryv, t);           /* rz5 ryv^T */
    matmul3(ryv, t, z5);         /* ryv (rz5 ryv^T) */
    real rx2[9] = {1, 0, 0, 0, -1, 0, 0, 0, -1};
    real ry2[9] = {-1, 0, 0, 0, 1, 0, 0, 0, -1};
    real r3[9] = {0, 0, 1, 1, 0, 0, 0, 1, 0};
    for (int i = 0; i < 60 * 9; ++i) r[i] = 0;
    r[0] = 1; r[4] = 1; r[8] = 1;
    for (int i = 1; i < 5; ++i) matmul3(r + 9 * (i - 1), z5, r + 9 * i);
    for (int i = 5; i < 10; ++i) matmul3(rx2, r + 9 * (i - 5), r + 9 * i);
    for (int i = 10; i < 20; ++i) matmul3(ry2, r + 9 * (i - 10), r + 9 * i);
    for (int i = 20; i < 40; ++i) matmul3(r3, r + 9 * (i - 20), r + 9 * i);
    for (int i = 40; i < 60; ++i) matmul3(r3, r + 9 * (i - 20), r + 9 * i);
}

/* icos_matrix_i3.f90 / icos_matrix_i5.f90: the I2 operators conjugated by a
 * rotation about x by -w (2f-3f angle) or about y by v (2f-5f angle):
 * r_i = rot0 (r_i rot0^T). */
static void o_icos_matrix_conj(real* r, int kind)
{
    real d2r = r_pi() / (real)180.0;
    real rot0[9];
    o_icos_matrix_i2(r);
    if (kind == 3) {
        real w = (real)acos((real)cos((real)36.0 * d2r) / (real)sin((real)60.0 * d2r));
        real c = (real)cos(-w), s = (real)sin(-w);
        real m[9] = {1, 0, 0, 0, c, s, 0, -s, c};
        for (int i = 0; i < 9; ++i) rot0[i] = m[i];
    } else {
        real v = (real)acos((real)cos((real)60.0 * d2r) / (real)sin((real)36.0 * d2r));
        real m[9] = {(real)cos(v), 0, (real)sin(v), 0, 1, 0, -(real)sin(v), 0, (real)cos(v)};
        for (int i = 0; i < 9; ++i) rot0[i] = m[i];
    }
    for (int i = 0; i < 60; ++i) {
        real t[9];
        mm3t(r + 9 * i, rot0, t);      /* r_i rot0^T */
        matmul3(rot0, t, r + 9 * i);   /* rot0 (r_i rot0^T) */
    }
}

/* local_reconstruct/SRC/0calc_initialvalues.f90:33-62: kind 2, 3, 5 for
 * I2 / I3 / I5 (60 operators), 0 for C_matrix with N operators.  Returns
 * the count. */
EXPORT int o_local_reconstruct_ops(int kind, int N, real* r)
{
    if (kind == 2) { o_icos_matrix_i2(r); return 60; }
    if (kind == 3 || kind == 5) { o_icos_matrix_conj(r, kind); return 60; }
    o_c_matrix(N, r);
    return N;
}

/* ------------------------------------------------------------------ */
/* local_box                                                            */
/* ------------------------------------------------------------------ */

/* Eularmatrix.f90: euler_mat's ZYZ matrix with deg2rad = 4 atan2(1,1)/180. */
static void o_eularmatrix(real th0, real ph0, real om0, real* r)
{
    real d2r = (real)4.0 * (real)atan2((real)1.0, (real)1.0) / (real)180.0;
    real th = th0 * d2r, ph = ph0 * d2r, om = om0 * d2r;
    real cp = (real)cos(ph), ct = (real)cos(th), co = (real)cos(om);
    real sp = (real)sin(ph), st = (real)sin(th), so = (real)sin(om);
    r[0] = cp * ct * co - sp * so; r[1] = -cp * ct * so - sp * co; r[2] = cp * st;
    r[3] = sp * ct * co + cp * so; r[4] = -sp * ct * so + cp * co; r[5] = sp * st;
    r[6] = -st * co;               r[7] = st * so;                 r[8] = ct;
}

/* equalort.f90: the orientation of S_i R(theta, phi, omega). */
static void o_equalort_one(real th, real ph, real om, const real* S, real* ang)
{
    real rad2deg = (real)45.0 / (real)atan2((real)1.0, (real)1.0);
    real E[9], r[9];
    o_eularmatrix(th, ph, om, E);
    matmul3(S, E, r);
    if (r[8] >= (real)1.0) ang[0] = 0;
    else if (r[8] < (real)-1.0) ang[0] = (real)180.0;
    else ang[0] = (real)acos(r[8]) * rad2deg;
    ang[1] = (real)atan2(r[5], r[2]) * rad2deg;
    ang[2] = (real)atan2(r[7], -r[6]) * rad2deg;
}

/* loc_box_coordinate.f90:16-41 for one particle: per operator the local
 * orientation, the sub-particle centre R^T x0 + nx/2 + (x, y) split into
 * integer (hx0, ky0) and fraction (local x, y), and the height
 * (z - nx/2) apix.  out: nops x {lth, lph, lom, hx0, ky0, lx, ly, lz}. */
EXPORT void o_local_box_coords(real th, real ph, real om, real x, real y, int nops, const real* ops, const real* x0,
                               int nx, real apix, real* out)
{
    for (int i = 0; i < nops; ++i) {
        real ang[3], R[9], c[3];
        o_equalort_one(th, ph, om, ops + 9 * i, ang);
        o_eularmatrix(ang[0], ang[1], ang[2], R);
        mattvec(R, x0, c);
        for (int k = 0; k < 3; ++k) c[k] = c[k] + (real)(nx / 2);
        c[0] = c[0] + x;
        c[1] = c[1] + y;
        int hx0 = (int)c[0], ky0 = (int)c[1];
        real* o = out + 8 * i;
        o[0] = ang[0]; o[1] = ang[1]; o[2] = ang[2];
        o[3] = (real)hx0; o[4] = (real)ky0;
        o[5] = c[0] - (real)hx0; o[6] = c[1] - (real)ky0;
        o[7] = (c[2] - (real)(nx / 2)) * apix;
    }
}

/* img2d_centshift.f90 (local_box): checkerboard, forward FFT (b = +Im), the
 * phase ramp of (x0, y0) without the nx/2 recentring, backward FFT, real
 * part, checkerboard -- a circular sub-pixel shift of the image. */
EXPORT void o_img2d_realshift(real* a, int nx, real x0, real y0)
{
    long N = (long)nx * nx;
    real* b = (real*)calloc(N, sizeof(real));
    real pi = r_pi();
    o_checker2d(a, nx);
    o_img2d_fft(a, b, nx, 1, 1);
    for (long i = 0; i < N; ++i) b[i] = -b[i];
    for (long i = 0; i < N; ++i) {
        int ky = (int)(i / nx) - nx / 2, hx = (int)(i % nx) - nx / 2;
        real phase = (real)2.0 * pi * (x0 * (real)hx + y0 * (real)ky) / (real)nx;
        real fr = a[i] * (real)cos(phase) + b[i] * (real)sin(phase);
        real fi = -a[i] * (real)sin(phase) + b[i] * (real)cos(phase);
        a[i] = fr;
        b[i] = fi;
    }
    o_img2d_fft(a, b, nx, -1, 1);
    o_checker2d(a, nx);
    free(b);
}

/* loc_box.f90:50-91 for one particle: the image shifted by -(x, y) and
 * normalised; per operator the orientation equalort gives, c = R^T x0 with
 * R that orientation's matrix, the sub x sub box whose corner is
 * int(c) + nx/2 - sub/2 (int truncates toward zero; one periodic wrap per
 * axis, as img2d_local_box), normalised.  boxes: nops x sub^2;
 * meta: nops x {lth, lph, lom, c1 - int(c1), c2 - int(c2), c3 apix}. */
EXPORT void o_local_box_stack(const real* img, int nx, real th, real ph, real om, real x, real y, int nops,
                              const real* ops, const real* x0, int sub, real apix, real* boxes, real* meta)
{
    long N = (long)nx * nx;
    real* a = (real*)malloc(sizeof(real) * N);
    for (long i = 0; i < N; ++i) a[i] = img[i];
    o_img2d_realshift(a, nx, -x, -y);
    o_img2d_norm(a, nx);
    for (int i = 0; i < nops; ++i) {
        real ang[3], R[9], c[3];
        o_equalort_one(th, ph, om, ops + 9 * i, ang);
        o_eularmatrix(ang[0], ang[1], ang[2], R);
        mattvec(R, x0, c);
        int hx0 = (int)c[0], ky0 = (int)c[1];
        real* m = meta + 6 * i;
        m[0] = ang[0]; m[1] = ang[1]; m[2] = ang[2];
        m[3] = c[0] - (real)hx0; m[4] = c[1] - (real)ky0;
        m[5] = c[2] * apix;
        real* b = boxes + (long)sub * sub * i;
        o_local_box(a, nx, hx0 + nx / 2, ky0 + nx / 2, b, sub);
        o_img2d_norm(b, sub);
    }
    free(a);
}
