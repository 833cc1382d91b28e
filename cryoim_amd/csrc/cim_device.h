// Device helpers shared by the kernels: CTF, centring phase, reductions.
#pragma once
#include <hip/hip_runtime.h>

#include "cim_internal.h"

namespace cim {

// Float-float numbers (hi + lo, |lo| <= ulp(hi) / 2): about 48 significant bits
// from float32 arithmetic and fused multiply-adds, for the few per-pixel
// phases whose arguments reach hundreds of radians.
struct ff32 {
    float h, l;
};

__host__ __device__ inline ff32 ff_from_double(double d) {
    const float h = (float)d;
    return ff32{h, (float)(d - (double)h)};
}

__device__ __forceinline__ ff32 ff_quick(float s, float e) {
    const float t = __fadd_rn(s, e);
    return ff32{t, __fsub_rn(e, __fsub_rn(t, s))};
}

// a + b exactly as a float-float (TwoSum)
__device__ __forceinline__ ff32 ff_two_sum(float a, float b) {
    const float s = __fadd_rn(a, b);
    const float bp = __fsub_rn(s, a);
    return ff32{s, __fadd_rn(__fsub_rn(a, __fsub_rn(s, bp)), __fsub_rn(b, bp))};
}

// a * b exactly as a float-float (TwoProd by fused multiply-add)
__device__ __forceinline__ ff32 ff_two_prod(float a, float b) {
    const float p = __fmul_rn(a, b);
    return ff32{p, __builtin_fmaf(a, b, -p)};
}

__device__ __forceinline__ ff32 ff_mul(ff32 x, ff32 y) {
    const ff32 p = ff_two_prod(x.h, y.h);
    return ff_quick(p.h, __fadd_rn(p.l, __fadd_rn(__fmul_rn(x.h, y.l), __fmul_rn(x.l, y.h))));
}

__device__ __forceinline__ ff32 ff_add(ff32 x, ff32 y) {
    const ff32 s = ff_two_sum(x.h, y.h);
    return ff_quick(s.h, __fadd_rn(s.l, __fadd_rn(x.l, y.l)));
}

// 1 / m for an integer m < 2^24, as a float-float (one Newton step)
__device__ __forceinline__ ff32 ff_recip_int(int m) {
    const float mf = (float)m;
    const float q = 1.0f / mf;
    return ff32{q, __fmul_rn(__builtin_fmaf(-q, mf, 1.0f), q)};
}

// x (float-float, |x| < 2^20) reduced to [-pi, pi] modulo 2 pi, as a float
__device__ __forceinline__ float ff_reduce_2pi(ff32 x) {
    const float two_pi_h = 6.28318548202514648f, two_pi_l = -1.74845553e-7f;   // 2 pi = h + l
    const float k = __builtin_rintf(x.h * 0.159154943091895336f);
    // x.h - k two_pi_h is exact: its bits lie between ulp(x.h) and 2^3
    const float r = __builtin_fmaf(-k, two_pi_h, x.h);
    return __fadd_rn(__fsub_rn(r, __fmul_rn(k, two_pi_l)), x.l);
}

struct CtfConst {
    ff32 a;          // pi lambda / apix^2: the defocus term's factor (phase = a df s - b s^2, s = n / nx^2)
    ff32 b;          // (pi / 2) Cs lambda^3 / apix^4
    float ampwg;     // amplitude contrast
    float sq;        // sqrt(1 - ampwg^2)
    float df_scale;  // 1e4 (um) or 1 (A)
    int corr_cycles; // defocus series: C steps each side (0: one defocus)
    float corr_step; // step of the series in the df unit
};

inline CtfConst make_ctf_const(const cim_params* p) {
    CtfConst c;
    // getCTF2d.f90:12-13 (wavelength, Cs in A), folded with the pixel size on the host in float64
    const double lam = 0.387818 / std::sqrt((double)p->vol * (1.0 + 0.978459e-3 * (double)p->vol));
    const double cs = (double)p->cs * 1.0e7;
    const double pi = 3.14159265358979323846, ap2 = (double)p->apix * (double)p->apix;
    c.a = ff_from_double(pi * lam / ap2);
    c.b = ff_from_double(0.5 * pi * cs * lam * lam * lam / (ap2 * ap2));
    c.ampwg = p->ampwg;
    c.sq = (float)std::sqrt(1.0 - (double)p->ampwg * p->ampwg);
    c.df_scale = p->df_scale;
    c.corr_cycles = p->corr_cycles;
    c.corr_step = p->corr_step;
    return c;
}

// getCTF2d (icosprocess/SRC/getCTF2d.f90:17-27) at centred frequency (hx, ky),
// in float32 and float-float arithmetic.  With n = hx^2 + ky^2, s = n / nx^2
// (the squared frequency times apix^2) and phi = atan2(ky, hx), the
// reference's defocus df1 cos^2(phi - ast) + df2 sin^2(phi - ast) is
// dfm + dfd cos 2(phi - ast), and n cos 2(phi - ast) = (hx^2 - ky^2) cos 2ast
// + 2 hx ky sin 2ast, so no angle of the pixel is formed; the phase
// a (df s) - b s^2 (about 100 rad at high frequency) is carried as a
// float-float and reduced modulo 2 pi before its sine and cosine.  Its
// error is about 2e-7 rad (3e-6 rad with astigmatism, from cos 2ast and
// sin 2ast in float32).
__device__ __forceinline__ float ctf_value(int hx, int ky, int nx, float df1, float df2, float astig_deg,
                                           const CtfConst& c) {
    const int n = hx * hx + ky * ky;
    if (n == 0) return -c.ampwg;     // phase 0: -(sq sin 0 + ampwg cos 0)
    const ff32 inv = ff_recip_int(nx * nx);
    const ff32 u = ff_mul(ff32{(float)n, 0.f}, inv);
    // dfm = (df1 + df2) / 2 and dfd = (df1 - df2) / 2 in A, as float-floats
    ff32 dfm = ff_two_sum(df1, df2), dfd = ff_two_sum(df1, -df2);
    dfm = ff_mul(ff32{0.5f * dfm.h, 0.5f * dfm.l}, ff32{c.df_scale, 0.f});
    dfd = ff_mul(ff32{0.5f * dfd.h, 0.5f * dfd.l}, ff32{c.df_scale, 0.f});
    float s2a = 0.f, c2a = 1.f;
    if (astig_deg != 0.f) sincosf(astig_deg * 0.0349065850398865915f, &s2a, &c2a);
    const ff32 q = ff_add(ff_two_prod((float)(hx * hx - ky * ky), c2a), ff_two_prod((float)(2 * hx * ky), s2a));
    const ff32 dfs = ff_add(ff_mul(dfm, u), ff_mul(ff_mul(dfd, q), inv));
    const ff32 ph1 = ff_mul(c.a, dfs), ph2 = ff_mul(c.b, ff_mul(u, u));
    const float ph = ff_reduce_2pi(ff_add(ph1, ff32{-ph2.h, -ph2.l}));
    float sn, co;
    sincosf(ph, &sn, &co);
    return -(c.sq * sn + c.ampwg * co);
}

// CTF factors of one inserted sample: fa, fb take ca and wg takes cw.
// Without a defocus series ca = ctf and cw = ctf^2.  With C = corr_cycles > 0
// (reconstruct_corrCTF.f90:80-88, 137-145): ca = the sum over k = -C..C of
// the CTF at df + k*step (defocus summed in float32 as there), and
// cw = (2C+1) ctf_C^2 -- every k adds the square of ctf2d, the last CTF the
// series computed (k = +C), not of its own.
__device__ __forceinline__ void ctf_factors(int hx, int ky, int nx, const float* pr, const CtfConst& c, float& ca,
                                            float& cw) {
    if (c.corr_cycles <= 0) {
        const float v = ctf_value(hx, ky, nx, pr[CIM_P_DF1], pr[CIM_P_DF2], pr[CIM_P_ASTIG], c);
        ca = v;
        cw = v * v;
        return;
    }
    float s = 0.f, last = 0.f;
    for (int k = -c.corr_cycles; k <= c.corr_cycles; ++k) {
        const float d1 = pr[CIM_P_DF1] + c.corr_step * (float)k, d2 = pr[CIM_P_DF2] + c.corr_step * (float)k;
        last = ctf_value(hx, ky, nx, d1, d2, pr[CIM_P_ASTIG], c);
        s += last;
    }
    ca = s;
    cw = (float)(2 * c.corr_cycles + 1) * (last * last);
}

// exp(-i * 2 pi * ((x0 + s*nx/2) hx + (y0 + s*nx/2) ky) / nx) with s = recentre.
// The nx/2 part contributes exactly (-1)^(hx+ky); the rest, (x0 hx + y0 ky) / nx
// turns, is formed as a float-float (exact products, one rounding of about
// 2^-48 relative) and reduced to [-1/2, 1/2] turn before sincospi, so the
// large-argument rounding of the float formula in img2d_FT_centshift.f90:11
// does not enter (error about 1e-7 rad).
__device__ __forceinline__ float2 shift_phase(int hx, int ky, int nx, float x0, float y0, int recentre) {
    const ff32 t = ff_mul(ff_add(ff_two_prod(x0, (float)hx), ff_two_prod(y0, (float)ky)), ff_recip_int(nx));
    const float f = __fadd_rn(__fsub_rn(t.h, __builtin_rintf(t.h)), t.l);   // turns, |f| <= 1/2 + tiny
    float fs, fc;
    sincospif(2.0f * f, &fs, &fc);
    // even nx: (nx/2)(hx+ky)/nx = (hx+ky)/2 turns, an odd sum flips the sign
    if (recentre && ((hx + ky) & 1)) {
        fc = -fc;
        fs = -fs;
    }
    return make_float2(fc, fs);
}

// Apply the shift of img2d_FT_centshift: fr = a cos + b sin, fi = -a sin + b cos.
__device__ __forceinline__ float2 apply_shift(float2 v, float2 cs) {
    return make_float2(v.x * cs.x + v.y * cs.y, -v.x * cs.y + v.y * cs.x);
}

// euler_mat.f90:3-21 in float32: R row-major, R[3*i+j] = RotMat(i+1, j+1).
__device__ __forceinline__ void euler_f(float th0, float ph0, float om0, float* R) {
    const float d2r = 0.0174532925199432957692f;   // atan2(1,1)/45
    float th = th0 * d2r, ph = ph0 * d2r, om = om0 * d2r;
    float c1 = cosf(ph), c2 = cosf(th), c3 = cosf(om), s1 = sinf(ph), s2 = sinf(th), s3 = sinf(om);
    R[0] = c1 * c2 * c3 - s1 * s3; R[1] = -c1 * c2 * s3 - c3 * s1; R[2] = c1 * s2;
    R[3] = c2 * c3 * s1 + c1 * s3; R[4] = -c2 * s1 * s3 + c1 * c3; R[5] = s1 * s2;
    R[6] = -c3 * s2;               R[7] = s2 * s3;                 R[8] = c2;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum of a double (blockDim multiple of 64, <= 1024).
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) r += sh[i];
    return r;
}

// Block-wide compensated sum of each thread's float-float partial (blockDim a
// multiple of 64, <= 1024; sh holds 16 entries): float32 arithmetic only, in a
// fixed order (an xor tree over the lanes -- ff_add is commutative, so every
// lane holds the same value -- then the waves in order), so the result does
// not depend on timing.  Every thread returns the total.
__device__ __forceinline__ ff32 block_sum_ff(ff32 v, ff32* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = ff_add(v, ff32{__shfl_xor(v.h, o, 64), __shfl_xor(v.l, o, 64)});
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    ff32 r = sh[0];
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 1; i < nw; ++i) r = ff_add(r, sh[i]);
    return r;
}

// Sum of nblk partials {sum hi, lo, count hi, lo} (4 floats each, written by
// one block each) in a fixed order: out = {sum hi, lo, count hi, lo}.  One
// block of 256 threads.  (static: each translation unit launches its own copy)
static __global__ void __launch_bounds__(256) k_sum_partials(const float* __restrict__ part, int nblk,
                                                      float* __restrict__ out) {
    __shared__ ff32 sh[4];
    ff32 s{0.f, 0.f}, c{0.f, 0.f};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        s = ff_add(s, ff32{part[4 * b], part[4 * b + 1]});
        c = ff_add(c, ff32{part[4 * b + 2], part[4 * b + 3]});
    }
    s = block_sum_ff(s, sh);
    c = block_sum_ff(c, sh);
    if (threadIdx.x == 0) {
        out[0] = s.h;
        out[1] = s.l;
        out[2] = c.h;
        out[3] = c.l;
    }
}


// x - a for a float-float a, rounded to float
__device__ __forceinline__ float ff_sub_from(float x, ff32 a) { return __fsub_rn(__fsub_rn(x, a.h), a.l); }

// Image preparation ([norm] -> [mask] -> [norm] -> checkerboard) of one pixel
// from its image's sums (k_prep_real): the same arithmetic wherever a value is
// formed, so the preparation pass and a transform that forms the values as it
// reads the image give the same bits.
struct PrepStats {
    ff32 avg;         // mean before the mask (0 unless do_norm == 1)
    float inv_dev;    // 1 / deviation before the mask (1 unless do_norm == 1)
    float ring_mean;  // mean over the soft edge of the mask
    ff32 avg2;        // mean after the mask (0 unless do_norm == 2)
    float inv_dev2;   // 1 / deviation after the mask (1 unless do_norm == 2)
    float x0, y0;     // mask centre (pixels)
};
struct PrepCfg {
    int do_norm, do_mask;
    float R1, R2;     // soft edge of the mask: (R1, R2]
};
constexpr int PREP_HALFW = 6;   // width of the soft edge

__host__ __device__ inline PrepCfg prep_cfg(int box, int do_norm, int do_mask, float mask_r) {
    PrepCfg c{do_norm, do_mask, 0.f, 0.f};
    if (do_mask) {
        c.R1 = mask_r - (float)(PREP_HALFW / 2);
        if (c.R1 <= 0.f) c.R1 = 0.f;
        c.R2 = mask_r + (float)(PREP_HALFW / 2);
        if (c.R2 > (float)(box / 2)) c.R2 = (float)(box / 2);
    }
    return c;
}

// [norm] -> [mask] of pixel (ix, iy) with value x.  Contraction is pinned to
// each expression, so the value does not depend on the kernel it is inlined in.
__device__ __forceinline__ float prep_masked(float x, int ix, int iy, const PrepStats& st, const PrepCfg& c) {
#pragma clang fp contract(on)
    const float pi = 3.14159265358979323846f;
    float v = ff_sub_from(x, st.avg) * st.inv_dev;
    if (c.do_mask) {
        float dy = -st.y0 + (float)iy, dx = -st.x0 + (float)ix;
        float ra = sqrtf(dy * dy + dx * dx);
        if (ra > c.R1 && ra <= c.R2) {
            float edge = (1.0f + cosf(pi * (ra - c.R1) / (float)PREP_HALFW)) / 2.0f;
            v = v * edge + (1.0f - edge) * st.ring_mean;
        }
        if (ra > c.R2) v = st.ring_mean;
    }
    return v;
}

// the whole preparation of pixel (ix, iy): masked, [norm], checkerboard sign
__device__ __forceinline__ float prep_value(float x, int ix, int iy, const PrepStats& st, const PrepCfg& c) {
#pragma clang fp contract(on)
    float v = prep_masked(x, ix, iy, st, c);
    if (c.do_norm == 2) v = ff_sub_from(v, st.avg2) * st.inv_dev2;
    if ((ix + iy) & 1) v = -v;
    return v;
}

}  // namespace cim
