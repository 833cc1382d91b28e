// Device helpers shared by the kernels: CTF, centring phase, reductions.
#pragma once
#include <hip/hip_runtime.h>

#include "cim_internal.h"

namespace cim {

struct CtfConst {
    double lam;      // electron wavelength (A)
    double cs;       // Cs in A
    float ampwg;     // amplitude contrast
    float sq;        // sqrt(1 - ampwg^2)
    float apix;
    float df_scale;  // 1e4 (um) or 1 (A)
    int corr_cycles; // defocus series: C steps each side (0: one defocus)
    float corr_step; // step of the series in the df unit
};

inline CtfConst make_ctf_const(const cim_params* p) {
    CtfConst c;
    // getCTF2d.f90:12-13
    c.lam = 0.387818 / std::sqrt((double)p->vol * (1.0 + 0.978459e-3 * (double)p->vol));
    c.cs = (double)p->cs * 1.0e7;
    c.ampwg = p->ampwg;
    c.sq = (float)std::sqrt(1.0 - (double)p->ampwg * p->ampwg);
    c.apix = p->apix;
    c.df_scale = p->df_scale;
    c.corr_cycles = p->corr_cycles;
    c.corr_step = p->corr_step;
    return c;
}

// getCTF2d (icosprocess/SRC/getCTF2d.f90:17-27) at centred frequency
// (hx, ky); the phase is evaluated in float64 so that its ~100 rad
// argument keeps full precision.
__device__ __forceinline__ float ctf_value(int hx, int ky, int nx, float df1, float df2, float astig_deg,
                                           const CtfConst& c) {
    const double PI = 3.14159265358979323846;
    double ph = atan2((double)ky, (double)hx);
    if (ph < 0) ph += 2.0 * PI;
    ph -= (double)astig_deg * (PI / 180.0);
    double cp = cos(ph), sp = sin(ph);
    double df = (double)df1 * c.df_scale * cp * cp + (double)df2 * c.df_scale * sp * sp;
    double res = sqrt((double)(hx * hx + ky * ky)) / (double)nx / (double)c.apix;
    double r2 = res * res;
    double phase = 2.0 * PI * (-c.cs * c.lam * c.lam * c.lam * r2 * r2 / 4.0 + df * c.lam * r2 / 2.0);
    double s, co;
    sincos(phase, &s, &co);
    return (float)(-((double)c.sq * s + (double)c.ampwg * co));
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
// The nx/2 part contributes exactly (-1)^(hx+ky); the rest is reduced in
// float64 before sincospi so the large-argument rounding of the float
// formula in img2d_FT_centshift.f90:11 does not enter.
__device__ __forceinline__ float2 shift_phase(int hx, int ky, int nx, float x0, float y0, int recentre) {
    double t = ((double)x0 * hx + (double)y0 * ky) / (double)nx;   // turns
    t -= floor(t);
    double s, c;
    sincospi(2.0 * t, &s, &c);
    float fc = (float)c, fs = (float)s;
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

}  // namespace cim
