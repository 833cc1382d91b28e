// Device building blocks of the batched FFTs (fft.hip) shared with kernels
// that run transforms inside their own workgroups: mixed-radix (8, 4, 6, 2, 3,
// 5) self-sorting Stockham passes over L lines held in LDS.
#pragma once
#include "cim_internal.h"

namespace cim {

// Radices of one length n, with the reciprocals ("magic numbers") that turn
// the passes' divisions by run-time values into a multiply-high: for
// x * d < 2^32, x / d = umulhi(x, ceil(2^32 / d)) (d > 1).
struct FacList {
    int nfac;
    int fac[16];
    unsigned ns_m[16];   // Ns = fac[0] * ... * fac[f-1] before pass f
    unsigned nb_m[16];   // n / fac[f]
    unsigned n_m;        // n
};

inline unsigned magic_u32(unsigned d) { return d > 1 ? (unsigned)((0x100000000ull + d - 1) / d) : 0u; }

__device__ __forceinline__ unsigned fdiv(unsigned x, unsigned d, unsigned m) { return d > 1 ? __umulhi(x, m) : x; }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by i*s (s = +1 or -1)
__device__ __forceinline__ float2 cmuli(float2 a, float s) { return make_float2(-a.y * s, a.x * s); }

// One Stockham pass of radix R over all L lines held in LDS.
template <int R>
__device__ void stockham_pass(const float2* __restrict__ in, float2* __restrict__ out, int n, int L,
                              int Ns, unsigned ns_m, unsigned nb_m, const float2* __restrict__ tw, float sgn) {
    const int nb = n / R;
    const int tstep = n / (Ns * R);
    for (int b = threadIdx.x; b < L * nb; b += blockDim.x) {
        const int line = (int)fdiv((unsigned)b, (unsigned)nb, nb_m);
        const int j = b - line * nb;
        const float2* src = in + line * n;
        float2* dst = out + line * n;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[j + r * nb];
        const int jq = (int)fdiv((unsigned)j, (unsigned)Ns, ns_m);
        const int jm = j - jq * Ns;
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = tw[r * jm * tstep];
                w.y *= -sgn;   // table holds exp(-2 pi i k/n); sgn=-1 keeps it
                v[r] = cmul(v[r], w);
            }
        }
        float2 y[R];
        if (R == 8) {
            // two 4-point DFTs (even and odd inputs), joined by w^k, w = exp(sgn 2 pi i / 8)
            const float c = 0.70710678118654752440f;
            float2 e[4], o[4];
            {
                float2 a0 = cadd(v[0], v[4]), a1 = csub(v[0], v[4]);
                float2 b0 = cadd(v[2], v[6]), b1 = cmuli(csub(v[2], v[6]), sgn);
                e[0] = cadd(a0, b0); e[2] = csub(a0, b0); e[1] = cadd(a1, b1); e[3] = csub(a1, b1);
            }
            {
                float2 a0 = cadd(v[1], v[5]), a1 = csub(v[1], v[5]);
                float2 b0 = cadd(v[3], v[7]), b1 = cmuli(csub(v[3], v[7]), sgn);
                o[0] = cadd(a0, b0); o[2] = csub(a0, b0); o[1] = cadd(a1, b1); o[3] = csub(a1, b1);
            }
            // w = c (1 + i sgn), w^2 = i sgn, w^3 = c (-1 + i sgn)
            const float2 t1 = make_float2(c * (o[1].x - sgn * o[1].y), c * (o[1].y + sgn * o[1].x));
            const float2 t2 = cmuli(o[2], sgn);
            const float2 t3 = make_float2(c * (-o[3].x - sgn * o[3].y), c * (-o[3].y + sgn * o[3].x));
            y[0] = cadd(e[0], o[0]); y[4] = csub(e[0], o[0]);
            y[1] = cadd(e[1], t1); y[5] = csub(e[1], t1);
            y[2] = cadd(e[2], t2); y[6] = csub(e[2], t2);
            y[3] = cadd(e[3], t3); y[7] = csub(e[3], t3);
        } else if (R == 6) {
            // two 3-point DFTs (even and odd inputs), joined by w^k, w = exp(sgn pi i / 3)
            const float h = 0.86602540378443864676f;
            float2 e[3], o[3];
            {
                float2 t1 = cadd(v[2], v[4]);
                float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
                float2 t3 = cmuli(cscale(csub(v[2], v[4]), h), sgn);
                e[0] = cadd(v[0], t1); e[1] = cadd(t2, t3); e[2] = csub(t2, t3);
            }
            {
                float2 t1 = cadd(v[3], v[5]);
                float2 t2 = make_float2(v[1].x - 0.5f * t1.x, v[1].y - 0.5f * t1.y);
                float2 t3 = cmuli(cscale(csub(v[3], v[5]), h), sgn);
                o[0] = cadd(v[1], t1); o[1] = cadd(t2, t3); o[2] = csub(t2, t3);
            }
            // w = 1/2 + i sgn h, w^2 = -1/2 + i sgn h
            const float2 t1 = make_float2(0.5f * o[1].x - sgn * h * o[1].y, 0.5f * o[1].y + sgn * h * o[1].x);
            const float2 t2 = make_float2(-0.5f * o[2].x - sgn * h * o[2].y, -0.5f * o[2].y + sgn * h * o[2].x);
            y[0] = cadd(e[0], o[0]); y[3] = csub(e[0], o[0]);
            y[1] = cadd(e[1], t1); y[4] = csub(e[1], t1);
            y[2] = cadd(e[2], t2); y[5] = csub(e[2], t2);
        } else if (R == 2) {
            y[0] = cadd(v[0], v[1]);
            y[1] = csub(v[0], v[1]);
        } else if (R == 4) {
            float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
            float2 b0 = cadd(v[1], v[3]), b1 = cmuli(csub(v[1], v[3]), sgn);
            y[0] = cadd(a0, b0);
            y[2] = csub(a0, b0);
            y[1] = cadd(a1, b1);
            y[3] = csub(a1, b1);
        } else if (R == 3) {
            const float h = 0.86602540378443864676f;
            float2 t1 = cadd(v[1], v[2]);
            float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
            float2 t3 = cmuli(cscale(csub(v[1], v[2]), h), sgn);
            y[0] = cadd(v[0], t1);
            y[1] = cadd(t2, t3);
            y[2] = csub(t2, t3);
        } else if (R == 5) {
            const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
            const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
            float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
            float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
            float2 p1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
            float2 p2 = make_float2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
            float2 q1 = cmuli(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y), sgn);
            float2 q2 = cmuli(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y), sgn);
            y[0] = make_float2(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
            y[1] = cadd(p1, q1);
            y[4] = csub(p1, q1);
            y[2] = cadd(p2, q2);
            y[3] = csub(p2, q2);
        }
        const int d0 = jq * Ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[d0 + r * Ns] = y[r];
    }
}

// All passes of fl over the L lines in b0 (scratch b1); returns the buffer
// holding the result in natural order.  Ends with a barrier.
__device__ __forceinline__ float2* lds_fft(float2* b0, float2* b1, int n, int L, const FacList& fl,
                                          const float2* __restrict__ tw, float sgn) {
    int Ns = 1;
    float2* in = b0;
    float2* out = b1;
    for (int f = 0; f < fl.nfac; ++f) {
        const int R = fl.fac[f];
        const unsigned nsm = fl.ns_m[f], nbm = fl.nb_m[f];
        if (R == 8) stockham_pass<8>(in, out, n, L, Ns, nsm, nbm, tw, sgn);
        else if (R == 4) stockham_pass<4>(in, out, n, L, Ns, nsm, nbm, tw, sgn);
        else if (R == 6) stockham_pass<6>(in, out, n, L, Ns, nsm, nbm, tw, sgn);
        else if (R == 2) stockham_pass<2>(in, out, n, L, Ns, nsm, nbm, tw, sgn);
        else if (R == 3) stockham_pass<3>(in, out, n, L, Ns, nsm, nbm, tw, sgn);
        else stockham_pass<5>(in, out, n, L, Ns, nsm, nbm, tw, sgn);
        __syncthreads();
        Ns *= R;
        float2* t = in;
        in = out;
        out = t;
    }
    return in;
}

inline FacList fac_list(const FftPlan* pl) {
    FacList fl{};
    fl.nfac = pl->nfac;
    unsigned ns = 1;
    for (int i = 0; i < pl->nfac; ++i) {
        fl.fac[i] = pl->fac[i];
        fl.ns_m[i] = magic_u32(ns);
        fl.nb_m[i] = magic_u32((unsigned)(pl->n / pl->fac[i]));
        ns *= (unsigned)pl->fac[i];
    }
    fl.n_m = magic_u32((unsigned)pl->n);
    return fl;
}

}  // namespace cim
