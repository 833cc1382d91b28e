// Batched complex FFTs for gfx950: mixed radix (4, 2, 3, 5) self-sorting
// Stockham passes in LDS.  A workgroup loads L lines of length n (rows:
// contiguous; columns: L neighbouring columns, so every row of the tile is
// one coalesced L*8-byte segment), runs all passes in LDS and writes back.
// 2D = row pass + column pass; 3D = x, y, z passes.  Replaces the FFTW3
// calls of icosprocess/SRC/FFTW.f90:179-219 with the same conventions:
// forward (exponent sign -1) scaled 1/N, backward (sign +1) unscaled,
// optional conjugated output (img2d_FFT.f90:19).
#include <cmath>

#include "cim_internal.h"

namespace cim {

struct FacList {
    int nfac;
    int fac[16];
};

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
                              int Ns, const float2* __restrict__ tw, float sgn) {
    const int nb = n / R;
    const int tstep = n / (Ns * R);
    for (int b = threadIdx.x; b < L * nb; b += blockDim.x) {
        const int line = b / nb;
        const int j = b - line * nb;
        const float2* src = in + line * n;
        float2* dst = out + line * n;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[j + r * nb];
        const int jm = j % Ns;
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = tw[r * jm * tstep];
                w.y *= -sgn;   // table holds exp(-2 pi i k/n); sgn=-1 keeps it
                v[r] = cmul(v[r], w);
            }
        }
        float2 y[R];
        if (R == 2) {
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
        const int d0 = (j / Ns) * Ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[d0 + r * Ns] = y[r];
    }
}

// Lines: line l starts at (l / inner) * ostride + (l % inner) * istride,
// elements are es apart.  L lines per workgroup.
__global__ void __launch_bounds__(256) k_fft_lines(float2* __restrict__ data, const float2* __restrict__ tw,
                                                   int n, FacList fl, long long nlines, long long inner,
                                                   long long istride, long long ostride, long long es,
                                                   int L, float sgn, float scale, int conj) {
    extern __shared__ float2 lds[];
    float2* b0 = lds;
    float2* b1 = lds + (size_t)L * n;
    const long long l0 = (long long)blockIdx.x * L;
    const int nl = (int)((nlines - l0) < L ? (nlines - l0) : L);
    // load
    if (es == 1) {
        for (int idx = threadIdx.x; idx < nl * n; idx += blockDim.x) {
            int li = idx / n, e = idx - li * n;
            long long l = l0 + li;
            long long base = (l / inner) * ostride + (l % inner) * istride;
            b0[li * n + e] = data[base + e];
        }
    } else {
        for (int idx = threadIdx.x; idx < nl * n; idx += blockDim.x) {
            int e = idx / nl, li = idx - e * nl;
            long long l = l0 + li;
            long long base = (l / inner) * ostride + (l % inner) * istride;
            b0[li * n + e] = data[base + (long long)e * es];
        }
    }
    for (int idx = nl * n + threadIdx.x; idx < L * n; idx += blockDim.x) b0[idx] = make_float2(0.f, 0.f);
    __syncthreads();
    int Ns = 1;
    float2* in = b0;
    float2* out = b1;
    for (int f = 0; f < fl.nfac; ++f) {
        const int R = fl.fac[f];
        if (R == 4) stockham_pass<4>(in, out, n, L, Ns, tw, sgn);
        else if (R == 2) stockham_pass<2>(in, out, n, L, Ns, tw, sgn);
        else if (R == 3) stockham_pass<3>(in, out, n, L, Ns, tw, sgn);
        else stockham_pass<5>(in, out, n, L, Ns, tw, sgn);
        __syncthreads();
        Ns *= R;
        float2* t = in;
        in = out;
        out = t;
    }
    const float cs = conj ? -scale : scale;
    if (es == 1) {
        for (int idx = threadIdx.x; idx < nl * n; idx += blockDim.x) {
            int li = idx / n, e = idx - li * n;
            long long l = l0 + li;
            long long base = (l / inner) * ostride + (l % inner) * istride;
            float2 v = in[li * n + e];
            data[base + e] = make_float2(v.x * scale, v.y * cs);
        }
    } else {
        for (int idx = threadIdx.x; idx < nl * n; idx += blockDim.x) {
            int e = idx / nl, li = idx - e * nl;
            long long l = l0 + li;
            long long base = (l / inner) * ostride + (l % inner) * istride;
            float2 v = in[li * n + e];
            data[base + (long long)e * es] = make_float2(v.x * scale, v.y * cs);
        }
    }
}

int fft_plan(cim_ctx* ctx, int n, FftPlan** out) {
    auto it = ctx->plans.find(n);
    if (it != ctx->plans.end()) {
        *out = &it->second;
        return CIM_OK;
    }
    if (n < 2 || n > 1024) return set_error(ctx, CIM_E_ARG, "FFT length %d outside [2, 1024]", n);
    FftPlan pl;
    pl.n = n;
    int m = n;
    while (m % 4 == 0) { pl.fac[pl.nfac++] = 4; m /= 4; }
    while (m % 2 == 0) { pl.fac[pl.nfac++] = 2; m /= 2; }
    while (m % 3 == 0) { pl.fac[pl.nfac++] = 3; m /= 3; }
    while (m % 5 == 0) { pl.fac[pl.nfac++] = 5; m /= 5; }
    if (m != 1) return set_error(ctx, CIM_E_ARG, "FFT length %d has a prime factor other than 2, 3, 5", n);
    std::vector<float2> h(n);
    for (int k = 0; k < n; ++k) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    CIM_HIP(hipMalloc(&pl.tw, sizeof(float2) * n));
    CIM_HIP(hipMemcpy(pl.tw, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    ctx->plans[n] = pl;
    *out = &ctx->plans[n];
    return CIM_OK;
}

static int lines_pass(cim_ctx* ctx, FftPlan* pl, float2* data, long long nlines, long long inner,
                      long long istride, long long ostride, long long es, int dir, float scale, int conj) {
    const int n = pl->n;
    int L;
    if (es == 1) L = n >= 1024 ? 1 : (n >= 512 ? 2 : (n >= 256 ? 4 : (n >= 128 ? 8 : 16)));
    else L = n <= 128 ? 16 : (n <= 512 ? 8 : 4);
    FacList fl;
    fl.nfac = pl->nfac;
    for (int i = 0; i < pl->nfac; ++i) fl.fac[i] = pl->fac[i];
    const long long blocks = (nlines + L - 1) / L;
    const size_t lds = sizeof(float2) * 2 * (size_t)L * n;
    const float sgn = dir == 1 ? -1.f : 1.f;
    hipLaunchKernelGGL(k_fft_lines, dim3((unsigned)blocks), dim3(256), lds, ctx->stream, data, pl->tw, n, fl,
                       nlines, inner, istride, ostride, es, L, sgn, scale, conj);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

int fft2d_dev(cim_ctx* ctx, float2* data, int n, int batch, int dir, int conj_out) {
    FftPlan* pl = nullptr;
    int rc = fft_plan(ctx, n, &pl);
    if (rc) return rc;
    TimedLaunch tl(ctx, "fft2d");
    const long long n2 = (long long)n * n;
    const float scale = dir == 1 ? (float)(1.0 / (double)n2) : 1.f;
    rc = lines_pass(ctx, pl, data, (long long)batch * n, 1, 0, n, 1, dir, 1.f, 0);   // rows
    if (rc) return rc;
    return lines_pass(ctx, pl, data, (long long)batch * n, n, 1, n2, n, dir, scale, conj_out);  // columns
}

int fft3d_dev(cim_ctx* ctx, float2* data, int n, int dir, int conj_out) {
    FftPlan* pl = nullptr;
    int rc = fft_plan(ctx, n, &pl);
    if (rc) return rc;
    TimedLaunch tl(ctx, "fft3d");
    const long long n2 = (long long)n * n;
    const float scale = dir == 1 ? (float)(1.0 / ((double)n2 * n)) : 1.f;
    rc = lines_pass(ctx, pl, data, n2, 1, 0, n, 1, dir, 1.f, 0);           // x
    if (rc) return rc;
    rc = lines_pass(ctx, pl, data, n2, n, 1, n2, n, dir, 1.f, 0);          // y
    if (rc) return rc;
    return lines_pass(ctx, pl, data, n2, n2, 1, 0, n2, dir, scale, conj_out);  // z
}

}  // namespace cim
