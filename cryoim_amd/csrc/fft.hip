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
#include "cim_fft.h"

namespace cim {

// Lines: line l starts at (l / inner) * ostride + (l % inner) * istride,
// elements are es apart.  L lines per workgroup (L a power of 2, <= 16);
// each line's start is computed once, the element loops divide by n with a
// multiply-high (rows) or index with shifts (columns).
__global__ void __launch_bounds__(256) k_fft_lines(float2* __restrict__ data, const float2* __restrict__ tw,
                                                   int n, FacList fl, long long nlines, long long inner,
                                                   long long istride, long long ostride, long long es,
                                                   int L, float sgn, float scale, int conj) {
    extern __shared__ float2 lds[];
    __shared__ long long lbase[16];
    float2* b0 = lds;
    float2* b1 = lds + (size_t)L * n;
    const long long l0 = (long long)blockIdx.x * L;
    const int nl = (int)((nlines - l0) < L ? (nlines - l0) : L);
    if (threadIdx.x < nl) {
        const long long l = l0 + threadIdx.x;
        lbase[threadIdx.x] = (l / inner) * ostride + (l % inner) * istride;
    }
    __syncthreads();
    const int lgL = __ffs(L) - 1;
    // load
    if (es == 1) {
        for (int idx = threadIdx.x; idx < nl * n; idx += blockDim.x) {
            const int li = (int)fdiv((unsigned)idx, (unsigned)n, fl.n_m), e = idx - li * n;
            b0[idx] = data[lbase[li] + e];
        }
    } else {
        for (int idx = threadIdx.x; idx < L * n; idx += blockDim.x) {
            const int e = idx >> lgL, li = idx & (L - 1);
            if (li < nl) b0[li * n + e] = data[lbase[li] + (long long)e * es];
        }
    }
    for (int idx = nl * n + threadIdx.x; idx < L * n; idx += blockDim.x) b0[idx] = make_float2(0.f, 0.f);
    __syncthreads();
    const float2* in = lds_fft(b0, b1, n, L, fl, tw, sgn);
    const float cs = conj ? -scale : scale;
    if (es == 1) {
        for (int idx = threadIdx.x; idx < nl * n; idx += blockDim.x) {
            const int li = (int)fdiv((unsigned)idx, (unsigned)n, fl.n_m), e = idx - li * n;
            const float2 v = in[idx];
            data[lbase[li] + e] = make_float2(v.x * scale, v.y * cs);
        }
    } else {
        for (int idx = threadIdx.x; idx < L * n; idx += blockDim.x) {
            const int e = idx >> lgL, li = idx & (L - 1);
            if (li >= nl) continue;
            const float2 v = in[li * n + e];
            data[lbase[li] + (long long)e * es] = make_float2(v.x * scale, v.y * cs);
        }
    }
}

// Factors (radix passes) and twiddles exp(-2 pi i k / n) of one length, on the
// host; pl.tw is left for the caller to point at device twiddles.
int fft_plan_host(cim_ctx* ctx, int n, FftPlan& pl, std::vector<float2>& tw) {
    if (n < 2 || n > 1024) return set_error(ctx, CIM_E_ARG, "FFT length %d outside [2, 1024]", n);
    pl = FftPlan();
    pl.n = n;
    int m = n;
    // fewest passes: radix 8 first (384 = 8 x 8 x 6: three passes through LDS)
    while (m % 8 == 0) { pl.fac[pl.nfac++] = 8; m /= 8; }
    while (m % 4 == 0) { pl.fac[pl.nfac++] = 4; m /= 4; }
    while (m % 6 == 0) { pl.fac[pl.nfac++] = 6; m /= 6; }
    while (m % 2 == 0) { pl.fac[pl.nfac++] = 2; m /= 2; }
    while (m % 3 == 0) { pl.fac[pl.nfac++] = 3; m /= 3; }
    while (m % 5 == 0) { pl.fac[pl.nfac++] = 5; m /= 5; }
    if (m != 1) return set_error(ctx, CIM_E_ARG, "FFT length %d has a prime factor other than 2, 3, 5", n);
    tw.resize(n);
    for (int k = 0; k < n; ++k) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    return CIM_OK;
}

int fft_plan(cim_ctx* ctx, int n, FftPlan** out) {
    auto it = ctx->plans.find(n);
    if (it != ctx->plans.end()) {
        *out = &it->second;
        return CIM_OK;
    }
    FftPlan pl;
    std::vector<float2> h;
    if (int rc = fft_plan_host(ctx, n, pl, h)) return rc;
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
    const FacList fl = fac_list(pl);
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
