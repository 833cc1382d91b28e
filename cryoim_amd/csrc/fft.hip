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
#include "cim_device.h"

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

// ---- forward 2D transform pruned to a centred crop ---------------------------
//
// The transform of each n x n image is needed only on the crop of Dc columns
// and rows starting at index o0 (the centred frequencies o0 - n/2 ..
// o0 + Dc - 1 - n/2).  The row pass transforms every row but keeps the Dc
// outputs of the crop's columns (rows x Dc, to tmp); the column pass
// transforms those Dc columns and keeps the crop's Dc rows, scaled (1/n^2,
// conjugated when asked) and, with parts, multiplied by the centre-shift
// phase of each image's record (shift_phase, as k_centshift).  Every kept
// value goes through the same passes and arithmetic as in fft2d_dev, so it
// has the same bits.

// rows: line l = image l / n, row l % n of data; outputs o0 .. o0 + Dc - 1 to tmp[l * Dc ..]
__global__ void __launch_bounds__(256) k_fft_rows_crop(const float2* __restrict__ data, const float2* __restrict__ tw,
                                                       int n, FacList fl, long long nlines, int L, int o0, int Dc,
                                                       float2* __restrict__ tmp) {
    extern __shared__ float2 lds[];
    float2* b0 = lds;
    float2* b1 = lds + (size_t)L * n;
    const long long l0 = (long long)blockIdx.x * L;
    const int nl = (int)((nlines - l0) < L ? (nlines - l0) : L);
    for (int idx = threadIdx.x; idx < nl * n; idx += blockDim.x) b0[idx] = data[l0 * n + idx];
    for (int idx = nl * n + threadIdx.x; idx < L * n; idx += blockDim.x) b0[idx] = make_float2(0.f, 0.f);
    __syncthreads();
    const float2* in = lds_fft(b0, b1, n, L, fl, tw, -1.f);
    for (int idx = threadIdx.x; idx < nl * Dc; idx += blockDim.x) {
        const int li = idx / Dc, e = idx - li * Dc;
        tmp[(l0 + li) * Dc + e] = in[li * n + o0 + e];
    }
}

// The same row pass over real images, each value prepared as it is read
// (prep_value from its image's sums): no prepared copy of the images.
__global__ void __launch_bounds__(256) k_fft_rows_crop_real(const float* __restrict__ imgs,
                                                            const PrepStats* __restrict__ stats, PrepCfg cfg,
                                                            const float2* __restrict__ tw, int n, FacList fl,
                                                            long long nlines, int L, int o0, int Dc,
                                                            float2* __restrict__ tmp) {
    extern __shared__ float2 lds[];
    float2* b0 = lds;
    float2* b1 = lds + (size_t)L * n;
    const long long l0 = (long long)blockIdx.x * L;
    const int nl = (int)((nlines - l0) < L ? (nlines - l0) : L);
    for (int li = 0; li < L; ++li) {
        const long long l = l0 + li;
        if (li < nl) {
            const long long img = l / n;
            const int iy = (int)(l - img * n);
            const PrepStats st = stats[img];
            const float* row = imgs + l * n;
            for (int ix = threadIdx.x; ix < n; ix += blockDim.x)
                b0[li * n + ix] = make_float2(prep_value(row[ix], ix, iy, st, cfg), 0.f);
        } else {
            for (int ix = threadIdx.x; ix < n; ix += blockDim.x) b0[li * n + ix] = make_float2(0.f, 0.f);
        }
    }
    __syncthreads();
    const float2* in = lds_fft(b0, b1, n, L, fl, tw, -1.f);
    for (int idx = threadIdx.x; idx < nl * Dc; idx += blockDim.x) {
        const int li = idx / Dc, e = idx - li * Dc;
        tmp[(l0 + li) * Dc + e] = in[li * n + o0 + e];
    }
}

// columns: line l = image l / Dc, column l % Dc of tmp (elements Dc apart);
// outputs oy0 .. oy0 + Dy - 1 to crop[image][row][column] (the columns are those
// from index ox0 of the row pass)
__global__ void __launch_bounds__(256) k_fft_cols_crop(const float2* __restrict__ tmp, const float2* __restrict__ tw,
                                                       int n, FacList fl, long long nlines, int L, int ox0, int Dc,
                                                       int oy0, int Dy, float scale, int conj,
                                                       const float* __restrict__ parts, float2* __restrict__ crop) {
    extern __shared__ float2 lds[];
    float2* b0 = lds;
    float2* b1 = lds + (size_t)L * n;
    const long long l0 = (long long)blockIdx.x * L;
    const int nl = (int)((nlines - l0) < L ? (nlines - l0) : L);
    const int lgL = __ffs(L) - 1;
    for (int idx = threadIdx.x; idx < L * n; idx += blockDim.x) {
        const int e = idx >> lgL, li = idx & (L - 1);
        if (li < nl) {
            const long long l = l0 + li, img = l / Dc;
            b0[li * n + e] = tmp[(img * n + e) * Dc + (l - img * Dc)];
        } else {
            b0[li * n + e] = make_float2(0.f, 0.f);
        }
    }
    __syncthreads();
    const float2* in = lds_fft(b0, b1, n, L, fl, tw, -1.f);
    const float cs = conj ? -scale : scale;
    const int c = n / 2;
    for (int idx = threadIdx.x; idx < L * Dy; idx += blockDim.x) {
        const int e = idx >> lgL, li = idx & (L - 1);
        if (li >= nl) continue;
        const long long l = l0 + li, img = l / Dc;
        const int col = (int)(l - img * Dc);
        const float2 v = in[li * n + oy0 + e];
        float2 o = make_float2(v.x * scale, v.y * cs);
        if (parts) {
            const float* pr = parts + img * CIM_PART_STRIDE;
            o = apply_shift(o, shift_phase(ox0 + col - c, oy0 + e - c, n, pr[CIM_P_X], pr[CIM_P_Y], 1));
        }
        crop[(img * Dy + e) * Dc + col] = o;
    }
}

// Forward 2D transform of batch images (data, n x n each, left unchanged) on
// the crop of columns ox0 .. ox0 + Dx - 1 and rows oy0 .. oy0 + Dy - 1
// (crop[image][row][column], frequency = index - n/2), centre-shifted by the
// records parts when given.  tmp: batch * n * Dx float2.
int fft2d_crop_dev(cim_ctx* ctx, const float2* data, int n, int batch, int ox0, int Dx, int oy0, int Dy,
                   int conj_out, const float* parts, float2* tmp, float2* crop) {
    if (ox0 < 0 || Dx <= 0 || ox0 + Dx > n || oy0 < 0 || Dy <= 0 || oy0 + Dy > n)
        return set_error(ctx, CIM_E_ARG, "crop columns %d+%d, rows %d+%d outside %d", ox0, Dx, oy0, Dy, n);
    FftPlan* pl = nullptr;
    int rc = fft_plan(ctx, n, &pl);
    if (rc) return rc;
    TimedLaunch tl(ctx, "fft2d");
    const FacList fl = fac_list(pl);
    const int Lr = n >= 1024 ? 1 : (n >= 512 ? 2 : (n >= 256 ? 4 : (n >= 128 ? 8 : 16)));
    const long long nrows = (long long)batch * n;
    hipLaunchKernelGGL(k_fft_rows_crop, dim3((unsigned)((nrows + Lr - 1) / Lr)), dim3(256),
                       sizeof(float2) * 2 * (size_t)Lr * n, ctx->stream, data, pl->tw, n, fl, nrows, Lr, ox0, Dx, tmp);
    CIM_HIP(hipGetLastError());
    const int Lc = n <= 128 ? 16 : (n <= 512 ? 8 : 4);
    const long long ncols = (long long)batch * Dx;
    const float scale = (float)(1.0 / ((double)n * n));
    hipLaunchKernelGGL(k_fft_cols_crop, dim3((unsigned)((ncols + Lc - 1) / Lc)), dim3(256),
                       sizeof(float2) * 2 * (size_t)Lc * n, ctx->stream, tmp, pl->tw, n, fl, ncols, Lc, ox0, Dx, oy0,
                       Dy, scale, conj_out, parts, crop);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

int fft2d_crop_real_dev(cim_ctx* ctx, const float* imgs, const PrepStats* stats, int do_norm, int do_mask,
                        float mask_r, int n, int batch, int ox0, int Dx, int oy0, int Dy, int conj_out,
                        const float* parts, float2* tmp, float2* crop) {
    if (batch <= 0) return CIM_OK;
    if (ox0 < 0 || Dx < 1 || ox0 + Dx > n || oy0 < 0 || Dy < 1 || oy0 + Dy > n)
        return set_error(ctx, CIM_E_ARG, "fft2d crop [%d, %d) x [%d, %d) outside %d", ox0, ox0 + Dx, oy0, oy0 + Dy, n);
    FftPlan* pl = nullptr;
    int rc = fft_plan(ctx, n, &pl);
    if (rc) return rc;
    TimedLaunch tl(ctx, "fft2d");
    const FacList fl = fac_list(pl);
    const PrepCfg cfg = prep_cfg(n, do_norm, do_mask, mask_r);
    const int Lr = n >= 1024 ? 1 : (n >= 512 ? 2 : (n >= 256 ? 4 : (n >= 128 ? 8 : 16)));
    const long long nrows = (long long)batch * n;
    hipLaunchKernelGGL(k_fft_rows_crop_real, dim3((unsigned)((nrows + Lr - 1) / Lr)), dim3(256),
                       sizeof(float2) * 2 * (size_t)Lr * n, ctx->stream, imgs, stats, cfg, pl->tw, n, fl, nrows, Lr,
                       ox0, Dx, tmp);
    CIM_HIP(hipGetLastError());
    const int Lc = n <= 128 ? 16 : (n <= 512 ? 8 : 4);
    const long long ncols = (long long)batch * Dx;
    const float scale = (float)(1.0 / ((double)n * n));
    hipLaunchKernelGGL(k_fft_cols_crop, dim3((unsigned)((ncols + Lc - 1) / Lc)), dim3(256),
                       sizeof(float2) * 2 * (size_t)Lc * n, ctx->stream, tmp, pl->tw, n, fl, ncols, Lc, ox0, Dx, oy0,
                       Dy, scale, conj_out, parts, crop);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
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
