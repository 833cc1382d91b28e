// Per-particle preprocessing kernels: normalisation (img2d_norm.f90:1-22),
// soft circular mask (img2d_mask.f90:2-49), checkerboard (shift2d of
// ort_globesearch.f90:25-29), centre-shift phase ramp
// (img2d_FT_centshift.f90:2-18 / local_refine img2d_FTphase.f90:1-22) and
// the ring-amplitude fit FT_amp_expfit.f90:2-47.  One workgroup per image
// for the reductions; sums are carried in float64.
#include <cmath>

#include "cim_device.h"

namespace cim {

__global__ void k_interleave(const float* __restrict__ re, const float* __restrict__ im, float2* __restrict__ out,
                             long long count) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = make_float2(re[i], im ? im[i] : 0.f);
}

__global__ void k_deinterleave(const float2* __restrict__ in, float* __restrict__ re, float* __restrict__ im,
                               long long count) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        float2 v = in[i];
        re[i] = v.x;
        if (im) im[i] = v.y;
    }
}

int interleave(cim_ctx* ctx, const float* re, const float* im, float2* out, long long count) {
    hipLaunchKernelGGL(k_interleave, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, ctx->stream, re, im, out,
                       count);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

int deinterleave(cim_ctx* ctx, const float2* in, float* re, float* im, long long count) {
    hipLaunchKernelGGL(k_deinterleave, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, ctx->stream, in, re, im,
                       count);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

// [norm] -> [mask] -> [norm] -> checkerboard, real image in, complex (v, 0)
// out.  do_norm = 1 normalises before the mask (reconstruct.f90:79-92),
// 2 after it (local_reconstruct/SRC/reconstruct_substep.f90:93-94).
__global__ void __launch_bounds__(1024) k_prep_real(const float* __restrict__ imgs, const float* __restrict__ parts,
                                                    float2* __restrict__ out, int box, int do_norm, int do_mask,
                                                    float mask_r) {
    __shared__ double sh[32];
    const long long N = (long long)box * box;
    const float* img = imgs + (long long)blockIdx.x * N;
    float2* o = out + (long long)blockIdx.x * N;
    const float* pr = parts + (long long)blockIdx.x * CIM_PART_STRIDE;
    double avg = 0.0, inv_dev = 1.0;
    if (do_norm == 1) {
        double s = 0.0;
        for (long long i = threadIdx.x; i < N; i += blockDim.x) s += img[i];
        avg = block_sum(s, sh) / (double)N;
        double q = 0.0;
        for (long long i = threadIdx.x; i < N; i += blockDim.x) {
            double d = (double)img[i] - avg;
            q += d * d;
        }
        inv_dev = 1.0 / sqrt(block_sum(q, sh) / (double)N);
    }
    float R1 = 0.f, R2 = 0.f, ring_mean = 0.f;
    float x0 = 0.f, y0 = 0.f;
    const int halfw = 6;
    if (do_mask) {
        R1 = mask_r - (float)(halfw / 2);
        if (R1 <= 0.f) R1 = 0.f;
        R2 = mask_r + (float)(halfw / 2);
        if (R2 > (float)(box / 2)) R2 = (float)(box / 2);
        x0 = pr[CIM_P_MASKX] + (float)(box / 2);
        y0 = pr[CIM_P_MASKY] + (float)(box / 2);
        double s = 0.0, c = 0.0;
        for (long long i = threadIdx.x; i < N; i += blockDim.x) {
            int iy = (int)(i / box), ix = (int)(i - (long long)iy * box);
            float dy = -y0 + (float)iy, dx = -x0 + (float)ix;
            float ra = sqrtf(dy * dy + dx * dx);
            if (ra > R1 && ra <= R2) {
                s += ((double)img[i] - avg) * inv_dev;
                c += 1.0;
            }
        }
        double ssum = block_sum(s, sh);
        double cnt = block_sum(c, sh);
        ring_mean = cnt != 0.0 ? (float)(ssum / cnt) : 0.f;
    }
    const float pi = 3.14159265358979323846f;
    auto masked = [&](long long i) -> float {
        int iy = (int)(i / box), ix = (int)(i - (long long)iy * box);
        float v = (float)(((double)img[i] - avg) * inv_dev);
        if (do_mask) {
            float dy = -y0 + (float)iy, dx = -x0 + (float)ix;
            float ra = sqrtf(dy * dy + dx * dx);
            if (ra > R1 && ra <= R2) {
                float edge = (1.0f + cosf(pi * (ra - R1) / (float)halfw)) / 2.0f;
                v = v * edge + (1.0f - edge) * ring_mean;
            }
            if (ra > R2) v = ring_mean;
        }
        return v;
    };
    double avg2 = 0.0, inv_dev2 = 1.0;
    if (do_norm == 2) {
        double s = 0.0;
        for (long long i = threadIdx.x; i < N; i += blockDim.x) s += masked(i);
        avg2 = block_sum(s, sh) / (double)N;
        double q = 0.0;
        for (long long i = threadIdx.x; i < N; i += blockDim.x) {
            double d = (double)masked(i) - avg2;
            q += d * d;
        }
        inv_dev2 = 1.0 / sqrt(block_sum(q, sh) / (double)N);
    }
    for (long long i = threadIdx.x; i < N; i += blockDim.x) {
        int iy = (int)(i / box), ix = (int)(i - (long long)iy * box);
        float v = masked(i);
        if (do_norm == 2) v = (float)(((double)v - avg2) * inv_dev2);
        if ((ix + iy) & 1) v = -v;
        o[i] = make_float2(v, 0.f);
    }
}

int preprocess_images(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev, const float* parts_dev,
                      float2* ft_out, int do_norm, int do_mask, int conj_out) {
    if (n <= 0) return CIM_OK;
    {
        TimedLaunch tl(ctx, "prep");
        hipLaunchKernelGGL(k_prep_real, dim3(n), dim3(1024), 0, ctx->stream, images_dev, parts_dev, ft_out, p->box,
                           do_norm, do_mask, p->imgmask);
        CIM_HIP(hipGetLastError());
    }
    return fft2d_dev(ctx, ft_out, p->box, n, 1, conj_out);
}

__global__ void k_centshift(float2* __restrict__ ft, int box, const float* __restrict__ parts, int recentre) {
    const long long N = (long long)box * box;
    const int img = blockIdx.y;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float* pr = parts + (long long)img * CIM_PART_STRIDE;
    int ky = (int)(i / box) - box / 2, hx = (int)(i % box) - box / 2;
    float2 cs = shift_phase(hx, ky, box, pr[CIM_P_X], pr[CIM_P_Y], recentre);
    float2* f = ft + (long long)img * N + i;
    *f = apply_shift(*f, cs);
}

int centre_shift(cim_ctx* ctx, float2* ft, int box, int n, const float* parts_dev, int recentre) {
    if (n <= 0) return CIM_OK;
    const long long N = (long long)box * box;
    hipLaunchKernelGGL(k_centshift, dim3((unsigned)((N + 255) / 256), n), dim3(256), 0, ctx->stream, ft, box,
                       parts_dev, recentre);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

// FT_amp_expfit (FT_amp_expfit.f90:2-47) of nitems centred transforms:
// item i is at ft + i*stride with rows of ld elements and the zero
// frequency at (cidx, cidx).
__global__ void __launch_bounds__(256) k_expfit(const float2* __restrict__ ft, long long stride, int ld, int cidx,
                                                int minR, int maxR, float* __restrict__ fcut) {
    extern __shared__ double acc[];   // Favg[nr], resLN[nr], Num[nr]
    const int nr = maxR - minR + 1;
    double* Favg = acc;
    double* resLN = acc + nr;
    double* Num = acc + 2 * nr;
    for (int k = threadIdx.x; k < 3 * nr; k += blockDim.x) acc[k] = 0.0;
    __syncthreads();
    const int item = blockIdx.x;
    const float2* f = ft + (long long)item * stride;
    const int w = 2 * maxR + 1;
    const int total = (maxR + 1) * w;
    for (int q = threadIdx.x; q < total; q += blockDim.x) {
        int ky = q / w, hx = q % w - maxR;
        float rad = sqrtf((float)(hx * hx + ky * ky));
        int ir = (int)(rad + 0.5f);
        if (ir >= minR && ir <= maxR) {
            float2 v = f[(long long)(ky + cidx) * ld + (hx + cidx)];
            double amp = sqrt((double)v.x * v.x + (double)v.y * v.y);
            atomicAdd(&Favg[ir - minR], amp);
            atomicAdd(&resLN[ir - minR], (double)rad);
            atomicAdd(&Num[ir - minR], 1.0);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a11 = 0, a12 = 0, a22 = 0, b1 = 0, b2 = 0;
        for (int k = 0; k < nr; ++k) {
            double r = resLN[k], F = 0.0;
            if (Num[k] > 0 && Favg[k] > 0) {
                r = resLN[k] / Num[k];
                F = log(Favg[k] / Num[k]);
            }
            resLN[k] = r;
            a11 += r * r;
            a12 += r;
            a22 += 1.0;
            b1 += r * F;
            b2 += F;
        }
        double x2 = (b2 - b1 * a12 / a11) / (a22 - a12 * a12 / a11);
        double x1 = (b1 - a12 * x2) / a11;
        float* out = fcut + (long long)item * nr;
        for (int k = 0; k < nr; ++k) out[k] = (float)(exp(resLN[k] * x1 + x2) * 0.8);
    }
}

int amp_expfit_general(cim_ctx* ctx, const float2* ft, long long stride, int ld, int cidx, int nitems, int minR,
                       int maxR, float* fcut_dev) {
    if (nitems <= 0) return CIM_OK;
    const int nr = maxR - minR + 1;
    TimedLaunch tl(ctx, "expfit");
    hipLaunchKernelGGL(k_expfit, dim3(nitems), dim3(256), sizeof(double) * 3 * nr, ctx->stream, ft, stride, ld, cidx,
                       minR, maxR, fcut_dev);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

int amp_expfit(cim_ctx* ctx, const float2* ft, int box, int n, int minR, int maxR, float* fcut_dev) {
    return amp_expfit_general(ctx, ft, (long long)box * box, box, box / 2, n, minR, maxR, fcut_dev);
}

}  // namespace cim
