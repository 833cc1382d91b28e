// Per-particle preprocessing kernels: normalisation (img2d_norm.f90:1-22),
// soft circular mask (img2d_mask.f90:2-49), checkerboard (shift2d of
// ort_globesearch.f90:25-29), centre-shift phase ramp
// (img2d_FT_centshift.f90:2-18 / local_refine img2d_FTphase.f90:1-22) and
// the ring-amplitude fit FT_amp_expfit.f90:2-47.  One workgroup per image
// for the reductions; sums are carried in float64.
#include <cmath>
#include <cstdlib>

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
// 2 after it (local_reconstruct/SRC/reconstruct_substep.f90:93-94).  Means,
// deviations and the ring mean are compensated float32 sums (block_sum_ff).
// With stats, the kernel stops after the sums and leaves them there (the
// row transform then forms each value as it reads the image, prep_value).
__global__ void __launch_bounds__(1024) k_prep_real(const float* __restrict__ imgs, const float* __restrict__ parts,
                                                    float2* __restrict__ out, int box, int do_norm, int do_mask,
                                                    float mask_r, PrepStats* __restrict__ stats) {
    __shared__ ff32 sh[16];
    const long long N = (long long)box * box;
    const float* img = imgs + (long long)blockIdx.x * N;
    const float* pr = parts + (long long)blockIdx.x * CIM_PART_STRIDE;
    const float invN = 1.0f / (float)N;
    const PrepCfg cfg = prep_cfg(box, do_norm, do_mask, mask_r);
    PrepStats st{};
    st.avg = ff32{0.f, 0.f};
    st.inv_dev = 1.f;
    st.avg2 = ff32{0.f, 0.f};
    st.inv_dev2 = 1.f;
    if (do_norm == 1) {
        ff32 s{0.f, 0.f};
        for (long long i = threadIdx.x; i < N; i += blockDim.x) s = ff_add(s, ff32{img[i], 0.f});
        s = block_sum_ff(s, sh);
        st.avg = ff_mul(s, ff_recip_int((int)N));
        ff32 q{0.f, 0.f};
        for (long long i = threadIdx.x; i < N; i += blockDim.x) {
            const float d = ff_sub_from(img[i], st.avg);
            q = ff_add(q, ff_two_prod(d, d));
        }
        q = block_sum_ff(q, sh);
        st.inv_dev = 1.0f / sqrtf(q.h * invN);
    }
    st.ring_mean = 0.f;
    st.x0 = st.y0 = 0.f;
    if (do_mask) {
        st.x0 = pr[CIM_P_MASKX] + (float)(box / 2);
        st.y0 = pr[CIM_P_MASKY] + (float)(box / 2);
        ff32 s{0.f, 0.f};
        int c = 0;
        for (long long i = threadIdx.x; i < N; i += blockDim.x) {
            int iy = (int)(i / box), ix = (int)(i - (long long)iy * box);
            float dy = -st.y0 + (float)iy, dx = -st.x0 + (float)ix;
            float ra = sqrtf(dy * dy + dx * dx);
            if (ra > cfg.R1 && ra <= cfg.R2) {
                s = ff_add(s, ff_two_prod(ff_sub_from(img[i], st.avg), st.inv_dev));
                ++c;
            }
        }
        s = block_sum_ff(s, sh);
        const ff32 cnt = block_sum_ff(ff32{(float)c, 0.f}, sh);
        st.ring_mean = cnt.h != 0.f ? __fdiv_rn(s.h, cnt.h) : 0.f;
    }
    if (do_norm == 2) {
        ff32 s{0.f, 0.f};
        for (long long i = threadIdx.x; i < N; i += blockDim.x) {
            const int iy = (int)(i / box), ix = (int)(i - (long long)iy * box);
            s = ff_add(s, ff32{prep_masked(img[i], ix, iy, st, cfg), 0.f});
        }
        s = block_sum_ff(s, sh);
        st.avg2 = ff_mul(s, ff_recip_int((int)N));
        ff32 q{0.f, 0.f};
        for (long long i = threadIdx.x; i < N; i += blockDim.x) {
            const int iy = (int)(i / box), ix = (int)(i - (long long)iy * box);
            const float d = ff_sub_from(prep_masked(img[i], ix, iy, st, cfg), st.avg2);
            q = ff_add(q, ff_two_prod(d, d));
        }
        q = block_sum_ff(q, sh);
        st.inv_dev2 = 1.0f / sqrtf(q.h * invN);
    }
    if (stats) {
        if (threadIdx.x == 0) stats[blockIdx.x] = st;
        return;
    }
    float2* o = out + (long long)blockIdx.x * N;
    for (long long i = threadIdx.x; i < N; i += blockDim.x) {
        const int iy = (int)(i / box), ix = (int)(i - (long long)iy * box);
        o[i] = make_float2(prep_value(img[i], ix, iy, st, cfg), 0.f);
    }
}

// [norm] -> [mask] -> [norm] -> checkerboard of n images into ft_out (complex,
// before the transform)
int prep_images_dev(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev, const float* parts_dev,
                    float2* ft_out, int do_norm, int do_mask) {
    if (n <= 0) return CIM_OK;
    TimedLaunch tl(ctx, "prep");
    hipLaunchKernelGGL(k_prep_real, dim3(n), dim3(1024), 0, ctx->stream, images_dev, parts_dev, ft_out, p->box,
                       do_norm, do_mask, p->imgmask, (PrepStats*)nullptr);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

bool prep_fused() {
    const char* e = getenv("CIM_PREP_FUSED");
    return !e || atoi(e) != 0;
}

// The sums of the same preparation, per image, for a transform that forms
// the values as it reads the images (fft2d_crop_real_dev).
int prep_stats_dev(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev, const float* parts_dev,
                   int do_norm, int do_mask, PrepStats* stats) {
    if (n <= 0) return CIM_OK;
    TimedLaunch tl(ctx, "prep");
    hipLaunchKernelGGL(k_prep_real, dim3(n), dim3(1024), 0, ctx->stream, images_dev, parts_dev, (float2*)nullptr,
                       p->box, do_norm, do_mask, p->imgmask, stats);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

int preprocess_images(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev, const float* parts_dev,
                      float2* ft_out, int do_norm, int do_mask, int conj_out) {
    if (n <= 0) return CIM_OK;
    if (int rc = prep_images_dev(ctx, p, n, images_dev, parts_dev, ft_out, do_norm, do_mask)) return rc;
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
// frequency at (cidx, cidx).  tpr consecutive threads (a power of two, at
// most 8, so a group sits in one wave) own ring minR + k: thread j of the
// group walks the half-plane rows ky >= 0 with ky = j mod tpr and, in each,
// the hx interval that can round to the ring's radius (int(sqrt(hx^2 +
// ky^2) + 0.5), tested exactly as the reference bins), summing |F| and the
// radii in compensated float32; the group's partial sums are combined in a
// fixed tree order, and the least-squares line through (radius, ln mean
// |F|) is solved in float32.
__global__ void __launch_bounds__(256) k_expfit(const float2* __restrict__ ft, long long stride, int ld, int cidx,
                                                int minR, int maxR, int tpr, float* __restrict__ fcut) {
    extern __shared__ float racc[];   // per ring: mean radius, ln mean |F| (or 0)
    __shared__ float fit[2];          // slope and intercept of the line
    const int nr = maxR - minR + 1;
    float* resLN = racc;
    float* lnF = racc + nr;
    const int item = blockIdx.x;
    const float2* f = ft + (long long)item * stride;
    const int j = threadIdx.x & (tpr - 1), rings_per_pass = blockDim.x / tpr;
    for (int k0 = 0; k0 < nr; k0 += rings_per_pass) {
        const int k = k0 + threadIdx.x / tpr;
        const int ir = minR + k;
        ff32 famp{0.f, 0.f}, frad{0.f, 0.f};
        int num = 0;
        if (k < nr) {
            const float rlo = fmaxf((float)ir - 0.5f, 0.f), rhi = (float)ir + 0.5f;
            for (int ky = j; ky <= min(ir + 1, maxR); ky += tpr) {
                // |hx| with hx^2 + ky^2 in [rlo^2, rhi^2), widened by one step on each side
                const float a2 = rlo * rlo - (float)(ky * ky), b2 = rhi * rhi - (float)(ky * ky);
                if (b2 < 0.f) break;
                const int h0 = a2 > 0.f ? max((int)sqrtf(a2) - 1, 0) : 0;
                const int h1 = min((int)sqrtf(b2) + 1, maxR);
                const float2* row = f + (long long)(ky + cidx) * ld + cidx;
                for (int side = 0; side < 2; ++side) {
                    for (int ah = h0; ah <= h1; ++ah) {
                        const int hx = side ? -ah : ah;
                        if (side && ah == 0) continue;
                        const float rad = sqrtf((float)(hx * hx + ky * ky));
                        if ((int)(rad + 0.5f) != ir) continue;
                        const float2 v = row[hx];
                        famp = ff_add(famp, ff32{sqrtf(__builtin_fmaf(v.x, v.x, v.y * v.y)), 0.f});
                        frad = ff_add(frad, ff32{rad, 0.f});
                        ++num;
                    }
                }
            }
        }
        // group sums: after the step with offset o, thread j holds threads j .. j + 2o - 1
        for (int o = 1; o < tpr; o <<= 1) {
            const ff32 pa{__shfl_down(famp.h, o, 64), __shfl_down(famp.l, o, 64)};
            const ff32 pr{__shfl_down(frad.h, o, 64), __shfl_down(frad.l, o, 64)};
            const int pn = __shfl_down(num, o, 64);
            famp = ff_add(famp, pa);
            frad = ff_add(frad, pr);
            num += pn;
        }
        if (k < nr && j == 0) {
            float r = frad.h, F = 0.f;
            if (num > 0 && famp.h > 0.f) {
                r = __fdiv_rn(frad.h, (float)num);
                F = logf(__fdiv_rn(famp.h, (float)num));
            }
            resLN[k] = r;
            lnF[k] = F;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ff32 a11{0.f, 0.f}, a12{0.f, 0.f}, b1{0.f, 0.f}, b2{0.f, 0.f};
        for (int k = 0; k < nr; ++k) {
            const float r = resLN[k], F = lnF[k];
            a11 = ff_add(a11, ff_two_prod(r, r));
            a12 = ff_add(a12, ff32{r, 0.f});
            b1 = ff_add(b1, ff_two_prod(r, F));
            b2 = ff_add(b2, ff32{F, 0.f});
        }
        const float a22 = (float)nr;
        const float x2 = (b2.h - b1.h * a12.h / a11.h) / (a22 - a12.h * a12.h / a11.h);
        const float x1 = (b1.h - a12.h * x2) / a11.h;
        fit[0] = x1;
        fit[1] = x2;
    }
    __syncthreads();
    float* out = fcut + (long long)item * nr;
    for (int k = threadIdx.x; k < nr; k += blockDim.x) out[k] = expf(resLN[k] * fit[0] + fit[1]) * 0.8f;
}

int amp_expfit_general(cim_ctx* ctx, const float2* ft, long long stride, int ld, int cidx, int nitems, int minR,
                       int maxR, float* fcut_dev) {
    if (nitems <= 0) return CIM_OK;
    const int nr = maxR - minR + 1;
    TimedLaunch tl(ctx, "expfit");
    // threads per ring: 256 threads over the rings, a power of two from 1 to 8
    int tpr = 8;
    while (tpr > 1 && tpr * nr > 256) tpr >>= 1;
    hipLaunchKernelGGL(k_expfit, dim3(nitems), dim3(256), sizeof(float) * 2 * nr, ctx->stream, ft, stride, ld, cidx,
                       minR, maxR, tpr, fcut_dev);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

int amp_expfit(cim_ctx* ctx, const float2* ft, int box, int n, int minR, int maxR, float* fcut_dev) {
    return amp_expfit_general(ctx, ft, (long long)box * box, box, box / 2, n, minR, maxR, fcut_dev);
}

}  // namespace cim
