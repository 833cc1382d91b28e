// Centre finding of icosprocess mode 3.
//
// calc_center.f90:2-113 (with a model) and calc_center_selfrot180.f90:1-95
// (without): each particle image is cross-correlated with the CTF-modulated
// central-slice projection of the model at the particle's orientation, or
// with itself rotated by 180 degrees, and the correlation peak is located
// to sub-pixel precision by the log-quadric fit of guass_regression.f90.
//
// Every image-sized step runs over a batch.  The reference's img2d_FFT calls
// become fft2d_dev calls with conj_out = 1 (img2d_FFT stores b = -Im in both
// directions, so each forward/backward pair rotates the image by 180 degrees
// about (n/2, n/2); the same calls reproduce that).  The elementwise steps
// between transforms are fused into one kernel each; per-image reductions
// (mask ring mean, mean and deviation, argmax and the fit) run one
// workgroup per image.
#include <cmath>

#include "cim_device.h"

namespace cim {

int upload_parts(cim_ctx* ctx, int n, const float* parts);
int check_params_public(cim_ctx* ctx, const cim_params* p);

namespace {

constexpr int kHalfw = 6;   // cosine edge width of img2d_mask / img3d_mask

__device__ __forceinline__ float edge_cos(float ra, float R1) {
    const float pi = 3.14159265358979323846f;
    return (1.0f + cosf(pi * (ra - R1) / (float)kHalfw)) / 2.0f;
}

// ---- model: img3d_mask, checkerboard, forward 3D FFT, checkerboard ------

// Shell sum and count of img3d_mask.f90:26-34: each block's compensated
// float32 partial (sum, count) goes to part[4 blockIdx.x ..]; k_sum_partials
// adds them in block order.
__global__ void __launch_bounds__(256) k_mask3d_sum(const float* __restrict__ v, int n, float R1, float R2,
                                                    float* __restrict__ part) {
    __shared__ ff32 sh[4];
    const long long n2 = (long long)n * n, N = n2 * n;
    const float c = (float)(n / 2);
    ff32 s{0.f, 0.f};
    int cnt = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        const float iz = (float)(i / n2), iy = (float)((i / n) % n), ix = (float)(i % n);
        const float zz = (-c + iz) * (-c + iz), yy = (-c + iy) * (-c + iy), xx = (-c + ix) * (-c + ix);
        const float ra = sqrtf(xx + yy + zz);
        if (ra > R1 && ra <= R2) {
            s = ff_add(s, ff32{v[i], 0.f});
            ++cnt;
        }
    }
    s = block_sum_ff(s, sh);
    const ff32 ct = block_sum_ff(ff32{(float)cnt, 0.f}, sh);
    if (threadIdx.x == 0) {
        float* q = part + 4 * (long long)blockIdx.x;
        q[0] = s.h;
        q[1] = s.l;
        q[2] = ct.h;
        q[3] = ct.l;
    }
}

// img3d_mask.f90:36-49 (when R2 > 0), then (-1)^(x+y+z), as complex (v, 0).
__global__ void k_mask3d_apply(const float* __restrict__ v, int n, float R1, float R2, const float* __restrict__ acc,
                               float2* __restrict__ out) {
    const long long n2 = (long long)n * n, N = n2 * n;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int iz = (int)(i / n2), iy = (int)((i / n) % n), ix = (int)(i % n);
    float d = v[i];
    if (R2 > 0.f) {
        const float c = (float)(n / 2);
        const float zz = (-c + iz) * (-c + iz), yy = (-c + iy) * (-c + iy), xx = (-c + ix) * (-c + ix);
        const float ra = sqrtf(xx + yy + zz);
        // acc: the shell's sum (hi, lo) and count (hi, lo); the count enters as a float, as there
        const float summ = acc[2] != 0.f ? __fdiv_rn(__fadd_rn(acc[0], acc[1]), acc[2]) : 0.f;
        if (ra > R1 && ra <= R2) {
            const float e = edge_cos(ra, R1);
            d = __builtin_fmaf(1.0f - e, summ, d * e);
        }
        if (ra > R2) d = summ;
    }
    if ((ix + iy + iz) & 1) d = -d;
    out[i] = make_float2(d, 0.f);
}

// Second shift3dComplex and img3di = -aimag (calc_center.f90:38-41).
__global__ void k_checker3d_conj(float2* __restrict__ d, int n) {
    const long long n2 = (long long)n * n, N = n2 * n;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float2 v = d[i];
    if (((i / n2) + (i / n) % n + i % n) & 1) v = make_float2(-v.x, -v.y);
    d[i] = make_float2(v.x, -v.y);
}

// ---- per-image stage: value -> [mask] -> [norm] -> [checkerboard] -------

struct StageArgs {
    const float* real_in;   // box^2 reals per image, or null: take Re(cplx)
    float2* cplx;           // per image box^2 complex: input (real_in null) and output
    float2* rot;            // null, or the 180-degree rotation of the output
    const float* parts;     // records of the images of this launch
    int box;
    int in_checker;         // input multiplied by (-1)^(x+y)
    int mask_mode;          // 0 none, 1 centred on (CIM_P_X, CIM_P_Y), 2 on the origin
    float mask_r;
    int do_norm;            // img2d_norm after the mask
    int out_checker;        // output multiplied by (-1)^(x+y)
};

struct MaskGeom {
    float R1, R2, x0, y0;
    __device__ __forceinline__ float ra(int ix, int iy) const {
        const float yy = (-y0 + (float)iy) * (-y0 + (float)iy), xx = (-x0 + (float)ix) * (-x0 + (float)ix);
        return sqrtf(yy + xx);
    }
};

// One workgroup per image.  Passes over the image recompute the value from
// the input (ring mean, then mean, then deviation, then the write), so the
// stage works in place on cplx.
__global__ void __launch_bounds__(1024) k_stage(StageArgs a) {
    __shared__ ff32 sh[16];
    const int box = a.box;
    const long long N = (long long)box * box;
    const long long base = (long long)blockIdx.x * N;
    const float* pr = a.parts + (long long)blockIdx.x * CIM_PART_STRIDE;
    auto value = [&](long long i) -> float {
        float v = a.real_in ? a.real_in[base + i] : a.cplx[base + i].x;
        if (a.in_checker && (((i / box) + (i % box)) & 1)) v = -v;
        return v;
    };
    // img2d_mask.f90:11-34
    MaskGeom g{0.f, 0.f, 0.f, 0.f};
    float summ = 0.f;
    if (a.mask_mode) {
        g.R1 = a.mask_r - (float)(kHalfw / 2);
        if (g.R1 <= 0.f) g.R1 = 0.f;
        g.R2 = a.mask_r + (float)(kHalfw / 2);
        if (g.R2 > (float)(box / 2)) g.R2 = (float)(box / 2);
        const float ox = a.mask_mode == 1 ? pr[CIM_P_X] : 0.f, oy = a.mask_mode == 1 ? pr[CIM_P_Y] : 0.f;
        g.x0 = ox + (float)(box / 2);
        g.y0 = oy + (float)(box / 2);
        ff32 s{0.f, 0.f};
        int c = 0;
        for (long long i = threadIdx.x; i < N; i += blockDim.x) {
            const float ra = g.ra((int)(i % box), (int)(i / box));
            if (ra > g.R1 && ra <= g.R2) {
                s = ff_add(s, ff32{value(i), 0.f});
                ++c;
            }
        }
        s = block_sum_ff(s, sh);
        const ff32 ct = block_sum_ff(ff32{(float)c, 0.f}, sh);
        summ = ct.h != 0.f ? __fdiv_rn(__fadd_rn(s.h, s.l), ct.h) : 0.f;
    }
    auto masked = [&](long long i) -> float {
        float v = value(i);
        if (a.mask_mode) {
            const float ra = g.ra((int)(i % box), (int)(i / box));
            if (ra > g.R1 && ra <= g.R2) {
                const float e = edge_cos(ra, g.R1);
                v = __builtin_fmaf(1.0f - e, summ, v * e);
            }
            if (ra > g.R2) v = summ;
        }
        return v;
    };
    // img2d_norm.f90:1-22 (compensated float32 sums)
    float avg = 0.f, dev = 1.f;
    if (a.do_norm) {
        ff32 s{0.f, 0.f};
        for (long long i = threadIdx.x; i < N; i += blockDim.x) s = ff_add(s, ff32{masked(i), 0.f});
        s = block_sum_ff(s, sh);
        const ff32 m = ff_mul(s, ff_recip_int((int)N));
        avg = __fadd_rn(m.h, m.l);
        ff32 q{0.f, 0.f};
        for (long long i = threadIdx.x; i < N; i += blockDim.x) {
            const float d = masked(i) - avg;
            q = ff_add(q, ff_two_prod(d, d));
        }
        q = block_sum_ff(q, sh);
        dev = sqrtf(ff_mul(q, ff_recip_int((int)N)).h);
    }
    __syncthreads();   // every pass has read the input before it is overwritten
    for (long long i = threadIdx.x; i < N; i += blockDim.x) {
        float w = masked(i);
        if (a.do_norm) w = (w - avg) / dev;
        const int ky = (int)(i / box), hx = (int)(i % box);
        if (a.out_checker && ((hx + ky) & 1)) w = -w;
        if (a.rot) {
            // calc_center_selfrot180.f90:37-48: row/column 0 of a1 cleared,
            // a2(n-ky, n-hx) = a1(ky, hx); a2's row/column 0 stay zero
            if (hx == 0 || ky == 0) {
                a.cplx[base + i] = make_float2(0.f, 0.f);
                a.rot[base + i] = make_float2(0.f, 0.f);
            } else {
                a.cplx[base + i] = make_float2(w, 0.f);
                a.rot[base + (long long)(box - ky) * box + (box - hx)] = make_float2(w, 0.f);
            }
        } else {
            a.cplx[base + i] = make_float2(w, 0.f);
        }
    }
}

// ---- Fourier-space elementwise steps ------------------------------------

// img2d_lowpass.f90:24-30: a = a * H / n^2 with H = exp(-r^2 / (2 lp^2)).
__global__ void k_lowpass_mul(float2* __restrict__ x, int box, long long total, float lowpass) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long long N = (long long)box * box;
    const long long j = i % N;
    const int ky = (int)(j / box) - box / 2, hx = (int)(j % box) - box / 2;
    const float r2 = (float)(hx * hx + ky * ky);
    const float H = expf(-r2 / lowpass);
    const float n2 = (float)N;
    float2 v = x[i];
    x[i] = make_float2(v.x * H / n2, v.y * H / n2);
}

// img2d_applyCTF.f90:19-20 with getCTF2d of each particle's record.
__global__ void k_ctf_mul(float2* __restrict__ x, int box, const float* __restrict__ parts, CtfConst c) {
    const long long N = (long long)box * box;
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const float* pr = parts + (long long)blockIdx.y * CIM_PART_STRIDE;
    const int ky = (int)(j / box) - box / 2, hx = (int)(j % box) - box / 2;
    const float f = ctf_value(hx, ky, box, pr[CIM_P_DF1], pr[CIM_P_DF2], pr[CIM_P_ASTIG], c);
    const long long i = (long long)blockIdx.y * N + j;
    const float2 v = x[i];
    x[i] = make_float2(v.x * f, v.y * f);
}

// cross_correlation.f90:15-22: x1 <- (a1 a2 + b1 b2, b1 a2 - a1 b2) (-1)^(x+y).
__global__ void k_cc_mul(float2* __restrict__ x1, const float2* __restrict__ x2, int box, long long total) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long long j = i % ((long long)box * box);
    const float2 p = x1[i], q = x2[i];
    float re = p.x * q.x + p.y * q.y;
    float im = p.y * q.x - p.x * q.y;
    if (((j / box) + (j % box)) & 1) {
        re = -re;
        im = -im;
    }
    x1[i] = make_float2(re, im);
}

// ---- central slice of the model FT (project2d_FT3dTo2d.f90:24-87) -------

// Sample (ix, iy) of image blockIdx.y: the rotated grid point, trilinear
// interpolation of {Re, -Im} inside radius n/2-2 (0 outside), multiplied
// by (-1)^(x+y) ready for the backward transform.
__global__ void __launch_bounds__(256) k_center_slice(const float2* __restrict__ model, int box,
                                                      const float* __restrict__ parts, float2* __restrict__ out) {
    const long long N = (long long)box * box;
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const float* pr = parts + (long long)blockIdx.y * CIM_PART_STRIDE;
    float R[9];
    euler_f(pr[CIM_P_THETA], pr[CIM_P_PHI], pr[CIM_P_OMEGA], R);
    const int ky0 = (int)(j / box), hx0 = (int)(j % box);
    const float X = (float)(hx0 - box / 2), Y = (float)(ky0 - box / 2);
    float x = R[0] * X + R[1] * Y, y = R[3] * X + R[4] * Y, z = R[6] * X + R[7] * Y;
    float2 v = make_float2(0.f, 0.f);
    // radius test of :39 written so that a NaN orientation reads nothing
    if (sqrtf(x * x + y * y + z * z) <= (float)(box / 2 - 2)) {
        x += (float)(box / 2);
        y += (float)(box / 2);
        z += (float)(box / 2);
        const int hx = (int)x, ky = (int)y, lz = (int)z;
        const float xb = x - (float)hx, yb = y - (float)ky, zb = z - (float)lz;
        const long long c = (long long)lz * N + (long long)ky * box + hx;
        const float2 t1 = model[c], t2 = model[c + 1], t3 = model[c + box], t4 = model[c + box + 1];
        const float2 t5 = model[c + N], t6 = model[c + N + 1], t7 = model[c + N + box], t8 = model[c + N + box + 1];
        const float w1 = (1.f - xb) * (1.f - yb) * (1.f - zb), w2 = xb * (1.f - yb) * (1.f - zb);
        const float w3 = (1.f - xb) * yb * (1.f - zb), w4 = xb * yb * (1.f - zb);
        const float w5 = (1.f - xb) * (1.f - yb) * zb, w6 = xb * (1.f - yb) * zb;
        const float w7 = (1.f - xb) * yb * zb, w8 = xb * yb * zb;
        v.x = w1 * t1.x + w2 * t2.x + w3 * t3.x + w4 * t4.x + w5 * t5.x + w6 * t6.x + w7 * t7.x + w8 * t8.x;
        v.y = w1 * t1.y + w2 * t2.y + w3 * t3.y + w4 * t4.y + w5 * t5.y + w6 * t6.y + w7 * t7.y + w8 * t8.y;
        if ((hx0 + ky0) & 1) v = make_float2(-v.x, -v.y);
    }
    out[(long long)blockIdx.y * N + j] = v;
}

// ---- peak: first maximum, 3x3 log-quadric fit, shift and PR --------------

// calc_center.f90:70-101 / calc_center_selfrot180.f90:58-89 on Re(x) of
// image blockIdx.x.  A neighbour index past a row end reads the adjacent
// row (flat indexing, as the reference); one outside the image reads 0.
__global__ void __launch_bounds__(1024) k_center_peak(const float2* __restrict__ x, int box, int halve,
                                                      float* __restrict__ parts, float* __restrict__ cc_out) {
    __shared__ float sv[16];
    __shared__ long long si[16];
    const long long N = (long long)box * box;
    const float2* img = x + (long long)blockIdx.x * N;
    float best = -INFINITY;
    long long bi = N;
    for (long long i = threadIdx.x; i < N; i += blockDim.x) {
        const float v = img[i].x;
        if (v > best) {
            best = v;
            bi = i;
        }
    }
    // (value desc, index asc) reduction: maxloc's first maximum
    for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(best, o, 64);
        const long long oi = __shfl_xor(bi, o, 64);
        if (ov > best || (ov == best && oi < bi)) {
            best = ov;
            bi = oi;
        }
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        sv[w] = best;
        si[w] = bi;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int nw = (blockDim.x + 63) >> 6;
    for (int k = 1; k < nw; ++k)
        if (sv[k] > best || (sv[k] == best && si[k] < bi)) {
            best = sv[k];
            bi = si[k];
        }
    if (bi >= N) bi = 0;   // all NaN: maxloc returns the first element
    const int iy0 = (int)(bi / box), ix0 = (int)(bi % box);
    float lz[3][3];   // log z(i, j) at [j+1][i+1]
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) {
            const long long k = (long long)(iy0 + j) * box + ix0 + i;
            lz[j + 1][i + 1] = logf((k >= 0 && k < N) ? img[k].x : 0.f);
        }
    // guass_regression.f90:9-53
    float c10 = 0.f, c01 = 0.f, c11 = 0.f, c20 = 0.f, c02 = 0.f, c00 = 0.f;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) c10 += (float)i * lz[j + 1][i + 1];
    c10 /= 6.0f;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) c01 += (float)j * lz[j + 1][i + 1];
    c01 /= 6.0f;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) c11 += (float)(i * j) * lz[j + 1][i + 1];
    c11 /= 4.0f;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) c20 += (float)(3 * i * i - 2) * lz[j + 1][i + 1];
    c20 /= 6.0f;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) c02 += (float)(3 * j * j - 2) * lz[j + 1][i + 1];
    c02 /= 6.0f;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) c00 += (float)(5 - 3 * i * i - 3 * j * j) * lz[j + 1][i + 1];
    c00 /= 9.0f;
    const float den = 4.0f * c20 * c02 - c11 * c11;
    const float dx = (c11 * c01 - 2.0f * c10 * c02) / den;
    const float dy = (c11 * c10 - 2.0f * c01 * c20) / den;
    const float cc = expf(c00 - c20 * dx * dx - c11 * dx * dy - c02 * dy * dy);
    float cx = (float)(ix0 - box / 2) + dx, cy = (float)(iy0 - box / 2) + dy;
    if (halve) {
        cx = cx / 2.0f;
        cy = cy / 2.0f;
    }
    if (fabsf(cx) >= (float)box / 16.0f) cx = 0.f;
    if (fabsf(cy) >= (float)box / 16.0f) cy = 0.f;
    float ccl = cc;
    if (ccl >= 1.0f) ccl = 1.0f;
    if (ccl <= -1.0f) ccl = -1.0f;
    float* pr = parts + (long long)blockIdx.x * CIM_PART_STRIDE;
    pr[CIM_P_X] = cx;
    pr[CIM_P_Y] = cy;
    pr[CIM_P_PR] = acosf(ccl) * 45.0f / atan2f(1.0f, 1.0f);
    if (cc_out) cc_out[blockIdx.x] = cc;
}

}  // namespace

static int launch_stage(cim_ctx* ctx, int n, const StageArgs& a) {
    TimedLaunch tl(ctx, "cen_stage");
    hipLaunchKernelGGL(k_stage, dim3(n), dim3(1024), 0, ctx->stream, a);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

static int lowpass(cim_ctx* ctx, float2* x, int box, int n, int lp) {
    int rc = fft2d_dev(ctx, x, box, n, 1, 1);
    if (rc) return rc;
    const long long total = (long long)n * box * box;
    {
        TimedLaunch tl(ctx, "cen_elem");
        hipLaunchKernelGGL(k_lowpass_mul, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, x, box,
                           total, 2.0f * (float)(lp * lp));
        CIM_HIP(hipGetLastError());
    }
    return fft2d_dev(ctx, x, box, n, -1, 1);
}

// One chunk of n particles already on the device.  parts_dev is updated in
// place (x, y, PR); cc_dev receives the unclipped fitted peak.
static int center_chunk(cim_ctx* ctx, const cim_params* p, int method, int n, const float* img_dev, float* parts_dev,
                        float* cc_dev) {
    const int box = p->box;
    const long long N = (long long)box * box;
    float2* x1 = ctx->cen.x1.as<float2>();
    float2* x2 = ctx->cen.x2.as<float2>();
    const int mask = p->mask_on ? 1 : 0;
    int rc;
    if (method == CIM_CENTER_MODEL) {
        // a1: img2d_mask at (x, y), img2d_lowpass, img2d_norm, forward FFT
        StageArgs s{img_dev, x1, nullptr, parts_dev, box, 0, mask ? 1 : 0, p->imgmask, 0, 1};
        if ((rc = launch_stage(ctx, n, s))) return rc;
        if ((rc = lowpass(ctx, x1, box, n, p->maxR))) return rc;
        s = StageArgs{nullptr, x1, nullptr, parts_dev, box, 1, 0, 0.f, 1, 0};
        if ((rc = launch_stage(ctx, n, s))) return rc;
        if ((rc = fft2d_dev(ctx, x1, box, n, 1, 1))) return rc;
        // a2: projection, CTF, img2d_mask at the origin, lowpass, norm, forward FFT
        {
            TimedLaunch tl(ctx, "cen_slice");
            hipLaunchKernelGGL(k_center_slice, dim3((unsigned)((N + 255) / 256), n), dim3(256), 0, ctx->stream,
                               ctx->cen.model.as<float2>(), box, parts_dev, x2);
            CIM_HIP(hipGetLastError());
        }
        if ((rc = fft2d_dev(ctx, x2, box, n, -1, 1))) return rc;
        if (p->apply_ctf) {
            // project2d's a = Re * s and applyCTF's a * s cancel: the input is Re
            s = StageArgs{nullptr, x2, nullptr, parts_dev, box, 0, 0, 0.f, 0, 0};
            if ((rc = launch_stage(ctx, n, s))) return rc;
            if ((rc = fft2d_dev(ctx, x2, box, n, 1, 1))) return rc;
            {
                TimedLaunch tl(ctx, "cen_elem");
                hipLaunchKernelGGL(k_ctf_mul, dim3((unsigned)((N + 255) / 256), n), dim3(256), 0, ctx->stream, x2,
                                   box, parts_dev, make_ctf_const(p));
                CIM_HIP(hipGetLastError());
            }
            if ((rc = fft2d_dev(ctx, x2, box, n, -1, 1))) return rc;
        }
        s = StageArgs{nullptr, x2, nullptr, parts_dev, box, 1, mask ? 2 : 0, p->imgmask, 0, 1};
        if ((rc = launch_stage(ctx, n, s))) return rc;
        if ((rc = lowpass(ctx, x2, box, n, p->maxR))) return rc;
        s = StageArgs{nullptr, x2, nullptr, parts_dev, box, 1, 0, 0.f, 1, 0};
        if ((rc = launch_stage(ctx, n, s))) return rc;
        if ((rc = fft2d_dev(ctx, x2, box, n, 1, 1))) return rc;
    } else {
        // img2d_lowpass, img2d_mask at the origin, img2d_norm, 180-degree copy
        StageArgs s{img_dev, x1, nullptr, parts_dev, box, 0, 0, 0.f, 0, 1};
        if ((rc = launch_stage(ctx, n, s))) return rc;
        if ((rc = lowpass(ctx, x1, box, n, p->maxR))) return rc;
        s = StageArgs{nullptr, x1, x2, parts_dev, box, 1, mask ? 2 : 0, p->imgmask, 1, 0};
        if ((rc = launch_stage(ctx, n, s))) return rc;
        if ((rc = fft2d_dev(ctx, x1, box, n, 1, 1))) return rc;
        if ((rc = fft2d_dev(ctx, x2, box, n, 1, 1))) return rc;
    }
    const long long total = (long long)n * N;
    {
        TimedLaunch tl(ctx, "cen_elem");
        hipLaunchKernelGGL(k_cc_mul, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, x1, x2, box,
                           total);
        CIM_HIP(hipGetLastError());
    }
    if ((rc = fft2d_dev(ctx, x1, box, n, -1, 1))) return rc;
    {
        TimedLaunch tl(ctx, "cen_peak");
        hipLaunchKernelGGL(k_center_peak, dim3(n), dim3(1024), 0, ctx->stream, x1, box,
                           method == CIM_CENTER_SELFROT180 ? 1 : 0, parts_dev, cc_dev);
        CIM_HIP(hipGetLastError());
    }
    return CIM_OK;
}

// Particles per chunk: two float2 working images each, about 512 MB in all.
static int center_chunk_size(int box) {
    const long long per = 2LL * box * box * (long long)sizeof(float2);
    long long c = (512LL << 20) / per;
    if (c < 1) c = 1;
    if (c > 1024) c = 1024;
    return (int)c;
}

static int center_run(cim_ctx* ctx, const cim_params* p, int method, int n, const float* images, bool images_on_dev,
                      float* parts, float* cc) {
    if (!ctx) return CIM_E_ARG;
    int rc = check_params_public(ctx, p);
    if (rc) return rc;
    if (!images || !parts) return set_error(ctx, CIM_E_ARG, "NULL images or parts");
    if (method != CIM_CENTER_MODEL && method != CIM_CENTER_SELFROT180)
        return set_error(ctx, CIM_E_ARG, "method %d is neither CIM_CENTER_MODEL nor CIM_CENTER_SELFROT180", method);
    if (method == CIM_CENTER_MODEL && (!ctx->cen.ready || ctx->cen.box != p->box))
        return set_error(ctx, CIM_E_STATE, "cim_center_model has not been called for box %d", p->box);
    if (p->mask_on && !(p->imgmask > 0.f)) return set_error(ctx, CIM_E_ARG, "mask radius must be > 0");
    if (n <= 0) return CIM_OK;
    CIM_HIP(hipSetDevice(ctx->device));
    const int box = p->box;
    const long long N = (long long)box * box;
    const int chunk = center_chunk_size(box);
    const int cn = n < chunk ? n : chunk;
    CIM_ALLOC(ctx->cen.x1, sizeof(float2) * N * cn);
    CIM_ALLOC(ctx->cen.x2, sizeof(float2) * N * cn);
    CIM_ALLOC(ctx->cen.parts, sizeof(float) * CIM_PART_STRIDE * (size_t)n);
    CIM_ALLOC(ctx->cen.cc, sizeof(float) * (size_t)n);
    if (!images_on_dev) CIM_ALLOC(ctx->cen.img, sizeof(float) * N * cn);
    float* pd = ctx->cen.parts.as<float>();
    CIM_HIP(hipMemcpyAsync(pd, parts, sizeof(float) * CIM_PART_STRIDE * (size_t)n, hipMemcpyHostToDevice,
                           ctx->stream));
    for (int i0 = 0; i0 < n; i0 += chunk) {
        const int m = n - i0 < chunk ? n - i0 : chunk;
        const float* src = images + (long long)i0 * N;
        if (!images_on_dev) {
            CIM_HIP(hipMemcpyAsync(ctx->cen.img.ptr, src, sizeof(float) * N * m, hipMemcpyHostToDevice, ctx->stream));
            src = ctx->cen.img.as<float>();
        }
        rc = center_chunk(ctx, p, method, m, src, pd + (long long)i0 * CIM_PART_STRIDE, ctx->cen.cc.as<float>() + i0);
        if (rc) return rc;
    }
    CIM_HIP(hipMemcpyAsync(parts, pd, sizeof(float) * CIM_PART_STRIDE * (size_t)n, hipMemcpyDeviceToHost,
                           ctx->stream));
    if (cc)
        CIM_HIP(hipMemcpyAsync(cc, ctx->cen.cc.ptr, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

}  // namespace cim

using namespace cim;

extern "C" {

int cim_center_model(cim_ctx* ctx, const cim_params* p, const float* model) {
    if (!ctx) return CIM_E_ARG;
    if (!p || !model) return set_error(ctx, CIM_E_ARG, "NULL argument");
    const int n = p->box;
    if (n < 8 || n > 1024 || (n & 1)) return set_error(ctx, CIM_E_ARG, "model box %d unsupported", n);
    if (p->mask_on && !(p->imgmask > 0.f)) return set_error(ctx, CIM_E_ARG, "mask radius must be > 0");
    CIM_HIP(hipSetDevice(ctx->device));
    const long long N = (long long)n * n * n;
    ctx->cen.ready = false;
    CIM_ALLOC(ctx->img, sizeof(float) * N);
    CIM_HIP(hipMemcpyAsync(ctx->img.ptr, model, sizeof(float) * N, hipMemcpyHostToDevice, ctx->stream));
    CIM_ALLOC(ctx->cen.model, sizeof(float2) * N);
    // shell sum partials of k_mask3d_sum's 1024 blocks, then their total (sum hi, lo, count hi, lo)
    CIM_ALLOC(ctx->cen.acc, sizeof(float) * 4 * (1024 + 1));
    CIM_HIP(hipMemsetAsync(ctx->cen.acc.ptr, 0, sizeof(float) * 4 * (1024 + 1), ctx->stream));
    float* acc_total = ctx->cen.acc.as<float>() + 4 * 1024;
    // img3d_mask.f90:11-17 radii; R2 = 0 switches the mask off
    float R1 = 0.f, R2 = 0.f;
    if (p->mask_on) {
        R1 = p->imgmask - (float)(kHalfw / 2);
        if (R1 <= 0.f) R1 = 0.f;
        R2 = p->imgmask + (float)(kHalfw / 2);
        if (R2 > (float)(n / 2)) R2 = (float)(n / 2);
        hipLaunchKernelGGL(k_mask3d_sum, dim3(1024), dim3(256), 0, ctx->stream, ctx->img.as<float>(), n, R1, R2,
                           ctx->cen.acc.as<float>());
        CIM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, ctx->stream, ctx->cen.acc.as<float>(), 1024,
                           acc_total);
        CIM_HIP(hipGetLastError());
    }
    const unsigned nb = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(k_mask3d_apply, dim3(nb), dim3(256), 0, ctx->stream, ctx->img.as<float>(), n, R1, R2,
                       acc_total, ctx->cen.model.as<float2>());
    CIM_HIP(hipGetLastError());
    int rc = fft3d_dev(ctx, ctx->cen.model.as<float2>(), n, 1, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_checker3d_conj, dim3(nb), dim3(256), 0, ctx->stream, ctx->cen.model.as<float2>(), n);
    CIM_HIP(hipGetLastError());
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    ctx->cen.box = n;
    ctx->cen.ready = true;
    return CIM_OK;
}

int cim_center(cim_ctx* ctx, const cim_params* p, int method, int n, const float* images, float* parts, float* cc) {
    return center_run(ctx, p, method, n, images, false, parts, cc);
}

int cim_center_dev(cim_ctx* ctx, const cim_params* p, int method, int n, const void* images_dev, float* parts,
                   float* cc) {
    return center_run(ctx, p, method, n, static_cast<const float*>(images_dev), true, parts, cc);
}

}  // extern "C"
