// Half-map FSC and the tnf merges.
//
// cim_fsc_map: reconstruct_FSC.f90:10-197 (icosprocess) and
//   merge_reconstruct_FSC.f90:9-210 (merge_icosprocess_subrec) -- shell FSC
//   of the odd/even tnf sums, FSC-weighted merge, fold, inverse FFT.
// cim_merge_map: merge_reconstruct.f90:14-157 -- the plain merge of partial
//   tnf sums accumulated with cim_rec_add_tnf.
//
// The tnf volumes are (2R+1)^2 (R+1) floats, lz slowest, x fastest, covering
// hx, ky in [-R, R] and lz in [0, R].
#include <cmath>
#include <vector>

#include "cim_device.h"

namespace cim {

int rec_merge_dev(cim_ctx* ctx, float* out_dev, int nicosavg, float imgmask);

// Weight sums + 1, normalised sums, and the per-shell FP64 sums of
// a1 a2 + b1 b2, |1|^2, |2|^2 over ir = int(r) <= R
// (reconstruct_FSC.f90:30-52).  in / nrm: six arrays of n floats
// (a1, b1, w1, a2, b2, w2).  Shell bins are gathered per workgroup in LDS.
__global__ void __launch_bounds__(256) k_fsc_prep(const float* __restrict__ in, float* __restrict__ nrm,
                                                  double* __restrict__ sums, int maxR, long long n) {
    extern __shared__ double bins[];   // 3 (maxR + 1)
    const int nb = 3 * (maxR + 1);
    for (int i = threadIdx.x; i < nb; i += blockDim.x) bins[i] = 0.0;
    __syncthreads();
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) {
        const float w1 = in[2 * n + id] + 1.f, w2 = in[5 * n + id] + 1.f;
        const float a1 = in[id] / w1, b1 = in[n + id] / w1;
        const float a2 = in[3 * n + id] / w2, b2 = in[4 * n + id] / w2;
        nrm[id] = a1;
        nrm[n + id] = b1;
        nrm[2 * n + id] = w1;
        nrm[3 * n + id] = a2;
        nrm[4 * n + id] = b2;
        nrm[5 * n + id] = w2;
        const int t = 2 * maxR + 1;
        const int lz = (int)(id / ((long long)t * t));
        const int ky = (int)((id / t) % t) - maxR, hx = (int)(id % t) - maxR;
        const int ir = (int)sqrtf((float)(hx * hx + ky * ky + lz * lz));
        if (ir <= maxR) {
            atomicAdd(&bins[ir], (double)a1 * a2 + (double)b1 * b2);
            atomicAdd(&bins[maxR + 1 + ir], (double)a1 * a1 + (double)b1 * b1);
            atomicAdd(&bins[2 * (maxR + 1) + ir], (double)a2 * a2 + (double)b2 * b2);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        if (bins[i] != 0.0) atomicAdd(&sums[i], bins[i]);
}

__device__ __forceinline__ bool in_shell(int hx, int ky, int lz, int maxR) {
    return (int)sqrtf((float)(hx * hx + ky * ky + lz * lz)) <= maxR;
}

// FSC-weighted merge (reconstruct_FSC.f90:65-89), written as {fa, fb, wg}
// with row pitch 2R+1 for the fold.  The weighting loop's else-branch
// (:82-84) zeroes fa, fb at the index left from the previous in-sphere voxel,
// i.e. every in-sphere voxel whose successor in the (lz, ky, hx) loop order
// lies outside the sphere ends with fa = fb = 0; that is reproduced.
__global__ void k_fsc_merge(const float* __restrict__ nrm, const float* __restrict__ wt, float4* __restrict__ out,
                            int maxR, long long n) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const float a1 = nrm[id], b1 = nrm[n + id], w1 = nrm[2 * n + id];
    const float a2 = nrm[3 * n + id], b2 = nrm[4 * n + id], w2 = nrm[5 * n + id];
    float a = __fadd_rn(__fmul_rn(a1, w1), __fmul_rn(a2, w2));
    float b = __fadd_rn(__fmul_rn(b1, w1), __fmul_rn(b2, w2));
    float w = __fadd_rn(__fadd_rn(w1, w2), -2.f);
    const int t = 2 * maxR + 1;
    const int lz = (int)(id / ((long long)t * t));
    const int ky = (int)((id / t) % t) - maxR, hx = (int)(id % t) - maxR;
    if (in_shell(hx, ky, lz, maxR)) {
        const int ir = (int)sqrtf((float)(hx * hx + ky * ky + lz * lz));
        a = __fmul_rn(a, wt[ir]);
        b = __fmul_rn(b, wt[ir]);
        w = __fmul_rn(w, wt[ir]);
        int nx = hx + 1, ny = ky, nz = lz;
        if (nx > maxR) {
            nx = -maxR;
            if (++ny > maxR) {
                ny = -maxR;
                ++nz;
            }
        }
        if (nz > maxR || !in_shell(nx, ny, nz, maxR)) {
            a = 0.f;
            b = 0.f;
        }
    }
    out[id] = make_float4(a, b, __fadd_rn(w, 1.f), 0.f);
}

static int fsc_map(cim_ctx* ctx, int box, int maxR, float imgmask, int variant, int nicosavg,
                   const float* const in[6], float* fsc, float* map) {
    const int t = 2 * maxR + 1;
    const long long n = (long long)t * t * (maxR + 1);
    CIM_ALLOC(ctx->fsc_in, sizeof(float) * 6 * n);
    CIM_ALLOC(ctx->fsc_nrm, sizeof(float) * 6 * n);
    CIM_ALLOC(ctx->fsc_acc, sizeof(float4) * n);
    CIM_ALLOC(ctx->fsc_sums, sizeof(double) * 3 * (maxR + 1));
    CIM_ALLOC(ctx->fsc_wt, sizeof(float) * (maxR + 1));
    float* din = ctx->fsc_in.as<float>();
    for (int k = 0; k < 6; ++k)
        CIM_HIP(hipMemcpyAsync(din + k * n, in[k], sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    double* dsum = ctx->fsc_sums.as<double>();
    std::vector<double> zeros(3 * (maxR + 1), 0.0);
    CIM_HIP(hipMemcpyAsync(dsum, zeros.data(), sizeof(double) * zeros.size(), hipMemcpyHostToDevice, ctx->stream));
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_fsc_prep, dim3(nb), dim3(256), sizeof(double) * 3 * (maxR + 1), ctx->stream, din,
                       ctx->fsc_nrm.as<float>(), dsum, maxR, n);
    CIM_HIP(hipGetLastError());
    std::vector<double> S(3 * (maxR + 1));
    CIM_HIP(hipMemcpyAsync(S.data(), dsum, sizeof(double) * S.size(), hipMemcpyDeviceToHost, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    // FSC(ir) = sum / sqrt((1e-10 + sum1)(1e-10 + sum2)); weight sqrt(2 FSC / (1 + FSC)), 0 if FSC <= 0
    // (reconstruct_FSC.f90:37-62)
    std::vector<float> wt(maxR + 1);
    for (int r = 0; r <= maxR; ++r) {
        const double s1 = 1.0e-10 + S[maxR + 1 + r], s2 = 1.0e-10 + S[2 * (maxR + 1) + r];
        const float f = (float)(S[r] / std::sqrt(s1 * s2));
        if (fsc) fsc[r] = f;
        wt[r] = f <= 0.f ? 0.f : std::sqrt(2.f * f / (1.f + f));
    }
    CIM_HIP(hipMemcpyAsync(ctx->fsc_wt.ptr, wt.data(), sizeof(float) * wt.size(), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_fsc_merge, dim3(nb), dim3(256), 0, ctx->stream, ctx->fsc_nrm.as<float>(),
                       ctx->fsc_wt.as<float>(), ctx->fsc_acc.as<float4>(), maxR, n);
    CIM_HIP(hipGetLastError());
    const long long nvox = (long long)box * box * box;
    CIM_ALLOC(ctx->out, sizeof(float) * nvox);
    // reconstruct_FSC transforms cmplx(fa, fb) backward (:173-179);
    // merge_reconstruct_FSC transforms it forward with FFT1d (:181-185),
    // whose real part is the backward transform of cmplx(fa, -fb).
    const float isg = variant == CIM_FSC_RECONSTRUCT ? 1.f : -1.f;
    const float rmask = variant == CIM_FSC_RECONSTRUCT ? -1.f : imgmask;
    int rc = fold_to_map_dev(ctx, ctx->fsc_acc.as<float4>(), nullptr, t, box, maxR, 1, isg, rmask,
                             ctx->out.as<float>());
    if (rc) return rc;
    // realspaceavg: icosprocess's img3d_icosavg on the map (reconstruct_FSC.f90:185),
    // or merge's on the first (box-8)^3 values taken as a cube
    // (merge_reconstruct_FSC.f90:202, after maphead set FFTsize = box-8)
    if (nicosavg == 1) {
        rc = variant == CIM_FSC_RECONSTRUCT
                 ? icosavg_dev(ctx, ctx->out.as<float>(), nvox, box, imgmask, CIM_ICOSAVG_FULL)
                 : icosavg_dev(ctx, ctx->out.as<float>(), nvox, box - 8, imgmask, CIM_ICOSAVG_MERGE);
        if (rc) return rc;
    }
    // merge_reconstruct_FSC writes (box-8) planes of (box-8)^2 values from the
    // start of the flat array (:204-206, after maphead set FFTsize = box-8)
    const long long nout = variant == CIM_FSC_RECONSTRUCT ? nvox : (long long)(box - 8) * (box - 8) * (box - 8);
    CIM_HIP(hipMemcpyAsync(map, ctx->out.ptr, sizeof(float) * nout, hipMemcpyDeviceToHost, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

}  // namespace cim

using namespace cim;

extern "C" {

int cim_fsc_map(cim_ctx* ctx, int box, int maxR, float imgmask, int variant, int nicosavg, const float* odd_fa,
                const float* odd_fb, const float* odd_wg, const float* eve_fa, const float* eve_fb, const float* eve_wg,
                float* fsc, float* map) {
    if (!ctx) return CIM_E_ARG;
    if (!odd_fa || !odd_fb || !odd_wg || !eve_fa || !eve_fb || !eve_wg || !map)
        return set_error(ctx, CIM_E_ARG, "NULL argument");
    if (variant != CIM_FSC_RECONSTRUCT && variant != CIM_FSC_MERGE)
        return set_error(ctx, CIM_E_ARG, "variant %d is neither CIM_FSC_RECONSTRUCT nor CIM_FSC_MERGE", variant);
    if (box < 16 || (box & 1)) return set_error(ctx, CIM_E_ARG, "box %d must be even and >= 16", box);
    if (maxR < 1 || maxR > box / 2 - 1)
        return set_error(ctx, CIM_E_ARG, "maxR %d outside [1, box/2-1] for box %d", maxR, box);
    CIM_HIP(hipSetDevice(ctx->device));
    const float* in[6] = {odd_fa, odd_fb, odd_wg, eve_fa, eve_fb, eve_wg};
    return fsc_map(ctx, box, maxR, imgmask, variant, nicosavg, in, fsc, map);
}

int cim_merge_map(cim_ctx* ctx, int nicosavg, float imgmask, float* map) {
    if (!ctx) return CIM_E_ARG;
    if (!map) return set_error(ctx, CIM_E_ARG, "NULL map");
    if (!ctx->rec.active) return set_error(ctx, CIM_E_STATE, "cim_rec_begin has not been called");
    if (ctx->rec.box < 16) return set_error(ctx, CIM_E_ARG, "box %d too small for the (box-8)^3 crop", ctx->rec.box);
    CIM_HIP(hipSetDevice(ctx->device));
    const long long m3 = (long long)(ctx->rec.box - 8) * (ctx->rec.box - 8) * (ctx->rec.box - 8);
    CIM_ALLOC(ctx->out, sizeof(float) * m3);
    int rc = rec_merge_dev(ctx, ctx->out.as<float>(), nicosavg, imgmask);
    if (rc) return rc;
    CIM_HIP(hipMemcpyAsync(map, ctx->out.ptr, sizeof(float) * m3, hipMemcpyDeviceToHost, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

}  // extern "C"
