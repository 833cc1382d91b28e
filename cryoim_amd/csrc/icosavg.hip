// Real-space icosahedral averaging of a map (realspaceavg=y).
//
// CIM_ICOSAVG_FULL: icosprocess/SRC/img3d_icosavg.f90:2-126 -- every voxel
//   within r0 of the centre becomes the mean over the operators (the 60 of
//   the icosahedral group, or the cyclic group of sym) of the trilinearly
//   interpolated map at S x; voxels beyond r0 become 0.
// CIM_ICOSAVG_MERGE: merge_icosprocess_subrec/SRC/img3d_icosavg.f90:1-91 --
//   the same mean over the quarter lz, ky >= 0 only, then the copies of its
//   second loop (to (-hx, -ky, lz), and for lz > 0 to (-hx, ky, -lz) and
//   (hx, -ky, -lz)) in the reference's loop order and flat indexing.
//
// Both treat the first n^3 values of the map as an n^3 cube; n may be
// smaller than the map's own box (merge_reconstruct_FSC.f90:202 averages
// after maphead has set FFTsize = box-8), and taps may then read past n^3:
// every tap reads the flat index the reference reads, within the len values
// the caller passes (0 outside them).
#include <vector>

#include "cim_device.h"

namespace cim {

namespace {

// Sum over the operators for voxels of the averaged region (all n^3, or the
// quarter lz in [0, n/2), ky in [0, n/2), hx in [-n/2, n/2)).  The
// operators (row-major float, the reference's order, at most 60) are
// staged in LDS.
__global__ void __launch_bounds__(256) k_icosavg_sum(const float* __restrict__ in, long long len,
                                                     float* __restrict__ acc, int n, float r0, int quarter,
                                                     const float* __restrict__ ops, int nops) {
    __shared__ float S[60 * 9];
    for (int i = threadIdx.x; i < 9 * nops; i += blockDim.x) S[i] = ops[i];
    __syncthreads();
    const long long n2 = (long long)n * n;
    const long long total = quarter ? (long long)(n / 2) * (n / 2) * n : n2 * n;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int hx, ky, lz;
    if (quarter) {
        hx = (int)(i % n) - n / 2;
        ky = (int)((i / n) % (n / 2));
        lz = (int)(i / ((long long)n * (n / 2)));
    } else {
        hx = (int)(i % n) - n / 2;
        ky = (int)((i / n) % n) - n / 2;
        lz = (int)(i / n2) - n / 2;
    }
    const float r = sqrtf((float)(hx * hx + ky * ky + lz * lz));
    if (r > r0) return;
    const float fx = (float)hx, fy = (float)ky, fz = (float)lz, c = (float)(n / 2);
    float a = 0.f;
    for (int s = 0; s < nops; ++s) {
        const float* M = S + 9 * s;
        const float x = M[0] * fx + M[1] * fy + M[2] * fz + c;
        const float y = M[3] * fx + M[4] * fy + M[5] * fz + c;
        const float z = M[6] * fx + M[7] * fy + M[8] * fz + c;
        const int ix = (int)x, iy = (int)y, iz = (int)z;
        const float ax = x - (float)ix, ay = y - (float)iy, az = z - (float)iz;
        const long long n1 = (long long)iz * n2 + (long long)iy * n + ix;
        auto tap = [&](long long k) -> float { return (k >= 0 && k < len) ? in[k] : 0.f; };
        // img3d1(i3) + t1 + ... + t8, left to right (img3d_icosavg.f90:75-83)
        a = a + (1.0f - ax) * (1.0f - ay) * (1.0f - az) * tap(n1);
        a = a + ax * (1.0f - ay) * (1.0f - az) * tap(n1 + 1);
        a = a + (1.0f - ax) * ay * (1.0f - az) * tap(n1 + n);
        a = a + ax * ay * (1.0f - az) * tap(n1 + n + 1);
        a = a + (1.0f - ax) * (1.0f - ay) * az * tap(n1 + n2);
        a = a + ax * (1.0f - ay) * az * tap(n1 + 1 + n2);
        a = a + (1.0f - ax) * ay * az * tap(n1 + n + n2);
        a = a + ax * ay * az * tap(n1 + n + 1 + n2);
    }
    acc[(long long)(lz + n / 2) * n2 + (long long)(ky + n / 2) * n + (hx + n / 2)] = a;
}

// The second loop of merge's img3d_icosavg, one thread per slice lz >= 0.
// A slice reads only its own quarter voxels, and the only writes that land
// on them come from the same slice (the ky = 0 row and the x-index-n wrap of
// hx = -n/2 into the next row); writes from different slices never meet.
// So each slice replays its part of the loop in order, independently.
__global__ void k_icosavg_copy(float* __restrict__ acc, int n, float r0) {
    const int lz = blockIdx.x * blockDim.x + threadIdx.x;
    if (lz >= n / 2) return;
    const long long n2 = (long long)n * n, N = n2 * n;
    const int h = n / 2;
    for (int ky = 0; ky < h; ++ky)
        for (int hx = -h; hx < h; ++hx) {
            const float r = sqrtf((float)(hx * hx + ky * ky + lz * lz));
            if (r > r0) continue;
            const float v = acc[(long long)(lz + h) * n2 + (long long)(ky + h) * n + (hx + h)];
            long long t = (long long)(lz + h) * n2 + (long long)(-ky + h) * n + (-hx + h);
            if (t >= 0 && t < N) acc[t] = v;
            if (lz > 0) {
                t = (long long)(-lz + h) * n2 + (long long)(ky + h) * n + (-hx + h);
                if (t >= 0 && t < N) acc[t] = v;
                t = (long long)(-lz + h) * n2 + (long long)(-ky + h) * n + (hx + h);
                if (t >= 0 && t < N) acc[t] = v;
            }
        }
}

__global__ void k_icosavg_div(const float* __restrict__ acc, float* __restrict__ out, long long N, float nsym) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) out[i] = acc[i] / nsym;
}

}  // namespace

int icosavg_dev(cim_ctx* ctx, float* map_dev, long long len, int n, float r0, int variant, int nops,
                const double* ops) {
    if (variant != CIM_ICOSAVG_FULL && variant != CIM_ICOSAVG_MERGE)
        return set_error(ctx, CIM_E_ARG, "variant %d is neither CIM_ICOSAVG_FULL nor CIM_ICOSAVG_MERGE", variant);
    if (n < 4 || n > 1024 || (n & 1)) return set_error(ctx, CIM_E_ARG, "cube side %d unsupported", n);
    const long long N = (long long)n * n * n;
    if (len < N) return set_error(ctx, CIM_E_ARG, "map of %lld values holds no %d^3 cube", len, n);
    if (ops && variant == CIM_ICOSAVG_MERGE)
        return set_error(ctx, CIM_E_ARG, "merge's average is icosahedral only (ops must be NULL)");
    if (ops && (nops < 1 || nops > 60)) return set_error(ctx, CIM_E_ARG, "nops %d outside [1, 60]", nops);
    std::vector<double> d;
    if (ops) {
        d.assign(ops, ops + 9 * (size_t)nops);
    } else {
        nops = 60;
        d.resize(60 * 9);
        cim_icos60_ref_order(d.data());
    }
    std::vector<float> f(d.begin(), d.end());
    CIM_ALLOC(ctx->avg_ops, sizeof(float) * f.size());
    CIM_HIP(hipMemcpyAsync(ctx->avg_ops.ptr, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice, ctx->stream));
    CIM_ALLOC(ctx->avg_acc, sizeof(float) * N);
    float* acc = ctx->avg_acc.as<float>();
    CIM_HIP(hipMemsetAsync(acc, 0, sizeof(float) * N, ctx->stream));
    const int quarter = variant == CIM_ICOSAVG_MERGE;
    const long long total = quarter ? (long long)(n / 2) * (n / 2) * n : N;
    {
        TimedLaunch tl(ctx, "icosavg");
        hipLaunchKernelGGL(k_icosavg_sum, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, map_dev,
                           len, acc, n, r0, quarter, ctx->avg_ops.as<float>(), nops);
        CIM_HIP(hipGetLastError());
    }
    if (quarter) {
        hipLaunchKernelGGL(k_icosavg_copy, dim3((unsigned)((n / 2 + 63) / 64)), dim3(64), 0, ctx->stream, acc, n, r0);
        CIM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_icosavg_div, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream, acc, map_dev, N,
                       (float)nops);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

}  // namespace cim

using namespace cim;

extern "C" {

int cim_icosavg(cim_ctx* ctx, float* map, long long len, int n, float r0, int variant, int nops,
                const double* ops) {
    if (!ctx) return CIM_E_ARG;
    if (!map) return set_error(ctx, CIM_E_ARG, "NULL map");
    if (len <= 0 || len > (1LL << 31)) return set_error(ctx, CIM_E_ARG, "map length %lld unsupported", len);
    CIM_HIP(hipSetDevice(ctx->device));
    CIM_ALLOC(ctx->out, sizeof(float) * len);
    CIM_HIP(hipMemcpyAsync(ctx->out.ptr, map, sizeof(float) * len, hipMemcpyHostToDevice, ctx->stream));
    int rc = icosavg_dev(ctx, ctx->out.as<float>(), len, n, r0, variant, nops, ops);
    if (rc) return rc;
    const long long N = (long long)n * n * n;
    CIM_HIP(hipMemcpyAsync(map, ctx->out.ptr, sizeof(float) * N, hipMemcpyDeviceToHost, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

int cim_icosavg_dev(cim_ctx* ctx, void* map_dev, long long len, int n, float r0, int variant, int nops,
                    const double* ops) {
    if (!ctx) return CIM_E_ARG;
    if (!map_dev) return set_error(ctx, CIM_E_ARG, "NULL map");
    CIM_HIP(hipSetDevice(ctx->device));
    return icosavg_dev(ctx, static_cast<float*>(map_dev), len, n, r0, variant, nops, ops);
}

}  // extern "C"
