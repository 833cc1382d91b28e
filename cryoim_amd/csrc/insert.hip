// Direct Fourier inversion: symmetry-expanded trilinear gridding insert
// (icosprocess/SRC/reconstruct.f90:103-147, reconstruct_substep.f90:101-145),
// fold/normalise (reconstruct.f90:151-237) and the final inverse 3D FFT
// with checkerboard (reconstruct.f90:240-247).
//
// Accumulator: one float4 {fa, fb, wg, -} per voxel so the three atomics of
// a corner hit one 16-byte line; it covers x, y in [-maxR, maxR+1] and
// z in [0, maxR+1] (every corner a sample of radius <= maxR can reach), which
// is both the full-box variant restricted to where it is ever written and a
// superset of the compact tnf block.
#include <cmath>

#include "cim_device.h"

namespace cim {

// Compose M = S_o * R(theta, phi, omega) in float64 for every particle and
// operator; stored as float[9] row-major.
__global__ void k_compose(const float* __restrict__ parts, int n, const double* __restrict__ ops, int nops,
                          float* __restrict__ M) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n * nops) return;
    int p = id / nops, o = id % nops;
    const float* pr = parts + (long long)p * CIM_PART_STRIDE;
    const double d2r = 3.14159265358979323846 / 180.0;
    double th = pr[CIM_P_THETA] * d2r, ph = pr[CIM_P_PHI] * d2r, om = pr[CIM_P_OMEGA] * d2r;
    double c1 = cos(ph), c2 = cos(th), c3 = cos(om), s1 = sin(ph), s2 = sin(th), s3 = sin(om);
    double R[9] = {c1 * c2 * c3 - s1 * s3, -c1 * c2 * s3 - c3 * s1, c1 * s2,
                   c2 * c3 * s1 + c1 * s3, -c2 * s1 * s3 + c1 * c3, s1 * s2,
                   -c3 * s2,               s2 * s3,                 c2};
    const double* S = ops + 9 * o;
    float* out = M + (long long)id * 9;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            out[3 * i + j] = (float)(S[3 * i] * R[j] + S[3 * i + 1] * R[3 + j] + S[3 * i + 2] * R[6 + j]);
}

struct InsertGeom {
    int box, maxR, dxy, nops;
    float off;     // added to x, y before truncation (box/2 full, maxR compact)
    float offz;    // added to z (box/2 full, 0 compact)
};

// One thread per (half-plane sample, particle); loops over the operators.
// ft: centred, centre-shifted transforms.  ctf: either computed per sample
// from the particle record (ctf_img == nullptr && apply_ctf) or read from
// ctf_img (n*box*box), or 1.
__global__ void __launch_bounds__(256) k_insert(float4* __restrict__ acc, const float2* __restrict__ ft,
                                                const int2* __restrict__ samples, int nsamples,
                                                const float* __restrict__ M, const float* __restrict__ parts,
                                                const float* __restrict__ ctf_img, int apply_ctf, CtfConst cc,
                                                InsertGeom g) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (s >= nsamples) return;
    const int2 hk = samples[s];
    const long long N = (long long)g.box * g.box;
    const long long i2 = (long long)(hk.y + g.box / 2) * g.box + (hk.x + g.box / 2);
    const float2 v = ft[(long long)p * N + i2];
    float c = 1.f;
    if (ctf_img) c = ctf_img[(long long)p * N + i2];
    else if (apply_ctf) {
        const float* pr = parts + (long long)p * CIM_PART_STRIDE;
        c = ctf_value(hk.x, hk.y, g.box, pr[CIM_P_DF1], pr[CIM_P_DF2], pr[CIM_P_ASTIG], cc);
    }
    const float va = v.x * c, vb = v.y * c, vw = c * c;
    const float fh = (float)hk.x, fk = (float)hk.y;
    const float* Mp = M + (long long)p * g.nops * 9;
    const long long dxy = g.dxy, dxy2 = dxy * dxy;
    for (int o = 0; o < g.nops; ++o) {
        const float* m = Mp + 9 * o;
        float x = m[0] * fh + m[1] * fk;
        float y = m[3] * fh + m[4] * fk;
        float z = m[6] * fh + m[7] * fk;
        float sb = vb;
        if (z < 0.f) {
            x = -x; y = -y; z = -z;
            sb = -vb;
        }
        x += g.off;
        y += g.off;
        z += g.offz;
        const int lx = (int)x, ly = (int)y, lz = (int)z;
        const float ax = x - (float)lx, ay = y - (float)ly, az = z - (float)lz;
        const int gx = lx - (int)g.off + g.maxR, gy = ly - (int)g.off + g.maxR, gz = lz - (int)g.offz;
        float4* b = acc + ((long long)gz * dxy2 + (long long)gy * dxy + gx);
        const float w00 = (1.f - ax) * (1.f - ay), w10 = ax * (1.f - ay), w01 = (1.f - ax) * ay, w11 = ax * ay;
        const float wz0 = 1.f - az, wz1 = az;
        const float w[8] = {w00 * wz0, w10 * wz0, w01 * wz0, w11 * wz0, w00 * wz1, w10 * wz1, w01 * wz1, w11 * wz1};
        const long long offs[8] = {0, 1, dxy, dxy + 1, dxy2, dxy2 + 1, dxy2 + dxy, dxy2 + dxy + 1};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float4* q = b + offs[j];
            unsafeAtomicAdd(&q->x, va * w[j]);
            unsafeAtomicAdd(&q->y, sb * w[j]);
            unsafeAtomicAdd(&q->z, vw * w[j]);
        }
    }
}

__global__ void k_fill_acc(float4* __restrict__ acc, long long count, float wg0) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) acc[i] = make_float4(0.f, 0.f, wg0, 0.f);
}

// Fold + normalise + Hermitian completion, written as the complex input of
// the backward FFT: cmplx(fa, -fb) (reconstruct.f90:241).
__global__ void k_fold(const float4* __restrict__ acc, float2* __restrict__ vol, int box, int maxR, int dxy) {
    const long long n2 = (long long)box * box;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n2 * box) return;
    const int Z = (int)(id / n2), Y = (int)((id / box) % box), X = (int)(id % box);
    const int c = box / 2;
    const int h = X - c, k = Y - c, l = Z - c;
    float2 outv = make_float2(0.f, 0.f);
    if (abs(h) <= maxR && abs(k) <= maxR && abs(l) <= maxR) {
        const float dr = sqrtf((float)(h * h + k * k + l * l));
        if (dr <= (float)maxR) {
            const long long dxy2 = (long long)dxy * dxy;
            auto at = [&](int hh, int kk, int ll) -> float4 {
                return acc[(long long)ll * dxy2 + (long long)(kk + maxR) * dxy + (hh + maxR)];
            };
            const float s = ((h + k + l) & 1) ? -1.f : 1.f;
            if (l != 0) {
                // l > 0 stored directly; l < 0 is the Friedel mate of (-h,-k,-l)
                const int sg = l > 0 ? 1 : -1;
                float4 a = at(sg * h, sg * k, sg * l);
                const double A = (double)a.x / (double)a.z, B = (double)a.y / (double)a.z;
                outv = make_float2((float)A * s, (l > 0 ? -(float)B : (float)B) * s);
            } else if (h == 0 && k == 0) {
                float4 a = at(0, 0, 0);
                outv = make_float2((float)((double)a.x / (double)a.z), 0.f);
            } else {
                // n1: k > 0, or k == 0 and h > 0; n2 = -n1
                const bool is_n1 = (k > 0) || (k == 0 && h > 0);
                const int h1 = is_n1 ? h : -h, k1 = is_n1 ? k : -k;
                float4 a1 = at(h1, k1, 0), a2 = at(-h1, -k1, 0);
                const double W = (double)a1.z + (double)a2.z;
                const double A = ((double)a1.x + (double)a2.x) / W;
                const double B = ((double)a1.y - (double)a2.y) / W;
                outv = make_float2((float)A * s, (is_n1 ? -(float)B : (float)B) * s);
            }
        }
    }
    vol[id] = outv;
}

__global__ void k_checker_real(const float2* __restrict__ vol, float* __restrict__ map, int box) {
    const long long n2 = (long long)box * box;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n2 * box) return;
    const int Z = (int)(id / n2), Y = (int)((id / box) % box), X = (int)(id % box);
    const float v = vol[id].x;
    map[id] = ((X + Y + Z) & 1) ? -v : v;
}

// Copy between the accumulator and tnf order ((2R+1)^2 (R+1), lz slowest).
__global__ void k_tnf_io(float4* __restrict__ acc, float* __restrict__ fa, float* __restrict__ fb,
                         float* __restrict__ wg, int maxR, int dxy, int add_into_acc) {
    const int t = 2 * maxR + 1;
    const long long total = (long long)t * t * (maxR + 1);
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total) return;
    const int lz = (int)(id / ((long long)t * t));
    const int ky = (int)((id / t) % t), hx = (int)(id % t);
    float4* a = acc + ((long long)lz * dxy * dxy + (long long)ky * dxy + hx);
    if (add_into_acc) {
        a->x += fa[id];
        a->y += fb[id];
        a->z += wg[id];
    } else {
        float4 v = *a;
        fa[id] = v.x;
        fb[id] = v.y;
        wg[id] = v.z;
    }
}

// ---- host side -----------------------------------------------------------

int rec_begin(cim_ctx* ctx, int box, int maxR, int layout, float rcut, float wg_init, int nops, const double* ops) {
    if (box < 8 || (box & 1)) return set_error(ctx, CIM_E_ARG, "box %d must be even and >= 8", box);
    if (maxR < 1 || maxR > box / 2 - 2)
        return set_error(ctx, CIM_E_ARG, "maxR %d outside [1, box/2-2] for box %d", maxR, box);
    if (layout != 0 && layout != 1) return set_error(ctx, CIM_E_ARG, "layout must be 0 (full) or 1 (compact)");
    if (layout == 1 && rcut > (float)(maxR - 2))
        return set_error(ctx, CIM_E_ARG, "compact layout needs rcut <= maxR-2 (reconstruct_substep.f90:104)");
    if (rcut > (float)maxR) return set_error(ctx, CIM_E_ARG, "rcut %.2f > maxR %d", rcut, maxR);
    if (nops < 1 || nops > 360) return set_error(ctx, CIM_E_ARG, "nops %d outside [1, 360]", nops);
    RecState& r = ctx->rec;
    r.box = box;
    r.maxR = maxR;
    r.layout = layout;
    r.rcut = rcut;
    r.wg_init = wg_init;
    r.nops = nops;
    r.dxy = 2 * maxR + 2;
    r.dz = maxR + 2;
    r.inserted = 0;
    const long long nvox = (long long)r.dxy * r.dxy * r.dz;
    CIM_ALLOC(r.acc, sizeof(float4) * nvox);
    hipLaunchKernelGGL(k_fill_acc, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, ctx->stream, r.acc.as<float4>(),
                       nvox, wg_init);
    CIM_HIP(hipGetLastError());
    std::vector<double> hops(9 * nops);
    if (ops) {
        for (int i = 0; i < 9 * nops; ++i) hops[i] = ops[i];
    } else {
        if (nops != 60) return set_error(ctx, CIM_E_ARG, "ops == NULL requires nops == 60 (icosahedral)");
        cim_icos60_ref_order(hops.data());
    }
    CIM_ALLOC(r.ops, sizeof(double) * 9 * nops);
    CIM_HIP(hipMemcpyAsync(r.ops.ptr, hops.data(), sizeof(double) * 9 * nops, hipMemcpyHostToDevice, ctx->stream));
    // half-plane sample list in the loop order of reconstruct.f90:103-106
    std::vector<int2> smp;
    for (int ky = -maxR; ky <= maxR; ++ky)
        for (int hx = 0; hx <= maxR; ++hx) {
            float dr = std::sqrt((float)(hx * hx + ky * ky));
            if (dr <= rcut) smp.push_back(make_int2(hx, ky));
        }
    r.nsamples = (int)smp.size();
    CIM_ALLOC(r.samples, sizeof(int2) * smp.size());
    CIM_HIP(hipMemcpyAsync(r.samples.ptr, smp.data(), sizeof(int2) * smp.size(), hipMemcpyHostToDevice, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    r.active = true;
    return CIM_OK;
}

// Insert n particles whose centred, centre-shifted FTs are in ft (device).
int rec_insert_ft_dev(cim_ctx* ctx, int n, const float2* ft, const float* parts_dev, const float* ctf_img_dev,
                      int apply_ctf, const cim_params* p) {
    RecState& r = ctx->rec;
    if (!r.active) return set_error(ctx, CIM_E_STATE, "cim_rec_begin has not been called");
    if (n <= 0) return CIM_OK;
    CIM_ALLOC(ctx->aux2, sizeof(float) * 9 * (size_t)n * r.nops);
    {
        const int tot = n * r.nops;
        hipLaunchKernelGGL(k_compose, dim3((tot + 255) / 256), dim3(256), 0, ctx->stream, parts_dev, n,
                           r.ops.as<double>(), r.nops, ctx->aux2.as<float>());
        CIM_HIP(hipGetLastError());
    }
    InsertGeom g;
    g.box = r.box;
    g.maxR = r.maxR;
    g.dxy = r.dxy;
    g.nops = r.nops;
    g.off = r.layout == 0 ? (float)(r.box / 2) : (float)r.maxR;
    g.offz = r.layout == 0 ? (float)(r.box / 2) : 0.f;
    CtfConst cc{};
    if (p) cc = make_ctf_const(p);
    TimedLaunch tl(ctx, "insert");
    hipLaunchKernelGGL(k_insert, dim3((r.nsamples + 255) / 256, n), dim3(256), 0, ctx->stream, r.acc.as<float4>(), ft,
                       r.samples.as<int2>(), r.nsamples, ctx->aux2.as<float>(), parts_dev, ctf_img_dev, apply_ctf, cc,
                       g);
    CIM_HIP(hipGetLastError());
    r.inserted += n;
    ctx->counters["insert_samples"] += (long long)n * r.nsamples;
    ctx->counters["insert_particles"] += n;
    return CIM_OK;
}

int rec_tnf_io(cim_ctx* ctx, float* fa, float* fb, float* wg, int add) {
    RecState& r = ctx->rec;
    if (!r.active) return set_error(ctx, CIM_E_STATE, "cim_rec_begin has not been called");
    const int t = 2 * r.maxR + 1;
    const long long total = (long long)t * t * (r.maxR + 1);
    CIM_ALLOC(ctx->out, sizeof(float) * 3 * total);
    float* d = ctx->out.as<float>();
    if (add) {
        CIM_HIP(hipMemcpyAsync(d, fa, sizeof(float) * total, hipMemcpyHostToDevice, ctx->stream));
        CIM_HIP(hipMemcpyAsync(d + total, fb, sizeof(float) * total, hipMemcpyHostToDevice, ctx->stream));
        CIM_HIP(hipMemcpyAsync(d + 2 * total, wg, sizeof(float) * total, hipMemcpyHostToDevice, ctx->stream));
    }
    hipLaunchKernelGGL(k_tnf_io, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, r.acc.as<float4>(), d,
                       d + total, d + 2 * total, r.maxR, r.dxy, add);
    CIM_HIP(hipGetLastError());
    if (!add) {
        CIM_HIP(hipMemcpyAsync(fa, d, sizeof(float) * total, hipMemcpyDeviceToHost, ctx->stream));
        CIM_HIP(hipMemcpyAsync(fb, d + total, sizeof(float) * total, hipMemcpyDeviceToHost, ctx->stream));
        CIM_HIP(hipMemcpyAsync(wg, d + 2 * total, sizeof(float) * total, hipMemcpyDeviceToHost, ctx->stream));
    }
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

int rec_finish_dev(cim_ctx* ctx, float* map_dev) {
    RecState& r = ctx->rec;
    if (!r.active) return set_error(ctx, CIM_E_STATE, "cim_rec_begin has not been called");
    const long long nvox = (long long)r.box * r.box * r.box;
    CIM_ALLOC(ctx->ft2, sizeof(float2) * nvox);
    float2* vol = ctx->ft2.as<float2>();
    {
        TimedLaunch tl(ctx, "fold");
        hipLaunchKernelGGL(k_fold, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, ctx->stream,
                           r.acc.as<float4>(), vol, r.box, r.maxR, r.dxy);
        CIM_HIP(hipGetLastError());
    }
    int rc = fft3d_dev(ctx, vol, r.box, -1, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_checker_real, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, ctx->stream, vol, map_dev,
                       r.box);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

}  // namespace cim
