// local_box: sub-particle coordinates and boxes around one point of the
// particle (local_x, local_y, local_z) for every symmetry operator.
//
// cim_local_box_coords -- loc_box_coordinate.f90:16-41: per particle and
//   operator, the orientation of S R(theta, phi, omega) (equalort.f90), the
//   centre R_k^T x0 + box/2 + (x, y) of the sub-particle in the image split
//   into integer part and fraction, and its height (z - box/2) apix.
// cim_local_box_stack -- loc_box.f90:50-91: the particle image shifted by
//   -(x, y) through its transform (img2d_centshift.f90), then per operator
//   the sub x sub box whose corner is int(R_k^T x0) + box/2 - sub/2
//   (periodic), normalised (img2d_norm.f90).
//
// One thread per (particle, operator) evaluates the geometry in float32 in
// the reference's order of operations (no contraction).  The stack path
// runs load+checkerboard, a batched forward 2D FFT, the phase ramp, the
// backward FFT, then one workgroup per (particle, operator) cuts and
// normalises its box from the shifted image, which stays in L2 while the 60
// boxes that overlap it are cut.  The whole-image normalisation of
// loc_box.f90:52 is an affine map of the image, which the per-box
// normalisation removes again, so it is not run.
#include <algorithm>
#include <vector>

#include "cim_device.h"

namespace cim {

namespace {

struct LboxGeomArgs {
    const float* pr;     // [n][5] theta phi omega x y
    const float* ops;    // [nops][9] row-major
    float* rec;          // [n*nops][CIM_LBOX_REC]
    int n, nops, box, stack;
    float x0, y0, z0, apix;
    float d2r, r2d;      // Eularmatrix.f90:5 and equalort.f90:8, evaluated in float32 on the host
};

// Eularmatrix.f90:1-15 (row-major, r[3*i+j] = r(i+1, j+1)).
__device__ __forceinline__ void eularmatrix(float th0, float ph0, float om0, float d2r, float* r) {
#pragma clang fp contract(off)
    const float th = th0 * d2r, ph = ph0 * d2r, om = om0 * d2r;
    const float cp = cosf(ph), ct = cosf(th), co = cosf(om), sp = sinf(ph), st = sinf(th), so = sinf(om);
    r[0] = cp * ct * co - sp * so;  r[1] = -cp * ct * so - sp * co;  r[2] = cp * st;
    r[3] = sp * ct * co + cp * so;  r[4] = -sp * ct * so + cp * co;  r[5] = sp * st;
    r[6] = -st * co;                r[7] = st * so;                  r[8] = ct;
}

__global__ void __launch_bounds__(256) k_lbox_geom(LboxGeomArgs a) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n * a.nops) return;
    const int i = t / a.nops, k = t % a.nops;
    const float* P = a.pr + 5 * i;
    const float* S = a.ops + 9 * k;
    float E[9], r[9], R[9];
    eularmatrix(P[0], P[1], P[2], a.d2r, E);
    // equalort.f90:10-21: r = S E, then its Euler angles
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) r[3 * p + q] = S[3 * p] * E[q] + S[3 * p + 1] * E[3 + q] + S[3 * p + 2] * E[6 + q];
    float th;
    if (r[8] >= 1.0f) th = 0.f;
    else if (r[8] < -1.0f) th = 180.0f;
    else th = acosf(r[8]) * a.r2d;
    const float ph = atan2f(r[5], r[2]) * a.r2d;
    const float om = atan2f(r[7], -r[6]) * a.r2d;
    eularmatrix(th, ph, om, a.d2r, R);
    // matmul(transpose(rot), x0)
    float c[3];
    for (int q = 0; q < 3; ++q) c[q] = R[q] * a.x0 + R[3 + q] * a.y0 + R[6 + q] * a.z0;
    float* o = a.rec + (long long)CIM_LBOX_REC * t;
    const float h = (float)(a.box / 2);
    if (a.stack) {
        // loc_box.f90:62-68
        const int hx0 = (int)c[0], ky0 = (int)c[1];
        o[3] = (float)hx0;
        o[4] = (float)ky0;
        o[5] = c[0] - (float)hx0;
        o[6] = c[1] - (float)ky0;
        o[7] = c[2] * a.apix;
    } else {
        // loc_box_coordinate.f90:25-33
        for (int q = 0; q < 3; ++q) c[q] = c[q] + h;
        c[0] = c[0] + P[3];
        c[1] = c[1] + P[4];
        const int hx0 = (int)c[0], ky0 = (int)c[1];
        o[3] = (float)hx0;
        o[4] = (float)ky0;
        o[5] = c[0] - (float)hx0;
        o[6] = c[1] - (float)ky0;
        o[7] = (c[2] - h) * a.apix;
    }
    o[0] = th;
    o[1] = ph;
    o[2] = om;
}

// img2d_centshift.f90:13-16: checkerboarded image as complex data.
__global__ void k_lbox_load(const float* __restrict__ img, float2* __restrict__ x, int box, long long total) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int q = (int)(t % ((long long)box * box));
    const int s = ((q / box) + (q % box)) & 1;
    const float v = img[t];
    x[t] = make_float2(s ? -v : v, 0.f);
}

// img2d_centshift.f90:21-29 on the (Re, +Im) transform: multiply by
// exp(-i phase), phase = 2 pi (x0 hx + y0 ky) / box with (x0, y0) = -(x, y).
__global__ void k_lbox_phase(float2* __restrict__ x, int box, long long total, const float* __restrict__ pr,
                             float twopi) {
#pragma clang fp contract(off)
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const long long b2 = (long long)box * box;
    const int i = (int)(t / b2), q = (int)(t % b2);
    const int ky = q / box - box / 2, hx = q % box - box / 2;
    const float x0 = -pr[5 * i + 3], y0 = -pr[5 * i + 4];
    const float phase = twopi * (x0 * (float)hx + y0 * (float)ky) / (float)box;
    float s, c;
    sincosf(phase, &s, &c);
    const float2 v = x[t];
    x[t] = make_float2(v.x * c + v.y * s, -v.x * s + v.y * c);
}

// loc_box.f90:64-86 for one (particle, operator): the sub x sub box of the
// real, checkerboard-corrected shifted image whose corner is
// (hx0, ky0) + box/2 - sub/2 (one periodic wrap per axis, then a modulo that
// only matters for corners the host has already refused), normalised.
__global__ void __launch_bounds__(256) k_lbox_cut(const float2* __restrict__ x, const float* __restrict__ rec,
                                                  int box, int sub, int nops, float* __restrict__ out) {
    __shared__ double sh[4];
    const int t = blockIdx.x;
    const int i = t / nops;
    const float* o = rec + (long long)CIM_LBOX_REC * t;
    const int hx0 = (int)o[3] + box / 2 - sub / 2, ky0 = (int)o[4] + box / 2 - sub / 2;
    const float2* img = x + (long long)i * box * box;
    float* dst = out + (long long)t * sub * sub;
    const int m = sub * sub;
    auto value = [&](int q) -> float {
        int ky = q / sub + ky0, hx = q % sub + hx0;
        if (hx < 0) hx += box;
        if (hx >= box) hx -= box;
        if (ky < 0) ky += box;
        if (ky >= box) ky -= box;
        hx = ((hx % box) + box) % box;
        ky = ((ky % box) + box) % box;
        const float v = img[(long long)ky * box + hx].x;
        return ((hx + ky) & 1) ? -v : v;
    };
    float s = 0.f;
    for (int q = threadIdx.x; q < m; q += blockDim.x) s += value(q);
    const float avg = (float)(block_sum((double)s, sh) / (double)m);
    float d = 0.f;
    for (int q = threadIdx.x; q < m; q += blockDim.x) {
        const float v = value(q) - avg;
        d += v * v;
    }
    const float dev = (float)sqrt(block_sum((double)d, sh) / (double)m);
    for (int q = threadIdx.x; q < m; q += blockDim.x) dst[q] = (value(q) - avg) / dev;
}

struct LboxSetup {
    std::vector<float> pr, ops;
    LboxGeomArgs a{};
};

int lbox_check(cim_ctx* ctx, int n, const float* parts, int nops, const double* ops, const float* x0, int box) {
    if (n < 0) return set_error(ctx, CIM_E_ARG, "negative particle count");
    if (n > 0 && !parts) return set_error(ctx, CIM_E_ARG, "NULL parts");
    if (!ops || !x0) return set_error(ctx, CIM_E_ARG, "NULL ops or x0");
    if (nops < 1 || nops > 4096) return set_error(ctx, CIM_E_ARG, "nops %d outside [1, 4096]", nops);
    if (box < 2 || box > 1 << 20) return set_error(ctx, CIM_E_ARG, "box %d unsupported", box);
    return CIM_OK;
}

// Upload the particles' angles and shifts and the operators, and fill the
// geometry arguments.
int lbox_setup(cim_ctx* ctx, int n, const float* parts, int nops, const double* ops, const float* x0, int box,
               int stack, float apix, LboxSetup& st) {
    st.pr.resize(5 * (size_t)std::max(n, 1));
    for (int i = 0; i < n; ++i) {
        const float* r = parts + (size_t)CIM_PART_STRIDE * i;
        st.pr[5 * i + 0] = r[CIM_P_THETA];
        st.pr[5 * i + 1] = r[CIM_P_PHI];
        st.pr[5 * i + 2] = r[CIM_P_OMEGA];
        st.pr[5 * i + 3] = r[CIM_P_X];
        st.pr[5 * i + 4] = r[CIM_P_Y];
    }
    st.ops.assign(ops, ops + 9 * (size_t)nops);
    CIM_ALLOC(ctx->lbox.pr, sizeof(float) * st.pr.size());
    CIM_ALLOC(ctx->lbox.ops, sizeof(float) * st.ops.size());
    CIM_ALLOC(ctx->lbox.rec, sizeof(float) * CIM_LBOX_REC * (size_t)std::max(n, 1) * nops);
    CIM_HIP(hipMemcpyAsync(ctx->lbox.pr.ptr, st.pr.data(), sizeof(float) * st.pr.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    CIM_HIP(hipMemcpyAsync(ctx->lbox.ops.ptr, st.ops.data(), sizeof(float) * st.ops.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    LboxGeomArgs& a = st.a;
    a.pr = ctx->lbox.pr.as<float>();
    a.ops = ctx->lbox.ops.as<float>();
    a.rec = ctx->lbox.rec.as<float>();
    a.n = n;
    a.nops = nops;
    a.box = box;
    a.stack = stack;
    a.x0 = x0[0];
    a.y0 = x0[1];
    a.z0 = x0[2];
    a.apix = apix;
    a.d2r = 4.0f * atan2f(1.0f, 1.0f) / 180.0f;
    a.r2d = 45.0f / atan2f(1.0f, 1.0f);
    return CIM_OK;
}

int lbox_geom(cim_ctx* ctx, const LboxGeomArgs& a) {
    const long long T = (long long)a.n * a.nops;
    if (T == 0) return CIM_OK;
    TimedLaunch tl(ctx, "lbox_geom");
    hipLaunchKernelGGL(k_lbox_geom, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, ctx->stream, a);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

// The stack path for n images already on the device; boxes_dev receives
// n*nops*sub^2 floats and ctx->lbox.rec the records.  Particles are taken
// in chunks whose complex working images fit chunk_bytes.
int lbox_stack_dev(cim_ctx* ctx, int n, const float* images_dev, const float* parts, int box, int nops,
                   const double* ops, const float* x0, int sub, float apix, float* boxes_dev, float* rec_host) {
    int rc = lbox_check(ctx, n, parts, nops, ops, x0, box);
    if (rc) return rc;
    if (sub < 2 || sub > box) return set_error(ctx, CIM_E_ARG, "sub_FFTsize %d outside [2, box %d]", sub, box);
    const double r0 = std::sqrt((double)x0[0] * x0[0] + (double)x0[1] * x0[1] + (double)x0[2] * x0[2]);
    if (r0 > (double)box - 2.0)
        return set_error(ctx, CIM_E_ARG, "|(local_x, local_y, local_z)| = %g reaches past the %d-pixel image", r0, box);
    FftPlan* pl = nullptr;
    rc = fft_plan(ctx, box, &pl);
    if (rc) return rc;
    LboxSetup st;
    rc = lbox_setup(ctx, n, parts, nops, ops, x0, box, 1, apix, st);
    if (rc) return rc;
    rc = lbox_geom(ctx, st.a);
    if (rc) return rc;
    const long long b2 = (long long)box * box;
    const int chunk = (int)std::max(1LL, std::min((long long)n, (512LL << 20) / (8 * b2)));
    if (n > 0) CIM_ALLOC(ctx->lbox.x, sizeof(float2) * b2 * chunk);
    float2* x = ctx->lbox.x.as<float2>();
    const float twopi = 2.0f * (4.0f * atan2f(1.0f, 1.0f));
    for (int i0 = 0; i0 < n; i0 += chunk) {
        const int nb = std::min(chunk, n - i0);
        const long long total = b2 * nb;
        hipLaunchKernelGGL(k_lbox_load, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream,
                           images_dev + b2 * i0, x, box, total);
        CIM_HIP(hipGetLastError());
        rc = fft2d_dev(ctx, x, box, nb, 1, 0);
        if (rc) return rc;
        {
            TimedLaunch tl(ctx, "lbox_phase");
            hipLaunchKernelGGL(k_lbox_phase, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, x,
                               box, total, st.a.pr + 5LL * i0, twopi);
            CIM_HIP(hipGetLastError());
        }
        rc = fft2d_dev(ctx, x, box, nb, -1, 0);
        if (rc) return rc;
        {
            TimedLaunch tl(ctx, "lbox_cut");
            hipLaunchKernelGGL(k_lbox_cut, dim3((unsigned)(nb * nops)), dim3(256), 0, ctx->stream, x,
                               st.a.rec + (long long)CIM_LBOX_REC * nops * i0, box, sub, nops,
                               boxes_dev + (long long)sub * sub * nops * i0);
            CIM_HIP(hipGetLastError());
        }
    }
    if (rec_host && n > 0)
        CIM_HIP(hipMemcpyAsync(rec_host, st.a.rec, sizeof(float) * CIM_LBOX_REC * (size_t)n * nops,
                               hipMemcpyDeviceToHost, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

}  // namespace

}  // namespace cim

using namespace cim;

extern "C" {

int cim_local_box_coords(cim_ctx* ctx, int n, const float* parts, int nops, const double* ops, const float* x0,
                         int box, float apix, float* rec) {
    if (!ctx) return CIM_E_ARG;
    int rc = lbox_check(ctx, n, parts, nops, ops, x0, box);
    if (rc) return rc;
    if (n > 0 && !rec) return set_error(ctx, CIM_E_ARG, "NULL rec");
    CIM_HIP(hipSetDevice(ctx->device));
    LboxSetup st;
    rc = lbox_setup(ctx, n, parts, nops, ops, x0, box, 0, apix, st);
    if (rc) return rc;
    rc = lbox_geom(ctx, st.a);
    if (rc) return rc;
    if (n > 0)
        CIM_HIP(hipMemcpyAsync(rec, st.a.rec, sizeof(float) * CIM_LBOX_REC * (size_t)n * nops, hipMemcpyDeviceToHost,
                               ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

int cim_local_box_stack(cim_ctx* ctx, int n, const float* images, const float* parts, int box, int nops,
                        const double* ops, const float* x0, int sub, float apix, float* boxes, float* rec) {
    if (!ctx) return CIM_E_ARG;
    if (n > 0 && (!images || !boxes)) return set_error(ctx, CIM_E_ARG, "NULL images or boxes");
    CIM_HIP(hipSetDevice(ctx->device));
    const long long b2 = (long long)box * box, s2 = (long long)sub * sub;
    // host batches bounded to about 1 GiB of device images + boxes
    const long long per = 4 * (b2 + s2 * std::max(nops, 1));
    const int batch = (int)std::max(1LL, std::min((long long)std::max(n, 1), (1LL << 30) / std::max(per, 1LL)));
    if (n == 0) return lbox_stack_dev(ctx, 0, nullptr, parts, box, nops, ops, x0, sub, apix, nullptr, rec);
    CIM_ALLOC(ctx->img, sizeof(float) * b2 * batch);
    CIM_ALLOC(ctx->out, sizeof(float) * s2 * nops * batch);
    for (int i0 = 0; i0 < n; i0 += batch) {
        const int nb = std::min(batch, n - i0);
        CIM_HIP(hipMemcpyAsync(ctx->img.ptr, images + b2 * i0, sizeof(float) * b2 * nb, hipMemcpyHostToDevice,
                               ctx->stream));
        int rc = lbox_stack_dev(ctx, nb, ctx->img.as<float>(), parts + (size_t)CIM_PART_STRIDE * i0, box, nops, ops,
                                x0, sub, apix, ctx->out.as<float>(),
                                rec ? rec + (size_t)CIM_LBOX_REC * nops * i0 : nullptr);
        if (rc) return rc;
        CIM_HIP(hipMemcpyAsync(boxes + s2 * nops * i0, ctx->out.ptr, sizeof(float) * s2 * nops * nb,
                               hipMemcpyDeviceToHost, ctx->stream));
    }
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

int cim_local_box_stack_dev(cim_ctx* ctx, int n, const void* images_dev, const float* parts, int box, int nops,
                            const double* ops, const float* x0, int sub, float apix, void* boxes_dev, float* rec) {
    if (!ctx) return CIM_E_ARG;
    if (n > 0 && (!images_dev || !boxes_dev)) return set_error(ctx, CIM_E_ARG, "NULL images or boxes");
    CIM_HIP(hipSetDevice(ctx->device));
    return lbox_stack_dev(ctx, n, static_cast<const float*>(images_dev), parts, box, nops, ops, x0, sub, apix,
                          static_cast<float*>(boxes_dev), rec);
}

}  // extern "C"
