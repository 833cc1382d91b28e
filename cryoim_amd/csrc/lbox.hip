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
// shifts each image with two kernels, rows then columns, each running the
// forward transform, the phase ramp and the backward transform of its axis
// in LDS (one read and one write of the image per axis), then one
// workgroup per (particle, operator) cuts and normalises its box from the
// shifted image, which stays in L2 while the particle's overlapping boxes
// are cut.  The whole-image normalisation of loc_box.f90:52 is an affine
// map of the image, which the per-box normalisation removes again, so it
// is not run.
#include <algorithm>
#include <mutex>
#include <vector>

#include "cim_device.h"
#include "cim_fft.h"

namespace cim {

// Working memory of cim_local_box_stack_noalloc, one copy per GPU for the
// life of the process: the FFT twiddles, the operators, and per pass of
// particles their angles and shifts, records, complex rows and shifted
// images.  Host code uses the address hipGetSymbolAddress returns for the
// current device (lbox_scratch_setup), never the array's name.
constexpr size_t LBOX_SCRATCH_BYTES = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[LBOX_SCRATCH_BYTES];
constexpr int LBOX_SCRATCH_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[LBOX_SCRATCH_DEVICES] = {};

// Called by cim_init with ctx->device current: resolves this device's copy of
// the scratch array once.
int lbox_scratch_setup(cim_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= LBOX_SCRATCH_DEVICES)
        return set_error(ctx, CIM_E_ARG, "device %d: no scratch slot", ctx->device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[ctx->device]) return CIM_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(ctx, CIM_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = p;
    return CIM_OK;
}

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

// img2d_centshift.f90:13-34 along one axis for L lines (one workgroup):
// the checkerboard factor (-1)^e, forward transform (scaled 1/n), the phase
// ramp exp(-i 2 pi s (k - n/2) / n) with s = -x (rows) or -y (columns) of
// the line's particle, backward transform, (-1)^e again.  The reference's
// 2D operator is the product of these row and column operators (the
// checkerboard, both transforms and the ramp all factor by axis), so rows
// then columns give its shifted image up to rounding.
//   COLS = 0: rows of the real images img -> complex rows of z;
//   COLS = 1: columns of z -> the real part, into out.
// Line l is row (or column) l % n of particle l / n.
template <int COLS>
__global__ void __launch_bounds__(256) k_lbox_shift(const float* __restrict__ img, float2* __restrict__ z,
                                                    float* __restrict__ out, const float2* __restrict__ tw, int n,
                                                    FacList fl, long long nlines, int L,
                                                    const float* __restrict__ pr, float twopi) {
#pragma clang fp contract(off)
    extern __shared__ float2 lds[];
    __shared__ long long lbase[16];   // first element of each line (COLS: column c of particle p)
    __shared__ float lsh[16];         // -x (rows) or -y (columns) of each line's particle
    float2* b0 = lds;
    float2* b1 = lds + (size_t)L * n;
    const long long l0 = (long long)blockIdx.x * L;
    const int nl = (int)((nlines - l0) < L ? (nlines - l0) : L);
    const long long n2 = (long long)n * n;
    if (threadIdx.x < nl) {
        const long long l = l0 + threadIdx.x, p = l / n;
        lbase[threadIdx.x] = COLS ? p * n2 + (l - p * n) : l * n;
        lsh[threadIdx.x] = -pr[5 * p + (COLS ? 4 : 3)];
    }
    __syncthreads();
    const int lgL = __ffs(L) - 1;
    if (!COLS) {
        for (int idx = threadIdx.x; idx < nl * n; idx += blockDim.x) {
            const int li = (int)fdiv((unsigned)idx, (unsigned)n, fl.n_m), e = idx - li * n;
            const float v = img[lbase[li] + e];
            b0[idx] = make_float2((e & 1) ? -v : v, 0.f);
        }
    } else {
        // L neighbouring columns: each row of the tile is one coalesced segment
        for (int idx = threadIdx.x; idx < L * n; idx += blockDim.x) {
            const int e = idx >> lgL, li = idx & (L - 1);
            if (li >= nl) continue;
            const float2 v = z[lbase[li] + (long long)e * n];
            b0[li * n + e] = (e & 1) ? make_float2(-v.x, -v.y) : v;
        }
    }
    for (int idx = nl * n + threadIdx.x; idx < L * n; idx += blockDim.x) b0[idx] = make_float2(0.f, 0.f);
    __syncthreads();
    float2* f = lds_fft(b0, b1, n, L, fl, tw, -1.f);
    const float inv = 1.0f / (float)n;
    for (int idx = threadIdx.x; idx < nl * n; idx += blockDim.x) {
        const int li = (int)fdiv((unsigned)idx, (unsigned)n, fl.n_m), k = idx - li * n;
        const float phase = twopi * (lsh[li] * (float)(k - n / 2)) / (float)n;
        float sn, cs;
        sincosf(phase, &sn, &cs);
        const float2 v = make_float2(f[idx].x * inv, f[idx].y * inv);
        f[idx] = make_float2(v.x * cs + v.y * sn, -v.x * sn + v.y * cs);
    }
    __syncthreads();
    const float2* r = lds_fft(f, f == b0 ? b1 : b0, n, L, fl, tw, 1.f);
    if (!COLS) {
        for (int idx = threadIdx.x; idx < nl * n; idx += blockDim.x) {
            const int li = (int)fdiv((unsigned)idx, (unsigned)n, fl.n_m), e = idx - li * n;
            const float2 v = r[idx];
            z[lbase[li] + e] = (e & 1) ? make_float2(-v.x, -v.y) : v;
        }
    } else {
        for (int idx = threadIdx.x; idx < L * n; idx += blockDim.x) {
            const int e = idx >> lgL, li = idx & (L - 1);
            if (li >= nl) continue;
            const float v = r[li * n + e].x;
            out[lbase[li] + (long long)e * n] = (e & 1) ? -v : v;
        }
    }
}

// loc_box.f90:64-86 for one (particle, operator): the sub x sub box of the
// shifted image whose corner is (hx0, ky0) + box/2 - sub/2 (one periodic
// wrap per axis, then a modulo that only matters for corners the host has
// already refused), normalised: mean, then the rms about it, then the
// normalised values (three passes; the box stays in L2).
__global__ void __launch_bounds__(256) k_lbox_cut(const float* __restrict__ x, const float* __restrict__ rec,
                                                  int box, int sub, int nops, float* __restrict__ out) {
    __shared__ ff32 sh[4];
    const int t = blockIdx.x;
    const int i = t / nops;
    const float* o = rec + (long long)CIM_LBOX_REC * t;
    const int hx0 = (int)o[3] + box / 2 - sub / 2, ky0 = (int)o[4] + box / 2 - sub / 2;
    const float* img = x + (long long)i * box * box;
    float* dst = out + (long long)t * sub * sub;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6, ny = blockDim.x >> 6;
    auto wrap = [&](int v) {
        if (v < 0) v += box;
        if (v >= box) v -= box;
        return ((v % box) + box) % box;
    };
    float s = 0.f;
    for (int r = ty; r < sub; r += ny) {
        const float* row = img + (long long)wrap(ky0 + r) * box;
        for (int c = tx; c < sub; c += 64) s += row[wrap(hx0 + c)];
    }
    const ff32 sa = ff_mul(block_sum_ff(ff32{s, 0.f}, sh), ff_recip_int(sub * sub));
    const float avg = __fadd_rn(sa.h, sa.l);
    float d = 0.f;
    for (int r = ty; r < sub; r += ny) {
        const float* row = img + (long long)wrap(ky0 + r) * box;
        for (int c = tx; c < sub; c += 64) {
            const float v = row[wrap(hx0 + c)] - avg;
            d += v * v;
        }
    }
    const float dev = sqrtf(ff_mul(block_sum_ff(ff32{d, 0.f}, sh), ff_recip_int(sub * sub)).h);
    for (int r = ty; r < sub; r += ny) {
        const float* row = img + (long long)wrap(ky0 + r) * box;
        for (int c = tx; c < sub; c += 64) dst[(long long)r * sub + c] = (row[wrap(hx0 + c)] - avg) / dev;
    }
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
    const int chunk = (int)std::max(1LL, std::min((long long)n, (512LL << 20) / (12 * b2)));
    if (n > 0) {
        CIM_ALLOC(ctx->lbox.x, sizeof(float2) * b2 * chunk);
        CIM_ALLOC(ctx->lbox.r, sizeof(float) * b2 * chunk);
    }
    float2* z = ctx->lbox.x.as<float2>();
    float* rimg = ctx->lbox.r.as<float>();
    const float twopi = 2.0f * (4.0f * atan2f(1.0f, 1.0f));
    const FacList fl = fac_list(pl);
    // lines per workgroup as the batched transforms choose them
    const int Lr = box >= 1024 ? 1 : (box >= 512 ? 2 : (box >= 256 ? 4 : (box >= 128 ? 8 : 16)));
    const int Lc = box <= 128 ? 16 : (box <= 512 ? 8 : 4);
    for (int i0 = 0; i0 < n; i0 += chunk) {
        const int nb = std::min(chunk, n - i0);
        const long long nlines = (long long)nb * box;
        const float* prc = st.a.pr + 5LL * i0;
        {
            TimedLaunch tl(ctx, "lbox_rows");
            hipLaunchKernelGGL(k_lbox_shift<0>, dim3((unsigned)((nlines + Lr - 1) / Lr)), dim3(256),
                               sizeof(float2) * 2 * Lr * box, ctx->stream, images_dev + b2 * i0, z, nullptr, pl->tw,
                               box, fl, nlines, Lr, prc, twopi);
            CIM_HIP(hipGetLastError());
        }
        {
            TimedLaunch tl(ctx, "lbox_cols");
            hipLaunchKernelGGL(k_lbox_shift<1>, dim3((unsigned)((nlines + Lc - 1) / Lc)), dim3(256),
                               sizeof(float2) * 2 * Lc * box, ctx->stream, nullptr, z, rimg, pl->tw, box, fl, nlines,
                               Lc, prc, twopi);
            CIM_HIP(hipGetLastError());
        }
        {
            TimedLaunch tl(ctx, "lbox_cut");
            hipLaunchKernelGGL(k_lbox_cut, dim3((unsigned)(nb * nops)), dim3(256), 0, ctx->stream, rimg,
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

// The stack path with every working buffer in this device's scratch array:
// particles in passes that fit it, each pass uploading its angles and
// shifts, computing its records and cutting its boxes.  No device memory is
// allocated.  Calls that overlap on one GPU share the array and race.
int lbox_stack_noalloc(cim_ctx* ctx, int n, const float* images_dev, const float* parts, int box, int nops,
                       const double* ops, const float* x0, int sub, float apix, float* boxes_dev, float* rec_host) {
    int rc = lbox_check(ctx, n, parts, nops, ops, x0, box);
    if (rc) return rc;
    if (sub < 2 || sub > box) return set_error(ctx, CIM_E_ARG, "sub_FFTsize %d outside [2, box %d]", sub, box);
    const double r0 = std::sqrt((double)x0[0] * x0[0] + (double)x0[1] * x0[1] + (double)x0[2] * x0[2]);
    if (r0 > (double)box - 2.0)
        return set_error(ctx, CIM_E_ARG, "|(local_x, local_y, local_z)| = %g reaches past the %d-pixel image", r0, box);
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < LBOX_SCRATCH_DEVICES)
            base = static_cast<unsigned char*>(g_scratch_ptr[ctx->device]);
    }
    if (!base) return set_error(ctx, CIM_E_STATE, "scratch of device %d not set up", ctx->device);
    FftPlan pl;
    std::vector<float2> tw;
    if ((rc = fft_plan_host(ctx, box, pl, tw))) return rc;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const long long b2 = (long long)box * box;
    const size_t fixed = al(sizeof(float2) * box) + al(sizeof(float) * 9 * (size_t)nops);
    // per particle: angles and shifts, records, complex rows, shifted image (+ alignment slack)
    const size_t per = sizeof(float) * (5 + (size_t)CIM_LBOX_REC * nops) + (sizeof(float2) + sizeof(float)) * b2;
    if (fixed + per + 4 * 256 > LBOX_SCRATCH_BYTES)
        return set_error(ctx, CIM_E_ARG,
                         "box %d with %d operators needs %zu bytes per particle: more than the %zu-byte scratch", box,
                         nops, fixed + per + 4 * 256, LBOX_SCRATCH_BYTES);
    const int chunk = (int)std::max<long long>(1, std::min<long long>(std::max(n, 1),
                                                   (long long)((LBOX_SCRATCH_BYTES - fixed - 4 * 256) / per)));
    unsigned char* q = base;
    float2* d_tw = reinterpret_cast<float2*>(q);
    q += al(sizeof(float2) * box);
    float* d_ops = reinterpret_cast<float*>(q);
    q += al(sizeof(float) * 9 * (size_t)nops);
    float* d_pr = reinterpret_cast<float*>(q);
    q += al(sizeof(float) * 5 * (size_t)chunk);
    float* d_rec = reinterpret_cast<float*>(q);
    q += al(sizeof(float) * CIM_LBOX_REC * (size_t)chunk * nops);
    float2* z = reinterpret_cast<float2*>(q);
    q += al(sizeof(float2) * b2 * chunk);
    float* rimg = reinterpret_cast<float*>(q);
    q += al(sizeof(float) * b2 * chunk);
    if ((size_t)(q - base) > LBOX_SCRATCH_BYTES)
        return set_error(ctx, CIM_E_ARG, "scratch layout of %zu bytes exceeds %zu", (size_t)(q - base),
                         LBOX_SCRATCH_BYTES);
    LboxSetup st;
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
    CIM_HIP(hipMemcpyAsync(d_tw, tw.data(), sizeof(float2) * box, hipMemcpyHostToDevice, ctx->stream));
    CIM_HIP(hipMemcpyAsync(d_ops, st.ops.data(), sizeof(float) * st.ops.size(), hipMemcpyHostToDevice, ctx->stream));
    pl.tw = d_tw;
    const FacList fl = fac_list(&pl);
    const float twopi = 2.0f * (4.0f * atan2f(1.0f, 1.0f));
    const int Lr = box >= 1024 ? 1 : (box >= 512 ? 2 : (box >= 256 ? 4 : (box >= 128 ? 8 : 16)));
    const int Lc = box <= 128 ? 16 : (box <= 512 ? 8 : 4);
    for (int i0 = 0; i0 < n; i0 += chunk) {
        const int nb = std::min(chunk, n - i0);
        CIM_HIP(hipMemcpyAsync(d_pr, st.pr.data() + 5 * (size_t)i0, sizeof(float) * 5 * nb, hipMemcpyHostToDevice,
                               ctx->stream));
        LboxGeomArgs a{};
        a.pr = d_pr;
        a.ops = d_ops;
        a.rec = d_rec;
        a.n = nb;
        a.nops = nops;
        a.box = box;
        a.stack = 1;
        a.x0 = x0[0];
        a.y0 = x0[1];
        a.z0 = x0[2];
        a.apix = apix;
        a.d2r = 4.0f * atan2f(1.0f, 1.0f) / 180.0f;
        a.r2d = 45.0f / atan2f(1.0f, 1.0f);
        if ((rc = lbox_geom(ctx, a))) return rc;
        const long long nlines = (long long)nb * box;
        hipLaunchKernelGGL(k_lbox_shift<0>, dim3((unsigned)((nlines + Lr - 1) / Lr)), dim3(256),
                           sizeof(float2) * 2 * Lr * box, ctx->stream, images_dev + b2 * i0, z, nullptr, d_tw, box, fl,
                           nlines, Lr, d_pr, twopi);
        CIM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_lbox_shift<1>, dim3((unsigned)((nlines + Lc - 1) / Lc)), dim3(256),
                           sizeof(float2) * 2 * Lc * box, ctx->stream, nullptr, z, rimg, d_tw, box, fl, nlines, Lc,
                           d_pr, twopi);
        CIM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_lbox_cut, dim3((unsigned)(nb * nops)), dim3(256), 0, ctx->stream, rimg, d_rec, box, sub,
                           nops, boxes_dev + (long long)sub * sub * nops * i0);
        CIM_HIP(hipGetLastError());
        if (rec_host)
            CIM_HIP(hipMemcpyAsync(rec_host + (size_t)CIM_LBOX_REC * nops * i0, d_rec,
                                   sizeof(float) * CIM_LBOX_REC * (size_t)nb * nops, hipMemcpyDeviceToHost,
                                   ctx->stream));
        // the next pass rewrites the angles, records and images of this one
        CIM_HIP(hipStreamSynchronize(ctx->stream));
    }
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

int cim_local_box_stack_noalloc(cim_ctx* ctx, int n, const void* images_dev, const float* parts, int box, int nops,
                                const double* ops, const float* x0, int sub, float apix, void* boxes_dev, float* rec) {
    if (!ctx) return CIM_E_ARG;
    if (n > 0 && (!images_dev || !boxes_dev)) return set_error(ctx, CIM_E_ARG, "NULL images or boxes");
    CIM_HIP(hipSetDevice(ctx->device));
    return lbox_stack_noalloc(ctx, n, static_cast<const float*>(images_dev), parts, box, nops, ops, x0, sub, apix,
                              static_cast<float*>(boxes_dev), rec);
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
