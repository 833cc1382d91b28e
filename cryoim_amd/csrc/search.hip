// icosprocess common-lines phase-residual search on gfx950.
//
// Global search (ort_globesearch.f90:34-134 with cal_crosslin_PR_omega.f90):
// one workgroup per (AU grid point, particle).  Each thread owns one (or
// VPT) in-plane angle omega, so the phase-residual sums of
// cal_crosslin_PR_omega.f90:48-69 accumulate in registers.  The particle
// transform inside radius maxR+2 is staged once per workgroup in LDS as a
// disk (rows of varying length), so every bilinear tap of
// img2d_fill_lin_omegearray.f90:18-36 is an LDS read; the template lines of
// img2d_fill_line.f90 are sampled cooperatively into LDS per chunk of
// common lines.  The argmin over (omega, phi, theta) keeps Fortran minloc
// semantics (first minimum in column-major order) by reducing
// (value, linear index) pairs.
//
// Local refinement (ort_refinement.f90:69-95): one workgroup per particle
// runs the whole coordinate descent; the 11 candidates of a cycle are scored
// together by cal_crosslin_PR (cal_crosslin_PR.f90:22-68) with the trial
// centre shift applied to the taps.
#include <cmath>
#include <map>
#include <type_traits>

#include "cim_device.h"

namespace cim {

int upload_parts(cim_ctx* ctx, int n, const float* parts);
int check_params_public(cim_ctx* ctx, const cim_params* p);

// ---- geometry shared by both kernels (float32, as the reference) ---------

// icos_crosslin_angle.f90:11-34 for one operator S (row-major float[9]).
__device__ __forceinline__ void crosslin_one(const float* E1, const float* E2, const float* S, float* a1, float* a2) {
    const float r2d = 57.2957795130823208768f;
    float z1[3] = {E1[2], E1[5], E1[8]};
    float z2[3] = {E2[2], E2[5], E2[8]};
    float z2a[3];
    for (int i = 0; i < 3; ++i) z2a[i] = S[3 * i] * z2[0] + S[3 * i + 1] * z2[1] + S[3 * i + 2] * z2[2];
    float v[3] = {z1[1] * z2a[2] - z1[2] * z2a[1], z1[2] * z2a[0] - z1[0] * z2a[2], z1[0] * z2a[1] - z1[1] * z2a[0]};
    float vv = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (vv < 0.5f) {
        *a1 = 1000.f;
        *a2 = 1000.f;
        return;
    }
    v[0] /= vv; v[1] /= vv; v[2] /= vv;
    float v10x = E1[0] * v[0] + E1[3] * v[1] + E1[6] * v[2];
    float v10y = E1[1] * v[0] + E1[4] * v[1] + E1[7] * v[2];
    float t[3];
    for (int i = 0; i < 3; ++i) t[i] = S[i] * v[0] + S[3 + i] * v[1] + S[6 + i] * v[2];
    float v20x = E2[0] * t[0] + E2[3] * t[1] + E2[6] * t[2];
    float v20y = E2[1] * t[0] + E2[4] * t[1] + E2[7] * t[2];
    *a1 = atan2f(v10y, v10x) * r2d;
    *a2 = atan2f(v20y, v20x) * r2d;
}

// atan(y/x) for y >= 0, any x, in [0, pi]: octant reduction to t in [0, 1]
// and t * P(t^2), P of degree 7 fitted to atan(t)/t on [0, 1] (weighted
// least squares, max abs error 2e-7 rad evaluated in float32).  With
// y = |Im(z1 conj z2)| and x = Re(z1 conj z2) this is the phase difference
// of two samples folded into [0, pi], which cal_crosslin_PR_omega.f90:57-60
// computes from two atan2 calls.
__device__ __forceinline__ float atan2_upper(float y, float x) {
#pragma clang fp contract(on)
    const float ax = fabsf(x);
    const float mn = fminf(ax, y), mx = fmaxf(ax, y);
    const float t = mx > 0.f ? mn * __builtin_amdgcn_rcpf(mx) : 0.f;
    const float s = t * t;
    float q = -0.004780458752065897f;
    q = fmaf(q, s, 0.02455713413655758f);
    q = fmaf(q, s, -0.05990472808480263f);
    q = fmaf(q, s, 0.09942759573459625f);
    q = fmaf(q, s, -0.14029419422149658f);
    q = fmaf(q, s, 0.19971375167369843f);
    q = fmaf(q, s, -0.3333209455013275f);
    q = fmaf(q, s, 0.9999999403953552f);
    float r = t * q;
    r = y > ax ? 1.57079632679489661923f - r : r;
    return x < 0.f ? 3.14159265358979323846f - r : r;
}

__device__ __forceinline__ float fold_phase(float d) {
    const float pi = 3.14159265358979323846f, tpi = 6.28318530717958647692f;
    d = fabsf(d);
    if (d >= tpi) d -= tpi;
    if (d > pi) d = tpi - d;
    return d;
}

// Bilinear sample (img2d_fill_line.f90:17-31) of a crop with rows of Dc
// elements whose index 0 is relative coordinate -(R+1).
__device__ __forceinline__ float2 crop_bilin(const float2* __restrict__ crop, int Dc, int R, float x, float y,
                                             int ic) {
#pragma clang fp contract(on)
    int lx = (int)x, ly = (int)y;
    float ax = x - (float)lx, ay = y - (float)ly;
    int cx = lx - ic + R + 1, cy = ly - ic + R + 1;
    const float2* r0 = crop + (long long)cy * Dc + cx;
    const float2* r1 = r0 + Dc;
    float2 a = r0[0], b = r0[1], c = r1[0], d = r1[1];
    float w00 = (1.f - ax) * (1.f - ay), w10 = ax * (1.f - ay), w01 = (1.f - ax) * ay, w11 = ax * ay;
    return make_float2(w00 * a.x + w10 * b.x + w01 * c.x + w11 * d.x, w00 * a.y + w10 * b.y + w01 * c.y + w11 * d.y);
}

// ---- crops ----------------------------------------------------------------

// Crop rel [-(R+1), R+2] around the centre of n centred FTs, optionally
// (shift_parts) centre-shifted by the phase of img2d_FT_centshift of record
// item, as k_centshift would (the same shift_phase, so the same bits) but on
// the crop's pixels only.  Without use_ctf, workgroup row y is item y and
// reads src[y].  With use_ctf (template x CTF), row y is particle record y:
// its CTF is evaluated once per pixel and multiplies the T template sources
// src[0 .. T-1] into items y*T + t.
__global__ void k_crop(const float2* __restrict__ src, long long src_stride, int box, int R, int Dc,
                       float2* __restrict__ dst, int T, const float* __restrict__ ctf_parts, CtfConst cc, int use_ctf,
                       const float* __restrict__ shift_parts) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Dc * Dc) return;
    const int cy = q / Dc, cx = q % Dc;
    const int c = box / 2;
    const int hx = cx - R - 1, ky = cy - R - 1;
    const int ax = hx + c, ay = ky + c;
    const bool inside = ax >= 0 && ax < box && ay >= 0 && ay < box;
    const long long si = (long long)ay * box + ax;
    auto shifted = [&](float2 v, int item) {
        if (!shift_parts) return v;
        const float* pr = shift_parts + (long long)item * CIM_PART_STRIDE;
        return apply_shift(v, shift_phase(hx, ky, box, pr[CIM_P_X], pr[CIM_P_Y], 1));
    };
    if (!use_ctf) {
        const int item = blockIdx.y;
        float2 v = inside ? src[(long long)item * src_stride + si] : make_float2(0.f, 0.f);
        dst[(long long)item * Dc * Dc + q] = shifted(v, item);
        return;
    }
    const int p = blockIdx.y;
    const float* pr = ctf_parts + (long long)p * CIM_PART_STRIDE;
    const float cv = ctf_value(hx, ky, box, pr[CIM_P_DF1], pr[CIM_P_DF2], pr[CIM_P_ASTIG], cc);
    for (int t = 0; t < T; ++t) {
        const int item = p * T + t;
        float2 v = inside ? src[(long long)t * src_stride + si] : make_float2(0.f, 0.f);
        v = shifted(v, item);
        v.x *= cv;
        v.y *= cv;
        dst[(long long)item * Dc * Dc + q] = v;
    }
}

// The centre shift of img2d_FT_centshift applied in place to n crops of
// Dc x Dc centred frequencies (crop index i <-> frequency i - R - 1), as
// k_centshift would on the full transform (the same shift_phase).
__global__ void k_crop_shift(float2* __restrict__ crop, int box, int R, int Dc, const float* __restrict__ parts) {
    const int item = blockIdx.y;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Dc * Dc) return;
    const int hx = q % Dc - R - 1, ky = q / Dc - R - 1;
    const float* pr = parts + (long long)item * CIM_PART_STRIDE;
    float2* f = crop + (long long)item * Dc * Dc + q;
    *f = apply_shift(*f, shift_phase(hx, ky, box, pr[CIM_P_X], pr[CIM_P_Y], 1));
}

// ---- global search kernel -------------------------------------------------

struct SearchArgs {
    const float2* pcrop;      // n x Dc^2 particle crops (centre-shifted)
    const float4* pcrop4;     // the same, tap-paired
    const float* fcut0;       // n x nr
    const float2* tcrop;      // template crops; particle p, template t at tcrop + p*tc_pstride + t*Dc^2
    long long tc_pstride;
    const float* tfcut;       // Fcut of template t for particle p at tfcut + p*tf_pstride + t*nr
    long long tf_pstride;
    const float* tang;        // T x 3 template angles
    int T;
    const float* pts;         // npts x 2 (theta, phi)
    const int* pbase;         // npts: linear index of (omega=0, phi, theta) in PR_array
    int npts;
    const float* ops;         // 60 x 9 float
    int nops;
    const int* rows;          // nrows x 3: lo, hi, base (disk layout)
    int nrows, disk_total;
    int R, minR, nr, Dc, Nomega, box, chunk;
    float step;
    float* outv;              // n x npts
    int* outi;
    float* prfull;            // optional n x npts x Nomega
    unsigned long long* nsamples;   // optional: evaluated (line, radius, omega) samples
    unsigned long long* npass;      // optional (k_cl_rescore): samples whose particle amplitude passed its gate
    int nbatch;                     // k_cl_rescore: particles of the launch
    // optional (k_cl_search): score the particles fb_list[s] for s < *fb_count, grid.y at a time,
    // results of list entry s at outv/outi[s * gridDim.x + point]
    const int* fb_list;
    const int* fb_count;
    const int* cand;          // optional n x ncand AU point indices: score only these (grid.x = ncand)
    int ncand;
    const int* kwin;          // optional n x ncand first column of a window of wincols columns (mod ncol)
    int wincols;
    int wsplit;               // workgroups per window (grid.x = ncand * wsplit), each wincols / wsplit columns
};

// Direction (cos, sin) of omega column k on a common line at in-plane angle a1 (deg).
__device__ __forceinline__ void column_dir(float a1, float step, int k, float& cs, float& sn) {
#pragma clang fp contract(on)
    const float d2r = 0.0174532925199432957692f;
    const float om = (a1 - step * (float)k) * d2r;
    cs = cosf(om);
    sn = sinf(om);
}

// Bilinear sample position of column (cs, sn) at radius FR about the centre cf:
// lower tap (lx, ly) and fractions.
__device__ __forceinline__ void sample_pos(float FR, float cs, float sn, float cf, int& lx, int& ly, float& ax,
                                           float& ay) {
#pragma clang fp contract(on)
    const float x = FR * cs + cf, y = FR * sn + cf;
    lx = (int)x;
    ly = (int)y;
    // x, y > 0 (the crop's centre lies more than the radius inside): x - (int)x
    // = x - floor(x), exact, which v_fract_f32 returns in one instruction
    ax = __builtin_amdgcn_fractf(x);
    ay = __builtin_amdgcn_fractf(y);
}

// One sample of a column from its four particle taps against the template
// line record tv = {re, im, |F| (-1 gated), particle Fcut f0}: amplitude gate,
// folded phase residual and weight added to the column's sums (and, PAIR, to
// those of the Friedel column with the conjugate sample).  Every kernel that
// scores columns calls this, so a column's bits do not depend on the kernel.
// npass counts the samples that pass the particle gate.
template <bool PAIR>
__device__ __forceinline__ void score_sample(float2 q00, float2 q10, float2 q01, float2 q11, float ax, float ay,
                                             float4 tv, float f0, float& num, float& den, float& num2, float& den2,
                                             unsigned int& npass) {
#pragma clang fp contract(on)
    const float w00 = (1.f - ax) * (1.f - ay), w10 = ax * (1.f - ay);
    const float w01 = (1.f - ax) * ay, w11 = ax * ay;
    const float zx = w00 * q00.x + w10 * q10.x + w01 * q01.x + w11 * q11.x;
    const float zy = w00 * q00.y + w10 * q10.y + w01 * q01.y + w11 * q11.y;
    const float amp1 = __builtin_amdgcn_sqrtf(zx * zx + zy * zy);
    if (amp1 >= f0) {
        ++npass;
        const float w = 0.5f * (amp1 + tv.z);
        const float pxy = zx * tv.y, pyx = zy * tv.x, pxx = zx * tv.x, pyy = zy * tv.y;
        num += atan2_upper(fabsf(pxy - pyx), pxx + pyy) * w;
        den += w;
        if (PAIR) {
            // conjugate sample of column k + Nomega/2
            num2 += atan2_upper(fabsf(pxy + pyx), pxx - pyy) * w;
            den2 += w;
        }
    }
}

// Template line record of line j, radius index fr: the template x CTF crop
// sampled along the common line at alpha2 (deg), gated by its Fcut.
__device__ __forceinline__ float4 template_line(const float2* __restrict__ tc, const float* __restrict__ tf, int Dc,
                                                int R, int minR, int fr, float alpha2, float cf, int ic, float fc0) {
#pragma clang fp contract(on)
    const float d2r = 0.0174532925199432957692f;
    float4 res = make_float4(0.f, 0.f, -1.f, fc0);
    const float FR = (float)(minR + fr);
    const float ang = alpha2 * d2r;
    float x = FR * cosf(ang) + cf, y = FR * sinf(ang) + cf;
    float2 z = crop_bilin(tc, Dc, R, x, y, ic);
    float amp2 = sqrtf(z.x * z.x + z.y * z.y);
    if (amp2 >= tf[fr]) res = make_float4(z.x, z.y, amp2, fc0);
    return res;
}

// Particle taps: MODE 0 reads the float2 disk staged in LDS, MODE 1 reads
// the tap-paired crop (float4 {F(x,y), F(x+1,y)}) through the caches, two
// 16-byte loads per bilinear sample.
//
// PAIR: when the omega grid closes on itself (Nomega even and
// step*Nomega/2 = 180 deg), column k and column k + Nomega/2 sample the
// particle transform at mirrored points about the origin.  The transform of
// a real image (also after the linear-phase centre shift) is Hermitian, so
// the second sample is the conjugate of the first: one interpolation feeds
// both columns.
template <int VPT, int MODE, bool PAIR>
__device__ __forceinline__ void cl_search_one(const SearchArgs& a, char* sm, const int p, const long long oslot) {
    float2* disk = reinterpret_cast<float2*>(sm);
    size_t off = MODE == 0 ? sizeof(float2) * (size_t)a.disk_total : 0;
    // template line record per (line, radius): re, im, |F| (-1 if gated out), particle Fcut
    float4* tline = reinterpret_cast<float4*>(sm + off);
    off += sizeof(float4) * (size_t)a.chunk * a.nr;
    float* alpha1 = reinterpret_cast<float*>(sm + off);
    float* alpha2 = alpha1 + 64;
    float* fc0 = alpha2 + 64;
    int* rbase = reinterpret_cast<int*>(fc0 + a.nr);
    float* redv = reinterpret_cast<float*>(rbase + a.nrows);
    int* redi = reinterpret_cast<int*>(redv + 16);

    const int split = a.kwin ? a.wsplit : 1;
    const int slot = blockIdx.x / split, part = blockIdx.x - slot * split;
    const int pt = a.cand ? a.cand[(long long)p * a.ncand + slot] : slot;
    const int tid = threadIdx.x, bd = blockDim.x;
    const float2* pc = a.pcrop + (long long)p * a.Dc * a.Dc;
    const float4* pc4 = a.pcrop4 + (long long)p * a.Dc * a.Dc;
    const int ic = a.box / 2;
    const float cf = (float)ic;
    const float d2r = 0.0174532925199432957692f;
    const int ncol = PAIR ? a.Nomega / 2 : a.Nomega;
    // columns this workgroup scores: all, or a window starting at kbase (wrapping); the
    // arithmetic of a column does not depend on which workgroup or lane computes it
    const int nact = a.kwin ? a.wincols / split : ncol;
    int kbase = a.kwin ? a.kwin[(long long)p * a.ncand + slot] + part * nact : 0;
    if (kbase >= ncol) kbase -= ncol;

    for (int i = tid; i < a.nr; i += bd) fc0[i] = a.fcut0[(long long)p * a.nr + i];
    if (MODE == 0) {
        for (int r = tid; r < a.nrows; r += bd) rbase[r] = a.rows[3 * r + 2];
        for (int r = 0; r < a.nrows; ++r) {
            const int lo = a.rows[3 * r], hi = a.rows[3 * r + 1], base = a.rows[3 * r + 2];
            for (int x = lo + tid; x <= hi; x += bd) disk[base + x] = pc[(long long)r * a.Dc + (x + a.R + 1)];
        }
    }
    const float theta = a.pts[2 * pt], phi = a.pts[2 * pt + 1];
    float E1[9];
    euler_f(theta, phi, 0.f, E1);

    float num[VPT], den[VPT], num2[VPT], den2[VPT];
#pragma unroll
    for (int v = 0; v < VPT; ++v) { num[v] = 0.f; den[v] = 0.f; num2[v] = 0.f; den2[v] = 0.f; }
    unsigned long long pairs = 0;   // (line, radius) pairs that pass the template gate
    unsigned int npass_unused = 0;
    __syncthreads();

    for (int t = 0; t < a.T; ++t) {
        const float2* tc = a.tcrop + (long long)p * a.tc_pstride + (long long)t * a.Dc * a.Dc;
        const float* tf = a.tfcut + (long long)p * a.tf_pstride + (long long)t * a.nr;
        for (int j = tid; j < a.nops; j += bd) {
            float E2[9];
            euler_f(a.tang[3 * t], a.tang[3 * t + 1], a.tang[3 * t + 2], E2);
            crosslin_one(E1, E2, a.ops + 9 * j, &alpha1[j], &alpha2[j]);
        }
        __syncthreads();
        for (int j0 = 0; j0 < a.nops; j0 += a.chunk) {
            const int jn = min(a.chunk, a.nops - j0);
            for (int q = tid; q < jn * a.nr; q += bd) {
                const int jj = q / a.nr, fr = q - jj * a.nr;
                const int j = j0 + jj;
                tline[q] = (alpha1[j] < 999.f && alpha2[j] < 999.f)
                               ? template_line(tc, tf, a.Dc, a.R, a.minR, fr, alpha2[j], cf, ic, fc0[fr])
                               : make_float4(0.f, 0.f, -1.f, fc0[fr]);
            }
            __syncthreads();
            for (int jj = 0; jj < jn; ++jj) {
                const int j = j0 + jj;
                const float a1 = alpha1[j];
                if (!(a1 < 999.f && alpha2[j] < 999.f)) continue;
                float cs[VPT], sn[VPT];
#pragma unroll
                for (int v = 0; v < VPT; ++v) {
                    int k = kbase + tid + v * bd;
                    if (k >= ncol) k -= ncol;
                    column_dir(a1, a.step, k, cs[v], sn[v]);
                }
                const float4* tl = tline + jj * a.nr;
                for (int fr = 0; fr < a.nr; ++fr) {
                    const float4 tv = tl[fr];
                    if (tv.z < 0.f) continue;
                    ++pairs;
                    const float FR = (float)(a.minR + fr);
                    const float f0 = tv.w;
#pragma unroll
                    for (int v = 0; v < VPT; ++v) {
                        if (tid + v * bd >= nact) continue;
                        int lx, ly;
                        float ax, ay;
                        sample_pos(FR, cs[v], sn[v], cf, lx, ly, ax, ay);
                        float2 q00, q10, q01, q11;
                        if (MODE == 0) {
                            const int r = ly - ic + a.R + 1;
                            const int i0 = rbase[r] + (lx - ic), i1 = rbase[r + 1] + (lx - ic);
                            q00 = disk[i0]; q10 = disk[i0 + 1]; q01 = disk[i1]; q11 = disk[i1 + 1];
                        } else {
                            const float4* r0 = pc4 + (long long)(ly - ic + a.R + 1) * a.Dc + (lx - ic + a.R + 1);
                            const float4 u = r0[0], w4 = r0[a.Dc];
                            q00 = make_float2(u.x, u.y); q10 = make_float2(u.z, u.w);
                            q01 = make_float2(w4.x, w4.y); q11 = make_float2(w4.z, w4.w);
                        }
                        score_sample<PAIR>(q00, q10, q01, q11, ax, ay, tv, f0, num[v], den[v], num2[v], den2[v],
                                           npass_unused);
                    }
                }
            }
            __syncthreads();
        }
    }

    // PR per omega, then the first minimum of this (theta, phi) column
    const float r2d = 57.2957795130823208768f;
    float best = 3.0e38f;
    int bidx = 0x7fffffff;
#pragma unroll
    for (int v = 0; v < VPT; ++v) {
        int k = kbase + tid + v * bd;
        if (k >= ncol) k -= ncol;
        if (tid + v * bd < nact) {
#pragma unroll
            for (int h = 0; h < (PAIR ? 2 : 1); ++h) {
                const float nn = h ? num2[v] : num[v], dd = h ? den2[v] : den[v];
                const int kk = k + h * ncol;
                const float pr = dd > 1.0e-10f ? nn / dd * r2d : 999.f;
                if (a.prfull && !a.cand) a.prfull[((long long)p * a.npts + pt) * a.Nomega + kk] = pr;
                const int li = a.pbase[pt] + kk;
                if (pr < best || (pr == best && li < bidx)) { best = pr; bidx = li; }
            }
        }
    }
    // wave reduction (over the lanes the workgroup has: bd is a multiple of 64 or a
    // power of two below it)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        if (o >= bd) continue;
        const float ov = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bidx, o, 64);
        if (ov < best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
    }
    const int w = tid >> 6;
    if ((tid & 63) == 0) { redv[w] = best; redi[w] = bidx; }
    __syncthreads();
    if (tid == 0) {
        const int nw = (bd + 63) >> 6;
        for (int i = 1; i < nw; ++i)
            if (redv[i] < best || (redv[i] == best && redi[i] < bidx)) { best = redv[i]; bidx = redi[i]; }
        a.outv[oslot] = best;
        a.outi[oslot] = bidx;
        if (a.nsamples) atomicAdd(a.nsamples, pairs * (unsigned long long)a.Nomega);
    }
}

template <int VPT, int MODE, bool PAIR>
__global__ void __launch_bounds__(1024) k_cl_search(SearchArgs a) {
    extern __shared__ float4 smem4[];
    char* sm = reinterpret_cast<char*>(smem4);
    if (!a.fb_list) {
        cl_search_one<VPT, MODE, PAIR>(a, sm, blockIdx.y, (long long)blockIdx.y * gridDim.x + blockIdx.x);
        return;
    }
    // listed particles, grid.y of them at a time (the count is known on the device only)
    const int cnt = *a.fb_count;
    for (int s = blockIdx.y; s < cnt; s += gridDim.y) {
        if (s != (int)blockIdx.y) __syncthreads();
        cl_search_one<VPT, MODE, PAIR>(a, sm, a.fb_list[s], (long long)s * gridDim.x + blockIdx.x);
    }
}

// The exhaustive fallback decision of the tabulated search, per particle:
// the rank of its exact minimum among its candidates (counted in rank[0..4]:
// < 4, < 8, < 16, < 24, more) and, with windows, whether it lies within 2
// columns of its window's edge.  A particle whose minimum ranks fb_rank or
// deeper, or lies at an edge, is appended to list (count in *cnt, order
// free: each listed particle's result goes back to its own index).
//
// The candidates are ncand per particle: the topk AU points of the table
// ranking, then (ncand > topk) second windows -- AU points whose table score
// has another minimum away from the first, re-scored around it.  The exact
// minimum belongs to the first candidate of its AU point whose window holds
// its column; a minimum found in a second window is counted in rank[5] and
// sends the particle to the fallback only from its window's edge.
__global__ void k_fb_select(const int* __restrict__ fi, const int* __restrict__ cand, const int* __restrict__ kwin,
                            const int* __restrict__ pbase, int nb, int topk, int ncand, int nomega, int ncol,
                            int wincols, int fb_rank, unsigned long long* __restrict__ rank, int* __restrict__ list,
                            int* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int li = fi[i];
    const int base = li - li % nomega;
    const int kb = (li % nomega) % ncol;
    int r = ncand, off = 0;
    for (int c = 0; c < ncand; ++c) {
        if (pbase[cand[(long long)i * ncand + c]] != base) continue;
        if (kwin) {
            const int o = ((kb - kwin[(long long)i * ncand + c]) % ncol + ncol) % ncol;
            if (o >= wincols) continue;
            off = o;
        }
        r = c;
        break;
    }
    atomicAdd(&rank[r < 4 ? 0 : r < 8 ? 1 : r < 16 ? 2 : r < 24 ? 3 : r < topk ? 4 : r < ncand ? 5 : 4], 1ull);
    const bool edge = kwin && r < ncand && (off < 2 || off > wincols - 3);
    if ((r < topk && r >= fb_rank) || r >= ncand || edge) list[atomicAdd(cnt, 1)] = i;
}

// Each listed particle's exhaustive minimum (first minimum by value, then
// linear index, over its npts points) replaces its tabulated result.
__global__ void k_fb_merge(const float* __restrict__ v, const int* __restrict__ ix, int npts,
                           const int* __restrict__ list, const int* __restrict__ cnt, float* __restrict__ ov,
                           int* __restrict__ oi) {
    __shared__ float sv[4];
    __shared__ int si[4];
    const int s = blockIdx.x;
    if (s >= *cnt) return;
    float best = 3.0e38f;
    int bidx = 0x7fffffff;
    for (int i = threadIdx.x; i < npts; i += blockDim.x) {
        const float x = v[(long long)s * npts + i];
        const int k = ix[(long long)s * npts + i];
        if (x < best || (x == best && k < bidx)) { best = x; bidx = k; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float xv = __shfl_xor(best, o, 64);
        const int xi = __shfl_xor(bidx, o, 64);
        if (xv < best || (xv == best && xi < bidx)) { best = xv; bidx = xi; }
    }
    if ((threadIdx.x & 63) == 0) { sv[threadIdx.x >> 6] = best; si[threadIdx.x >> 6] = bidx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i)
            if (sv[i] < best || (sv[i] == best && si[i] < bidx)) { best = sv[i]; bidx = si[i]; }
        ov[list[s]] = best;
        oi[list[s]] = bidx;
    }
}

// Device records in: n records copied into the search's working copy, each
// masked around its own centre (ort_globesearch.f90:50).
__global__ void k_records_in(const float* __restrict__ src, int n, float* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * CIM_PART_STRIDE) return;
    const int k = i % CIM_PART_STRIDE;
    const int b = i - k;
    dst[i] = k == CIM_P_MASKX ? src[b + CIM_P_X] : k == CIM_P_MASKY ? src[b + CIM_P_Y] : src[i];
}

// Device records out: the search's result (first minimum's linear index and
// PR) as theta, phi, omega and PR of each record.
__global__ void k_records_out(const float* __restrict__ fv, const int* __restrict__ fi, int n, int plane, int nomega,
                              int np, float step, float* __restrict__ parts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int idx = fi[i];
    const int it = idx / plane, ip = (idx % plane) / nomega - np, io = idx % nomega;
    float* q = parts + (long long)i * CIM_PART_STRIDE;
    q[CIM_P_THETA] = 69.0f + step * (float)it;
    q[CIM_P_PHI] = step * (float)ip;
    q[CIM_P_OMEGA] = step * (float)io;
    q[CIM_P_PR] = fv[i];
}

// Fill n 64-bit words (counters) with zero: a kernel, so that the work stays
// capturable in a graph.
__global__ void k_zero_u64(unsigned long long* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Exact re-score of the tabulated search's candidates, CPW candidates per
// wave: lane l scores column kwin[c] + l % W of candidate c = l / W (and its
// Friedel column), W = 64 / CPW columns around the omega of the candidate's
// table minimum.  Each column takes the same terms in the same order through
// the same arithmetic as in k_cl_search (template_line, sample_pos,
// score_sample), so its bits equal the plain search's; the lanes of the
// candidates differ only in which lines and radii they skip.  The template
// line records that pass the template gate are compacted per (candidate,
// line) in radius order, so each lane walks only its own candidate's
// records.  Particle taps are read tap-paired through the caches; the
// workgroups of one particle all go to one XCD (workgroup L runs on XCD
// L mod 8: L mod 8 picks the particle within a group of 8, (L / 8) mod G the
// candidate group), so its crops stay in that XCD's L2.  The first minimum of
// each candidate's columns goes to outv/outi[p * ncand + candidate].
template <int CPW, bool PAIR>
__global__ void __launch_bounds__(64) k_cl_rescore(SearchArgs a) {
    constexpr int W = 64 / CPW;
    extern __shared__ float4 smem4[];
    float4* tline = smem4;                                           // CPW x chunk x nr, compacted
    float* alpha1 = reinterpret_cast<float*>(tline + (size_t)CPW * a.chunk * a.nr);   // CPW x 64
    float* alpha2 = alpha1 + CPW * 64;
    float* fc0 = alpha2 + CPW * 64;                                   // nr
    int* tcnt = reinterpret_cast<int*>(fc0 + a.nr);                  // CPW x chunk: records kept
    unsigned short* tfr = reinterpret_cast<unsigned short*>(tcnt + CPW * a.chunk);   // their radius indices
    const int lane = threadIdx.x, c = lane / W, col = lane - c * W;
    const int G = (a.ncand + CPW - 1) / CPW;
    const long long L = blockIdx.x;
    const int p = (int)(L & 7) + 8 * (int)(L / (8LL * G)), g0 = (int)((L >> 3) % G) * CPW;
    if (p >= a.nbatch) return;
    const bool act = g0 + c < a.ncand;
    const int slot = act ? g0 + c : g0;
    const long long cbase = (long long)p * a.ncand;
    const int pt = a.cand[cbase + slot];
    const float2* tcp = a.tcrop + (long long)p * a.tc_pstride;
    const float* tfp = a.tfcut + (long long)p * a.tf_pstride;
    const float4* pc4 = a.pcrop4 + (long long)p * a.Dc * a.Dc;
    const int ic = a.box / 2;
    const float cf = (float)ic;
    const float d2r = 0.0174532925199432957692f;
    const int c0 = a.R + 1 - ic;                    // crop index of centred frequency 0, minus ic
    const unsigned rowb = (unsigned)a.Dc << 4;      // bytes per crop row
    const int ncol = PAIR ? a.Nomega / 2 : a.Nomega;
    int k = a.kwin[cbase + slot] + col;
    if (k >= ncol) k -= ncol;
    for (int i = lane; i < a.nr; i += 64) fc0[i] = a.fcut0[(long long)p * a.nr + i];
    float num = 0.f, den = 0.f, num2 = 0.f, den2 = 0.f;
    unsigned long long pairs = 0;
    unsigned int npass = 0;
    for (int t = 0; t < a.T; ++t) {
        const float2* tc = tcp + (long long)t * a.Dc * a.Dc;
        const float* tf = tfp + (long long)t * a.nr;
        __syncthreads();
        for (int q = lane; q < CPW * a.nops; q += 64) {
            const int cc = q / a.nops, j = q - cc * a.nops;
            const int ptc = a.cand[cbase + min(g0 + cc, a.ncand - 1)];
            float E1[9], E2[9];
            euler_f(a.pts[2 * ptc], a.pts[2 * ptc + 1], 0.f, E1);
            euler_f(a.tang[3 * t], a.tang[3 * t + 1], a.tang[3 * t + 2], E2);
            crosslin_one(E1, E2, a.ops + 9 * j, &alpha1[cc * 64 + j], &alpha2[cc * 64 + j]);
        }
        __syncthreads();
        for (int j0 = 0; j0 < a.nops; j0 += a.chunk) {
            const int jn = min(a.chunk, a.nops - j0);
            // one (candidate, line) row at a time: the lanes take radii, the records that
            // pass the template gate are packed in radius order
            for (int row = 0; row < CPW * jn; ++row) {
                const int cc = row / jn, jj = row - cc * jn;
                const int j = j0 + jj;
                const float al1 = alpha1[cc * 64 + j], al2 = alpha2[cc * 64 + j];
                const bool valid = al1 < 999.f && al2 < 999.f;
                float4* dst = tline + (cc * a.chunk + jj) * a.nr;
                unsigned short* dfr = tfr + (cc * a.chunk + jj) * a.nr;
                int kept = 0;
                for (int f0 = 0; f0 < a.nr; f0 += 64) {
                    const int fr = f0 + lane;
                    float4 tv = make_float4(0.f, 0.f, -1.f, 0.f);
                    if (valid && fr < a.nr) tv = template_line(tc, tf, a.Dc, a.R, a.minR, fr, al2, cf, ic, fc0[fr]);
                    const bool ok = tv.z >= 0.f;
                    const unsigned long long m = __ballot(ok);
                    if (ok) {
                        const int at = kept + __popcll(m & ((1ull << lane) - 1ull));
                        dst[at] = tv;
                        dfr[at] = (unsigned short)fr;
                    }
                    kept += __popcll(m);
                }
                if (lane == 0) tcnt[cc * a.chunk + jj] = kept;
            }
            __syncthreads();
            for (int jj = 0; jj < jn; ++jj) {
                const int j = j0 + jj;
                const float a1 = alpha1[c * 64 + j];
                if (!(a1 < 999.f && alpha2[c * 64 + j] < 999.f)) continue;
                float cs, sn;
                column_dir(a1, a.step, k, cs, sn);
                const float4* tl = tline + (c * a.chunk + jj) * a.nr;
                const unsigned short* tr = tfr + (c * a.chunk + jj) * a.nr;
                const int nk = tcnt[c * a.chunk + jj];
                pairs += nk;
                if (nk == 0) continue;
                // software pipeline over pairs of records: the taps of one record are requested
                // before the other is scored.  An odd count ends with a repeat of the last record
                // whose gate is +inf, so it adds nothing.
                int lx, ly;
                float axA, ayA, axB, ayB;
                // taps at a 32-bit byte offset from the particle's crop (under 2^32 bytes), so the loads
                // take a scalar base and a vector offset
                auto taps = [&](int tx, int ty, float4& u, float4& w) {
                    const unsigned off = (__umul24((unsigned)(ty + c0), (unsigned)a.Dc) + (unsigned)(tx + c0)) << 4;
                    const char* b = reinterpret_cast<const char*>(pc4);
                    u = *reinterpret_cast<const float4*>(b + off);
                    w = *reinterpret_cast<const float4*>(b + (off + rowb));
                };
                sample_pos((float)(a.minR + (int)tr[0]), cs, sn, cf, lx, ly, axA, ayA);
                float4 uA, wA;
                taps(lx, ly, uA, wA);
                for (int i = 0; i < nk; i += 2) {
                    const int ib = min(i + 1, nk - 1);
                    sample_pos((float)(a.minR + (int)tr[ib]), cs, sn, cf, lx, ly, axB, ayB);
                    float4 uB, wB;
                    taps(lx, ly, uB, wB);
                    const float4 tva = tl[i];
                    score_sample<PAIR>(make_float2(uA.x, uA.y), make_float2(uA.z, uA.w), make_float2(wA.x, wA.y),
                                       make_float2(wA.z, wA.w), axA, ayA, tva, tva.w, num, den, num2, den2, npass);
                    sample_pos((float)(a.minR + (int)tr[min(i + 2, nk - 1)]), cs, sn, cf, lx, ly, axA, ayA);
                    taps(lx, ly, uA, wA);
                    const float4 tvb = tl[ib];
                    score_sample<PAIR>(make_float2(uB.x, uB.y), make_float2(uB.z, uB.w), make_float2(wB.x, wB.y),
                                       make_float2(wB.z, wB.w), axB, ayB, tvb, i + 1 < nk ? tvb.w : INFINITY, num,
                                       den, num2, den2, npass);
                }
            }
            __syncthreads();
        }
    }
    const float r2d = 57.2957795130823208768f;
    float best = 3.0e38f;
    int bidx = 0x7fffffff;
#pragma unroll
    for (int h = 0; h < (PAIR ? 2 : 1); ++h) {
        const float nn = h ? num2 : num, dd = h ? den2 : den;
        const float pr = dd > 1.0e-10f ? nn / dd * r2d : 999.f;
        const int li = a.pbase[pt] + k + h * ncol;
        if (pr < best || (pr == best && li < bidx)) { best = pr; bidx = li; }
    }
    // first minimum over the W lanes of each candidate
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) {
        const float ov = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bidx, o, 64);
        if (ov < best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
    }
    if (col == 0 && act) {
        a.outv[cbase + slot] = best;
        a.outi[cbase + slot] = bidx;
    }
    if (a.nsamples) {
        unsigned long long tot = act ? pairs * (PAIR ? 2ull : 1ull) : 0ull;
        unsigned long long tp = act ? npass * (PAIR ? 2ull : 1ull) : 0ull;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            tot += __shfl_xor(tot, o, 64);
            tp += __shfl_xor(tp, o, 64);
        }
        if (lane == 0) {
            atomicAdd(a.nsamples, tot);
            if (a.npass) atomicAdd(a.npass, tp);
        }
    }
}

// Tap-paired copy of the particle crops: {F(x, y), F(x+1, y)} per element.
__global__ void k_pair_taps(const float2* __restrict__ crop, float4* __restrict__ out, int Dc, int n) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)n * Dc * Dc;
    if (q >= total) return;
    const int cx = (int)(q % Dc);
    const float2 v0 = crop[q];
    const float2 v1 = cx + 1 < Dc ? crop[q + 1] : make_float2(0.f, 0.f);
    out[q] = make_float4(v0.x, v0.y, v1.x, v1.y);
}

// First minimum over the AU points of each particle.
__global__ void k_argmin(const float* __restrict__ v, const int* __restrict__ ix, int npts, float* __restrict__ ov,
                         int* __restrict__ oi) {
    __shared__ float sv[16];
    __shared__ int si[16];
    const int p = blockIdx.x;
    float best = 3.0e38f;
    int bidx = 0x7fffffff;
    for (int i = threadIdx.x; i < npts; i += blockDim.x) {
        const float x = v[(long long)p * npts + i];
        const int k = ix[(long long)p * npts + i];
        if (x < best || (x == best && k < bidx)) { best = x; bidx = k; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float xv = __shfl_xor(best, o, 64);
        const int xi = __shfl_xor(bidx, o, 64);
        if (xv < best || (xv == best && xi < bidx)) { best = xv; bidx = xi; }
    }
    if ((threadIdx.x & 63) == 0) { sv[threadIdx.x >> 6] = best; si[threadIdx.x >> 6] = bidx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i)
            if (sv[i] < best || (sv[i] == best && si[i] < bidx)) { best = sv[i]; bidx = si[i]; }
        ov[p] = best;
        oi[p] = bidx;
    }
}

// ---- tabulated search ------------------------------------------------------
//
// The phase residual of cal_crosslin_PR_omega.f90:48-69 sums, over common
// lines and radii, a term that depends only on the particle ray and the
// template ray the common line picks.  For each (particle, template) the
// tables hold that sum over radii for every pair of rays on a polar grid of
// step dpsi = searchstep / sub:
//   N(jT, jP) = sum_r [gates] |fold(phase_T - phase_P)| (|F_T| + |F_P|) / 2
//   D(jT, jP) = sum_r [gates] (|F_T| + |F_P|) / 2
// (gates: |F_T| >= template Fcut, |F_P| >= particle Fcut, as the reference).
// The score of (AU point, omega_k) is then sum N / sum D over the common
// lines, with the template ray round(alpha2 / dpsi) and the particle ray
// round(alpha1 / dpsi) - k * sub.  Rays are sampled exactly as the exact
// kernel samples them; the only approximation is the rounding of the
// common-line angles to the polar grid.

constexpr float TWO_PI_F = 6.28318530717958647692f;
constexpr float GATED = -1.0e30f;

// Polar rays: out[item][fr][j] = {phase, |F|/2 or GATED}, ray j in direction
// ray[j] = {cos, sin}(j dpsi deg), radius minR + fr, bilinear on the crop as
// the exact kernel.  One workgroup per (ring, item); with rmax, it also
// leaves the ring's largest ungated |F|/2 (0 when none) in rmax[item][fr].
__global__ void __launch_bounds__(256) k_polar(const float2* __restrict__ crops, long long crop_stride, int Dc,
                                               int R, int box, int minR, int nr, int npsi,
                                               const float2* __restrict__ ray, const float* __restrict__ gate,
                                               long long gate_stride, float2* __restrict__ out,
                                               float* __restrict__ rmax) {
    __shared__ float red[4];
    const int item = blockIdx.y, fr = blockIdx.x;
    const int ic = box / 2;
    const float cf = (float)ic;
    const float FR = (float)(minR + fr);
    const float g = gate[(long long)item * gate_stride + fr];
    const float2* crop = crops + (long long)item * crop_stride;
    float2* o = out + ((long long)item * nr + fr) * npsi;
    float m = 0.f;
    for (int j = threadIdx.x; j < npsi; j += blockDim.x) {
        const float2 cs = ray[j];
        const float x = FR * cs.x + cf, y = FR * cs.y + cf;
        const float2 z = crop_bilin(crop, Dc, R, x, y, ic);
        const float amp = sqrtf(z.x * z.x + z.y * z.y);
        const float h = amp >= g ? 0.5f * amp : GATED;
        o[j] = make_float2(atan2f(z.y, z.x), h);
        m = fmaxf(m, h);
    }
    if (!rmax) return;
#pragma unroll
    for (int k = 32; k > 0; k >>= 1) m = fmaxf(m, __shfl_xor(m, k, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) rmax[(long long)item * nr + fr] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// Direction {cos, sin} of ray j of the polar grid (angle j dpsi deg).
__global__ void k_rays(int npsi, float dpsi, float2* __restrict__ ray) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npsi) return;
    const float d2r = 0.0174532925199432957692f;
    const float ang = ((float)j * dpsi) * d2r;
    ray[j] = make_float2(cosf(ang), sinf(ang));
}

// bfloat16 of a finite non-negative float, rounded to nearest even.
__device__ __forceinline__ unsigned int bf16_bits(float x) {
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

// Residual tables, one 64 x 64 tile of (jT, jP) per workgroup, 4 x 4 per
// thread; radii staged through LDS in chunks.  table[item][jT][jP] holds N
// (low half) and D (high half) as bfloat16, item = p * T + t: the tables
// only rank AU points (their 8-bit mantissas perturb a score by ~1e-4 of
// itself, far below the effect of rounding the angles to the grid).
// Rays 180 deg apart carry conjugate samples (real images, CTF and the
// centre-shift ramp keep the transforms Hermitian), and turning both rays of
// a pair by 180 deg leaves |phase difference| and the weights unchanged:
// N(jT + npsi/2, jP + npsi/2) = N(jT, jP).  So only jT < npsi/2 is stored.
constexpr int TAB_RC = 16;
__global__ void __launch_bounds__(256) k_cl_table(const float2* __restrict__ ppol, const float2* __restrict__ tpol,
                                                  int tpol_per_particle, int T, int nr, int npsi,
                                                  unsigned int* __restrict__ table) {
    __shared__ float2 sT[TAB_RC][64];
    __shared__ float2 sP[TAB_RC][64];
    const int item = blockIdx.z, p = item / T, t = item - p * T;
    const float2* P = ppol + (long long)p * nr * npsi;
    const float2* Tm = tpol + (long long)(tpol_per_particle ? item : t) * nr * npsi;
    const int jP0 = blockIdx.x * 64, jT0 = blockIdx.y * 64;
    const int nrow = npsi / 2;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    float N[4][4], D[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) { N[i][j] = 0.f; D[i][j] = 0.f; }
    for (int r0 = 0; r0 < nr; r0 += TAB_RC) {
        for (int e = tid; e < TAB_RC * 64; e += 256) {
            const int rr = e >> 6, c = e & 63, r = r0 + rr;
            const int jt = jT0 + c, jp = jP0 + c;
            sT[rr][c] = (r < nr && jt < nrow) ? Tm[(long long)r * npsi + jt] : make_float2(0.f, GATED);
            sP[rr][c] = (r < nr && jp < npsi) ? P[(long long)r * npsi + jp] : make_float2(0.f, GATED);
        }
        __syncthreads();
#pragma unroll 4
        for (int rr = 0; rr < TAB_RC; ++rr) {
            float2 tv[4], pv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) tv[i] = sT[rr][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) pv[j] = sP[rr][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    float d = fabsf(tv[i].x - pv[j].x);
                    d = fminf(d, TWO_PI_F - d);
                    const float w = fmaxf(tv[i].y + pv[j].y, 0.f);
                    N[i][j] = fmaf(d, w, N[i][j]);
                    D[i][j] += w;
                }
        }
        __syncthreads();
    }
    unsigned int* out = table + (long long)item * nrow * npsi;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int jt = jT0 + ty * 4 + i;
        if (jt >= nrow) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int jp = jP0 + tx * 4 + j;
            if (jp < npsi) {
                const unsigned int e = bf16_bits(N[i][j]) | (bf16_bits(D[i][j]) << 16);
                out[(long long)jt * npsi + jp] = e;
            }
        }
    }
}

// The same tables on the matrix cores.  With g = 1 for an ungated sample
// (0 when gated) and a = g |F|/2, the weight of a ray pair at one radius is
// w = gT gP (aT + aP), so D = sum_r (aT gP + gT aP) is a product of a
// 2 nr-column template matrix and a particle matrix.  The folded phase
// difference has the cosine series |d| = pi/2 - (4/pi) sum_{k odd} cos(k d) / k^2
// (d wrapped to [-pi, pi]); cut after NH odd harmonics and with
// cos k(pT - pP) = cos kpT cos kpP + sin kpT sin kpP, N = pi/2 D + sum_r sum_k
// c_k (aT gP + gT aP) cos k(pT - pP), c_k = -4 / (pi k^2), is a product with
// 4 NH columns per radius.  Both run as v_mfma_f32_32x32x16_bf16 (bfloat16
// features, float32 sums).  The truncated series ranks AU points by a
// smoothed residual (NH = 1: by the weighted cosine of the phase
// difference); the exact re-score of the best candidates decides.
// One workgroup per 128 x 128 tile of (jT, jP), 4 waves of 64 x 64, each
// 2 x 2 tiles of 32 x 32; per chunk of TMR radii the 256 threads write the
// features of the tile's 128 template rows and 128 particle rays to LDS.
typedef __bf16 cim_bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 cim_bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 cim_bf16x2 __attribute__((ext_vector_type(2)));
typedef float cim_f32x16 __attribute__((ext_vector_type(16)));

template <int NH>
struct TabMfma {
    static constexpr int RC = NH <= 2 ? 8 : 4;     // radii per LDS chunk
    static constexpr int KH = RC * 4 * NH;         // harmonic features per chunk (a multiple of 32)
    static constexpr int KD = 16;                  // D features per chunk (2 RC, zero-padded to 16)
    static constexpr int PH = KH + 8;              // row pitch (bfloat16): 16 bytes of padding against bank conflicts
    static constexpr int PD = KD + 8;
};

template <int NH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_cl_table_mfma(const float2* __restrict__ ppol, const float2* __restrict__ tpol,
                                                       int tpol_per_particle, int T, int nr, int npsi,
                                                       unsigned int* __restrict__ table) {
    using C = TabMfma<NH>;
    __shared__ __attribute__((aligned(16))) __bf16 sHT[128][C::PH], sHP[128][C::PH];
    __shared__ __attribute__((aligned(16))) __bf16 sDT[128][C::PD], sDP[128][C::PD];
    const int item = blockIdx.z, p = item / T, t = item - p * T;
    const float2* P = ppol + (long long)p * nr * npsi;
    const float2* Tm = tpol + (long long)(tpol_per_particle ? item : t) * nr * npsi;
    const int jP0 = blockIdx.x * 128, jT0 = blockIdx.y * 128;
    const int nrow = npsi / 2;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
    // this thread's template row (tid < 128) or particle ray
    const bool tside = tid < 128;
    const int jj = tside ? jT0 + tid : jP0 + tid - 128;
    const bool live = tside ? jj < nrow : jj < npsi;
    const float2* src = tside ? Tm : P;
    __bf16* hrow = tside ? &sHT[tid][0] : &sHP[tid - 128][0];
    __bf16* drow = tside ? &sDT[tid][0] : &sDP[tid - 128][0];
    cim_f32x16 accH[2][2], accD[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) { accH[i][j][e] = 0.f; accD[i][j][e] = 0.f; }
    // rays of the next chunk are loaded before the current chunk's products, so their
    // latency overlaps the matrix work
    float2 nxt[C::RC];
#pragma unroll
    for (int rr = 0; rr < C::RC; ++rr)
        nxt[rr] = (live && rr < nr) ? src[(long long)rr * npsi + jj] : make_float2(0.f, GATED);
    // waves 0-1 hold template rows, waves 2-3 particle rays: the side is uniform per
    // wave, so each wave runs only its own side's feature code
    const bool tside_w = __builtin_amdgcn_readfirstlane(tid) < 128;
    auto features = [&](auto side) {
        constexpr bool TS = decltype(side)::value;
#pragma unroll
        for (int rr = 0; rr < C::RC; ++rr) {
            const float2 v = nxt[rr];
            const float g = v.y > 0.5f * GATED ? 1.f : 0.f;
            const float a = g * v.y;
            // D: template [a, g], particle [g, a]
            cim_bf16x2 dv;
            dv[0] = (__bf16)(TS ? a : g);
            dv[1] = (__bf16)(TS ? g : a);
            *reinterpret_cast<cim_bf16x2*>(drow + 2 * rr) = dv;
            float c1, s1;
            __sincosf(v.x, &s1, &c1);
            const float c2 = c1 * c1 - s1 * s1, s2 = 2.f * c1 * s1;
            float ck = c1, sk = s1;
#pragma unroll
            for (int kk = 0; kk < NH; ++kk) {
                const int k = 2 * kk + 1;
                // template [a c, a s, g c, g s]; particle c_k [g c, g s, a c, a s]
                // (f x0 and f x1 first, as the products (f x0) c associate; 1 a = a)
                const float f = -1.2732395447351628f / (float)(k * k);
                const float f0 = TS ? a : f * g, f1 = TS ? g : f * a;
                cim_bf16x4 hv;
                hv[0] = (__bf16)(f0 * ck);
                hv[1] = (__bf16)(f0 * sk);
                hv[2] = (__bf16)(f1 * ck);
                hv[3] = (__bf16)(f1 * sk);
                *reinterpret_cast<cim_bf16x4*>(hrow + rr * 4 * NH + 4 * kk) = hv;
                const float cn = ck * c2 - sk * s2, sn = sk * c2 + ck * s2;
                ck = cn;
                sk = sn;
            }
        }
    };
    for (int r0 = 0; r0 < nr; r0 += C::RC) {
        if (tside_w) features(std::true_type{});
        else features(std::false_type{});
#pragma unroll
        for (int e = 2 * C::RC; e < C::KD; e += 2) *reinterpret_cast<cim_bf16x2*>(drow + e) = cim_bf16x2{};
        __syncthreads();
#pragma unroll
        for (int rr = 0; rr < C::RC; ++rr) {
            const int r = r0 + C::RC + rr;
            nxt[rr] = (live && r < nr) ? src[(long long)r * npsi + jj] : make_float2(0.f, GATED);
        }
        const int lr = lane & 31, lk = 8 * (lane >> 5);
#pragma unroll
        for (int kb = 0; kb < C::KH / 16; ++kb) {
            cim_bf16x8 fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                fa[i] = *reinterpret_cast<const cim_bf16x8*>(&sHT[wr + 32 * i + lr][16 * kb + lk]);
                fb[i] = *reinterpret_cast<const cim_bf16x8*>(&sHP[wc + 32 * i + lr][16 * kb + lk]);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    accH[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], accH[i][j], 0, 0, 0);
        }
        {
            cim_bf16x8 fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                fa[i] = *reinterpret_cast<const cim_bf16x8*>(&sDT[wr + 32 * i + lr][lk]);
                fb[i] = *reinterpret_cast<const cim_bf16x8*>(&sDP[wc + 32 * i + lr][lk]);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    accD[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], accD[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // 32 x 32 result tile: register e of lane l is row 8 (e / 4) + 4 (l / 32) + e % 4, column l % 32
    unsigned int* out = table + (long long)item * nrow * npsi;
    const float half_pi = 1.5707963267948966f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int jp = jP0 + wc + 32 * j + (lane & 31);
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int jt = jT0 + wr + 32 * i + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                if (jt < nrow && jp < npsi) {
                    const float D = accD[i][j][e];
                    const float N = fmaf(half_pi, D, accH[i][j][e]);
                    const unsigned int v = bf16_bits(N) | (bf16_bits(D) << 16);
                    out[(long long)jt * npsi + jp] = v;
                }
            }
        }
}

struct GatherArgs {
    // n x T x (nray/2) x nray entries: {N, D} as bfloat16 (unsigned int), or as E4M3
    // (unsigned short) on the coarse grid
    const void* table;
    const unsigned int* lines;   // npts x T x 64: (table row << 16) | zero-omega column (k_cl_lines)
    const int* lcount;           // npts x T: lines of each (AU point, template)
    const int* pbase;
    int T, npts;
    int nray;                 // rays of the table's polar grid
    int sub;                  // table columns per omega column
    int ncol;                 // omega columns scored
    int ostep;                // linear-index omega steps per scored column
    int n;                    // particles
    float* outv;              // n x npts: best table score of the AU point
    int* outi;                // its linear index
    // AU points to score: all npts (plist null), or plist[p * plstride + j] for j < nlist
    // (plstride 0: one list for every particle), j < pcnt[p] when pcnt is given
    const int* plist;
    int plstride, nlist;
    const int* pcnt;
    // optional: with plist, the first omegas of a window per listed point: ncol columns in two
    // halves, from pwin and from pwin2 (mod nomega); null: columns 0 .. ncol-1
    const int* pwin;
    const int* pwin2;
    int nomega;
    // optional: the linear index of each point's best omega at least excl omega steps from its
    // first minimum (circularly, modulo exmod steps; 0: nomega), and with outv2 its score
    int* outi2;
    int excl;
    float* outv2;
    int exmod;
};

// Add the {N, D} bfloat16 pair of a table entry to the sums: N the low half,
// D the high half, each widened exactly to float and added with one rounding.
// (Summing with v_dot2c_f32_bf16 against register-held selectors was measured
// at the same speed: the gather waits on its table reads, not on its adds.)
__device__ __forceinline__ void acc_entry(unsigned int e, float& num, float& den) {
    num += __uint_as_float(e << 16);
    den += __uint_as_float(e & 0xffff0000u);
}

// The same for an E4M3 {N, D} pair of the coarse table (N the low byte).
__device__ __forceinline__ void acc_entry(unsigned short e, float& num, float& den) {
    const auto f = __builtin_amdgcn_cvt_pk_f32_fp8((int)e, false);
    num += f[0];
    den += f[1];
}

__device__ __forceinline__ int wrap_index(int j, int n) {
    j %= n;
    return j < 0 ? j + n : j;
}

// The common lines of every (AU point, template) pair on the tables' polar
// grid, which do not depend on the particle: the valid lines of the 60
// operators compacted in operator order, each as its table row (the
// template ray, folded to the stored half: both rays of a pair turned by
// 180 deg give the same entry) and the column of the particle ray at omega 0.
// One wave per pair.
__device__ __forceinline__ unsigned int line_entry(float a1, float a2, float dpsi, int npsi) {
    const int nrow = npsi >> 1;
    int jt = wrap_index((int)rintf(a2 / dpsi), npsi);
    int j0 = wrap_index((int)rintf(a1 / dpsi), npsi);
    if (jt >= nrow) {
        jt -= nrow;
        j0 += nrow;
        if (j0 >= npsi) j0 -= npsi;
    }
    return ((unsigned int)jt << 16) | (unsigned int)j0;
}

__global__ void __launch_bounds__(64) k_cl_lines(const float* __restrict__ pts, const float* __restrict__ tang,
                                                 const float* __restrict__ ops, int T, int nops, int npsi, float dpsi,
                                                 unsigned int* __restrict__ lines, unsigned int* __restrict__ lines2,
                                                 int* __restrict__ lcount) {
    const int pt = blockIdx.x, t = blockIdx.y, tid = threadIdx.x;
    bool ok = false;
    float a1 = 0.f, a2 = 0.f;
    if (tid < nops) {
        float E1[9], E2[9];
        euler_f(pts[2 * pt], pts[2 * pt + 1], 0.f, E1);
        euler_f(tang[3 * t], tang[3 * t + 1], tang[3 * t + 2], E2);
        crosslin_one(E1, E2, ops + 9 * tid, &a1, &a2);
        ok = a1 < 999.f && a2 < 999.f;
    }
    const unsigned long long mask = __ballot(ok);
    const long long at = ((long long)pt * T + t) * 64 + __popcll(mask & ((1ull << tid) - 1ull));
    if (ok) {
        lines[at] = line_entry(a1, a2, dpsi, npsi);
        // the same lines on the polar grid of twice the step (lines2 null: none)
        if (lines2) lines2[at] = line_entry(a1, a2, 2.f * dpsi, npsi >> 1);
    }
    if (tid == 0) lcount[(long long)pt * T + t] = __popcll(mask);
}

// Upper bound of the D entries of a residual table with the polar rays of
// one item: the sum over radii of the largest ungated |F|/2 on the ring
// (D(jT, jP) = sum_r aT gP + gT aP is at most the template's bound plus the
// particle's).  k_polar leaves each ring's largest value (or 0) in rmax;
// one thread per item adds them in ring order.
__global__ void k_ring_sum(const float* __restrict__ rmax, int nr, int nitems, float* __restrict__ out) {
    const int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nitems) return;
    const float* m = rmax + (long long)item * nr;
    float acc = 0.f;
    for (int r = 0; r < nr; ++r) acc += m[r];
    out[item] = acc;
}

// The coarse tables: the entries of the even template and particle rays (the
// residual tables of the polar grid of twice the step) as E4M3 {N, D}
// pairs, 2 bytes each, scaled by 128 / (the particle's D bound plus the
// largest of its templates'), so D <= 128 and N <= pi D stay below the
// format's 448.  One scale per particle: a score sums the entries of all
// its templates and is a ratio, so the scale cancels and is not stored.
__global__ void __launch_bounds__(256) k_coarse_table(const unsigned int* __restrict__ table,
                                                      const float* __restrict__ pbound,
                                                      const float* __restrict__ tbound, int tb_per_particle, int T,
                                                      int npsi, unsigned short* __restrict__ ctab) {
    const int item = blockIdx.z, p = item / T, i = blockIdx.y;
    const int nrow = npsi >> 1, n2 = npsi >> 1, nrow2 = npsi >> 2;
    const int jc = blockIdx.x * blockDim.x + threadIdx.x;
    if (jc >= n2) return;
    float tb = 0.f;
    for (int t = 0; t < T; ++t) tb = fmaxf(tb, tbound[tb_per_particle ? p * T + t : t]);
    const float bound = pbound[p] + tb;
    const float sc = bound > 0.f ? 128.f / bound : 0.f;
    const unsigned int e = table[((long long)item * nrow + 2 * i) * npsi + 2 * jc];
    const float N = fminf(__uint_as_float(e << 16) * sc, 448.f);
    const float D = fminf(__uint_as_float(e & 0xffff0000u) * sc, 448.f);
    ctab[((long long)item * nrow2 + i) * n2 + jc] = (unsigned short)__builtin_amdgcn_cvt_pk_fp8_f32(N, D, 0, false);
}

// First minimum (value, then linear index) over a workgroup of at most 4
// waves; every thread returns with it.
__device__ __forceinline__ void block_argmin(float& best, int& bidx, float* redv, int* redi) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bidx, o, 64);
        if (ov < best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
    }
    const int tid = threadIdx.x, nw = (int)(blockDim.x >> 6);
    if ((tid & 63) == 0) { redv[tid >> 6] = best; redi[tid >> 6] = bidx; }
    __syncthreads();
    best = redv[0];
    bidx = redi[0];
    for (int i = 1; i < nw; ++i)
        if (redv[i] < best || (redv[i] == best && redi[i] < bidx)) { best = redv[i]; bidx = redi[i]; }
    __syncthreads();
}

// One workgroup per (AU point, particle), threads over omega: every
// template's common lines (k_cl_lines) are summed in registers and the
// first minimum over omega kept (and, with outi2, a second one elsewhere).  Workgroups are dealt to the 8 XCDs in turn
// (workgroup L on XCD L mod 8), so L mod 8 picks the particle within a group
// of 8 and L / 8 walks its AU points: each XCD reads the tables of one
// particle at a time.  The line lists are the same for the whole workgroup
// (scalar loads), and the loads of 8 lines are issued together.  Entry:
// unsigned int (bfloat16 pairs) or unsigned short (the coarse E4M3 pairs).
template <int VPT, typename Entry>
__global__ void __launch_bounds__(256) k_cl_gather(GatherArgs a) {
    __shared__ float redv[4];
    __shared__ int redi[4];
    const long long L = blockIdx.x;
    const long long r8 = L >> 3;
    const int nl = a.plist ? a.nlist : a.npts;
    const int p = (int)(L & 7) + 8 * (int)(r8 / nl), slot = (int)(r8 % nl);
    if (p >= a.n) return;
    if (a.pcnt && slot >= a.pcnt[p]) return;
    const int pt = a.plist ? a.plist[(long long)p * a.plstride + slot] : slot;
    const int tid = threadIdx.x, bd = blockDim.x;
    const int npsi = a.nray, nrow = npsi >> 1;
    const int k0 = a.pwin ? a.pwin[(long long)p * a.plstride + slot] : 0;
    const int half = a.pwin ? a.ncol >> 1 : a.ncol;
    const int k1 = a.pwin ? a.pwin2[(long long)p * a.plstride + slot] : 0;
    float num[VPT], den[VPT];
    // column of omega k on a line with zero-angle column j0: j0 - k sub, plus npsi when
    // that is negative (columns past ncol read column j0 and are not used)
    int ks[VPT];
#pragma unroll
    for (int v = 0; v < VPT; ++v) {
        num[v] = 0.f;
        den[v] = 0.f;
        const int i = tid + v * bd;
        int k = i < half ? k0 + i : k1 + (i - half);
        k -= k >= a.nomega ? a.nomega : 0;
        ks[v] = i < a.ncol ? k * a.sub : 0;
    }
    for (int t = 0; t < a.T; ++t) {
        const unsigned int* lp = a.lines + ((long long)pt * a.T + t) * 64;
        const int nv = a.lcount[(long long)pt * a.T + t];
        const Entry* tab = static_cast<const Entry*>(a.table) + (long long)(p * a.T + t) * nrow * npsi;
        // entry of column k on the line: row base (scalar) + a 32-bit byte offset (vector)
        auto entry = [&](unsigned int pk, int v) -> Entry {
            const char* row = reinterpret_cast<const char*>(tab + (long long)(pk >> 16) * npsi);
            int c = (int)(pk & 0xffffu) - ks[v];
            c += c < 0 ? npsi : 0;
            return *reinterpret_cast<const Entry*>(row + (unsigned)c * (unsigned)sizeof(Entry));
        };
        constexpr int G = 8;
        int j = 0;
        for (; j + G <= nv; j += G) {
            Entry e[G][VPT];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const unsigned int pk = lp[j + g];
#pragma unroll
                for (int v = 0; v < VPT; ++v) e[g][v] = entry(pk, v);
            }
            // keep all G x VPT loads in flight before the first use
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int v = 0; v < VPT; ++v) acc_entry(e[g][v], num[v], den[v]);
        }
        for (; j < nv; ++j) {
            const unsigned int pk = lp[j];
#pragma unroll
            for (int v = 0; v < VPT; ++v) acc_entry(entry(pk, v), num[v], den[v]);
        }
    }
    const float r2d = 57.2957795130823208768f;
    float best = 3.0e38f;
    int bidx = 0x7fffffff;
#pragma unroll
    for (int v = 0; v < VPT; ++v) {
        const int i = tid + v * bd;
        if (i >= a.ncol) continue;
        int k = i < half ? k0 + i : k1 + (i - half);
        k -= k >= a.nomega ? a.nomega : 0;
        const float pr = den[v] > 1.0e-10f ? num[v] / den[v] * r2d : 999.f;
        const int li = a.pbase[pt] + k * a.ostep;
        if (pr < best || (pr == best && li < bidx)) { best = pr; bidx = li; }
    }
    block_argmin(best, bidx, redv, redi);
    if (tid == 0) {
        a.outv[(long long)p * a.npts + pt] = best;
        a.outi[(long long)p * a.npts + pt] = bidx;
    }
    if (!a.outi2) return;
    // the best omega at least excl steps from the first minimum (every thread holds it)
    const int kb = bidx - a.pbase[pt];
    const int em = a.exmod ? a.exmod : a.nomega;
    float best2 = 3.0e38f;
    int bidx2 = 0x7fffffff;
#pragma unroll
    for (int v = 0; v < VPT; ++v) {
        const int i = tid + v * bd;
        if (i >= a.ncol) continue;
        int k = i < half ? k0 + i : k1 + (i - half);
        k -= k >= a.nomega ? a.nomega : 0;
        int dk = abs(k * a.ostep - kb) % em;
        dk = min(dk, em - dk);
        if (dk < a.excl) continue;
        const float pr = den[v] > 1.0e-10f ? num[v] / den[v] * r2d : 999.f;
        const int li = a.pbase[pt] + k * a.ostep;
        if (pr < best2 || (pr == best2 && li < bidx2)) { best2 = pr; bidx2 = li; }
    }
    block_argmin(best2, bidx2, redv, redi);
    if (tid == 0) {
        a.outi2[(long long)p * a.npts + pt] = bidx2;
        if (a.outv2) a.outv2[(long long)p * a.npts + pt] = best2;
    }
}

// The K AU points of lowest table score per particle (ties: lower linear
// index), in increasing order, and for each the first column of a window of
// wincols columns (of ncol, the columns per omega half when paired) centred
// on the omega of its table minimum.
__global__ void k_fill_scores(float* __restrict__ v, int* __restrict__ ix, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        v[i] = 3.0e38f;
        ix[i] = 0x7fffffff;
    }
}

// Coarse-to-fine gather, second stage: the 3x3 grid neighbours of a
// particle's M best coarse points that have no score yet (every one of them,
// the coarse points included, when rescore_all: the coarse stage scored
// another grid), each once, as that particle's fine list (one workgroup per
// particle; npts <= 8192).  A point's parent is the best-ranked coarse point
// that lists it (an atomic minimum of the entry index in LDS, so the choice
// does not depend on timing); with fwin, the point takes its parent's omega
// window cwin.  The order of the list is not fixed, and nothing depends on it.
__global__ void __launch_bounds__(256) k_expand(const int* __restrict__ ccand, int M, const int* __restrict__ nbr,
                                                const float* __restrict__ v, int npts, int* __restrict__ flist,
                                                int fstride, int* __restrict__ fcnt, int rescore_all,
                                                const int* __restrict__ cwin, int* __restrict__ fwin,
                                                const int* __restrict__ cwin2, int* __restrict__ fwin2) {
    __shared__ int owner[8192];
    __shared__ int cnt;
    const int p = blockIdx.x, tid = threadIdx.x;
    for (int i = tid; i < npts; i += blockDim.x) owner[i] = 0x7fffffff;
    if (tid == 0) cnt = 0;
    __syncthreads();
    for (int e = tid; e < M * 9; e += blockDim.x) {
        const int c = ccand[(long long)p * M + e / 9];
        const int q = nbr[(long long)c * 9 + e % 9];
        if (q < 0 || (!rescore_all && v[(long long)p * npts + q] < 1.0e38f)) continue;
        atomicMin(&owner[q], e);
    }
    __syncthreads();
    for (int e = tid; e < M * 9; e += blockDim.x) {
        const int c = ccand[(long long)p * M + e / 9];
        const int q = nbr[(long long)c * 9 + e % 9];
        if (q < 0 || owner[q] != e) continue;
        const int at = atomicAdd(&cnt, 1);
        flist[(long long)p * fstride + at] = q;
        if (fwin) {
            fwin[(long long)p * fstride + at] = cwin[(long long)p * M + e / 9];
            fwin2[(long long)p * fstride + at] = cwin2[(long long)p * M + e / 9];
        }
    }
    __syncthreads();
    if (tid == 0) fcnt[p] = cnt;
}

// The first column of a window of w columns (of nomega) centred on the second
// minimum of each of a particle's M coarse winners (180 deg from the first
// when a point has none).
__global__ void k_second_win(const int* __restrict__ ccand, int M, const int* __restrict__ ix2,
                             const int* __restrict__ ix, int npts, int nomega, int w, int* __restrict__ kwin2) {
    const int p = blockIdx.x;
    for (int m = threadIdx.x; m < M; m += blockDim.x) {
        const long long q = (long long)p * npts + ccand[(long long)p * M + m];
        const int i2 = ix2[q];
        const int kb = i2 == 0x7fffffff ? ix[q] % nomega + nomega / 2 : i2 % nomega;
        kwin2[(long long)p * M + m] = ((kb - w / 2) % nomega + nomega) % nomega;
    }
}

// The K AU points of lowest table score per particle, any number of points:
// K rounds, each scanning the points not chosen yet.  Candidate kk of
// particle p goes to cand / kwin[p * cstride + coff + kk].
__global__ void k_topk_scan(const float* __restrict__ v, const int* __restrict__ ix, int npts, int K,
                            int* __restrict__ cand, int Nomega, int ncol, int wincols, int* __restrict__ kwin,
                            int cstride, int coff) {
    __shared__ int chosen[256];
    __shared__ float sv[16];
    __shared__ int si[16], sp[16];
    const int p = blockIdx.x, tid = threadIdx.x;
    for (int kk = 0; kk < K; ++kk) {
        float best = 3.0e38f;
        int bidx = 0x7fffffff, bpt = -1;
        for (int i = tid; i < npts; i += blockDim.x) {
            bool taken = false;
            for (int c = 0; c < kk; ++c) taken |= chosen[c] == i;
            if (taken) continue;
            const float x = v[(long long)p * npts + i];
            const int li = ix[(long long)p * npts + i];
            if (x < best || (x == best && li < bidx)) { best = x; bidx = li; bpt = i; }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float xv = __shfl_xor(best, o, 64);
            const int xi = __shfl_xor(bidx, o, 64);
            const int xp = __shfl_xor(bpt, o, 64);
            if (xv < best || (xv == best && xi < bidx)) { best = xv; bidx = xi; bpt = xp; }
        }
        if ((tid & 63) == 0) { sv[tid >> 6] = best; si[tid >> 6] = bidx; sp[tid >> 6] = bpt; }
        __syncthreads();
        if (tid == 0) {
            for (int i = 1; i < (int)(blockDim.x >> 6); ++i)
                if (sv[i] < best || (sv[i] == best && si[i] < bidx)) { best = sv[i]; bidx = si[i]; bpt = sp[i]; }
            chosen[kk] = bpt < 0 ? 0 : bpt;
            cand[(long long)p * cstride + coff + kk] = chosen[kk];
            const int kb = (bidx == 0x7fffffff ? 0 : bidx % Nomega) % ncol;
            kwin[(long long)p * cstride + coff + kk] = ((kb - wincols / 2) % ncol + ncol) % ncol;
        }
        __syncthreads();
    }
}

// The same selection with each thread's AU points (i = tid + 256 c, c < TOPK_PER) in
// registers; a round takes the block's lowest (value, linear index) among the
// points not yet chosen, and the owner thread retires it.
constexpr int TOPK_PER = 16;   // npts <= 4096 (else k_topk_scan)
__global__ void __launch_bounds__(256) k_topk(const float* __restrict__ v, const int* __restrict__ ix, int npts, int K,
                                              int* __restrict__ cand, int Nomega, int ncol, int wincols,
                                              int* __restrict__ kwin, int cstride, int coff) {
    __shared__ float sv[4];
    __shared__ int si[4], sp[4];
    __shared__ int win_pt, win_li;
    const int p = blockIdx.x, tid = threadIdx.x;
    float xv[TOPK_PER];
    int xl[TOPK_PER];
    bool live[TOPK_PER];
#pragma unroll
    for (int c = 0; c < TOPK_PER; ++c) {
        const int i = tid + c * 256;
        live[c] = i < npts;
        xv[c] = live[c] ? v[(long long)p * npts + i] : 3.0e38f;
        xl[c] = live[c] ? ix[(long long)p * npts + i] : 0x7fffffff;
    }
    for (int kk = 0; kk < K; ++kk) {
        float best = 3.0e38f;
        int bidx = 0x7fffffff, bpt = -1;
#pragma unroll
        for (int c = 0; c < TOPK_PER; ++c)
            if (live[c] && (xv[c] < best || (xv[c] == best && xl[c] < bidx))) {
                best = xv[c];
                bidx = xl[c];
                bpt = tid + c * 256;
            }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float ov = __shfl_xor(best, o, 64);
            const int oi = __shfl_xor(bidx, o, 64);
            const int op = __shfl_xor(bpt, o, 64);
            if (ov < best || (ov == best && oi < bidx)) { best = ov; bidx = oi; bpt = op; }
        }
        if ((tid & 63) == 0) { sv[tid >> 6] = best; si[tid >> 6] = bidx; sp[tid >> 6] = bpt; }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < 4; ++w)
                if (sv[w] < best || (sv[w] == best && si[w] < bidx)) { best = sv[w]; bidx = si[w]; bpt = sp[w]; }
            const int ch = bpt < 0 ? 0 : bpt;
            cand[(long long)p * cstride + coff + kk] = ch;
            const int kb = (bidx == 0x7fffffff ? 0 : bidx % Nomega) % ncol;
            kwin[(long long)p * cstride + coff + kk] = ((kb - wincols / 2) % ncol + ncol) % ncol;
            win_pt = bpt;
        }
        __syncthreads();
        const int w = win_pt;
        if (w >= 0 && (w & 255) == tid) {
#pragma unroll
            for (int c = 0; c < TOPK_PER; ++c)
                if (c == (w >> 8)) live[c] = false;
        }
    }
}

// ---- host: grid, disk layout, template bank ---------------------------------

struct SearchGrid {
    int Nt = 0, Np = 0, No = 0;
    std::vector<float> pts;   // theta, phi
    std::vector<int> base;
    std::vector<int> it, sip; // theta index and signed phi index of each point
};

// ort_globesearch.f90:16-21,68-82 evaluated in float32 like the reference.
static SearchGrid make_grid(float step) {
    SearchGrid g;
    g.Nt = (int)(21.0f / step);
    g.Np = (int)(32.0f / step);
    g.No = (int)(360.0f / step);
    const float d2r = atan2f(1.0f, 1.0f) / 45.0f;
    const float A1 = -sinf(31.717f * d2r) * cosf(69.09f * d2r), B1 = -cosf(31.717f * d2r) * cosf(69.09f * d2r);
    const float C1 = sinf(31.717f * d2r) * sinf(69.09f * d2r);
    const float A2 = -sinf(31.717f * d2r) * cosf(69.09f * d2r), B2 = cosf(31.717f * d2r) * cosf(69.09f * d2r);
    const float C2 = sinf(31.717f * d2r) * sinf(69.09f * d2r);
    for (int it = 0; it <= g.Nt; ++it) {
        const float theta = 69.0f + step * (float)it;
        float tmp = theta * d2r;
        const float z = cosf(tmp), z0 = sinf(tmp);
        for (int ip = 0; ip <= g.Np; ++ip) {
            const float phi = step * (float)ip;
            tmp = phi * d2r;
            const float x = z0 * cosf(tmp), y = z0 * sinf(tmp);
            const float z1 = -(A1 * x + B1 * y) / C1, z2 = -(A2 * x + B2 * y) / C2;
            if (!((z >= 0.f) && (z <= z1) && (z <= z2))) continue;
            for (int s = 0; s < (ip == 0 ? 1 : 2); ++s) {
                const int sip = s == 0 ? ip : -ip;
                g.pts.push_back(theta);
                g.pts.push_back(s == 0 ? phi : -phi);
                g.base.push_back(g.No * ((sip + g.Np) + (2 * g.Np + 1) * it));
                g.it.push_back(it);
                g.sip.push_back(sip);
            }
        }
    }
    return g;
}

// Coarse-to-fine lists of the tabulated search: the coarse points are the
// grid points with even theta and phi indices (the grid of twice the step),
// plus every point none of whose 3x3 grid neighbours is such a point (at the
// AU border), so every point is coarse or next to a coarse one.  nbr: for each
// point its 3x3 neighbours in the grid (-1 where outside the AU).
static void coarse_lists(const SearchGrid& g, std::vector<int>& clist, std::vector<int>& nbr) {
    const int npts = (int)g.it.size(), W = 2 * g.Np + 1;
    std::vector<int> id((size_t)(g.Nt + 1) * W, -1);
    for (int q = 0; q < npts; ++q) id[(size_t)g.it[q] * W + g.sip[q] + g.Np] = q;
    auto at = [&](int it, int sip) {
        return (it < 0 || it > g.Nt || sip < -g.Np || sip > g.Np) ? -1 : id[(size_t)it * W + sip + g.Np];
    };
    nbr.assign((size_t)npts * 9, -1);
    for (int q = 0; q < npts; ++q)
        for (int d = 0; d < 9; ++d) nbr[(size_t)q * 9 + d] = at(g.it[q] + d / 3 - 1, g.sip[q] + d % 3 - 1);
    std::vector<char> coarse(npts, 0), covered(npts, 0);
    auto take = [&](int q) {
        coarse[q] = 1;
        for (int d = 0; d < 9; ++d)
            if (nbr[(size_t)q * 9 + d] >= 0) covered[nbr[(size_t)q * 9 + d]] = 1;
    };
    for (int q = 0; q < npts; ++q)
        if (g.it[q] % 2 == 0 && g.sip[q] % 2 == 0) take(q);
    for (int q = 0; q < npts; ++q)
        if (!covered[q]) take(q);
    clist.clear();
    for (int q = 0; q < npts; ++q)
        if (coarse[q]) clist.push_back(q);
}

// Disk of taps reachable from radius <= R (plus a one-pixel margin) as rows
// lo..hi of relative x with LDS base offsets.
static void make_disk(int R, std::vector<int>& rows, int& total) {
    const int nrows = 2 * R + 4;
    rows.assign(3 * nrows, 0);
    total = 0;
    for (int r = 0; r < nrows; ++r) {
        const int ly = r - R - 1;
        double ymin = (ly - 1 <= 0 && 0 <= ly + 1) ? 0.0 : std::min(std::abs(ly - 1), std::abs(ly + 1));
        double xm = std::sqrt(std::max(0.0, (double)R * R - ymin * ymin));
        int lo = (int)std::floor(-xm) - 1, hi = (int)std::floor(xm) + 2;
        lo = std::max(lo, -R - 1);
        hi = std::min(hi, R + 2);
        rows[3 * r] = lo;
        rows[3 * r + 1] = hi;
        rows[3 * r + 2] = total - lo;
        total += hi - lo + 1;
    }
}

// The template images (T x box^2, in ctx->img) masked at their own shifts,
// transformed, centre-shifted and Fcut-fitted (load_template.f90:35-48).
static int prep_templates_img(cim_ctx* ctx, const cim_params* p, int T, const float* tort) {
    SearchState& s = ctx->srch;
    const int box = p->box;
    const int nr = p->maxR - p->minR + 1;
    const long long N = (long long)box * box;
    std::vector<float> parts((size_t)T * CIM_PART_STRIDE, 0.f);
    s.tang.assign(3 * T, 0.f);
    for (int t = 0; t < T; ++t) {
        float* q = &parts[(size_t)t * CIM_PART_STRIDE];
        q[CIM_P_THETA] = tort[5 * t];
        q[CIM_P_PHI] = tort[5 * t + 1];
        q[CIM_P_OMEGA] = tort[5 * t + 2];
        q[CIM_P_X] = tort[5 * t + 3];
        q[CIM_P_Y] = tort[5 * t + 4];
        q[CIM_P_MASKX] = tort[5 * t + 3];
        q[CIM_P_MASKY] = tort[5 * t + 4];
        for (int k = 0; k < 3; ++k) s.tang[3 * t + k] = tort[5 * t + k];
    }
    int rc = upload_parts(ctx, T, parts.data());
    if (rc) return rc;
    CIM_ALLOC(s.tft, sizeof(float2) * N * T);
    rc = preprocess_images(ctx, p, T, ctx->img.as<float>(), ctx->parts.as<float>(), s.tft.as<float2>(), 0,
                           p->mask_on, 1);
    if (rc) return rc;
    rc = centre_shift(ctx, s.tft.as<float2>(), box, T, ctx->parts.as<float>(), 1);
    if (rc) return rc;
    CIM_ALLOC(s.tfcut, sizeof(float) * nr * T);
    rc = amp_expfit(ctx, s.tft.as<float2>(), box, T, p->minR, p->maxR, s.tfcut.as<float>());
    if (rc) return rc;
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    s.box = box;
    s.T = T;
    s.minR = p->minR;
    s.maxR = p->maxR;
    s.ready = true;
    s.version += 1;
    return CIM_OK;
}

static int prep_templates(cim_ctx* ctx, const cim_params* p, int T, const float* images, const float* tort) {
    const long long N = (long long)p->box * p->box;
    CIM_ALLOC(ctx->img, sizeof(float) * N * T);
    CIM_HIP(hipMemcpyAsync(ctx->img.ptr, images, sizeof(float) * N * T, hipMemcpyHostToDevice, ctx->stream));
    return prep_templates_img(ctx, p, T, tort);
}

// project_template2d.f90:3-80 (and mode 0's project2d.f90): the real-space
// projection of the nx^3 model at each template's orientation.  One thread
// per (pixel, template) walks its voxel column: for each voxel strictly
// inside radius nx/2 - 1, X = R (ix, iy, iz) - nx/2 + nx/2 and the trilinear
// value of the model at X is added, iz ascending, in the reference's REAL*4
// order of operations (no contraction).  R: T row-major euler_mat matrices.
__global__ void k_project_template(const float* __restrict__ vol, int nx, const float* __restrict__ R,
                                   float* __restrict__ out) {
#pragma clang fp contract(off)
    const long long n2 = (long long)nx * nx;
    const int t = blockIdx.y;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n2) return;
    const int ix = (int)(idx % nx), iy = (int)(idx / nx), h = nx / 2;
    const float* r = R + 9 * t;
    const float x0 = (float)ix - (float)h, x1 = (float)iy - (float)h, lim = (float)(h - 1);
    float s = 0.f;
    for (int iz = 0; iz < nx; ++iz) {
        const float x2 = (float)iz - (float)h;
        if (sqrtf(x0 * x0 + x1 * x1 + x2 * x2) >= lim) continue;
        const float X = r[0] * x0 + r[1] * x1 + r[2] * x2 + (float)h;
        const float Y = r[3] * x0 + r[4] * x1 + r[5] * x2 + (float)h;
        const float Z = r[6] * x0 + r[7] * x1 + r[8] * x2 + (float)h;
        const int jx = (int)X, jy = (int)Y, jz = (int)Z;
        const float xb = X - (float)jx, yb = Y - (float)jy, zb = Z - (float)jz;
        const float* v = vol + (long long)jz * n2 + (long long)jy * nx + jx;
        s = s + (1.f - xb) * (1.f - yb) * (1.f - zb) * v[0];
        s = s + xb * (1.f - yb) * (1.f - zb) * v[1];
        s = s + (1.f - xb) * yb * (1.f - zb) * v[nx];
        s = s + xb * yb * (1.f - zb) * v[nx + 1];
        s = s + (1.f - xb) * (1.f - yb) * zb * v[n2];
        s = s + xb * (1.f - yb) * zb * v[n2 + 1];
        s = s + (1.f - xb) * yb * zb * v[n2 + nx];
        s = s + xb * yb * zb * v[n2 + nx + 1];
    }
    out[(long long)t * n2 + idx] = s;
}

// euler_mat.f90:3-21 in float32 on the host (R row-major).
static void euler_host(float th0, float ph0, float om0, float* R) {
    const float d2r = std::atan2(1.0f, 1.0f) / 45.0f;
    const float th = th0 * d2r, ph = ph0 * d2r, om = om0 * d2r;
    const float c1 = std::cos(ph), c2 = std::cos(th), c3 = std::cos(om);
    const float s1 = std::sin(ph), s2 = std::sin(th), s3 = std::sin(om);
    R[0] = c1 * c2 * c3 - s1 * s3; R[1] = -c1 * c2 * s3 - c3 * s1; R[2] = c1 * s2;
    R[3] = c2 * c3 * s1 + c1 * s3; R[4] = -c2 * s1 * s3 + c1 * c3; R[5] = s1 * s2;
    R[6] = -c3 * s2;               R[7] = s2 * s3;                 R[8] = c2;
}

// Project T templates from the model (host, box^3) into ctx->img.
static int project_templates(cim_ctx* ctx, int box, int T, const float* model, const float* tort) {
    const long long n3 = (long long)box * box * box, n2 = (long long)box * box;
    std::vector<float> R(9 * (size_t)T);
    for (int t = 0; t < T; ++t) euler_host(tort[5 * t], tort[5 * t + 1], tort[5 * t + 2], &R[9 * (size_t)t]);
    CIM_ALLOC(ctx->aux, sizeof(float) * n3);
    CIM_ALLOC(ctx->aux2, sizeof(float) * 9 * T);
    CIM_ALLOC(ctx->img, sizeof(float) * n2 * T);
    CIM_HIP(hipMemcpyAsync(ctx->aux.ptr, model, sizeof(float) * n3, hipMemcpyHostToDevice, ctx->stream));
    CIM_HIP(hipMemcpyAsync(ctx->aux2.ptr, R.data(), sizeof(float) * 9 * T, hipMemcpyHostToDevice, ctx->stream));
    TimedLaunch tl(ctx, "project");
    hipLaunchKernelGGL(k_project_template, dim3((unsigned)((n2 + 255) / 256), T), dim3(256), 0, ctx->stream,
                       ctx->aux.as<float>(), box, ctx->aux2.as<float>(), ctx->img.as<float>());
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}


// exhaustive fallback: listed particles scored concurrently by one launch
constexpr int FB_CONCURRENT = 8;

static int launch_search(cim_ctx* ctx, SearchArgs& a, int n, int lds_bytes, int mode, bool pair, int vpt, int bd) {
    dim3 grid(a.cand ? a.ncand * (a.kwin ? a.wsplit : 1) : a.npts, n);
#define CIM_LAUNCH(V, M, P)                                                                                      \
    do {                                                                                                         \
        CIM_HIP(hipFuncSetAttribute((const void*)k_cl_search<V, M, P>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                    lds_bytes));                                                                 \
        hipLaunchKernelGGL((k_cl_search<V, M, P>), grid, dim3(bd), lds_bytes, ctx->stream, a);                   \
    } while (0)
#define CIM_LAUNCH_V(M, P)                  \
    do {                                    \
        if (vpt == 1) CIM_LAUNCH(1, M, P);  \
        else if (vpt == 2) CIM_LAUNCH(2, M, P); \
        else if (vpt == 3) CIM_LAUNCH(3, M, P); \
        else CIM_LAUNCH(4, M, P);           \
    } while (0)
    if (mode == 0) {
        if (pair) CIM_LAUNCH_V(0, true);
        else CIM_LAUNCH_V(0, false);
    } else {
        if (pair) CIM_LAUNCH_V(1, true);
        else CIM_LAUNCH_V(1, false);
    }
#undef CIM_LAUNCH_V
#undef CIM_LAUNCH
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

static int block_for(int No, int vpt) {
    int bd = (No + vpt - 1) / vpt;
    return ((bd + 63) / 64) * 64;
}

// Autotuner: the launch shape -- omega values per thread (with it the
// workgroup size) and whether the particle disk is staged in LDS or read
// through the caches -- is chosen by timing one-particle launches of every
// candidate on the first batch, and cached per (maxR, Nomega, T).
// CIM_SEARCH_VPT=<1..4> and CIM_SEARCH_LDS=<0|1> force a choice.
struct SearchShape {
    int vpt;
    bool lds;
    int lds_bytes;
    bool pair;
};

static int tune_shape(cim_ctx* ctx, SearchArgs a, int lds_with, int lds_without, bool lds_fits, int No, bool pair,
                      SearchShape* out) {
    const int ncol = pair ? No / 2 : No;
    const char* ev = getenv("CIM_SEARCH_VPT");
    const char* el = getenv("CIM_SEARCH_LDS");
    char key[96];
    snprintf(key, sizeof(key), "shape:%d:%d:%d:%d", a.R, No, a.T, (int)pair);
    auto it = ctx->tuned.find(key);
    if (it != ctx->tuned.end() && !ev && !el) {
        out->vpt = it->second & 7;
        out->lds = (it->second >> 3) & 1;
        out->lds_bytes = out->lds ? lds_with : lds_without;
        out->pair = pair;
        return CIM_OK;
    }
    a.nsamples = nullptr;
    hipEvent_t e0, e1;
    CIM_HIP(hipEventCreate(&e0));
    CIM_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    SearchShape bs{0, false, 0, pair};
    for (int l = 1; l >= 0; --l) {
        if (l == 1 && !lds_fits) continue;
        if (el && atoi(el) != l) continue;
        for (int v = 1; v <= 4; ++v) {
            const int bd = block_for(ncol, v);
            if (bd > 1024 || (v > 1 && block_for(ncol, v) == block_for(ncol, v - 1))) continue;
            if (ev && atoi(ev) != v) continue;
            const int lb = l ? lds_with : lds_without;
            CIM_HIP(hipEventRecord(e0, ctx->stream));
            int rc = launch_search(ctx, a, 1, lb, l == 1 ? 0 : 1, pair, v, bd);
            if (rc) return rc;
            CIM_HIP(hipEventRecord(e1, ctx->stream));
            CIM_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            CIM_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (getenv("CIM_TUNE_VERBOSE")) fprintf(stderr, "cim tune: vpt %d lds %d block %d: %.3f ms\n", v, l, bd, ms);
            if (ms < best) {
                best = ms;
                bs = SearchShape{v, l == 1, lb, pair};
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (bs.vpt == 0) return set_error(ctx, CIM_E_ARG, "no launch shape fits Nomega %d", No);
    if (!ev && !el) ctx->tuned[key] = bs.vpt | (bs.lds ? 8 : 0);
    *out = bs;
    return CIM_OK;
}

// Global search of n particles already on the device (images_dev) with
// host records parts (in/out).  pr_array: optional, for n == 1.
constexpr int CIM_TABLE_NH_DEFAULT = 4;

// The residual tables of nb particles x T templates: the MFMA kernel with nh
// harmonics, or the exact VALU kernel (nh = 0).
int launch_cl_table(cim_ctx* ctx, int nh, const float2* ppol, const float2* tpol, int tpol_per_particle, int nb, int T,
                    int nr, int npsi, unsigned int* table) {
    if (nh == 0) {
        const unsigned nt = (unsigned)((npsi + 63) / 64), ntr = (unsigned)((npsi / 2 + 63) / 64);
        hipLaunchKernelGGL(k_cl_table, dim3(nt, ntr, nb * T), dim3(256), 0, ctx->stream, ppol, tpol,
                           tpol_per_particle, T, nr, npsi, table);
    } else {
        const dim3 grid((unsigned)((npsi + 127) / 128), (unsigned)((npsi / 2 + 127) / 128), (unsigned)(nb * T));
        if (nh == 1)
            hipLaunchKernelGGL(k_cl_table_mfma<1>, grid, dim3(256), 0, ctx->stream, ppol, tpol, tpol_per_particle, T,
                               nr, npsi, table);
        else if (nh == 2)
            hipLaunchKernelGGL(k_cl_table_mfma<2>, grid, dim3(256), 0, ctx->stream, ppol, tpol, tpol_per_particle, T,
                               nr, npsi, table);
        else if (nh == 4)
            hipLaunchKernelGGL(k_cl_table_mfma<4>, grid, dim3(256), 0, ctx->stream, ppol, tpol, tpol_per_particle, T,
                               nr, npsi, table);
        else if (nh == 8)
            hipLaunchKernelGGL(k_cl_table_mfma<8>, grid, dim3(256), 0, ctx->stream, ppol, tpol, tpol_per_particle, T,
                               nr, npsi, table);
        else
            return set_error(ctx, CIM_E_ARG, "CIM_TABLE_NH %d: use 0, 1, 2, 4 or 8", nh);
    }
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

// parts_dev: device records (n x CIM_PART_STRIDE) read and written on the
// stream with no host synchronisation, copy or allocation, so that the call
// can be captured in a graph; only for the tabulated search of one batch,
// after a host-records call with the same parameters has set the constant
// tables up.  parts (host records) is then unused.
// ucand, uomega (host, n x ucn): score only these AU points, each over a
// window of omega columns centred on the given column (the tables are not
// built and nothing falls back).
static int search_global_dev(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev, float* parts,
                             float* pr_array, float* parts_dev = nullptr, const int* ucand = nullptr,
                             const int* uomega = nullptr, int ucn = 0) {
    SearchState& s = ctx->srch;
    const bool devmode = parts_dev != nullptr;
    if (!s.ready) return set_error(ctx, CIM_E_STATE, "cim_search_templates has not been called");
    if (p->box != s.box || p->minR != s.minR || p->maxR != s.maxR)
        return set_error(ctx, CIM_E_ARG, "box/minR/maxR differ from the template bank");
    if (!(p->searchstep > 0.f)) return set_error(ctx, CIM_E_ARG, "searchstep must be > 0");
    if (n <= 0) return CIM_OK;
    const int box = p->box, R = p->maxR, nr = R - p->minR + 1, T = s.T;
    const int Dc = 2 * R + 4;
    const long long N = (long long)box * box;
    SearchGrid g = make_grid(p->searchstep);
    if (g.pts.empty()) return set_error(ctx, CIM_E_ARG, "search grid has no AU point for step %g", p->searchstep);
    const int npts = (int)g.base.size();
    if (g.No > 4096) return set_error(ctx, CIM_E_ARG, "searchstep %g too fine (Nomega > 4096)", p->searchstep);
    // Friedel pairing needs the omega grid to close: step * Nomega/2 == 180
    const bool pair = (g.No % 2 == 0) && std::fabs(p->searchstep * (double)(g.No / 2) - 180.0) < 1e-4 &&
                      getenv("CIM_SEARCH_NOPAIR") == nullptr;
    std::vector<int> rows;
    int disk_total = 0;
    make_disk(R, rows, disk_total);
    const int nrows = 2 * R + 4;
    const size_t misc = sizeof(float) * (64 + 64 + nr) + sizeof(int) * nrows + 16 * 8 + 64;
    const size_t lds_cap = 160 * 1024;
    const bool lds_fits = sizeof(float2) * (size_t)disk_total + misc + sizeof(float4) * 4 * nr <= lds_cap;
    size_t rem = lds_cap - misc - (lds_fits ? sizeof(float2) * (size_t)disk_total : 0);
    // one line-chunk size for both variants so the tuned kernels do the same work
    int chunk = (int)std::min<size_t>(60, rem / (sizeof(float4) * nr));
    if (chunk < 1) return set_error(ctx, CIM_E_ARG, "maxR %d too large for the LDS line buffer", R);
    const int lds_without = (int)(misc + sizeof(float4) * (size_t)chunk * nr);
    const int lds_with = lds_without + (int)(sizeof(float2) * (size_t)disk_total);

    // device copies of the small tables
    std::vector<float> ops(540);
    {
        double d[540];
        cim_icos60_ref_order(d);
        for (int i = 0; i < 540; ++i) ops[i] = (float)d[i];
    }
    SearchBatch& sb = ctx->srch.wk;
    // the constant tables depend on these; a device-records call needs them set up already
    char sig[160];
    snprintf(sig, sizeof(sig), "%d:%d:%d:%g:%d:%d:%d:%d:%lld", box, p->minR, R, (double)p->searchstep, T,
             p->search_method, p->search_topk, n, (long long)s.version);
    if (devmode) {
        if (ctx->srch.step_sig != sig)
            return set_error(ctx, CIM_E_STATE, "device-records search: call cim_search_global_dev with host records "
                                               "and the same parameters and count first");
        if (pr_array || p->search_method != CIM_SEARCH_TABLES)
            return set_error(ctx, CIM_E_ARG, "device-records search: tabulated search only");
    } else {
        CIM_ALLOC(sb.ops, sizeof(float) * 540);
        CIM_ALLOC(sb.rows, sizeof(int) * rows.size());
        CIM_ALLOC(sb.pts, sizeof(float) * g.pts.size());
        CIM_ALLOC(sb.pbase, sizeof(int) * g.base.size());
        CIM_ALLOC(sb.tang, sizeof(float) * 3 * T);
        CIM_HIP(hipMemcpyAsync(sb.ops.ptr, ops.data(), sizeof(float) * 540, hipMemcpyHostToDevice, ctx->stream));
        CIM_HIP(hipMemcpyAsync(sb.rows.ptr, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice,
                               ctx->stream));
        CIM_HIP(hipMemcpyAsync(sb.pts.ptr, g.pts.data(), sizeof(float) * g.pts.size(), hipMemcpyHostToDevice,
                               ctx->stream));
        CIM_HIP(hipMemcpyAsync(sb.pbase.ptr, g.base.data(), sizeof(int) * g.base.size(), hipMemcpyHostToDevice,
                               ctx->stream));
        CIM_HIP(hipMemcpyAsync(sb.tang.ptr, s.tang.data(), sizeof(float) * 3 * T, hipMemcpyHostToDevice, ctx->stream));
    }

    // tabulated search (CIM_SEARCH_TABLES): polar grid of searchstep / sub
    const bool tables = p->search_method == CIM_SEARCH_TABLES && !pr_array && !ucand;
    if (p->search_method != CIM_SEARCH_PLAIN && p->search_method != CIM_SEARCH_TABLES)
        return set_error(ctx, CIM_E_ARG, "search_method %d unknown", p->search_method);
    int sub = 1;
    if (const char* es = getenv("CIM_TABLE_SUB")) sub = std::max(1, std::min(4, atoi(es)));
    const int npsi = g.No * sub;
    const float dpsi = p->searchstep / (float)sub;
    const int topk = std::min(std::min(p->search_topk > 0 ? p->search_topk : 32, npts), 256);
    // residual tables on the matrix cores with NH odd harmonics of the folded phase difference
    // (CIM_TABLE_NH = 1, 2, 4 or 8), or the exact residual on the VALU (0)
    int table_nh = CIM_TABLE_NH_DEFAULT;
    if (const char* eh = getenv("CIM_TABLE_NH")) table_nh = atoi(eh);
    if (tables && g.No > 2048) return set_error(ctx, CIM_E_ARG, "searchstep %g too fine for the tables", p->searchstep);
    // the polar grid must close on itself with rays 180 deg apart on it
    if (tables && (std::fabs(p->searchstep * (double)g.No - 360.0) > 1e-3 || (npsi & 1)))
        return set_error(ctx, CIM_E_ARG, "CIM_SEARCH_TABLES needs 360/searchstep to be an even integer (step %g)",
                         p->searchstep);
    // fallback to the exhaustive search when the exact minimum ranks in the last quarter of the
    // candidates.  CIM_TABLE_FALLBACK=<r>: from rank r on; 0: never; negative: every particle.
    int fb_rank = topk - topk / 4;
    if (const char* ef = getenv("CIM_TABLE_FALLBACK")) {
        const int v = atoi(ef);
        fb_rank = v > 0 ? v : (v == 0 ? topk + 1 : 0);
    }
    // exact re-scoring window: wincols columns per (paired) omega half around the table minimum
    const int ncolx = pair ? g.No / 2 : g.No;
    // CIM_RESCORE_CPW candidates (1, 2, 4 or 8) share a wave, each with 64 / CPW columns; CPW = 1
    // scores windows of CIM_TABLE_WINDOW (64) columns, one wave per candidate
    int rescore_cpw = 4;
    if (const char* ec = getenv("CIM_RESCORE_CPW")) rescore_cpw = atoi(ec);
    if (rescore_cpw != 1 && rescore_cpw != 2 && rescore_cpw != 4 && rescore_cpw != 8) rescore_cpw = 4;
    int wincols = 64 / rescore_cpw;
    if (const char* ew = getenv("CIM_TABLE_WINDOW")) {
        wincols = atoi(ew) > 0 ? ((atoi(ew) + 63) / 64) * 64 : ncolx;
        rescore_cpw = 1;
    }
    if (wincols > ncolx) {
        wincols = ncolx;
        rescore_cpw = 1;
    }
    // second windows: besides the topk AU points, the nsecond best second minima of the table
    // scores over omega (at least a window from the point's first minimum) are re-scored too
    // (search_second; CIM_TABLE_SECOND=<S> overrides it)
    int nsecond = p->search_second;
    if (const char* e2 = getenv("CIM_TABLE_SECOND")) nsecond = atoi(e2);
    nsecond = tables && wincols < ncolx ? std::max(0, std::min(std::min(nsecond, npts), 256 - topk)) : 0;

    // the tables take 4 T npsi^2 bytes per particle (8.3 MB at T = 4, 0.5 deg): batches of
    // CIM_TABLE_BATCH (default 512, 4.2 GB) particles.  Large batches keep enough waves of
    // the one-wave re-score workgroups resident (32 per particle): measured at box 384,
    // the re-score takes 5.5 ms per 64 particles in batches of 64 and 2.75 ms in batches of 512
    int tbatch = 512;
    if (const char* eb = getenv("CIM_TABLE_BATCH")) tbatch = std::max(1, std::min(1024, atoi(eb)));
    const int batch = std::max(1, std::min(n, tables ? tbatch : 256));
    CtfConst cc = make_ctf_const(p);

    // coarse-to-fine table scores: the coarse points (every other theta and phi index,
    // plus the AU border points no such point neighbours) first, then the unscored 3x3
    // neighbours of each particle's coarseM best.  CIM_TABLE_COARSE=<M>; 0 scores every point.
    int coarseM = 64;
    // (k_topk_scan keeps the chosen points in a 256-entry LDS array)
    if (const char* ec = getenv("CIM_TABLE_COARSE")) coarseM = std::max(0, std::min(256, atoi(ec)));
    // the coarse stage searches the grid of twice the step in all three angles, from E4M3 tables of
    // the polar grid of twice the step (CIM_TABLE_COARSE_GRID=0: the step's omega and tables); the
    // fine stage then scores every omega column of the 3x3 neighbourhoods of the coarseM best, their
    // centres included
    bool coarse_grid = sub == 1 && npsi % 4 == 0;
    if (const char* ek = getenv("CIM_TABLE_COARSE_GRID")) coarse_grid = coarse_grid && atoi(ek) != 0;
    // with the coarse grid, the fine stage scores CIM_TABLE_FINE_WINDOW (128) omega columns of each
    // point: half around its parent's coarse minimum, half around the parent's best omega at least
    // a half window from it (a second minimum); 0: every omega
    int fine_win = 128;
    if (const char* ew = getenv("CIM_TABLE_FINE_WINDOW")) fine_win = std::max(0, atoi(ew)) & ~1;
    std::vector<int> clist, nbrv;
    bool c2f = false;
    if (tables && coarseM > 0 && npts <= 8192) {
        coarse_lists(g, clist, nbrv);
        c2f = (int)clist.size() > coarseM && (int)clist.size() >= topk;
    }
    if (tables) {
        // the common lines of every (AU point, template) on the polar grid, once per call
        CIM_ALLOC(sb.lines, sizeof(unsigned int) * (size_t)npts * T * 64 * 2);
        CIM_ALLOC(sb.lcount, sizeof(int) * (size_t)npts * T);
        hipLaunchKernelGGL(k_cl_lines, dim3(npts, T), dim3(64), 0, ctx->stream, sb.pts.as<float>(), sb.tang.as<float>(),
                           sb.ops.as<float>(), T, 60, npsi, dpsi, sb.lines.as<unsigned int>(),
                           coarse_grid ? sb.lines.as<unsigned int>() + (size_t)npts * T * 64 : nullptr,
                           sb.lcount.as<int>());
        CIM_HIP(hipGetLastError());
    }
    if (c2f) {
        CIM_ALLOC(sb.clist, sizeof(int) * clist.size());
        CIM_ALLOC(sb.nbr, sizeof(int) * nbrv.size());
        CIM_ALLOC(sb.ccand, sizeof(int) * (size_t)batch * coarseM);
        CIM_ALLOC(sb.ckwin, sizeof(int) * (size_t)batch * coarseM);
        CIM_ALLOC(sb.flist, sizeof(int) * (size_t)batch * coarseM * 9);
        CIM_ALLOC(sb.fwin, sizeof(int) * (size_t)batch * coarseM * 9);
        CIM_ALLOC(sb.fwin2, sizeof(int) * (size_t)batch * coarseM * 9);
        CIM_ALLOC(sb.ckwin2, sizeof(int) * (size_t)batch * coarseM);
        CIM_ALLOC(sb.fcnt, sizeof(int) * (size_t)batch);
        if (!devmode) {
            CIM_HIP(hipMemcpyAsync(sb.clist.ptr, clist.data(), sizeof(int) * clist.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
            CIM_HIP(hipMemcpyAsync(sb.nbr.ptr, nbrv.data(), sizeof(int) * nbrv.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
        }
    }
    if (devmode && n > batch)
        return set_error(ctx, CIM_E_ARG, "device-records search: %d particles exceed one batch (%d)", n, batch);
    for (int b0 = 0; b0 < n; b0 += batch) {
        const int nb = std::min(batch, n - b0);
        // records: mask around each particle's own centre (ort_globesearch.f90:50)
        int rc = CIM_OK;
        if (devmode) {
            CIM_ALLOC(ctx->parts, sizeof(float) * CIM_PART_STRIDE * (size_t)nb);
            hipLaunchKernelGGL(k_records_in, dim3((nb * CIM_PART_STRIDE + 255) / 256), dim3(256), 0, ctx->stream,
                               parts_dev + (size_t)b0 * CIM_PART_STRIDE, nb, ctx->parts.as<float>());
            CIM_HIP(hipGetLastError());
        } else {
            std::vector<float> pr(parts + (size_t)b0 * CIM_PART_STRIDE, parts + (size_t)(b0 + nb) * CIM_PART_STRIDE);
            for (int i = 0; i < nb; ++i) {
                pr[i * CIM_PART_STRIDE + CIM_P_MASKX] = pr[i * CIM_PART_STRIDE + CIM_P_X];
                pr[i * CIM_PART_STRIDE + CIM_P_MASKY] = pr[i * CIM_PART_STRIDE + CIM_P_Y];
            }
            rc = upload_parts(ctx, nb, pr.data());
            if (rc) return rc;
        }
        CIM_ALLOC(sb.fcut0, sizeof(float) * nr * nb);
        CIM_ALLOC(sb.pcrop, sizeof(float2) * (size_t)Dc * Dc * nb);
        // the global search masks unconditionally (ort_globesearch.f90:50, unlike mode 5)
        const int o0 = box / 2 - R - 1;
        if (o0 >= 0 && o0 + Dc <= box && prep_fused()) {
            // the transform is needed on the crop only: rows pruned to its columns, then its columns;
            // the images are prepared as the row pass reads them, from their sums
            CIM_ALLOC(ctx->prepst, sizeof(PrepStats) * (size_t)nb);
            rc = prep_stats_dev(ctx, p, nb, images_dev + (long long)b0 * N, ctx->parts.as<float>(), 0, 1,
                                ctx->prepst.as<PrepStats>());
            if (rc) return rc;
            CIM_ALLOC(sb.ftrow, sizeof(float2) * (size_t)nb * box * Dc);
            rc = fft2d_crop_real_dev(ctx, images_dev + (long long)b0 * N, ctx->prepst.as<PrepStats>(), 0, 1,
                                     p->imgmask, box, nb, o0, Dc, o0, Dc, 1, nullptr, sb.ftrow.as<float2>(),
                                     sb.pcrop.as<float2>());
            if (rc) return rc;
            rc = amp_expfit_general(ctx, sb.pcrop.as<float2>(), (long long)Dc * Dc, Dc, R + 1, nb, p->minR, R,
                                    sb.fcut0.as<float>());
            if (rc) return rc;
            hipLaunchKernelGGL(k_crop_shift, dim3((Dc * Dc + 255) / 256, nb), dim3(256), 0, ctx->stream,
                               sb.pcrop.as<float2>(), box, R, Dc, ctx->parts.as<float>());
            CIM_HIP(hipGetLastError());
        } else if (o0 >= 0 && o0 + Dc <= box) {
            // the same with a prepared copy of the images (CIM_PREP_FUSED=0)
            CIM_ALLOC(ctx->ft, sizeof(float2) * N * nb);
            float2* ft = ctx->ft.as<float2>();
            rc = prep_images_dev(ctx, p, nb, images_dev + (long long)b0 * N, ctx->parts.as<float>(), ft, 0, 1);
            if (rc) return rc;
            CIM_ALLOC(sb.ftrow, sizeof(float2) * (size_t)nb * box * Dc);
            rc = fft2d_crop_dev(ctx, ft, box, nb, o0, Dc, o0, Dc, 1, nullptr, sb.ftrow.as<float2>(),
                                sb.pcrop.as<float2>());
            if (rc) return rc;
            rc = amp_expfit_general(ctx, sb.pcrop.as<float2>(), (long long)Dc * Dc, Dc, R + 1, nb, p->minR, R,
                                    sb.fcut0.as<float>());
            if (rc) return rc;
            hipLaunchKernelGGL(k_crop_shift, dim3((Dc * Dc + 255) / 256, nb), dim3(256), 0, ctx->stream,
                               sb.pcrop.as<float2>(), box, R, Dc, ctx->parts.as<float>());
            CIM_HIP(hipGetLastError());
        } else {
            CIM_ALLOC(ctx->ft, sizeof(float2) * N * nb);
            float2* ft = ctx->ft.as<float2>();
            rc = preprocess_images(ctx, p, nb, images_dev + (long long)b0 * N, ctx->parts.as<float>(), ft, 0, 1, 1);
            if (rc) return rc;
            rc = amp_expfit(ctx, ft, box, nb, p->minR, R, sb.fcut0.as<float>());
            if (rc) return rc;
            // the centre shift is applied to the crop's pixels only
            hipLaunchKernelGGL(k_crop, dim3((Dc * Dc + 255) / 256, nb), dim3(256), 0, ctx->stream, ft, N, box, R, Dc,
                               sb.pcrop.as<float2>(), 1, (const float*)nullptr, cc, 0, ctx->parts.as<float>());
            CIM_HIP(hipGetLastError());
        }
        CIM_ALLOC(sb.pcrop4, sizeof(float4) * (size_t)Dc * Dc * nb);
        {
            const long long tot = (long long)Dc * Dc * nb;
            hipLaunchKernelGGL(k_pair_taps, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream,
                               sb.pcrop.as<float2>(), sb.pcrop4.as<float4>(), Dc, nb);
            CIM_HIP(hipGetLastError());
        }
        SearchArgs a{};
        if (p->apply_ctf) {
            CIM_ALLOC(sb.tcrop, sizeof(float2) * (size_t)Dc * Dc * nb * T);
            hipLaunchKernelGGL(k_crop, dim3((Dc * Dc + 255) / 256, nb), dim3(256), 0, ctx->stream,
                               s.tft.as<float2>(), N, box, R, Dc, sb.tcrop.as<float2>(), T, ctx->parts.as<float>(), cc,
                               1, (const float*)nullptr);
            CIM_HIP(hipGetLastError());
            CIM_ALLOC(sb.tfcut, sizeof(float) * nr * nb * T);
            rc = amp_expfit_general(ctx, sb.tcrop.as<float2>(), (long long)Dc * Dc, Dc, R + 1, nb * T, p->minR, R,
                                    sb.tfcut.as<float>());
            if (rc) return rc;
            a.tc_pstride = (long long)T * Dc * Dc;
            a.tf_pstride = (long long)T * nr;
        } else {
            CIM_ALLOC(sb.tcrop, sizeof(float2) * (size_t)Dc * Dc * T);
            hipLaunchKernelGGL(k_crop, dim3((Dc * Dc + 255) / 256, T), dim3(256), 0, ctx->stream, s.tft.as<float2>(), N,
                               box, R, Dc, sb.tcrop.as<float2>(), 1, (const float*)nullptr, cc, 0,
                               (const float*)nullptr);
            CIM_HIP(hipGetLastError());
            CIM_ALLOC(sb.tfcut, sizeof(float) * nr * T);
            CIM_HIP(hipMemcpyAsync(sb.tfcut.ptr, s.tfcut.ptr, sizeof(float) * nr * T, hipMemcpyDeviceToDevice,
                                   ctx->stream));
            a.tc_pstride = 0;
            a.tf_pstride = 0;
        }
        // the exact scores: one per AU point, or per (candidate, window part) (at most 4 x 256)
        const size_t nout = std::max((size_t)npts, (size_t)4 * 256);
        CIM_ALLOC(sb.outv, sizeof(float) * nout * nb);
        CIM_ALLOC(sb.outi, sizeof(int) * nout * nb);
        if (pr_array) CIM_ALLOC(sb.prfull, sizeof(float) * (size_t)npts * g.No * nb);
        a.pcrop = sb.pcrop.as<float2>();
        a.pcrop4 = sb.pcrop4.as<float4>();
        a.fcut0 = sb.fcut0.as<float>();
        a.tcrop = sb.tcrop.as<float2>();
        a.tfcut = sb.tfcut.as<float>();
        a.tang = sb.tang.as<float>();
        a.T = T;
        a.pts = sb.pts.as<float>();
        a.pbase = sb.pbase.as<int>();
        a.npts = npts;
        a.ops = sb.ops.as<float>();
        a.nops = 60;
        a.rows = sb.rows.as<int>();
        a.nrows = nrows;
        a.disk_total = disk_total;
        a.R = R;
        a.minR = p->minR;
        a.nr = nr;
        a.Dc = Dc;
        a.Nomega = g.No;
        a.box = box;
        a.chunk = chunk;
        a.step = p->searchstep;
        a.outv = sb.outv.as<float>();
        a.outi = sb.outi.as<int>();
        a.prfull = pr_array ? sb.prfull.as<float>() : nullptr;
        // counters: samples of exhaustive launches, re-score gate passes, re-score samples, the 5 rank
        // bins of k_fb_select
        CIM_ALLOC(sb.counter, 9 * sizeof(unsigned long long));
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, sb.counter.as<unsigned long long>(), 9);
        CIM_HIP(hipGetLastError());
        a.nsamples = sb.counter.as<unsigned long long>();
        a.npass = nullptr;
        SearchShape sh{};
        rc = tune_shape(ctx, a, lds_with, lds_without, lds_fits, g.No, pair, &sh);
        if (rc) return rc;
        ctx->counters["cl_search_vpt"] = sh.vpt;
        ctx->counters["cl_search_lds"] = sh.lds;
        ctx->counters["cl_search_pair"] = sh.pair;
        const int bd_exact = block_for(pair ? g.No / 2 : g.No, sh.vpt);
        int nscore = npts;   // AU points per particle whose exact scores the final argmin reads
        if (tables) {
            const int titems = p->apply_ctf ? nb * T : T;
            CIM_ALLOC(sb.ppol, sizeof(float2) * (size_t)nb * nr * npsi);
            CIM_ALLOC(sb.tpol, sizeof(float2) * (size_t)titems * nr * npsi);
            const bool cgrid = c2f && coarse_grid;
            // ring maxima of the particle and template rays (the coarse tables' D bounds)
            float* rmax = nullptr;
            if (cgrid) {
                CIM_ALLOC(sb.rmax, sizeof(float) * (size_t)(nb + titems) * nr);
                rmax = sb.rmax.as<float>();
            }
            {
                TimedLaunch tl(ctx, "cl_polar");
                CIM_ALLOC(sb.rays, sizeof(float2) * (size_t)npsi);
                hipLaunchKernelGGL(k_rays, dim3((npsi + 255) / 256), dim3(256), 0, ctx->stream, npsi, dpsi,
                                   sb.rays.as<float2>());
                CIM_HIP(hipGetLastError());
                hipLaunchKernelGGL(k_polar, dim3(nr, nb), dim3(256), 0, ctx->stream, sb.pcrop.as<float2>(),
                                   (long long)Dc * Dc, Dc, R, box, p->minR, nr, npsi, sb.rays.as<float2>(),
                                   sb.fcut0.as<float>(), (long long)nr, sb.ppol.as<float2>(), rmax);
                CIM_HIP(hipGetLastError());
                hipLaunchKernelGGL(k_polar, dim3(nr, titems), dim3(256), 0, ctx->stream, sb.tcrop.as<float2>(),
                                   (long long)Dc * Dc, Dc, R, box, p->minR, nr, npsi, sb.rays.as<float2>(),
                                   sb.tfcut.as<float>(), (long long)nr, sb.tpol.as<float2>(),
                                   rmax ? rmax + (size_t)nb * nr : nullptr);
                CIM_HIP(hipGetLastError());
            }
            CIM_ALLOC(sb.table, sizeof(unsigned int) * (size_t)nb * T * (npsi / 2) * npsi);
            {
                TimedLaunch tl(ctx, "cl_table");
                rc = launch_cl_table(ctx, table_nh, sb.ppol.as<float2>(), sb.tpol.as<float2>(), p->apply_ctf ? 1 : 0,
                                     nb, T, nr, npsi, sb.table.as<unsigned int>());
                if (rc) return rc;
            }
            if (cgrid) {
                // coarse E4M3 tables of the even rays, scaled per particle by the D bounds
                CIM_ALLOC(sb.bound, sizeof(float) * (size_t)(nb + titems));
                CIM_ALLOC(sb.ctab, sizeof(unsigned short) * (size_t)nb * T * (npsi / 4) * (npsi / 2));
                TimedLaunch tl(ctx, "cl_coarse_table");
                hipLaunchKernelGGL(k_ring_sum, dim3((nb + titems + 255) / 256), dim3(256), 0, ctx->stream,
                                   sb.rmax.as<float>(), nr, nb + titems, sb.bound.as<float>());
                CIM_HIP(hipGetLastError());
                hipLaunchKernelGGL(k_coarse_table, dim3((unsigned)((npsi / 2 + 255) / 256), npsi / 4, nb * T),
                                   dim3(256), 0, ctx->stream, sb.table.as<unsigned int>(), sb.bound.as<float>(),
                                   sb.bound.as<float>() + nb, p->apply_ctf ? 1 : 0, T, npsi,
                                   sb.ctab.as<unsigned short>());
                CIM_HIP(hipGetLastError());
            }
            GatherArgs ga{sb.table.ptr, sb.lines.as<unsigned int>(), sb.lcount.as<int>(), sb.pbase.as<int>(), T, npts,
                          npsi, sub, g.No, 1, nb, sb.outv.as<float>(), sb.outi.as<int>(), nullptr, 0, 0, nullptr,
                          nullptr, nullptr, g.No, nullptr, 0};
            const bool fwin = cgrid && fine_win > 0 && fine_win < g.No;
            // one workgroup of 64 x ceil(ncol / (64 VPT)) threads per (AU point, particle)
            auto gather = [&](int nlist, bool coarse) -> int {
                const int vg = (ga.ncol + 255) / 256;
                const int vpt = vg <= 1 ? 1 : vg == 2 ? 2 : vg == 3 ? 3 : vg == 4 ? 4 : vg <= 6 ? 6 : 8;
                const dim3 gg((unsigned)(8LL * nlist * ((nb + 7) / 8)));
                const dim3 blk((unsigned)(64 * ((ga.ncol + 64 * vpt - 1) / (64 * vpt))));
#define CIM_GATHER(V)                                                                                            \
    if (coarse) hipLaunchKernelGGL((k_cl_gather<V, unsigned short>), gg, blk, 0, ctx->stream, ga);             \
    else hipLaunchKernelGGL((k_cl_gather<V, unsigned int>), gg, blk, 0, ctx->stream, ga);
                if (vpt == 1) { CIM_GATHER(1) }
                else if (vpt == 2) { CIM_GATHER(2) }
                else if (vpt == 3) { CIM_GATHER(3) }
                else if (vpt == 4) { CIM_GATHER(4) }
                else if (vpt == 6) { CIM_GATHER(6) }
                else { CIM_GATHER(8) }
#undef CIM_GATHER
                CIM_HIP(hipGetLastError());
                return CIM_OK;
            };
            // second windows: the fine scores also keep each point's best omega at least a window
            // from its first minimum (in the re-score's column space), with its score
            auto second_minima = [&]() -> int {
                const long long ns = (long long)nb * npts;
                CIM_ALLOC(sb.outi2, sizeof(int) * (size_t)ns);
                CIM_ALLOC(sb.outv2, sizeof(float) * (size_t)ns);
                hipLaunchKernelGGL(k_fill_scores, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, ctx->stream,
                                   sb.outv2.as<float>(), sb.outi2.as<int>(), ns);
                CIM_HIP(hipGetLastError());
                ga.outi2 = sb.outi2.as<int>();
                ga.outv2 = sb.outv2.as<float>();
                ga.excl = wincols;
                ga.exmod = ncolx;
                return CIM_OK;
            };
            {
                TimedLaunch tl(ctx, "cl_gather");
                if (c2f) {
                    // coarse points, the M best per particle, their unscored neighbours
                    const long long ns = (long long)nb * npts;
                    hipLaunchKernelGGL(k_fill_scores, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, ctx->stream,
                                       sb.outv.as<float>(), sb.outi.as<int>(), ns);
                    CIM_HIP(hipGetLastError());
                    ga.plist = sb.clist.as<int>();
                    ga.plstride = 0;
                    ga.nlist = (int)clist.size();
                    if (cgrid) {
                        // the grid of twice the step: half the rays, every other omega
                        GatherArgs gc = ga;
                        ga.table = sb.ctab.ptr;
                        ga.lines = sb.lines.as<unsigned int>() + (size_t)npts * T * 64;
                        ga.nray = npsi / 2;
                        ga.ncol = g.No / 2;
                        ga.ostep = 2;
                        if (fwin) {
                            CIM_ALLOC(sb.outi2, sizeof(int) * (size_t)nb * npts);
                            ga.outi2 = sb.outi2.as<int>();
                            ga.excl = fine_win / 2;
                        }
                        if ((rc = gather(ga.nlist, true))) return rc;
                        ga = gc;
                    } else if ((rc = gather(ga.nlist, false))) {
                        return rc;
                    }
                    // the coarse winners, each with two windows of fine_win / 2 omega columns: around
                    // its coarse minimum and around its best omega away from it
                    hipLaunchKernelGGL(npts <= 256 * TOPK_PER ? k_topk : k_topk_scan, dim3(nb), dim3(256), 0,
                                       ctx->stream, sb.outv.as<float>(), sb.outi.as<int>(), npts, coarseM,
                                       sb.ccand.as<int>(), g.No, g.No, fwin ? fine_win / 2 : 0, sb.ckwin.as<int>(),
                                       coarseM, 0);
                    CIM_HIP(hipGetLastError());
                    if (fwin) {
                        hipLaunchKernelGGL(k_second_win, dim3(nb), dim3(64), 0, ctx->stream, sb.ccand.as<int>(),
                                           coarseM, sb.outi2.as<int>(), sb.outi.as<int>(), npts, g.No, fine_win / 2,
                                           sb.ckwin2.as<int>());
                        CIM_HIP(hipGetLastError());
                        ga.outi2 = nullptr;
                    }
                    hipLaunchKernelGGL(k_expand, dim3(nb), dim3(256), 0, ctx->stream, sb.ccand.as<int>(), coarseM,
                                       sb.nbr.as<int>(), sb.outv.as<float>(), npts, sb.flist.as<int>(), coarseM * 9,
                                       sb.fcnt.as<int>(), cgrid ? 1 : 0, sb.ckwin.as<int>(),
                                       fwin ? sb.fwin.as<int>() : nullptr, sb.ckwin2.as<int>(),
                                       fwin ? sb.fwin2.as<int>() : nullptr);
                    CIM_HIP(hipGetLastError());
                    ga.plist = sb.flist.as<int>();
                    ga.plstride = coarseM * 9;
                    ga.nlist = coarseM * 9;
                    ga.pcnt = sb.fcnt.as<int>();
                    if (fwin) {
                        ga.pwin = sb.fwin.as<int>();
                        ga.pwin2 = sb.fwin2.as<int>();
                        ga.ncol = fine_win;
                    }
                    if (nsecond && (rc = second_minima())) return rc;
                    if ((rc = gather(ga.nlist, false))) return rc;
                } else {
                    if (nsecond && (rc = second_minima())) return rc;
                    if ((rc = gather(npts, false))) return rc;
                }
            }
            const int ncand = topk + nsecond;
            CIM_ALLOC(sb.cand, sizeof(int) * (size_t)nb * ncand);
            CIM_ALLOC(sb.kwin, sizeof(int) * (size_t)nb * ncand);
            hipLaunchKernelGGL(npts <= 256 * TOPK_PER ? k_topk : k_topk_scan, dim3(nb), dim3(256), 0, ctx->stream,
                               sb.outv.as<float>(), sb.outi.as<int>(), npts, topk, sb.cand.as<int>(), g.No, ncolx,
                               wincols, sb.kwin.as<int>(), ncand, 0);
            CIM_HIP(hipGetLastError());
            if (nsecond) {
                // the nsecond best second minima, after the topk first ones
                hipLaunchKernelGGL(npts <= 256 * TOPK_PER ? k_topk : k_topk_scan, dim3(nb), dim3(256), 0,
                                   ctx->stream, sb.outv2.as<float>(), sb.outi2.as<int>(), npts, nsecond,
                                   sb.cand.as<int>(), g.No, ncolx, wincols, sb.kwin.as<int>(), ncand, topk);
                CIM_HIP(hipGetLastError());
                ga.outi2 = nullptr;
                ga.outv2 = nullptr;
            }
            ctx->counters["cl_table_pairs"] += (long long)nb * T * npsi * npsi * nr;
            if (!c2f) ctx->counters["cl_gather_points"] += (long long)nb * npts * g.No;
            a.cand = sb.cand.as<int>();
            a.ncand = ncand;
            a.prfull = nullptr;
            if (wincols < ncolx) {
                a.kwin = sb.kwin.as<int>();
                a.wincols = wincols;
                // each window on wsplit workgroups of wincols / wsplit columns
                // (CIM_RESCORE_SPLIT=<1|2|4>): more waves, but each lane-idle wave costs issue
                // slots; measured at box 384, 512 particles: 85 ms (1), 93 ms (2), 136 ms (4)
                a.wsplit = 1;
                if (const char* es = getenv("CIM_RESCORE_SPLIT")) a.wsplit = std::max(1, std::min(4, atoi(es)));
                while (a.wsplit > 1 && wincols % a.wsplit) --a.wsplit;
                // one-wave workgroups: a short line chunk keeps LDS small so many stay resident
                // (the chunking does not change any column's accumulation order)
                a.chunk = std::min(chunk, 6);
            }
            nscore = ncand * (a.kwin && rescore_cpw == 1 ? a.wsplit : 1);
        }
        if (ucand) {
            // the caller's candidates: AU point, and the omega column its window is centred on
            std::vector<int> hc((size_t)nb * ucn), hw((size_t)nb * ucn);
            for (size_t i = 0; i < hc.size(); ++i) {
                const int q = ucand[(size_t)b0 * ucn + i], kc = uomega[(size_t)b0 * ucn + i];
                if (q < 0 || q >= npts || kc < 0 || kc >= g.No)
                    return set_error(ctx, CIM_E_ARG, "candidate %zu: AU point %d (of %d) or omega %d (of %d)", i, q,
                                     npts, kc, g.No);
                hc[i] = q;
                hw[i] = ((kc % ncolx - wincols / 2) % ncolx + ncolx) % ncolx;
            }
            CIM_ALLOC(sb.cand, sizeof(int) * hc.size());
            CIM_ALLOC(sb.kwin, sizeof(int) * hw.size());
            CIM_HIP(hipMemcpyAsync(sb.cand.ptr, hc.data(), sizeof(int) * hc.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
            CIM_HIP(hipMemcpyAsync(sb.kwin.ptr, hw.data(), sizeof(int) * hw.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
            a.cand = sb.cand.as<int>();
            a.ncand = ucn;
            a.prfull = nullptr;
            a.kwin = sb.kwin.as<int>();
            a.wincols = wincols;
            a.wsplit = 1;
            a.chunk = std::min(chunk, 6);
            nscore = ucn;
        }
        {
            TimedLaunch tl(ctx, tables || ucand ? "cl_rescore" : "cl_search");
            // windows: one wave of single-column lanes in the same kernel variant
            if (a.kwin && rescore_cpw > 1) {
                // lines whose template records are staged per pass (CIM_RESCORE_CHUNK; the
                // chunking does not change any column's order of terms): the LDS per one-wave
                // workgroup sets how many of them stay resident.  One line: 7.6 KB at maxR 72, so
                // 4 waves per SIMD (a batch of 512 is 16 workgroups per CU, all resident at once);
                // measured at box 384: 6.4 ms per 512 particles against 7.9 (two lines) and 9.0 (three)
                int rchunk = 1;
                if (const char* er = getenv("CIM_RESCORE_CHUNK")) rchunk = std::max(1, std::min(6, atoi(er)));
                a.chunk = std::min(chunk, rchunk);
                a.nsamples = sb.counter.as<unsigned long long>() + 2;
                a.npass = sb.counter.as<unsigned long long>() + 1;
                const int lw = (int)(sizeof(float4) * (size_t)rescore_cpw * a.chunk * nr +
                                     sizeof(float) * (2 * 64 * rescore_cpw + nr) +
                                     sizeof(int) * (size_t)rescore_cpw * a.chunk +
                                     sizeof(unsigned short) * (size_t)rescore_cpw * a.chunk * nr);
                const dim3 grid((unsigned)(((a.ncand + rescore_cpw - 1) / rescore_cpw) * 8LL * ((nb + 7) / 8)));
                a.nbatch = nb;
#define CIM_RESCORE(C, P) hipLaunchKernelGGL((k_cl_rescore<C, P>), grid, dim3(64), lw, ctx->stream, a)
                if (rescore_cpw == 2) { if (sh.pair) CIM_RESCORE(2, true); else CIM_RESCORE(2, false); }
                else if (rescore_cpw == 4) { if (sh.pair) CIM_RESCORE(4, true); else CIM_RESCORE(4, false); }
                else { if (sh.pair) CIM_RESCORE(8, true); else CIM_RESCORE(8, false); }
#undef CIM_RESCORE
                CIM_HIP(hipGetLastError());
                a.nsamples = sb.counter.as<unsigned long long>();
                a.npass = nullptr;
            } else if (a.kwin) {
                const int lw = (int)(misc + sizeof(float4) * (size_t)a.chunk * nr) +
                               (sh.lds ? (int)(sizeof(float2) * (size_t)disk_total) : 0);
                rc = launch_search(ctx, a, nb, lw, sh.lds ? 0 : 1, sh.pair, 1, wincols / a.wsplit);
            }
            else rc = launch_search(ctx, a, nb, sh.lds_bytes, sh.lds ? 0 : 1, sh.pair, sh.vpt, bd_exact);
            if (rc) return rc;
        }
        CIM_ALLOC(sb.fv, sizeof(float) * nb);
        CIM_ALLOC(sb.fi, sizeof(int) * nb);
        hipLaunchKernelGGL(k_argmin, dim3(nb), dim3(256), 0, ctx->stream, sb.outv.as<float>(), sb.outi.as<int>(),
                           nscore, sb.fv.as<float>(), sb.fi.as<int>());
        CIM_HIP(hipGetLastError());
        if (tables && fb_rank <= topk) {
            // Fallback: when the exact minimum sits deep in the candidate list (or at its window's edge),
            // the tables ranked the particle poorly; score every AU point of it exactly instead (the
            // plain search).  Decided and run on the device: the listed particles go through one
            // exhaustive launch, FBS of them at a time.
            CIM_ALLOC(sb.fbl, sizeof(int) * (size_t)(nb + 2));
            int* fbcnt = sb.fbl.as<int>();
            int* fblist = fbcnt + 2;
            hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream,
                               reinterpret_cast<unsigned long long*>(fbcnt), 1);
            hipLaunchKernelGGL(k_fb_select, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream, sb.fi.as<int>(),
                               sb.cand.as<int>(), a.kwin, sb.pbase.as<int>(), nb, topk, a.ncand, g.No, ncolx, wincols,
                               fb_rank, sb.counter.as<unsigned long long>() + 3, fblist, fbcnt);
            CIM_HIP(hipGetLastError());
            SearchArgs a1 = a;
            a1.cand = nullptr;
            a1.ncand = 0;
            a1.kwin = nullptr;      // every omega column of every AU point
            a1.wsplit = 1;
            a1.fb_list = fblist;
            a1.fb_count = fbcnt;
            CIM_ALLOC(sb.cv, sizeof(float) * (size_t)npts * nb);
            CIM_ALLOC(sb.ci, sizeof(int) * (size_t)npts * nb);
            a1.outv = sb.cv.as<float>();
            a1.outi = sb.ci.as<int>();
            {
                TimedLaunch tl(ctx, "cl_fallback");
                rc = launch_search(ctx, a1, std::min(nb, FB_CONCURRENT), sh.lds_bytes, sh.lds ? 0 : 1, sh.pair,
                                   sh.vpt, bd_exact);
                if (rc) return rc;
            }
            hipLaunchKernelGGL(k_fb_merge, dim3(nb), dim3(256), 0, ctx->stream, sb.cv.as<float>(), sb.ci.as<int>(),
                               npts, fblist, fbcnt, sb.fv.as<float>(), sb.fi.as<int>());
            CIM_HIP(hipGetLastError());
        }
        const int plane = g.No * (2 * g.Np + 1);
        if (devmode) {
            hipLaunchKernelGGL(k_records_out, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream, sb.fv.as<float>(),
                               sb.fi.as<int>(), nb, plane, g.No, g.Np, p->searchstep,
                               parts_dev + (size_t)b0 * CIM_PART_STRIDE);
            CIM_HIP(hipGetLastError());
            continue;
        }
        std::vector<float> hv(nb);
        std::vector<int> hi(nb);
        CIM_HIP(hipMemcpyAsync(hv.data(), sb.fv.ptr, sizeof(float) * nb, hipMemcpyDeviceToHost, ctx->stream));
        CIM_HIP(hipMemcpyAsync(hi.data(), sb.fi.ptr, sizeof(int) * nb, hipMemcpyDeviceToHost, ctx->stream));
        std::vector<int> hfc;
        if (c2f) {
            hfc.resize(nb);
            CIM_HIP(hipMemcpyAsync(hfc.data(), sb.fcnt.ptr, sizeof(int) * nb, hipMemcpyDeviceToHost, ctx->stream));
        }
        std::vector<float> full;
        if (pr_array) {
            full.resize((size_t)npts * g.No);
            CIM_HIP(hipMemcpyAsync(full.data(), sb.prfull.ptr, sizeof(float) * full.size(), hipMemcpyDeviceToHost,
                                   ctx->stream));
        }
        int hfb = 0;
        if (tables && fb_rank <= topk)
            CIM_HIP(hipMemcpyAsync(&hfb, sb.fbl.ptr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        CIM_HIP(hipStreamSynchronize(ctx->stream));
        if (c2f) {
            long long scored = 0;
            for (int c : hfc) scored += c;
            const bool fwin = coarse_grid && fine_win > 0 && fine_win < g.No;
            ctx->counters["cl_gather_points"] += scored * (fwin ? fine_win : g.No) +
                                                 (long long)nb * (long long)clist.size() *
                                                     (coarse_grid ? g.No / 2 : g.No);
        }
        ctx->counters["cl_fallback_particles"] += hfb;
        unsigned long long hcount[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        CIM_HIP(hipMemcpyAsync(hcount, sb.counter.ptr, sizeof(hcount), hipMemcpyDeviceToHost, ctx->stream));
        CIM_HIP(hipStreamSynchronize(ctx->stream));
        ctx->counters["cl_search_samples"] += (long long)(hcount[0] + hcount[2]);
        ctx->counters["cl_rescore_samples"] += (long long)hcount[2];
        ctx->counters["cl_rescore_pass"] += (long long)hcount[1];
        if (tables && fb_rank <= topk) {
            // where each particle's exact minimum ranked among the table's candidates
            const char* bins[6] = {"cl_rank_lt4", "cl_rank_lt8", "cl_rank_lt16", "cl_rank_lt24", "cl_rank_ge24",
                                   "cl_rank_second"};
            for (int b = 0; b < 6; ++b) ctx->counters[bins[b]] += (long long)hcount[3 + b];
        }
        ctx->counters["cl_search_particles"] += nb;
        for (int i = 0; i < nb; ++i) {
            float* q = parts + (size_t)(b0 + i) * CIM_PART_STRIDE;
            const int idx = hi[i];
            const int it = idx / plane, ip = (idx % plane) / g.No - g.Np, io = idx % g.No;
            q[CIM_P_THETA] = 69.0f + p->searchstep * (float)it;
            q[CIM_P_PHI] = p->searchstep * (float)ip;
            q[CIM_P_OMEGA] = p->searchstep * (float)io;
            q[CIM_P_PR] = hv[i];
        }
        if (pr_array) {
            const size_t total = (size_t)plane * (g.Nt + 1);
            for (size_t i = 0; i < total; ++i) pr_array[i] = 1000.f;
            for (int pt = 0; pt < npts; ++pt)
                for (int k = 0; k < g.No; ++k) pr_array[g.base[pt] + k] = full[(size_t)pt * g.No + k];
        }
    }
    // the device-records search reuses the constant tables of the last tabulated host-records call
    if (!devmode) ctx->srch.step_sig = tables ? std::string(sig) : std::string();
    return CIM_OK;
}

int search_records_dev(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev, float* parts_dev) {
    return search_global_dev(ctx, p, n, images_dev, nullptr, nullptr, parts_dev);
}

int search_candidates_dev(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev, float* parts,
                          int ncand, const int* cand, const int* omega) {
    return search_global_dev(ctx, p, n, images_dev, parts, nullptr, nullptr, cand, omega, ncand);
}

// ---- local refinement (mode 5) ------------------------------------------------

struct RefineArgs {
    const float2* ft;        // n x box^2 particle FTs (masked, FFT, no centre shift)
    const float* fcut0;      // n x nr
    const float2* tft;       // template FTs; particle p, template t at tft + p*t_pstride + t*tstride
    long long t_pstride, tstride;
    int t_ld, t_c;           // template rows and centre index (box, box/2 or crop)
    const float* tfcut;      // p*tf_pstride + t*nr
    long long tf_pstride;
    const float* tang;
    int T;
    const float* ops;        // 60 x 9
    float* parts;            // n x CIM_PART_STRIDE in/out
    int box, minR, nr, ncycle;
    float dang0, dtran0;
};

// return_icos_asymmetry_unit.f90:14-31 (float32)
__device__ void return_au_f(const float* ops, float* th, float* ph, float* om) {
    const float d2r = 0.0174532925199432957692f;
    const float s31 = sinf(31.717f * d2r), c31 = cosf(31.717f * d2r), s69 = sinf(69.090f * d2r),
                c69 = cosf(69.090f * d2r);
    const float A1 = -s31 * c69, B1 = -c31 * c69, C1 = s31 * s69, A2 = -s31 * c69, B2 = c31 * c69, C2 = s31 * s69;
    float E[9];
    euler_f(*th, *ph, *om, E);
    float t = 0.f, p = 0.f, o = 0.f;
    for (int i = 0; i < 60; ++i) {
        const float* S = ops + 9 * i;
        float M[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) M[3 * r + c] = S[3 * r] * E[c] + S[3 * r + 1] * E[3 + c] + S[3 * r + 2] * E[6 + c];
        t = acosf(M[8]);
        p = atan2f(M[5], M[2]);
        o = atan2f(M[7], -M[6]);
        const float x = sinf(t) * cosf(p), y = sinf(t) * sinf(p), z = cosf(t);
        const float z1 = -(A1 * x + B1 * y) / C1, z2 = -(A2 * x + B2 * y) / C2;
        if ((z >= 0.f) && (z <= z1) && (z <= z2)) break;
    }
    *th = t / d2r;
    *ph = p / d2r;
    *om = o / d2r;
}

__global__ void __launch_bounds__(256) k_refine(RefineArgs a) {
    __shared__ float cand[11][5];
    __shared__ float al1[11][64], al2[11][64];
    __shared__ float acc_n[11][4], acc_d[11][4];
    __shared__ float cur[5];
    __shared__ float fc0s[520];
    const int p = blockIdx.x, tid = threadIdx.x;
    const int ic = a.box / 2;
    const float cf = (float)ic;
    const float d2r = 0.0174532925199432957692f;
    const float2* pf = a.ft + (long long)p * a.box * a.box;
    float* rec = a.parts + (long long)p * CIM_PART_STRIDE;
    for (int i = tid; i < a.nr; i += blockDim.x) fc0s[i] = a.fcut0[(long long)p * a.nr + i];
    if (tid < 5) cur[tid] = rec[tid];
    __syncthreads();
    float PR0 = 0.f;
    float f = 1.f;
    for (int cyc = 1; cyc <= a.ncycle; ++cyc) {
        const float dang = a.dang0 * f, dtran = a.dtran0 * f;
        f *= 0.8f;
        // ort_cent_11.f90:6-91
        if (tid == 0) {
            float th0 = cur[0], ph0 = cur[1], om0 = cur[2];
            return_au_f(a.ops, &th0, &ph0, &om0);
            for (int jj = 0; jj <= 10; ++jj) {
                float o[5] = {th0, ph0, om0, cur[3], cur[4]};
                switch (jj) {
                    case 1: o[0] = th0 - dang; break;
                    case 2: o[0] = th0 + dang; break;
                    case 3: o[1] = ph0 - dang; break;
                    case 4: o[1] = ph0 + dang; break;
                    case 5: o[2] = om0 - dang; break;
                    case 6: o[2] = om0 + dang; break;
                    case 7: o[3] = cur[3] - dtran; break;
                    case 8: o[3] = cur[3] + dtran; break;
                    case 9: o[4] = cur[4] - dtran; break;
                    case 10: o[4] = cur[4] + dtran; break;
                }
                cand[jj][0] = o[0]; cand[jj][1] = o[1]; cand[jj][2] = o[2]; cand[jj][3] = o[3]; cand[jj][4] = o[4];
            }
        }
        __syncthreads();
        if (tid >= 1 && tid <= 10) return_au_f(a.ops, &cand[tid][0], &cand[tid][1], &cand[tid][2]);
        __syncthreads();
        float num[11], den[11];
        for (int c = 0; c < 11; ++c) { num[c] = 0.f; den[c] = 0.f; }
        for (int t = 0; t < a.T; ++t) {
            const float2* tf = a.tft + (long long)p * a.t_pstride + (long long)t * a.tstride;
            const float* tfc = a.tfcut + (long long)p * a.tf_pstride + (long long)t * a.nr;
            for (int q = tid; q < 11 * 60; q += blockDim.x) {
                const int c = q / 60, j = q % 60;
                float E1[9], E2[9];
                euler_f(cand[c][0], cand[c][1], cand[c][2], E1);
                euler_f(a.tang[3 * t], a.tang[3 * t + 1], a.tang[3 * t + 2], E2);
                crosslin_one(E1, E2, a.ops + 9 * j, &al1[c][j], &al2[c][j]);
            }
            __syncthreads();
            // samples: (line j, radius) per thread, all 11 candidates
            for (int q = tid; q < 60 * a.nr; q += blockDim.x) {
                const int j = q / a.nr, fr = q % a.nr;
                const float FR = (float)(a.minR + fr);
                for (int c = 0; c < 11; ++c) {
                    const float A1 = al1[c][j], A2 = al2[c][j];
                    if (!(A1 < 999.f && A2 < 999.f)) continue;
                    // particle tap with the trial shift (cal_crosslin_PR.f90:22,41)
                    float x = FR * cosf(A1 * d2r) + cf, y = FR * sinf(A1 * d2r) + cf;
                    int lx = (int)x, ly = (int)y;
                    float ax = x - (float)lx, ay = y - (float)ly;
                    float2 zt[4];
                    const int dx[4] = {0, 1, 0, 1}, dy[4] = {0, 0, 1, 1};
                    for (int k = 0; k < 4; ++k) {
                        const int X = lx + dx[k], Y = ly + dy[k];
                        float2 v = pf[(long long)Y * a.box + X];
                        float2 cs = shift_phase(X - ic, Y - ic, a.box, cand[c][3], cand[c][4], 1);
                        zt[k] = apply_shift(v, cs);
                    }
                    const float w00 = (1.f - ax) * (1.f - ay), w10 = ax * (1.f - ay), w01 = (1.f - ax) * ay,
                                w11 = ax * ay;
                    float re1 = w00 * zt[0].x + w10 * zt[1].x + w01 * zt[2].x + w11 * zt[3].x;
                    float im1 = w00 * zt[0].y + w10 * zt[1].y + w01 * zt[2].y + w11 * zt[3].y;
                    // template tap
                    x = FR * cosf(A2 * d2r) + cf;
                    y = FR * sinf(A2 * d2r) + cf;
                    lx = (int)x;
                    ly = (int)y;
                    ax = x - (float)lx;
                    ay = y - (float)ly;
                    const int tx = lx - ic + a.t_c, ty = ly - ic + a.t_c;
                    const float2* r0 = tf + (long long)ty * a.t_ld + tx;
                    const float2* r1 = r0 + a.t_ld;
                    const float u00 = (1.f - ax) * (1.f - ay), u10 = ax * (1.f - ay), u01 = (1.f - ax) * ay,
                                u11 = ax * ay;
                    float re2 = u00 * r0[0].x + u10 * r0[1].x + u01 * r1[0].x + u11 * r1[1].x;
                    float im2 = u00 * r0[0].y + u10 * r0[1].y + u01 * r1[0].y + u11 * r1[1].y;
                    const float amp1 = sqrtf(re1 * re1 + im1 * im1), amp2 = sqrtf(re2 * re2 + im2 * im2);
                    if (amp1 >= fc0s[fr] && amp2 >= tfc[fr]) {
                        const float diff = fold_phase(atan2f(im2, re2) - atan2f(im1, re1));
                        const float w = 0.5f * (amp1 + amp2);
                        num[c] += diff * w;
                        den[c] += w;
                    }
                }
            }
            __syncthreads();
        }
        // block reduction of the 11 (num, den) pairs
        for (int c = 0; c < 11; ++c) {
            float nv = wave_sum(num[c]), dv = wave_sum(den[c]);
            if ((tid & 63) == 0) { acc_n[c][tid >> 6] = nv; acc_d[c][tid >> 6] = dv; }
        }
        __syncthreads();
        if (tid == 0) {
            float best = 0.f;
            int bi = 0;
            for (int c = 0; c < 11; ++c) {
                float nn = 0.f, dd = 0.f;
                for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { nn += acc_n[c][w]; dd += acc_d[c][w]; }
                const float pr = dd > 1.0e-10f ? nn / dd * 57.2957795130823208768f : 1000.f;
                if (c == 0 || pr < best) { best = pr; bi = c; }
            }
            PR0 = best;
            for (int k = 0; k < 5; ++k) cur[k] = cand[bi][k];
            acc_n[0][0] = best;
        }
        __syncthreads();
    }
    if (tid == 0) {
        for (int k = 0; k < 5; ++k) rec[k] = cur[k];
        rec[CIM_P_PR] = PR0;
    }
}

static int refine_local(cim_ctx* ctx, const cim_params* p, int n, const float* images, float* parts) {
    SearchState& s = ctx->srch;
    if (!s.ready) return set_error(ctx, CIM_E_STATE, "cim_search_templates has not been called");
    if (p->box != s.box || p->minR != s.minR || p->maxR != s.maxR)
        return set_error(ctx, CIM_E_ARG, "box/minR/maxR differ from the template bank");
    if (n <= 0) return CIM_OK;
    const int box = p->box, R = p->maxR, nr = R - p->minR + 1, T = s.T;
    if (nr > 520) return set_error(ctx, CIM_E_ARG, "too many radii (%d)", nr);
    const long long N = (long long)box * box;
    CIM_ALLOC(ctx->img, sizeof(float) * N * n);
    CIM_HIP(hipMemcpyAsync(ctx->img.ptr, images, sizeof(float) * N * n, hipMemcpyHostToDevice, ctx->stream));
    std::vector<float> pr(parts, parts + (size_t)n * CIM_PART_STRIDE);
    for (int i = 0; i < n; ++i) {
        pr[i * CIM_PART_STRIDE + CIM_P_MASKX] = pr[i * CIM_PART_STRIDE + CIM_P_X];
        pr[i * CIM_PART_STRIDE + CIM_P_MASKY] = pr[i * CIM_PART_STRIDE + CIM_P_Y];
    }
    int rc = upload_parts(ctx, n, pr.data());
    if (rc) return rc;
    CIM_ALLOC(ctx->ft, sizeof(float2) * N * n);
    float2* ft = ctx->ft.as<float2>();
    // mask only when imgmask is on (ort_refinement.f90:46), no centre shift (:50)
    rc = preprocess_images(ctx, p, n, ctx->img.as<float>(), ctx->parts.as<float>(), ft, 0, p->mask_on, 1);
    if (rc) return rc;
    SearchBatch& sb = ctx->srch.wk;
    CIM_ALLOC(sb.fcut0, sizeof(float) * nr * n);
    rc = amp_expfit(ctx, ft, box, n, p->minR, R, sb.fcut0.as<float>());
    if (rc) return rc;
    CtfConst cc = make_ctf_const(p);
    RefineArgs a{};
    const int Dc = 2 * R + 4;
    if (p->apply_ctf) {
        CIM_ALLOC(sb.tcrop, sizeof(float2) * (size_t)Dc * Dc * n * T);
        hipLaunchKernelGGL(k_crop, dim3((Dc * Dc + 255) / 256, n), dim3(256), 0, ctx->stream, s.tft.as<float2>(), N,
                           box, R, Dc, sb.tcrop.as<float2>(), T, ctx->parts.as<float>(), cc, 1, (const float*)nullptr);
        CIM_HIP(hipGetLastError());
        CIM_ALLOC(sb.tfcut, sizeof(float) * nr * n * T);
        rc = amp_expfit_general(ctx, sb.tcrop.as<float2>(), (long long)Dc * Dc, Dc, R + 1, n * T, p->minR, R,
                                sb.tfcut.as<float>());
        if (rc) return rc;
        a.tft = sb.tcrop.as<float2>();
        a.t_pstride = (long long)T * Dc * Dc;
        a.tstride = (long long)Dc * Dc;
        a.t_ld = Dc;
        a.t_c = R + 1;
        a.tfcut = sb.tfcut.as<float>();
        a.tf_pstride = (long long)T * nr;
    } else {
        a.tft = s.tft.as<float2>();
        a.t_pstride = 0;
        a.tstride = N;
        a.t_ld = box;
        a.t_c = box / 2;
        a.tfcut = s.tfcut.as<float>();
        a.tf_pstride = 0;
    }
    std::vector<float> ops(540);
    {
        double d[540];
        cim_icos60_ref_order(d);
        for (int i = 0; i < 540; ++i) ops[i] = (float)d[i];
    }
    CIM_ALLOC(sb.ops, sizeof(float) * 540);
    CIM_ALLOC(sb.tang, sizeof(float) * 3 * T);
    CIM_HIP(hipMemcpyAsync(sb.ops.ptr, ops.data(), sizeof(float) * 540, hipMemcpyHostToDevice, ctx->stream));
    CIM_HIP(hipMemcpyAsync(sb.tang.ptr, s.tang.data(), sizeof(float) * 3 * T, hipMemcpyHostToDevice, ctx->stream));
    a.ft = ft;
    a.fcut0 = sb.fcut0.as<float>();
    a.tang = sb.tang.as<float>();
    a.T = T;
    a.ops = sb.ops.as<float>();
    a.parts = ctx->parts.as<float>();
    a.box = box;
    a.minR = p->minR;
    a.nr = nr;
    a.ncycle = p->ncycle;
    a.dang0 = p->delta_angle;
    a.dtran0 = p->delta_trans;
    {
        TimedLaunch tl(ctx, "refine");
        hipLaunchKernelGGL(k_refine, dim3(n), dim3(256), 0, ctx->stream, a);
        CIM_HIP(hipGetLastError());
    }
    CIM_HIP(hipMemcpyAsync(pr.data(), ctx->parts.ptr, sizeof(float) * CIM_PART_STRIDE * n, hipMemcpyDeviceToHost,
                           ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i)
        for (int k : {CIM_P_THETA, CIM_P_PHI, CIM_P_OMEGA, CIM_P_X, CIM_P_Y, CIM_P_PR})
            parts[i * CIM_PART_STRIDE + k] = pr[i * CIM_PART_STRIDE + k];
    return CIM_OK;
}

}  // namespace cim

using namespace cim;

extern "C" {

int cim_debug_cl_table(cim_ctx* ctx, int nh, int np, int T, int tpol_per_particle, int nr, int npsi, const float* ppol,
                       const float* tpol, unsigned int* table) {
    if (!ctx) return CIM_E_ARG;
    if (np < 1 || T < 1 || nr < 1 || npsi < 2 || (npsi & 1) || !ppol || !tpol || !table)
        return set_error(ctx, CIM_E_ARG, "cim_debug_cl_table: bad sizes or NULL arrays");
    CIM_HIP(hipSetDevice(ctx->device));
    const size_t pb = sizeof(float2) * (size_t)np * nr * npsi;
    const size_t tb = sizeof(float2) * (size_t)(tpol_per_particle ? np * T : T) * nr * npsi;
    const size_t ob = sizeof(unsigned int) * (size_t)np * T * (npsi / 2) * npsi;
    CIM_ALLOC(ctx->img, pb);
    CIM_ALLOC(ctx->aux, tb);
    CIM_ALLOC(ctx->out, ob);
    CIM_HIP(hipMemcpyAsync(ctx->img.ptr, ppol, pb, hipMemcpyHostToDevice, ctx->stream));
    CIM_HIP(hipMemcpyAsync(ctx->aux.ptr, tpol, tb, hipMemcpyHostToDevice, ctx->stream));
    int rc = launch_cl_table(ctx, nh, ctx->img.as<float2>(), ctx->aux.as<float2>(), tpol_per_particle, np, T, nr, npsi,
                             ctx->out.as<unsigned int>());
    if (rc) return rc;
    CIM_HIP(hipMemcpyAsync(table, ctx->out.ptr, ob, hipMemcpyDeviceToHost, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

int cim_search_templates_model(cim_ctx* ctx, const cim_params* p, int T, const float* model, const float* tort) {
    if (!ctx) return CIM_E_ARG;
    int rc = check_params_public(ctx, p);
    if (rc) return rc;
    if (T < 1 || !model || !tort) return set_error(ctx, CIM_E_ARG, "need T >= 1 templates and a model");
    if (p->box < 4) return set_error(ctx, CIM_E_ARG, "box %d too small for a projection", p->box);
    CIM_HIP(hipSetDevice(ctx->device));
    if ((rc = project_templates(ctx, p->box, T, model, tort))) return rc;
    return prep_templates_img(ctx, p, T, tort);
}

int cim_project_templates(cim_ctx* ctx, int box, int T, const float* model, const float* angles, float* out) {
    if (!ctx) return CIM_E_ARG;
    if (box < 4 || T < 1 || !model || !angles || !out) return set_error(ctx, CIM_E_ARG, "bad sizes or NULL arrays");
    CIM_HIP(hipSetDevice(ctx->device));
    std::vector<float> tort(5 * (size_t)T, 0.f);
    for (int t = 0; t < T; ++t)
        for (int k = 0; k < 3; ++k) tort[5 * (size_t)t + k] = angles[3 * (size_t)t + k];
    int rc = project_templates(ctx, box, T, model, tort.data());
    if (rc) return rc;
    CIM_HIP(hipMemcpyAsync(out, ctx->img.ptr, sizeof(float) * (size_t)box * box * T, hipMemcpyDeviceToHost,
                           ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

int cim_search_templates(cim_ctx* ctx, const cim_params* p, int T, const float* images, const float* tort) {
    if (!ctx) return CIM_E_ARG;
    int rc = check_params_public(ctx, p);
    if (rc) return rc;
    if (T < 1 || !images || !tort) return set_error(ctx, CIM_E_ARG, "need T >= 1 templates");
    CIM_HIP(hipSetDevice(ctx->device));
    return prep_templates(ctx, p, T, images, tort);
}

int cim_search_global_dev(cim_ctx* ctx, const cim_params* p, int n, const void* images_dev, float* parts) {
    if (!ctx) return CIM_E_ARG;
    int rc = check_params_public(ctx, p);
    if (rc) return rc;
    if (!images_dev || !parts) return set_error(ctx, CIM_E_ARG, "NULL images or parts");
    CIM_HIP(hipSetDevice(ctx->device));
    return search_global_dev(ctx, p, n, (const float*)images_dev, parts, nullptr);
}

int cim_search_global(cim_ctx* ctx, const cim_params* p, int n, const float* images, float* parts) {
    if (!ctx) return CIM_E_ARG;
    int rc = check_params_public(ctx, p);
    if (rc) return rc;
    if (!images || !parts) return set_error(ctx, CIM_E_ARG, "NULL images or parts");
    if (n <= 0) return CIM_OK;
    CIM_HIP(hipSetDevice(ctx->device));
    const long long N = (long long)p->box * p->box;
    CIM_ALLOC(ctx->img, sizeof(float) * N * n);
    CIM_HIP(hipMemcpyAsync(ctx->img.ptr, images, sizeof(float) * N * n, hipMemcpyHostToDevice, ctx->stream));
    return search_global_dev(ctx, p, n, ctx->img.as<float>(), parts, nullptr);
}

int cim_search_pr_array(cim_ctx* ctx, const cim_params* p, const float* image, const float* part, float* pr_array) {
    if (!ctx) return CIM_E_ARG;
    int rc = check_params_public(ctx, p);
    if (rc) return rc;
    if (!image || !part || !pr_array) return set_error(ctx, CIM_E_ARG, "NULL argument");
    CIM_HIP(hipSetDevice(ctx->device));
    const long long N = (long long)p->box * p->box;
    CIM_ALLOC(ctx->img, sizeof(float) * N);
    CIM_HIP(hipMemcpyAsync(ctx->img.ptr, image, sizeof(float) * N, hipMemcpyHostToDevice, ctx->stream));
    std::vector<float> rec(part, part + CIM_PART_STRIDE);
    return search_global_dev(ctx, p, 1, ctx->img.as<float>(), rec.data(), pr_array);
}

int cim_refine_local(cim_ctx* ctx, const cim_params* p, int n, const float* images, float* parts) {
    if (!ctx) return CIM_E_ARG;
    int rc = check_params_public(ctx, p);
    if (rc) return rc;
    if (!images || !parts) return set_error(ctx, CIM_E_ARG, "NULL images or parts");
    CIM_HIP(hipSetDevice(ctx->device));
    return refine_local(ctx, p, n, images, parts);
}

}  // extern "C"
