// local_refine on gfx950: central-slice extraction from the resident 3D
// model transform and the full-plane phase residual, scored for the
// candidates of the symmetry stage and of every coordinate-descent cycle
// (local_refine/SRC/refine_PR.f90:126-233, FFT3d_slice.f90, cal_template_FT.f90,
// cal_PR_FT.f90, ort_cent_11.f90, icos_equalort.f90).
//
// Only the pixels the phase residual reads (radius <= maxRes) are sliced:
// one workgroup per sub-particle walks that disk, computing each trilinear
// slice sample once and scoring it for every candidate that shares the
// slice (the 5 shift-only candidates of a cycle share candidate 0's slice,
// as cal_template_FT.f90:9-22 does).
#include <array>
#include <cmath>
#include <string>
#include <cstring>

#include "cim_device.h"

namespace cim {

int upload_parts(cim_ctx* ctx, int n, const float* parts);
int check_params_public(cim_ctx* ctx, const cim_params* p);

// ---- model preparation: softmask (img3d_softmask.f90), checkerboard, FFT ----

// Shell sum and count of img3d_softmask.f90: each block's compensated float32
// partial (sum, count) goes to part[4 blockIdx.x ..]; k_sum_partials adds them.
__global__ void __launch_bounds__(256) k_softmask_sum(const float* __restrict__ v, int n, float inn, float out,
                                                      float* __restrict__ part) {
    __shared__ ff32 sh[4];
    const long long N = (long long)n * n * n, n2 = (long long)n * n;
    ff32 s{0.f, 0.f};
    int c = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        const int iz = (int)(i / n2) - n / 2, iy = (int)((i / n) % n) - n / 2, ix = (int)(i % n) - n / 2;
        const float r = sqrtf((float)(ix * ix + iy * iy + iz * iz));
        if (r > inn && r <= out) { s = ff_add(s, ff32{v[i], 0.f}); ++c; }
    }
    s = block_sum_ff(s, sh);
    const ff32 ct = block_sum_ff(ff32{(float)c, 0.f}, sh);
    if (threadIdx.x == 0) {
        float* q = part + 4 * (long long)blockIdx.x;
        q[0] = s.h;
        q[1] = s.l;
        q[2] = ct.h;
        q[3] = ct.l;
    }
}

__global__ void k_softmask_apply(const float* __restrict__ v, int n, float inn, float out,
                                 const float* __restrict__ acc, float2* __restrict__ dst) {
    const long long N = (long long)n * n * n, n2 = (long long)n * n;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    // acc: the shell's sum (hi, lo) and count (hi, lo): their ratio in float-float, rounded
    float summ = 0.f;
    if (acc[2] != 0.f) {
        const float q = __fdiv_rn(acc[0], acc[2]);
        // one correction step: (sum - q count) / count
        const ff32 qc = ff_mul(ff32{q, 0.f}, ff32{acc[2], acc[3]});
        summ = __fadd_rn(q, __fdiv_rn(__fadd_rn(__fsub_rn(acc[0], qc.h), __fsub_rn(acc[1], qc.l)), acc[2]));
    }
    const int Z = (int)(i / n2), Y = (int)((i / n) % n), X = (int)(i % n);
    const int iz = Z - n / 2, iy = Y - n / 2, ix = X - n / 2;
    const float r = sqrtf((float)(ix * ix + iy * iy + iz * iz));
    float d = v[i];
    if (r > inn && r <= out) {
        const float edge = (1.0f + cosf(3.14159265358979323846f * (r - inn) / 6.0f)) / 2.0f;
        d = d * edge + (1.0f - edge) * summ;
    }
    if (r > out) d = summ;
    if ((X + Y + Z) & 1) d = -d;            // shift3d before the transform (refine_PR.f90:50)
    dst[i] = make_float2(d, 0.f);
}

__global__ void k_checker3d(float2* __restrict__ v, int n) {
    const long long N = (long long)n * n * n, n2 = (long long)n * n;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int Z = (int)(i / n2), Y = (int)((i / n) % n), X = (int)(i % n);
    if ((X + Y + Z) & 1) v[i] = make_float2(-v[i].x, -v[i].y);
}

// ---- sub-particle preparation ---------------------------------------------------

// img2d_norm of the whole image, then img2d_local_box cut around (hx0, ky0)
// with wrap-around (refine_PR.f90:87-126, img2d_local_box.f90:6-17).
__global__ void __launch_bounds__(1024) k_cut_norm(const float* __restrict__ imgs, int n0, const float* __restrict__ parts,
                                                   int n, float* __restrict__ out) {
    __shared__ ff32 sh[16];
    const long long N0 = (long long)n0 * n0;
    const float* img = imgs + (long long)blockIdx.x * N0;
    const float* pr = parts + (long long)blockIdx.x * CIM_PART_STRIDE;
    // mean and rms in compensated float32
    ff32 s{0.f, 0.f};
    for (long long i = threadIdx.x; i < N0; i += blockDim.x) s = ff_add(s, ff32{img[i], 0.f});
    const ff32 avg = ff_mul(block_sum_ff(s, sh), ff_recip_int((int)N0));
    ff32 q{0.f, 0.f};
    for (long long i = threadIdx.x; i < N0; i += blockDim.x) {
        const float d = ff_sub_from(img[i], avg);
        q = ff_add(q, ff_two_prod(d, d));
    }
    const float inv = 1.0f / sqrtf(ff_mul(block_sum_ff(q, sh), ff_recip_int((int)N0)).h);
    const int ix0 = (int)pr[CIM_P_BOXX], iy0 = (int)pr[CIM_P_BOXY];
    float* o = out + (long long)blockIdx.x * n * n;
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) {
        int hx = i % n + ix0 - n / 2, ky = i / n + iy0 - n / 2;
        if (hx < 0) hx += n0;
        if (hx >= n0) hx -= n0;
        if (ky < 0) ky += n0;
        if (ky >= n0) ky -= n0;
        o[i] = ff_sub_from(img[(long long)ky * n0 + hx], avg) * inv;
    }
}

// The particle side of a phase-residual term: {|g|, atan2(g)}.
__device__ __forceinline__ float2 g_polar(float2 g) {
    return make_float2(sqrtf(g.x * g.x + g.y * g.y), atan2f(g.y, g.x));
}

// Particle values on the disk, g = FT * w_lp, kept as {|g|, atan2(g)}: the
// particle side of every phase-residual term (cal_PR_FT.f90:15-25), formed
// once here instead of once per candidate and cycle; and the template factor
// shift2d * CTF * w_lp (refine_PR.f90:43-48,133-134,151-152).
__global__ void k_lref_disk(const float2* __restrict__ ft, int n, const int2* __restrict__ disk, int nq,
                            const float* __restrict__ parts, int apply_ctf, CtfConst cc, float maxres,
                            float2* __restrict__ g, float* __restrict__ tfac) {
    const int p = blockIdx.y;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int2 hk = disk[q];
    const int X = hk.x + n / 2, Y = hk.y + n / 2;
    const float r2 = (float)(hk.x * hk.x + hk.y * hk.y) / 2.0f / (maxres * maxres);
    const float w = expf(-r2);
    const float2 v = ft[(long long)p * n * n + (long long)Y * n + X];
    g[(long long)p * nq + q] = g_polar(make_float2(v.x * w, v.y * w));
    float c = 1.f;
    if (apply_ctf) {
        const float* pr = parts + (long long)p * CIM_PART_STRIDE;
        c = ctf_value(hk.x, hk.y, n, pr[CIM_P_DF1], pr[CIM_P_DF2], pr[CIM_P_ASTIG], cc);
    }
    const float s = ((X + Y) & 1) ? -1.f : 1.f;
    tfac[(long long)p * nq + q] = s * c * w;
}

// ---- refinement ---------------------------------------------------------------

struct LrefArgs {
    const float2* model;   // n^3 model transform
    const float4* model4;  // its tap-paired copy (k_pair_model)
    int n;
    const int2* disk;      // nq pixels (hx, ky) centred, radius <= maxres
    int nq;
    const float2* g;       // particle values {|g|, atan2(g)}, per particle nq
    const float* tfac;     // template factor, per particle nq
    const float* symops;   // nsym x 9 (float) or null
    int nsym;
    int ncycle;
    float refine_angle, refine_shift, convF;
    float* parts;          // in/out records
};

// Trilinear central-slice sample of FFT3d_slice.f90:19-75 at image pixel (hx, ky).
__device__ __forceinline__ float2 slice_at(const float2* __restrict__ m, int n, const float* R, int hx, int ky) {
    const float fx = (float)hx, fy = (float)ky;
    float X = R[0] * fx + R[1] * fy, Y = R[3] * fx + R[4] * fy, Z = R[6] * fx + R[7] * fy;
    if (sqrtf(X * X + Y * Y + Z * Z) > (float)(n / 2 - 2)) return make_float2(0.f, 0.f);
    X += (float)(n / 2);
    Y += (float)(n / 2);
    Z += (float)(n / 2);
    const int ix = (int)X, iy = (int)Y, iz = (int)Z;
    const float xb = X - ix, yb = Y - iy, zb = Z - iz;
    const long long n2 = (long long)n * n;
    const float2* b = m + (long long)iz * n2 + (long long)iy * n + ix;
    const float2 t1 = b[0], t2 = b[1], t3 = b[n], t4 = b[n + 1];
    const float2 t5 = b[n2], t6 = b[n2 + 1], t7 = b[n2 + n], t8 = b[n2 + n + 1];
    const float w1 = (1.f - xb) * (1.f - yb) * (1.f - zb), w2 = xb * (1.f - yb) * (1.f - zb);
    const float w3 = (1.f - xb) * yb * (1.f - zb), w4 = xb * yb * (1.f - zb);
    const float w5 = (1.f - xb) * (1.f - yb) * zb, w6 = xb * (1.f - yb) * zb;
    const float w7 = (1.f - xb) * yb * zb, w8 = xb * yb * zb;
    return make_float2(w1 * t1.x + w2 * t2.x + w3 * t3.x + w4 * t4.x + w5 * t5.x + w6 * t6.x + w7 * t7.x + w8 * t8.x,
                       w1 * t1.y + w2 * t2.y + w3 * t3.y + w4 * t4.y + w5 * t5.y + w6 * t6.y + w7 * t7.y + w8 * t8.y);
}

// slice_at on the tap-paired model m4[i] = {m[i], m[i + 1]}: the same taps,
// weights and sums, from 4 loads of 16 bytes instead of 8 of 8 (the radius
// test keeps ix <= n - 2, so the pair never crosses a row).
__device__ __forceinline__ float2 slice_at_p(const float4* __restrict__ m4, int n, const float* R, int hx, int ky) {
    const float fx = (float)hx, fy = (float)ky;
    float X = R[0] * fx + R[1] * fy, Y = R[3] * fx + R[4] * fy, Z = R[6] * fx + R[7] * fy;
    if (sqrtf(X * X + Y * Y + Z * Z) > (float)(n / 2 - 2)) return make_float2(0.f, 0.f);
    X += (float)(n / 2);
    Y += (float)(n / 2);
    Z += (float)(n / 2);
    const int ix = (int)X, iy = (int)Y, iz = (int)Z;
    const float xb = X - ix, yb = Y - iy, zb = Z - iz;
    const long long n2 = (long long)n * n;
    const float4* b = m4 + (long long)iz * n2 + (long long)iy * n + ix;
    const float4 u1 = b[0], u3 = b[n], u5 = b[n2], u7 = b[n2 + n];
    const float2 t1 = make_float2(u1.x, u1.y), t2 = make_float2(u1.z, u1.w);
    const float2 t3 = make_float2(u3.x, u3.y), t4 = make_float2(u3.z, u3.w);
    const float2 t5 = make_float2(u5.x, u5.y), t6 = make_float2(u5.z, u5.w);
    const float2 t7 = make_float2(u7.x, u7.y), t8 = make_float2(u7.z, u7.w);
    const float w1 = (1.f - xb) * (1.f - yb) * (1.f - zb), w2 = xb * (1.f - yb) * (1.f - zb);
    const float w3 = (1.f - xb) * yb * (1.f - zb), w4 = xb * yb * (1.f - zb);
    const float w5 = (1.f - xb) * (1.f - yb) * zb, w6 = xb * (1.f - yb) * zb;
    const float w7 = (1.f - xb) * yb * zb, w8 = xb * yb * zb;
    return make_float2(w1 * t1.x + w2 * t2.x + w3 * t3.x + w4 * t4.x + w5 * t5.x + w6 * t6.x + w7 * t7.x + w8 * t8.x,
                       w1 * t1.y + w2 * t2.y + w3 * t3.y + w4 * t4.y + w5 * t5.y + w6 * t6.y + w7 * t7.y + w8 * t8.y);
}

// Tap-paired copy of an n^3 transform: out[i] = {m[i], m[i + 1]} (the last
// element repeats itself).
__global__ void k_pair_model(const float2* __restrict__ m, long long N, float4* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float2 a = m[i], b = m[i + 1 < N ? i + 1 : i];
    out[i] = make_float4(a.x, a.y, b.x, b.y);
}

// One term of cal_PR_FT.f90:15-25: weight |F1| + |F2|, phase difference
// folded into [0, pi].  A zero sample has phase atan2(0, 0) = 0 there.  The
// particle value g comes as gp = g_polar(g).
__device__ __forceinline__ void pr_term(float2 t, float2 gp, float& num, float& den) {
    const float a1 = sqrtf(t.x * t.x + t.y * t.y), a2 = gp.x, p2 = gp.y;
    float d;
    if (a1 > 0.f && a2 > 0.f) {
        const float p1 = atan2f(t.y, t.x);
        d = fabsf(p2 - p1);
    } else {
        d = fabsf(a1 > 0.f ? atan2f(t.y, t.x) : p2);
    }
    const float pi = 3.14159265358979323846f, tpi = 6.28318530717958647692f;
    if (d >= tpi) d -= tpi;
    if (d > pi) d = tpi - d;
    const float w = a1 + a2;
    num += d * w;
    den += w;
}

template <int NC>
__device__ void block_reduce_pairs(float* num, float* den, float (*sh)[16], float* out_pr) {
    const int tid = threadIdx.x;
    for (int c = 0; c < NC; ++c) {
        float a = wave_sum(num[c]), b = wave_sum(den[c]);
        if ((tid & 63) == 0) {
            sh[2 * c][tid >> 6] = a;
            sh[2 * c + 1][tid >> 6] = b;
        }
    }
    __syncthreads();
    if (tid < NC) {
        float a = 0.f, b = 0.f;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            a += sh[2 * tid][w];
            b += sh[2 * tid + 1][w];
        }
        out_pr[tid] = b > 1.0e-10f ? a / b * 57.2957795130823208768f : 1000.f;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_lref_refine(LrefArgs a) {
    __shared__ float red[22][16];
    __shared__ float prs[16];
    __shared__ float cand[11][5];
    __shared__ float cur[6];
    __shared__ float symang[3];
    // the rotations of candidates 0 and 5-10 of a cycle, formed once per workgroup (an array indexed
    // by the candidate loop in each thread would live in scratch memory)
    __shared__ float rot[7][9];
    const int p = blockIdx.x, tid = threadIdx.x;
    const float2* g = a.g + (long long)p * a.nq;
    const float* tf = a.tfac + (long long)p * a.nq;
    float* rec = a.parts + (long long)p * CIM_PART_STRIDE;
    if (tid < 5) cur[tid] = rec[tid];
    if (tid == 0) cur[5] = 1000.f;
    __syncthreads();

    // symmetry stage: icos_equalort candidates at the starting shift (refine_PR.f90:136-163)
    if (a.nsym > 0) {
        float best = 0.f;
        for (int s = 0; s < a.nsym; ++s) {
            if (tid == 0) {
                const float r2d = 57.2957795130823208768f;
                float E[9], M[9];
                euler_f(cur[0], -cur[1], 180.0f - cur[2], E);
                const float* S = a.symops + 9 * s;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) M[3 * i + j] = S[3 * i] * E[j] + S[3 * i + 1] * E[3 + j] + S[3 * i + 2] * E[6 + j];
                symang[0] = acosf(M[8]) * r2d;
                symang[1] = -atan2f(M[5], M[2]) * r2d;
                symang[2] = atan2f(M[7], M[6]) * r2d;
            }
            __syncthreads();
            float R[9];
            euler_f(symang[0], symang[1], symang[2], R);
            float num[1] = {0.f}, den[1] = {0.f};
            for (int q = tid; q < a.nq; q += blockDim.x) {
                const int2 hk = a.disk[q];
                float2 z = slice_at_p(a.model4, a.n, R, hk.x, hk.y);
                z = apply_shift(z, shift_phase(hk.x, hk.y, a.n, cur[3], cur[4], 0));
                pr_term(make_float2(z.x * tf[q], z.y * tf[q]), g[q], num[0], den[0]);
            }
            block_reduce_pairs<1>(num, den, red, prs);
            if (tid == 0 && (s == 0 || prs[0] < best)) {
                best = prs[0];
                cand[0][0] = symang[0];
                cand[0][1] = symang[1];
                cand[0][2] = symang[2];
                cur[5] = best;
            }
            __syncthreads();
        }
        if (tid == 0) {
            cur[0] = cand[0][0];
            cur[1] = cand[0][1];
            cur[2] = cand[0][2];
        }
        __syncthreads();
    }

    float f = 1.f;
    for (int cyc = 1; cyc <= a.ncycle; ++cyc) {
        const float da = a.refine_angle * f, dt = a.refine_shift * f;
        f *= a.convF;
        if (tid < 11) {
            float o[5] = {cur[0], cur[1], cur[2], cur[3], cur[4]};
            switch (tid) {   // local_refine/SRC/ort_cent_11.f90:6-84
                case 1: o[3] -= dt; break;
                case 2: o[3] += dt; break;
                case 3: o[4] -= dt; break;
                case 4: o[4] += dt; break;
                case 5: o[0] -= da; break;
                case 6: o[0] += da; break;
                case 7: o[1] -= da; break;
                case 8: o[1] += da; break;
                case 9: o[2] -= da; break;
                case 10: o[2] += da; break;
            }
            for (int k = 0; k < 5; ++k) cand[tid][k] = o[k];
        }
        __syncthreads();
        if (tid < 7) {
            const int c = tid ? tid + 4 : 0;
            euler_f(cand[c][0], cand[c][1], cand[c][2], rot[tid]);
        }
        __syncthreads();
        float num[11], den[11];
        for (int c = 0; c < 11; ++c) { num[c] = 0.f; den[c] = 0.f; }
        // candidates 0 and 5-10 keep the current shift (cand[c][3..4] are copies of cur[3..4]), so
        // one shift phase serves all seven
        for (int q = tid; q < a.nq; q += blockDim.x) {
            const int2 hk = a.disk[q];
            const float2 gq = g[q];
            const float t = tf[q];
            const float2 z0 = slice_at_p(a.model4, a.n, rot[0], hk.x, hk.y);
            const float2 ph0 = shift_phase(hk.x, hk.y, a.n, cand[0][3], cand[0][4], 0);
            for (int c = 0; c <= 4; ++c) {
                const float2 z = apply_shift(z0, c ? shift_phase(hk.x, hk.y, a.n, cand[c][3], cand[c][4], 0) : ph0);
                pr_term(make_float2(z.x * t, z.y * t), gq, num[c], den[c]);
            }
            for (int c = 5; c <= 10; ++c) {
                float2 z = slice_at_p(a.model4, a.n, rot[c - 4], hk.x, hk.y);
                z = apply_shift(z, ph0);
                pr_term(make_float2(z.x * t, z.y * t), gq, num[c], den[c]);
            }
        }
        block_reduce_pairs<11>(num, den, red, prs);
        if (tid == 0) {
            int bi = 0;
            for (int c = 1; c < 11; ++c)
                if (prs[c] < prs[bi]) bi = c;
            for (int k = 0; k < 5; ++k) cur[k] = cand[bi][k];
            cur[5] = prs[bi];
        }
        __syncthreads();
    }
    if (tid == 0) {
        for (int k = 0; k < 5; ++k) rec[k] = cur[k];
        rec[CIM_P_PR] = cur[5];
    }
}

// refine_particle_df.f90:100-135: at the sub-particle's local orientation
// and shift, the template t = slice * phase ramp * shift2d * w_lp against the
// particle for 401 defoci z_i = z0 + 4 (i - 200) A (df -/+ z_i by sign), the
// phase residual of cal_PR_FT for each, and the first minimum (minloc).  One
// workgroup per sub-particle; t staged in LDS once.  Records: local angles,
// x, y in; CIM_P_Z holds z0 in and the chosen z out, CIM_P_PR the residual.
__global__ void __launch_bounds__(256) k_lref_defocus(LrefArgs a, CtfConst cc, float sign) {
    extern __shared__ float2 tz[];
    __shared__ float red[2][16];
    __shared__ float prs[16];
    const int p = blockIdx.x, tid = threadIdx.x;
    const float2* g = a.g + (long long)p * a.nq;
    const float* tf = a.tfac + (long long)p * a.nq;
    float* rec = a.parts + (long long)p * CIM_PART_STRIDE;
    float R[9];
    euler_f(rec[CIM_P_THETA], rec[CIM_P_PHI], rec[CIM_P_OMEGA], R);
    const float x0 = rec[CIM_P_X], y0 = rec[CIM_P_Y], z0 = rec[CIM_P_Z];
    for (int q = tid; q < a.nq; q += blockDim.x) {
        const int2 hk = a.disk[q];
        float2 z = slice_at(a.model, a.n, R, hk.x, hk.y);
        z = apply_shift(z, shift_phase(hk.x, hk.y, a.n, x0, y0, 0));
        tz[q] = make_float2(z.x * tf[q], z.y * tf[q]);
    }
    __syncthreads();
    float best = 0.f, bz = z0;
    for (int i = 0; i <= 400; ++i) {
        const float zi = z0 + (float)(i - 200) * 4.0f;
        const float d1 = sign > 0.f ? rec[CIM_P_DF1] + zi : rec[CIM_P_DF1] - zi;
        const float d2 = sign > 0.f ? rec[CIM_P_DF2] + zi : rec[CIM_P_DF2] - zi;
        float num[1] = {0.f}, den[1] = {0.f};
        for (int q = tid; q < a.nq; q += blockDim.x) {
            const int2 hk = a.disk[q];
            const float c = ctf_value(hk.x, hk.y, a.n, d1, d2, rec[CIM_P_ASTIG], cc);
            const float2 t = tz[q];
            pr_term(make_float2(t.x * c, t.y * c), g[q], num[0], den[0]);
        }
        block_reduce_pairs<1>(num, den, red, prs);
        if (tid == 0 && (i == 0 || prs[0] < best)) {
            best = prs[0];
            bz = zi;
        }
    }
    if (tid == 0) {
        rec[CIM_P_PR] = best;
        rec[CIM_P_Z] = bz;
    }
}

// ---- symmetry operators of the local programs -------------------------------------

// local_refine/SRC/icos2f_sym_matc.f90 generators, evaluated in float32 as
// the reference does and returned in its order.
static void mm3f(const float* A, const float* B, float* C) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

static int local_symops(const char* sym, std::vector<float>& r) {
    std::string s(sym ? sym : "c1");
    for (auto& ch : s) ch = (char)tolower(ch);
    const float pi = 4.0f * atanf(1.0f);
    auto cyclic = [&](int N, float ph) {
        r.assign(9 * N, 0.f);
        float R1[9] = {cosf(ph), -sinf(ph), 0, sinf(ph), cosf(ph), 0, 0, 0, 1};
        memcpy(r.data(), R1, sizeof(R1));
        for (int i = 1; i < N; ++i) mm3f(&r[9 * (i - 1)], R1, &r[9 * i]);
        return N;
    };
    if (s == "asym") return cyclic(180, pi / 6.0f / 180.0f);
    if (s.size() >= 2 && s[0] == 'c') {
        const int N = atoi(s.c_str() + 1);
        if (N < 1 || N > 360) return -1;
        return cyclic(N, 2.0f * pi / (float)N);
    }
    if (s == "i2" || s == "i5") {
        const float d2r = pi / 180.0f;
        const float a5 = 72.0f * d2r;
        const float v = acosf(cosf(60.0f * d2r) / sinf(36.0f * d2r));
        float rz5[9] = {cosf(a5), -sinf(a5), 0, sinf(a5), cosf(a5), 0, 0, 0, 1};
        float ryv[9] = {cosf(v), 0, sinf(v), 0, 1, 0, -sinf(v), 0, cosf(v)};
        float ryvT[9] = {ryv[0], ryv[3], ryv[6], ryv[1], ryv[4], ryv[7], ryv[2], ryv[5], ryv[8]};
        float t1[9], t2[9];
        mm3f(rz5, ryvT, t1);
        mm3f(ryv, t1, t2);
        memcpy(rz5, t2, sizeof(t2));
        const float rx2[9] = {1, 0, 0, 0, -1, 0, 0, 0, -1};
        const float ry2[9] = {-1, 0, 0, 0, 1, 0, 0, 0, -1};
        const float r111[9] = {0, 0, 1, 1, 0, 0, 0, 1, 0};
        r.assign(9 * 60, 0.f);
        r[0] = r[4] = r[8] = 1.f;
        for (int i = 1; i < 5; ++i) mm3f(&r[9 * (i - 1)], rz5, &r[9 * i]);
        for (int i = 5; i < 10; ++i) mm3f(rx2, &r[9 * (i - 5)], &r[9 * i]);
        for (int i = 10; i < 20; ++i) mm3f(ry2, &r[9 * (i - 10)], &r[9 * i]);
        for (int i = 20; i < 40; ++i) mm3f(r111, &r[9 * (i - 20)], &r[9 * i]);
        for (int i = 40; i < 60; ++i) mm3f(r111, &r[9 * (i - 20)], &r[9 * i]);
        if (s == "i5") {
            const float rot0[9] = {cosf(v), 0, sinf(v), 0, 1, 0, -sinf(v), 0, cosf(v)};
            const float rT[9] = {rot0[0], rot0[3], rot0[6], rot0[1], rot0[4], rot0[7], rot0[2], rot0[5], rot0[8]};
            for (int i = 0; i < 60; ++i) {
                mm3f(&r[9 * i], rT, t1);
                mm3f(rot0, t1, &r[9 * i]);
            }
        }
        return 60;
    }
    if (s == "icos") {
        double d[540];
        cim_icos60_ref_order(d);
        r.assign(540, 0.f);
        for (int i = 0; i < 540; ++i) r[i] = (float)d[i];
        return 60;
    }
    return -1;
}

// ---- local_reconstruct's operators (0calc_initialvalues.f90:33-62), float32 ----

namespace {

using M3 = std::array<float, 9>;   // row-major

M3 mul3(const M3& A, const M3& B) {
    M3 C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
    return C;
}

M3 tr3(const M3& A) {
    return M3{A[0], A[3], A[6], A[1], A[4], A[7], A[2], A[5], A[8]};
}

// icos_matrix_i2.f90:1-56: products of the tilted 5-fold, the 2-folds about
// x and y and the 3-fold permutation, in the reference's order.
std::vector<M3> icos_matrix_i2() {
    const float d2r = 4.0f * atan2f(1.0f, 1.0f) / 180.0f;
    const float a5 = 72.0f * d2r;
    const float v = acosf(cosf(60.0f * d2r) / sinf(36.0f * d2r));
    const M3 rz5{cosf(a5), -sinf(a5), 0.f, sinf(a5), cosf(a5), 0.f, 0.f, 0.f, 1.f};
    const M3 ryv{cosf(v), 0.f, sinf(v), 0.f, 1.f, 0.f, -sinf(v), 0.f, cosf(v)};
    const M3 z5 = mul3(ryv, mul3(rz5, tr3(ryv)));
    const M3 rx2{1.f, 0.f, 0.f, 0.f, -1.f, 0.f, 0.f, 0.f, -1.f};
    const M3 ry2{-1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, -1.f};
    const M3 r3{0.f, 0.f, 1.f, 1.f, 0.f, 0.f, 0.f, 1.f, 0.f};
    std::vector<M3> r(60);
    r[0] = M3{1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
    for (int i = 1; i < 5; ++i) r[i] = mul3(r[i - 1], z5);
    for (int i = 5; i < 10; ++i) r[i] = mul3(rx2, r[i - 5]);
    for (int i = 10; i < 20; ++i) r[i] = mul3(ry2, r[i - 10]);
    for (int i = 20; i < 60; ++i) r[i] = mul3(r3, r[i - 20]);
    return r;
}

// local_reconstruct's operator set: I2, I3 (I2 conjugated by a rotation about
// x by -w, icos_matrix_i3.f90), I5 (about y by v, icos_matrix_i5.f90), or
// C_matrix (C_matrix.f90: r1 = rotation about z by 2 pi / N, r_i = r_(i-1) r1).
// Names as 0calc_initialvalues.f90 spells them ("C_10".."C_15"); -1 for
// others.  box_names: as local_box/SRC/0main.f90:80-122 spells them
// ("C10".."C15"), and a name it does not know keeps its default Nsym = 60
// and goes to C_matrix.
int local_reconstruct_ops(const char* sym, std::vector<M3>& r, bool box_names = false) {
    std::string s(sym ? sym : "");
    while (!s.empty() && s.back() == ' ') s.pop_back();
    const float d2r = 4.0f * atan2f(1.0f, 1.0f) / 180.0f;
    if (s == "i2" || s == "I2" || s == "i3" || s == "I3" || s == "i5" || s == "I5") {
        r = icos_matrix_i2();
        if (s == "i2" || s == "I2") return 60;
        M3 rot0;
        if (s == "i3" || s == "I3") {
            const float w = acosf(cosf(36.0f * d2r) / sinf(60.0f * d2r));
            rot0 = M3{1.f, 0.f, 0.f, 0.f, cosf(-w), sinf(-w), 0.f, -sinf(-w), cosf(-w)};
        } else {
            const float v = acosf(cosf(60.0f * d2r) / sinf(36.0f * d2r));
            rot0 = M3{cosf(v), 0.f, sinf(v), 0.f, 1.f, 0.f, -sinf(v), 0.f, cosf(v)};
        }
        for (auto& m : r) m = mul3(rot0, mul3(m, tr3(rot0)));
        return 60;
    }
    int N = 0;
    for (int n = 1; n <= 15; ++n) {
        const std::string lo = "c" + std::to_string(n);
        const std::string up = n < 10 || box_names ? "C" + std::to_string(n) : "C_" + std::to_string(n);
        if (s == lo || s == up) N = n;
    }
    if (N == 0 && box_names) N = 60;
    if (N == 0) return -1;
    const float pi = 4.0f * atan2f(1.0f, 1.0f);
    const float phi = 2.0f * pi / (float)N;
    r.assign(N, M3{});
    r[0] = M3{cosf(phi), -sinf(phi), 0.f, sinf(phi), cosf(phi), 0.f, 0.f, 0.f, 1.f};
    for (int i = 1; i < N; ++i) r[i] = mul3(r[i - 1], r[0]);
    return N;
}

}  // namespace

}  // namespace cim

using namespace cim;

extern "C" {

int cim_local_reconstruct_sym_ops(const char* sym, double* ops, int* nops) {
    std::vector<M3> r;
    const int n = local_reconstruct_ops(sym, r);
    if (n < 0) return CIM_E_ARG;
    if (nops) *nops = n;
    if (ops)
        for (int k = 0; k < n; ++k)
            for (int i = 0; i < 9; ++i) ops[9 * k + i] = r[k][i];
    return CIM_OK;
}

int cim_local_box_sym_ops(const char* sym, double* ops, int* nops) {
    std::vector<M3> r;
    const int n = local_reconstruct_ops(sym, r, true);
    if (nops) *nops = n;
    if (ops)
        for (int k = 0; k < n; ++k)
            for (int i = 0; i < 9; ++i) ops[9 * k + i] = r[k][i];
    return CIM_OK;
}

int cim_sym_ops(const char* sym, double* ops, int* nops) {
    std::vector<float> r;
    const int n = local_symops(sym, r);
    if (n < 0) return CIM_E_ARG;
    if (nops) *nops = n;
    if (ops)
        for (int i = 0; i < 9 * n; ++i) ops[i] = r[i];
    return CIM_OK;
}

int cim_lref_model(cim_ctx* ctx, const cim_params* p, const float* vol) {
    if (!ctx) return CIM_E_ARG;
    if (!p || !vol) return set_error(ctx, CIM_E_ARG, "NULL argument");
    const int n = p->box;
    if (n < 8 || n > 1024 || (n & 1)) return set_error(ctx, CIM_E_ARG, "model box %d unsupported", n);
    CIM_HIP(hipSetDevice(ctx->device));
    const long long N = (long long)n * n * n;
    CIM_ALLOC(ctx->img, sizeof(float) * N);
    CIM_HIP(hipMemcpyAsync(ctx->img.ptr, vol, sizeof(float) * N, hipMemcpyHostToDevice, ctx->stream));
    CIM_ALLOC(ctx->lref.model, sizeof(float2) * N);
    // shell sum partials of k_softmask_sum's 1024 blocks, then their total (sum hi, lo, count hi, lo)
    CIM_ALLOC(ctx->lref.acc, sizeof(float) * 4 * (1024 + 1));
    float* acc_total = ctx->lref.acc.as<float>() + 4 * 1024;
    float inn = p->imgmask - 3.f, out = p->imgmask + 3.f;
    if (inn <= 0.f) inn = 0.f;
    if (out > (float)(n / 2)) out = (float)(n / 2);
    hipLaunchKernelGGL(k_softmask_sum, dim3(1024), dim3(256), 0, ctx->stream, ctx->img.as<float>(), n, inn, out,
                       ctx->lref.acc.as<float>());
    CIM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, ctx->stream, ctx->lref.acc.as<float>(), 1024, acc_total);
    CIM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_softmask_apply, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream,
                       ctx->img.as<float>(), n, inn, out, acc_total, ctx->lref.model.as<float2>());
    CIM_HIP(hipGetLastError());
    int rc = fft3d_dev(ctx, ctx->lref.model.as<float2>(), n, 1, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_checker3d, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream,
                       ctx->lref.model.as<float2>(), n);
    CIM_HIP(hipGetLastError());
    CIM_ALLOC(ctx->lref.model4, sizeof(float4) * N);
    hipLaunchKernelGGL(k_pair_model, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream,
                       ctx->lref.model.as<float2>(), N, ctx->lref.model4.as<float4>());
    CIM_HIP(hipGetLastError());
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    ctx->lref.box = n;
    ctx->lref.ready = true;
    return CIM_OK;
}

}  // extern "C"

namespace cim {

// The particle side of refine_PR.f90:87-134 / refine_particle_df.f90:71-109 for
// n sub-particles: whole-image norm, local box cut, mask at the local shift,
// checkerboard, FFT (+Im), and on the disk of radius maxres the low-passed
// values g and the template factor shift2d * CTF (apply_ctf) * w_lp.
static int lref_prepare(cim_ctx* ctx, const cim_params* p, int n, const float* images, int box0, const float* parts,
                        float maxres, int apply_ctf, int* nq_out) {
    const int nb = p->box;
    const long long N0 = (long long)box0 * box0, Nb = (long long)nb * nb;
    // disk of pixels the phase residual reads (cal_PR_FT.f90:11-15, minRes = -1)
    std::vector<int2> disk;
    for (int ky = -nb / 2; ky < nb / 2; ++ky)
        for (int hx = -nb / 2; hx < nb / 2; ++hx)
            if (sqrtf((float)(hx * hx + ky * ky)) <= maxres) disk.push_back(make_int2(hx, ky));
    const int nq = (int)disk.size();
    LrefState& L = ctx->lref;
    CIM_ALLOC(L.disk, sizeof(int2) * nq);
    CIM_HIP(hipMemcpyAsync(L.disk.ptr, disk.data(), sizeof(int2) * nq, hipMemcpyHostToDevice, ctx->stream));
    CIM_ALLOC(ctx->img, sizeof(float) * N0 * n);
    CIM_HIP(hipMemcpyAsync(ctx->img.ptr, images, sizeof(float) * N0 * n, hipMemcpyHostToDevice, ctx->stream));
    int rc = upload_parts(ctx, n, parts);
    if (rc) return rc;
    CIM_ALLOC(ctx->aux, sizeof(float) * Nb * n);
    hipLaunchKernelGGL(k_cut_norm, dim3(n), dim3(1024), 0, ctx->stream, ctx->img.as<float>(), box0,
                       ctx->parts.as<float>(), nb, ctx->aux.as<float>());
    CIM_HIP(hipGetLastError());
    // mask around the sub-particle centre (local x0, y0): records with MASKX/Y = x, y
    std::vector<float> mrec(parts, parts + (size_t)n * CIM_PART_STRIDE);
    for (int i = 0; i < n; ++i) {
        mrec[i * CIM_PART_STRIDE + CIM_P_MASKX] = parts[i * CIM_PART_STRIDE + CIM_P_X];
        mrec[i * CIM_PART_STRIDE + CIM_P_MASKY] = parts[i * CIM_PART_STRIDE + CIM_P_Y];
    }
    CIM_ALLOC(L.mparts, sizeof(float) * CIM_PART_STRIDE * n);
    CIM_HIP(hipMemcpyAsync(L.mparts.ptr, mrec.data(), sizeof(float) * CIM_PART_STRIDE * n, hipMemcpyHostToDevice,
                           ctx->stream));
    CIM_ALLOC(ctx->ft, sizeof(float2) * Nb * n);
    rc = preprocess_images(ctx, p, n, ctx->aux.as<float>(), L.mparts.as<float>(), ctx->ft.as<float2>(), 0,
                           p->mask_on, 0);
    if (rc) return rc;
    CIM_ALLOC(L.g, sizeof(float2) * (size_t)nq * n);
    CIM_ALLOC(L.tfac, sizeof(float) * (size_t)nq * n);
    CtfConst cc = make_ctf_const(p);
    hipLaunchKernelGGL(k_lref_disk, dim3((nq + 255) / 256, n), dim3(256), 0, ctx->stream, ctx->ft.as<float2>(), nb,
                       L.disk.as<int2>(), nq, ctx->parts.as<float>(), apply_ctf, cc, maxres, L.g.as<float2>(),
                       L.tfac.as<float>());
    CIM_HIP(hipGetLastError());
    *nq_out = nq;
    return CIM_OK;
}

static int lref_check(cim_ctx* ctx, const cim_params* p, int n, const float* images, int box0, const float* parts,
                      float maxres) {
    if (!ctx) return CIM_E_ARG;
    int rc = check_params_public(ctx, p);
    if (rc) return rc;
    if (n > 0 && (!images || !parts)) return set_error(ctx, CIM_E_ARG, "NULL images or parts");
    if (!ctx->lref.ready || ctx->lref.box != p->box)
        return set_error(ctx, CIM_E_STATE, "cim_lref_model has not been called for box %d", p->box);
    if (box0 < p->box || box0 > 4096) return set_error(ctx, CIM_E_ARG, "image box %d < sub-particle box %d", box0, p->box);
    if (!(maxres > 0.f)) return set_error(ctx, CIM_E_ARG, "maxres must be > 0");
    return CIM_OK;
}

}  // namespace cim

extern "C" {

int cim_lref_refine(cim_ctx* ctx, const cim_params* p, int n, const float* images, int box0, float* parts,
                    const char* sym, float convF, float maxres) {
    int rc = lref_check(ctx, p, n, images, box0, parts, maxres);
    if (rc) return rc;
    if (n <= 0) return CIM_OK;
    CIM_HIP(hipSetDevice(ctx->device));
    std::vector<float> sops;
    int nsym = 0;
    std::string s(sym ? sym : "c1");
    if (s != "c1" && s != "C1") {
        nsym = local_symops(sym, sops);
        if (nsym < 0) return set_error(ctx, CIM_E_ARG, "unknown symmetry '%s'", sym);
    }
    LrefState& L = ctx->lref;
    if (nsym > 0) {
        CIM_ALLOC(L.symops, sizeof(float) * 9 * nsym);
        CIM_HIP(hipMemcpyAsync(L.symops.ptr, sops.data(), sizeof(float) * 9 * nsym, hipMemcpyHostToDevice, ctx->stream));
    }
    int nq = 0;
    if ((rc = lref_prepare(ctx, p, n, images, box0, parts, maxres, p->apply_ctf, &nq))) return rc;
    const int nb = p->box;
    LrefArgs a{};
    a.model = L.model.as<float2>();
    a.model4 = L.model4.as<float4>();
    a.n = nb;
    a.disk = L.disk.as<int2>();
    a.nq = nq;
    a.g = L.g.as<float2>();
    a.tfac = L.tfac.as<float>();
    a.symops = nsym > 0 ? L.symops.as<float>() : nullptr;
    a.nsym = nsym;
    a.ncycle = p->ncycle;
    a.refine_angle = p->delta_angle;
    a.refine_shift = p->delta_trans;
    a.convF = convF;
    a.parts = ctx->parts.as<float>();
    {
        TimedLaunch tl(ctx, "lref_refine");
        hipLaunchKernelGGL(k_lref_refine, dim3(n), dim3(256), 0, ctx->stream, a);
        CIM_HIP(hipGetLastError());
    }
    std::vector<float> outp((size_t)n * CIM_PART_STRIDE);
    CIM_HIP(hipMemcpyAsync(outp.data(), ctx->parts.ptr, sizeof(float) * CIM_PART_STRIDE * n, hipMemcpyDeviceToHost,
                           ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i)
        for (int k : {CIM_P_THETA, CIM_P_PHI, CIM_P_OMEGA, CIM_P_X, CIM_P_Y, CIM_P_PR})
            parts[i * CIM_PART_STRIDE + k] = outp[i * CIM_PART_STRIDE + k];
    ctx->counters["lref_particles"] += n;
    return CIM_OK;
}

int cim_lref_defocus(cim_ctx* ctx, const cim_params* p, int n, const float* images, int box0, float* parts,
                     float sign, float maxres) {
    int rc = lref_check(ctx, p, n, images, box0, parts, maxres);
    if (rc) return rc;
    if (!(sign == 1.f || sign == -1.f)) return set_error(ctx, CIM_E_ARG, "sign must be +1 or -1 (the sign key)");
    if (n <= 0) return CIM_OK;
    CIM_HIP(hipSetDevice(ctx->device));
    int nq = 0;
    if ((rc = lref_prepare(ctx, p, n, images, box0, parts, maxres, 0, &nq))) return rc;
    if ((size_t)nq * sizeof(float2) > 64 * 1024)
        return set_error(ctx, CIM_E_ARG, "maxres %g: %d disk pixels exceed the 64 KB template buffer", maxres, nq);
    LrefState& L = ctx->lref;
    LrefArgs a{};
    a.model = L.model.as<float2>();
    a.model4 = L.model4.as<float4>();
    a.n = p->box;
    a.disk = L.disk.as<int2>();
    a.nq = nq;
    a.g = L.g.as<float2>();
    a.tfac = L.tfac.as<float>();
    a.parts = ctx->parts.as<float>();
    CtfConst cc = make_ctf_const(p);
    {
        TimedLaunch tl(ctx, "lref_defocus");
        hipLaunchKernelGGL(k_lref_defocus, dim3(n), dim3(256), sizeof(float2) * nq, ctx->stream, a, cc, sign);
        CIM_HIP(hipGetLastError());
    }
    std::vector<float> outp((size_t)n * CIM_PART_STRIDE);
    CIM_HIP(hipMemcpyAsync(outp.data(), ctx->parts.ptr, sizeof(float) * CIM_PART_STRIDE * n, hipMemcpyDeviceToHost,
                           ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i)
        for (int k : {CIM_P_PR, CIM_P_Z}) parts[i * CIM_PART_STRIDE + k] = outp[i * CIM_PART_STRIDE + k];
    ctx->counters["lref_particles"] += n;
    return CIM_OK;
}

}  // extern "C"
