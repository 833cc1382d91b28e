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
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

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

// Position of sample (hx, ky) under operator matrix m (float32, the
// reference's S * R * (hx, ky, 0)), folded into z >= 0 (reconstruct.f90:117-122),
// as the accumulator index of its lower corner and the trilinear fractions.
// Both insert kernels use this, so they place every contribution identically.
__device__ __forceinline__ void place(const float* m, float fh, float fk, const InsertGeom& g, int& gx, int& gy,
                                      int& gz, float& ax, float& ay, float& az, bool& flip) {
    float x = m[0] * fh + m[1] * fk;
    float y = m[3] * fh + m[4] * fk;
    float z = m[6] * fh + m[7] * fk;
    flip = z < 0.f;
    if (flip) {
        x = -x; y = -y; z = -z;
    }
    x += g.off;
    y += g.off;
    z += g.offz;
    const int lx = (int)x, ly = (int)y, lz = (int)z;
    ax = x - (float)lx;
    ay = y - (float)ly;
    az = z - (float)lz;
    gx = lx - (int)g.off + g.maxR;
    gy = ly - (int)g.off + g.maxR;
    gz = lz - (int)g.offz;
}

// Position of sample (hx, ky) under m (the reference's S * R * (hx, ky, 0) in
// float32) with no fold, for the full-cube accumulator of the coset insert:
// index of the lower corner (coordinate + R on each axis) and the trilinear
// fractions x - floor(x), which are exact in float32.
__device__ __forceinline__ void place_cube(const float* m, float fh, float fk, int R, int& gx, int& gy, int& gz,
                                           float& ax, float& ay, float& az) {
    const float x = m[0] * fh + m[1] * fk;
    const float y = m[3] * fh + m[4] * fk;
    const float z = m[6] * fh + m[7] * fk;
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    ax = x - fx;
    ay = y - fy;
    az = z - fz;
    gx = (int)fx + R;
    gy = (int)fy + R;
    gz = (int)fz + R;
}

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
    float ca = 1.f, cw = 1.f;
    if (ctf_img) {
        ca = ctf_img[(long long)p * N + i2];
        cw = ca * ca;
    } else if (apply_ctf) {
        ctf_factors(hk.x, hk.y, g.box, parts + (long long)p * CIM_PART_STRIDE, cc, ca, cw);
    }
    const float va = v.x * ca, vb = v.y * ca, vw = cw;
    const float fh = (float)hk.x, fk = (float)hk.y;
    const float* Mp = M + (long long)p * g.nops * 9;
    const long long dxy = g.dxy, dxy2 = dxy * dxy;
    for (int o = 0; o < g.nops; ++o) {
        const float* m = Mp + 9 * o;
        int gx, gy, gz;
        float ax, ay, az;
        bool flip;
        place(m, fh, fk, g, gx, gy, gz, ax, ay, az, flip);
        const float sb = flip ? -vb : vb;
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
// the backward FFT.  isg = -1 writes cmplx(fa, -fb) (reconstruct.f90:241,
// merge_reconstruct.f90:133), isg = +1 cmplx(fa, fb) (reconstruct_FSC.f90:173).
// The sphere test is dr <= maxR (reconstruct.f90) or, with int_radius,
// int(dr) <= maxR (reconstruct_FSC.f90:101, merge_reconstruct.f90:42).
// acc holds {fa, fb, wg} for x, y in [-maxR, .] (row pitch dxy) and z >= 0.
__global__ void k_fold(const float4* __restrict__ acc, const float4* __restrict__ lo, float2* __restrict__ vol,
                       int box, int maxR, int dxy, int int_radius, float isg) {
    const long long n2 = (long long)box * box;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n2 * box) return;
    const int Z = (int)(id / n2), Y = (int)((id / box) % box), X = (int)(id % box);
    const int c = box / 2;
    const int h = X - c, k = Y - c, l = Z - c;
    float2 outv = make_float2(0.f, 0.f);
    if (abs(h) <= maxR && abs(k) <= maxR && abs(l) <= maxR) {
        const float dr = sqrtf((float)(h * h + k * k + l * l));
        if (int_radius ? ((int)dr <= maxR) : (dr <= (float)maxR)) {
            const long long dxy2 = (long long)dxy * dxy;
            // sums as double: hi + lo in the compensated mode
            auto at = [&](int hh, int kk, int ll) -> double4 {
                const long long i = (long long)ll * dxy2 + (long long)(kk + maxR) * dxy + (hh + maxR);
                const float4 h = acc[i];
                double4 r = make_double4(h.x, h.y, h.z, 0.0);
                if (lo) {
                    const float4 l = lo[i];
                    r.x += l.x;
                    r.y += l.y;
                    r.z += l.z;
                }
                return r;
            };
            const float s = ((h + k + l) & 1) ? -1.f : 1.f;
            if (l != 0) {
                // l > 0 stored directly; l < 0 is the Friedel mate of (-h,-k,-l)
                const int sg = l > 0 ? 1 : -1;
                const double4 a = at(sg * h, sg * k, sg * l);
                const double A = a.x / a.z, B = a.y / a.z;
                outv = make_float2((float)A * s, (l > 0 ? isg : -isg) * (float)B * s);
            } else if (h == 0 && k == 0) {
                const double4 a = at(0, 0, 0);
                outv = make_float2((float)(a.x / a.z), 0.f);
            } else {
                // n1: k > 0, or k == 0 and h > 0; n2 = -n1
                const bool is_n1 = (k > 0) || (k == 0 && h > 0);
                const int h1 = is_n1 ? h : -h, k1 = is_n1 ? k : -k;
                const double4 a1 = at(h1, k1, 0), a2 = at(-h1, -k1, 0);
                const double W = a1.z + a2.z;
                const double A = (a1.x + a2.x) / W;
                const double B = (a1.y - a2.y) / W;
                outv = make_float2((float)A * s, (is_n1 ? isg : -isg) * (float)B * s);
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

// Real part with the checkerboard inside the sphere |x - box/2| <= rmask,
// zero outside (merge_reconstruct_FSC.f90:188-197).
__global__ void k_checker_mask(const float2* __restrict__ vol, float* __restrict__ map, int box, float rmask) {
    const long long n2 = (long long)box * box;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n2 * box) return;
    const int Z = (int)(id / n2), Y = (int)((id / box) % box), X = (int)(id % box);
    const int c = box / 2;
    const float d = sqrtf((float)((X - c) * (X - c) + (Y - c) * (Y - c) + (Z - c) * (Z - c)));
    const float v = vol[id].x;
    map[id] = d <= rmask ? (((X + Y + Z) & 1) ? -v : v) : 0.f;
}

// The (box-8)^3 output block of merge_reconstruct.f90:143-149, whose plane
// index is i/(box-8)^2 + 4 with the 1-based i: the last voxel of each output
// plane is read from the next plane.
__global__ void k_merge_crop(const float* __restrict__ map, float* __restrict__ out, int box) {
    const long long M = box - 8, M2 = M * M;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;   // = i - 1
    if (k >= M2 * M) return;
    const long long lz = (k + 1) / M2 + 4, ky = (k % M2) / M + 4, hx = (k % M2) % M + 4;
    out[k] = map[lz * box * box + ky * box + hx];
}

// Copy between the accumulator and tnf order ((2R+1)^2 (R+1), lz slowest).
// For the compact layout the readout reproduces the reference's flat
// indexing of its (2R+1)^2 (R+1) array (reconstruct_substep.f90:133-136,
// reconstruct_half.f90:127-130): a corner at x = 2R+1 lands on x = 0 of the
// next row, one at y = 2R+1 on y = 0 of the next plane, and anything past
// the last plane is lost.  The accumulator keeps those corners in its extra
// x = 2R+1 / y = 2R+1 cells, so tnf cell f gathers every accumulator cell
// whose flat reference index is f (their weights minus the initial value).
__global__ void k_tnf_io(float4* __restrict__ acc, const float4* __restrict__ lo, float* __restrict__ fa,
                         float* __restrict__ fb, float* __restrict__ wg, int maxR, int dxy, int add_into_acc, int wrap,
                         float wg_init) {
    const int t = 2 * maxR + 1;
    const long long total = (long long)t * t * (maxR + 1);
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total) return;
    const int lz = (int)(id / ((long long)t * t));
    const int ky = (int)((id / t) % t), hx = (int)(id % t);
    const long long dxy2 = (long long)dxy * dxy;
    float4* a = acc + ((long long)lz * dxy2 + (long long)ky * dxy + hx);
    if (add_into_acc) {
        a->x += fa[id];
        a->y += fb[id];
        a->z += wg[id];
        return;
    }
    // compensated mode: the sum is hi + lo, rounded once to float
    auto cell = [&](long long i) {
        float4 e = acc[i];
        if (lo) {
            // one float add of hi and lo is their exact sum rounded once
            const float4 l = lo[i];
            e.x = __fadd_rn(e.x, l.x);
            e.y = __fadd_rn(e.y, l.y);
            e.z = __fadd_rn(e.z, l.z);
        }
        return e;
    };
    float4 v = cell((long long)lz * dxy2 + (long long)ky * dxy + hx);
    if (wrap) {
        // cells (x', y', z') with x' <= t, y' <= t and z' t^2 + y' t + x' = id
        auto add = [&](int x, int y, int z) {
            const float4 e = cell((long long)z * dxy2 + (long long)y * dxy + x);
            v.x += e.x;
            v.y += e.y;
            v.z += e.z - wg_init;
        };
        if (hx == 0 && ky >= 1) add(t, ky - 1, lz);
        if (hx == 0 && ky == 0 && lz >= 1) add(t, t - 1, lz - 1);
        if (ky == 0 && lz >= 1) add(hx, t, lz - 1);
        if (hx == 0 && ky == 1 && lz >= 1) add(t, t, lz - 1);
    }
    fa[id] = v.x;
    fb[id] = v.y;
    wg[id] = v.z;
}

// ---- brick-gather insert ----------------------------------------------------
//
// The accumulator is cut into 8x8x8 bricks; one workgroup owns one brick in
// LDS and gathers every (particle, operator) plane of the batch that passes
// within reach of it: for each such plane it enumerates the lattice points of
// the plane inside the brick's reach box, row by row, places the samples
// they stand for with the same place() as the scatter kernel, and adds the
// corners that fall inside the brick with LDS atomics.  The brick is then
// added to HBM once, so the insert needs no global atomics at all.

constexpr int BRICK = 8;

// Per-particle sample values {a*ca, b*ca, cw} (ctf_factors) on the (hx, ky) half
// plane grid hx in [0, maxR], ky in [-maxR, maxR].
// A sample is inserted iff hx <= hxmax[ky + maxR] (the host's radius test).
// shift: the transform is not centre-shifted yet; each sample takes the phase
// of img2d_FT_centshift here (shift_phase, recentred, as k_centshift would).
// Each thread stages VALS_PER_THREAD samples, 256 apart.  maxbits holds one
// slot per particle (slot p at maxbits[MAXV_STRIDE * (p + 1)], each in its own
// 128-byte line, so the blocks' maxima do not queue on one address): slot p
// collects the largest |value| of particle p (float bits; non-negative floats
// order as unsigned integers), and k_max_slots folds the slots into maxbits[0].
constexpr int VALS_PER_THREAD = 4;
constexpr int MAXV_STRIDE = 32;
__global__ void __launch_bounds__(256) k_insert_vals(FtView ft, int box, int maxR,
                                                     const int* __restrict__ hxmax, const float* __restrict__ parts,
                                                     const float* __restrict__ ctf_img, int apply_ctf, CtfConst cc,
                                                     int shift, float4* __restrict__ vals,
                                                     unsigned int* __restrict__ maxbits) {
    __shared__ float red[4];
    const int p = blockIdx.y;
    const int w = maxR + 1, total = (2 * maxR + 1) * w;
    const float* pr = parts + (long long)p * CIM_PART_STRIDE;
    float mx = 0.f;
#pragma unroll
    for (int k = 0; k < VALS_PER_THREAD; ++k) {
        const int q = (blockIdx.x * VALS_PER_THREAD + k) * 256 + threadIdx.x;
        if (q >= total) break;
        float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
        const int hx = q % w, ky = q / w - maxR;
        if (hx <= hxmax[ky + maxR]) {
            float2 v = ft.p[(long long)p * ft.stride + (long long)(ky + ft.oy) * ft.ld + (hx + ft.ox)];
            if (shift) v = apply_shift(v, shift_phase(hx, ky, box, pr[CIM_P_X], pr[CIM_P_Y], 1));
            float ca = 1.f, cw = 1.f;
            if (ctf_img) {
                const long long N = (long long)box * box;
                const long long i2 = (long long)(ky + box / 2) * box + (hx + box / 2);
                ca = ctf_img[(long long)p * N + i2];
                cw = ca * ca;
            } else if (apply_ctf) {
                ctf_factors(hx, ky, box, pr, cc, ca, cw);
            }
            out = make_float4(v.x * ca, v.y * ca, cw, 1.f);
        }
        vals[(long long)p * total + q] = out;
        mx = fmaxf(mx, fmaxf(fabsf(out.x), fmaxf(fabsf(out.y), fabsf(out.z))));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) mx = fmaxf(mx, red[i]);
        atomicMax(maxbits + MAXV_STRIDE * (p + 1), __float_as_uint(mx));
    }
}

// Clears the n particle slots of maxbits.
__global__ void k_clear_slots(unsigned int* maxbits, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) maxbits[MAXV_STRIDE * (p + 1)] = 0u;
}

// maxbits[0] = the largest of the n particle slots (one workgroup of 256).
__global__ void __launch_bounds__(256) k_max_slots(unsigned int* maxbits, int n) {
    __shared__ unsigned int red[4];
    unsigned int m = 0u;
    for (int p = threadIdx.x; p < n; p += 256) m = max(m, maxbits[MAXV_STRIDE * (p + 1)]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) maxbits[0] = max(max(red[0], red[1]), max(red[2], red[3]));
}

struct BrickArgs {
    float4* acc;
    float4* lo;               // compensated mode: TwoSum residues (else nullptr)
    const unsigned int* maxbits;   // largest |value| of the batch (float bits)
    const float4* vals;       // n x (2maxR+1)(maxR+1); w = 1 where the sample is inserted
    const float* M;           // n x nops x 9
    const int* bricks;        // nb x 3 brick origins (accumulator indices)
    const int* hxmax;         // 2maxR+1: sample (hx, ky) is inserted iff hx <= hxmax[ky + maxR]
    int n, dz;
    InsertGeom g;
};

//
// The brick's sums of a batch are exact: every contribution (rounded to a
// fixed-point grid of 2^-cap) is added as a 64-bit integer.  cap is the same
// for the whole launch: the largest |value| of the batch times the most
// contributions one voxel can take (32 per plane: a 2x2x2 cell around it
// meets at most ~12 lattice points of a plane, doubled on the hx = 0 column)
// stays below 2^62.  Integer LDS adds are fast where float LDS atomics are
// not, and the sums do not depend on the order of the adds.
// The terms are also kept below 2^51 in magnitude, so that fix64 converts
// them exactly.
__device__ __forceinline__ int brick_cap(const BrickArgs& a) {
    const float mx = fmaxf(__uint_as_float(*a.maxbits), 1e-30f);
    int e;
    (void)frexpf(mx, &e);          // mx < 2^e
    int cb = 0;
    while ((1ll << cb) < (long long)a.n * a.g.nops * 32) ++cb;
    return min(min(62 - e - cb, 51 - e), 100);
}

// x (|x| < 2^51) rounded to the nearest integer, ties to even, as a 64-bit
// two's complement value, in float32 and integer arithmetic: x = xh 2^24 + xr
// with xh = rint(x 2^-24) and xr = x - xh 2^24, which the fused multiply-add
// gives exactly (|xr| <= 2^23 and its bits lie within x's); xh 2^24 is an even
// integer, so rint(x) = xh 2^24 + rint(xr), ties included.
__device__ __forceinline__ unsigned long long fix64(float x) {
    const float xh = __builtin_rintf(x * 0x1p-24f);
    const float xr = __builtin_fmaf(-xh, 0x1p24f, x);
    const long long h = (long long)(int)xh;          // |xh| < 2^27
    return (unsigned long long)(h * (1ll << 24) + (long long)(int)__builtin_rintf(xr));
}

// s + c += x: TwoSum (Knuth) adds x to s and the add's rounding error to c.
__device__ __forceinline__ void two_sum_acc(float& s, float& c, float x) {
    const float t = __fadd_rn(s, x);
    const float xp = __fsub_rn(t, s);
    const float err = __fadd_rn(__fsub_rn(s, __fsub_rn(t, xp)), __fsub_rn(x, xp));
    s = t;
    c = __fadd_rn(c, err);
}

// A batch sum S of fixed-point terms (|S| < 2^63) times 2^-cap as three exact
// float32 parts, largest first: S = c2 2^42 + c1 2^21 + c0 with 0 <= c0, c1 < 2^21
// and |c2| < 2^21; sc = {2^(42-cap), 2^(21-cap), 2^-cap}.
__device__ __forceinline__ void fix_parts(long long S, const float* sc, float* part) {
    part[0] = (float)(int)(S >> 42) * sc[0];
    part[1] = (float)(int)((S >> 21) & 0x1fffff) * sc[1];
    part[2] = (float)(int)(S & 0x1fffff) * sc[2];
}

constexpr int BRICK_NV = BRICK * BRICK * BRICK;

// LDS slot of brick voxel i in the queue kernel: a rotation of each block of
// 32 slots by 5 x (block index), so the 64-bit atomics of a wave (corners of
// a patch of one plane) spread over more banks than with the plain layout,
// whatever the plane's orientation.
__device__ __forceinline__ int brick_slot(int i) { return (i & ~31) | ((i + 5 * (i >> 5)) & 31); }

// hi + lo += the batch sum S 2^-cap: its three exact parts (fix_parts) added
// largest first by TwoSum, the rounding errors into lo (compensated); or the
// parts' sum added to hi (plain).
__device__ __forceinline__ void flush_add(float& hi, float* lo, long long S, const float* sc) {
    float part[3];
    fix_parts(S, sc, part);
    if (lo) {
#pragma unroll
        for (int k = 0; k < 3; ++k) two_sum_acc(hi, *lo, part[k]);
    } else {
        float s = 0.f, c = 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k) two_sum_acc(s, c, part[k]);
        hi = __fadd_rn(hi, __fadd_rn(s, c));
    }
}

// The fixed-point scales of a flush: {2^(42-cap), 2^(21-cap), 2^-cap}.
__device__ __forceinline__ void flush_scales(int cap, float* sc) {
    sc[0] = ldexpf(1.f, 42 - cap);
    sc[1] = ldexpf(1.f, 21 - cap);
    sc[2] = ldexpf(1.f, -cap);
}

// One read-modify-write of the brick's exact batch sums into HBM (brick of
// BX x BY x BZ voxels at accumulator index (bx0, by0, bz0), x fastest; SWZ:
// voxel i at LDS slot brick_slot(i)).
template <int BX, int BY, int BZ, bool SWZ>
__device__ void brick_flush(const BrickArgs& a, const unsigned long long* sacc, int bx0, int by0, int bz0,
                            int cap) {
    constexpr int NV = BX * BY * BZ;
    static_assert(!SWZ || NV % 32 == 0, "slot rotation needs whole blocks of 32");
    const InsertGeom& g = a.g;
    const int tid = threadIdx.x;
    const long long dxy = g.dxy;
    float sc[3];
    flush_scales(cap, sc);
    for (int i = tid; i < NV; i += blockDim.x) {
        const int cx = i % BX, cy = (i / BX) % BY, cz = i / (BX * BY);
        const int X = bx0 + cx, Y = by0 + cy, Z = bz0 + cz;
        if (X >= g.dxy || Y >= g.dxy || Z >= a.dz) continue;
        const int si = SWZ ? brick_slot(i) : i;
        const long long sa = (long long)sacc[si], sb = (long long)sacc[NV + si], sw = (long long)sacc[2 * NV + si];
        if ((sa | sb | sw) == 0) continue;
        const long long vi = (long long)Z * dxy * dxy + (long long)Y * dxy + X;
        float4 v = a.acc[vi];
        if (a.lo) {
            // compensated: hi + lo carries the sum; TwoSum keeps the rounding error of each add
            float4 l = a.lo[vi];
            flush_add(v.x, &l.x, sa, sc);
            flush_add(v.y, &l.y, sb, sc);
            flush_add(v.z, &l.z, sw, sc);
            a.lo[vi] = l;
        } else {
            flush_add(v.x, nullptr, sa, sc);
            flush_add(v.y, nullptr, sb, sc);
            flush_add(v.z, nullptr, sw, sc);
        }
        a.acc[vi] = v;
    }
}


__global__ void __launch_bounds__(256) k_insert_brick(BrickArgs a) {
    constexpr int NV = BRICK * BRICK * BRICK;
    __shared__ unsigned long long sacc[3 * NV];
    __shared__ int list[1024];
    __shared__ int nlist;
    const int tid = threadIdx.x;
    const int bi = blockIdx.x;
    const int bx0 = a.bricks[3 * bi], by0 = a.bricks[3 * bi + 1], bz0 = a.bricks[3 * bi + 2];
    const InsertGeom& g = a.g;
    const int R = g.maxR;
    for (int i = tid; i < 3 * NV; i += blockDim.x) sacc[i] = 0ull;
    const int cap = brick_cap(a);
    const float scale = ldexpf(1.f, cap);
    // positions (before the index offsets) whose corners can reach the brick:
    // index gx covers x in [gx - R - 1, gx - R + 1), likewise y; z in [gz - 1, gz + 1)
    const float lo[3] = {(float)(bx0 - R - 1), (float)(by0 - R - 1), (float)(bz0 - 1)};
    const float hi[3] = {(float)(bx0 - R + BRICK), (float)(by0 - R + BRICK), (float)(bz0 + BRICK)};
    const float c[3] = {0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2])};
    const float reach = 0.5f * (float)(BRICK + 1) * 1.7320508f + 0.5f;
    const int w = R + 1;
    const int ntot = a.n * g.nops;
    const float fR = (float)R;
    for (int q0 = 0; q0 < ntot; q0 += 1024) {
        if (tid == 0) nlist = 0;
        __syncthreads();
        for (int q = q0 + tid; q < min(ntot, q0 + 1024); q += blockDim.x) {
            const float* m = a.M + (long long)q * 9;
            const float d = fabsf(m[2] * c[0] + m[5] * c[1] + m[8] * c[2]);   // plane normal = third column
            if (d <= reach) list[atomicAdd(&nlist, 1)] = q;
        }
        __syncthreads();
        const int nl = nlist;
        // Each wave takes its own plane.  The lattice points (h, k) of the plane whose position
        // x(h, k) = h col0 + k col1 lies in the brick's reach box are enumerated row by row: for row
        // k the three slab constraints give the exact h interval (widened by one).  A point with
        // h >= 0 stands for sample (h, k) when that sample is not flipped; a point with h <= 0 for
        // sample (-h, -k) when it is flipped (z < 0 fold).  So every sample is met once, at the
        // point where place() puts it (the h = 0 column: (0, k) and (0, -k) both land there).
        const int lane = tid & 63, wv = tid >> 6, nwv = blockDim.x >> 6;
        for (int li = wv; li < nl; li += nwv) {
            const int q = list[li];
            const int p = q / g.nops;
            const float* m = a.M + (long long)q * 9;
            float vmin = 3e38f, vmax = -3e38f;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const float X = (k & 1) ? hi[0] : lo[0], Y = (k & 2) ? hi[1] : lo[1], Z = (k & 4) ? hi[2] : lo[2];
                const float v = m[1] * X + m[4] * Y + m[7] * Z;
                vmin = fminf(vmin, v);
                vmax = fmaxf(vmax, v);
            }
            const int ky0 = max((int)floorf(vmin) - 1, -R), ky1 = min((int)ceilf(vmax) + 1, R);
            const int nrows = ky1 - ky0 + 1;
            if (nrows <= 0) continue;
            // row interval of lane r (rows <= 64: the box spans < 18 rows)
            int ha = 0, cnt = 0;
            if (lane < nrows) {
                const float kf = (float)(ky0 + lane);
                const float c0[3] = {m[0], m[3], m[6]}, c1[3] = {m[1], m[4], m[7]};
                float ta = -3e38f, tb = 3e38f;
                bool empty = false;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const float base = kf * c1[i];
                    if (fabsf(c0[i]) > 1e-6f) {
                        const float t1 = (lo[i] - base) / c0[i], t2 = (hi[i] - base) / c0[i];
                        ta = fmaxf(ta, fminf(t1, t2));
                        tb = fminf(tb, fmaxf(t1, t2));
                    } else if (!(base >= lo[i] - 1e-3f && base <= hi[i] + 1e-3f)) {
                        empty = true;
                    }
                }
                if (!empty && ta <= tb) {
                    ha = max((int)ceilf(ta) - 1, -R);
                    const int hb = min((int)floorf(tb) + 1, R);
                    cnt = max(hb - ha + 1, 0);
                }
            }
            int wmax = cnt;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) wmax = max(wmax, __shfl_xor(wmax, o, 64));
            if (wmax == 0) continue;
            int w2 = 1;
            while (w2 < wmax) w2 <<= 1;     // columns per row slot (<= 64)
            const int rpp = 64 / w2;          // rows per pass
            const float4* pv = a.vals + (long long)p * (2 * R + 1) * w;
            for (int r0 = 0; r0 < nrows; r0 += rpp) {
                const int r = r0 + lane / w2, col = lane % w2;
                const int rha = __shfl(ha, r < 64 ? r : 0, 64), rcnt = __shfl(cnt, r < 64 ? r : 0, 64);
                const bool have = r < nrows && col < rcnt;
                const int hp = rha + col, kp = ky0 + r;
                // candidate 0: the sample the point stands for; candidate 1: (0, -k) on the h = 0 column
#pragma unroll
                for (int cidx = 0; cidx < 2; ++cidx) {
                    bool ok = have && (cidx == 0 || hp == 0);
                    const bool mirror = cidx == 1 || hp < 0;
                    const int hx = mirror ? -hp : hp, ky = mirror ? -kp : kp;
                    ok = ok && hx >= 0 && hx <= R && ky >= -R && ky <= R;
                    const float4 vv = ok ? pv[(ky + R) * w + hx] : make_float4(0.f, 0.f, 0.f, 0.f);
                    if (!ok || vv.w == 0.f) continue;
                    int gx, gy, gz;
                    float ax, ay, az;
                    bool flip;
                    place(m, (float)hx, (float)ky, g, gx, gy, gz, ax, ay, az, flip);
                    if (flip != mirror) continue;
                    const int ox = gx - bx0, oy = gy - by0, oz = gz - bz0;
                    if (ox < -1 || ox >= BRICK || oy < -1 || oy >= BRICK || oz < -1 || oz >= BRICK) continue;
                    const float va = vv.x, vb = flip ? -vv.y : vv.y, vw = vv.z;
                    const float w00 = (1.f - ax) * (1.f - ay), w10 = ax * (1.f - ay), w01 = (1.f - ax) * ay,
                                w11 = ax * ay;
                    const float wz0 = 1.f - az, wz1 = az;
                    const float wt[8] = {w00 * wz0, w10 * wz0, w01 * wz0, w11 * wz0,
                                         w00 * wz1, w10 * wz1, w01 * wz1, w11 * wz1};
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int cx = ox + (j & 1), cy = oy + ((j >> 1) & 1), cz = oz + (j >> 2);
                        if (cx < 0 || cx >= BRICK || cy < 0 || cy >= BRICK || cz < 0 || cz >= BRICK) continue;
                        const int li3 = (cz * BRICK + cy) * BRICK + cx;
                        atomicAdd(&sacc[li3], fix64(va * wt[j] * scale));
                        atomicAdd(&sacc[NV + li3], fix64(vb * wt[j] * scale));
                        atomicAdd(&sacc[2 * NV + li3], fix64(vw * wt[j] * scale));
                    }
                }
            }
        }
        __syncthreads();
    }
    brick_flush<BRICK, BRICK, BRICK, false>(a, sacc, bx0, by0, bz0, cap);
}

// The same brick gather with the work of a wave packed into full lanes.
// Per plane, the lattice points of the rows (row intervals built as in
// k_insert_brick, over this kernel's brick) are numbered through a wave
// prefix sum, so each lane takes
// one point; the point's candidate samples are placed and tested, and the
// ones whose cell touches the brick go to a queue in LDS (its offsets in
// the brick, fractions, flip and value index).  Whenever the queue holds 64
// entries, 64 lanes add their 8 corners.  Every contribution is the same
// fixed-point term as in k_insert_brick and the sums are exact, so both
// kernels give the same bits, whatever the brick shape BX x BY x BZ (each
// at most 16) and the workgroup size NT.  Bricks larger than 8^3 meet fewer
// (plane, brick) pairs, whose row set-up otherwise costs as much as the
// points; the slot rotation (brick_slot) keeps their atomics' bank conflicts
// at the level of the small bricks.
constexpr int QCAP = 128;   // < 64 left over + 64 lanes: drained after every candidate pass

// CUBE: the coset insert's accumulator -- the full cube of coordinates
// [-R, R+1]^3 with no z fold, each plane a half disk (hx >= 0) placed by
// place_cube; the bricks' z origins are then cube indices too.
template <int BX, int BY, int BZ, int NT, int WPE = 1, bool CUBE = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) k_insert_brickq(BrickArgs a) {
    static_assert(BX <= 16 && BY <= 16 && BZ <= 16 && NT % 64 == 0 && NT <= 1024, "brick shape");
    constexpr int NV = BX * BY * BZ, NW = NT / 64;
    __shared__ unsigned long long sacc[3 * NV];
    __shared__ int list[1024];
    __shared__ int nlist;
    __shared__ int qpos[NW][QCAP];       // (ox+1) | (oy+1) << 5 | (oz+1) << 10 | flip << 15
    __shared__ unsigned qval[NW][QCAP];  // index into a.vals
    __shared__ float qax[NW][QCAP], qay[NW][QCAP], qaz[NW][QCAP];
    __shared__ int shx[2 * 512 + 1];     // hxmax (maxR <= 512)
    const int tid = threadIdx.x;
    const int bi = blockIdx.x;
    const int bx0 = a.bricks[3 * bi], by0 = a.bricks[3 * bi + 1], bz0 = a.bricks[3 * bi + 2];
    const InsertGeom& g = a.g;
    const int R = g.maxR;
    for (int i = tid; i < 3 * NV; i += NT) sacc[i] = 0ull;
    for (int i = tid; i <= 2 * R; i += NT) shx[i] = a.hxmax[i];
    const int cap = brick_cap(a);
    const float scale = ldexpf(1.f, cap);
    const int zsh = CUBE ? R : 0;
    const float lo[3] = {(float)(bx0 - R - 1), (float)(by0 - R - 1), (float)(bz0 - zsh - 1)};
    const float hi[3] = {(float)(bx0 - R + BX), (float)(by0 - R + BY), (float)(bz0 - zsh + BZ)};
    const float c[3] = {0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2])};
    // A point whose cell touches the brick lies in the reach box [lo, hi) (half-widths
    // (B+1)/2 about c), so its plane passes within sum_i half_i |n_i| of c (n the plane
    // normal, the third column); 0.01 covers the float error of the plane.
    const float hw[3] = {0.5f * (float)(BX + 1), 0.5f * (float)(BY + 1), 0.5f * (float)(BZ + 1)};
    const int w = R + 1;
    const int ntot = a.n * g.nops;
    const int lane = tid & 63, wv = tid >> 6;
    int qn = 0;   // entries in this wave's queue (wave-uniform)
    // add the corners of the last cnt entries, one entry per lane
    auto drain = [&](int cnt) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < cnt) {
            const int e = qn - cnt + lane;
            const int pk = qpos[wv][e];
            const int ox = (pk & 31) - 1, oy = ((pk >> 5) & 31) - 1, oz = ((pk >> 10) & 31) - 1;
            const bool flip = (pk >> 15) & 1;
            const float4 vv = a.vals[qval[wv][e]];
            const float ax = qax[wv][e], ay = qay[wv][e], az = qaz[wv][e];
            // values scaled once: (v * scale) * w equals (v * w) * scale for the power of two
            // scale (a product that differs is below 2^-126 and rounds to a zero term)
            const float va = vv.x * scale, vb = (flip ? -vv.y : vv.y) * scale, vw = vv.z * scale;
            const float w00 = (1.f - ax) * (1.f - ay), w10 = ax * (1.f - ay), w01 = (1.f - ax) * ay, w11 = ax * ay;
            const float wz0 = 1.f - az, wz1 = az;
            const float wt[8] = {w00 * wz0, w10 * wz0, w01 * wz0, w11 * wz0, w00 * wz1, w10 * wz1, w01 * wz1, w11 * wz1};
            const bool inx[2] = {ox >= 0, ox + 1 < BX}, iny[2] = {oy >= 0, oy + 1 < BY},
                       inz[2] = {oz >= 0, oz + 1 < BZ};
            const int base = (oz * BY + oy) * BX + ox;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (!(inx[j & 1] && iny[(j >> 1) & 1] && inz[j >> 2])) continue;
                const int li3 = brick_slot(base + (j & 1) + ((j >> 1) & 1) * BX + (j >> 2) * BX * BY);
                atomicAdd(&sacc[li3], fix64(va * wt[j]));
                atomicAdd(&sacc[NV + li3], fix64(vb * wt[j]));
                atomicAdd(&sacc[2 * NV + li3], fix64(vw * wt[j]));
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        qn -= cnt;
    };
    for (int q0 = 0; q0 < ntot; q0 += 1024) {
        if (tid == 0) nlist = 0;
        __syncthreads();
        for (int q = q0 + tid; q < min(ntot, q0 + 1024); q += NT) {
            const float* m = a.M + (long long)q * 9;
            const float d = fabsf(m[2] * c[0] + m[5] * c[1] + m[8] * c[2]);
            if (d <= (hw[0] + 0.01f) * fabsf(m[2]) + (hw[1] + 0.01f) * fabsf(m[5]) + (hw[2] + 0.01f) * fabsf(m[8]))
                list[atomicAdd(&nlist, 1)] = q;
        }
        __syncthreads();
        const int nl = nlist;
        for (int li = wv; li < nl; li += NW) {
            const int q = list[li];
            const int p = q / g.nops;
            const float* m = a.M + (long long)q * 9;
            // k = col1 . x over the reach box: its centre value -/+ the half-widths times |col1|
            const float kc = m[1] * c[0] + m[4] * c[1] + m[7] * c[2];
            const float kh = hw[0] * fabsf(m[1]) + hw[1] * fabsf(m[4]) + hw[2] * fabsf(m[7]);
            const float vmin = kc - kh, vmax = kc + kh;
            // Rows and per-row h intervals of the plane's points in the reach box, with a
            // margin of 0.01 lattice steps over the float error of the bounds (~1e-4); the
            // exact test below decides which points count, as in k_insert_brick, whose
            // wider intervals hold the same points.
            const int ky0 = max((int)floorf(vmin - 0.01f), -R), ky1 = min((int)ceilf(vmax + 0.01f), R);
            const int nrows = ky1 - ky0 + 1;
            if (nrows <= 0) continue;
            int ha = 0, cnt = 0;
            if (lane < nrows) {
                const float kf = (float)(ky0 + lane);
                const float c0[3] = {m[0], m[3], m[6]}, c1[3] = {m[1], m[4], m[7]};
                float ta = -3e38f, tb = 3e38f;
                bool empty = false;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const float base = kf * c1[i];
                    if (fabsf(c0[i]) > 1e-6f) {
                        const float rc = 1.0f / c0[i];
                        const float t1 = (lo[i] - base) * rc, t2 = (hi[i] - base) * rc;
                        ta = fmaxf(ta, fminf(t1, t2));
                        tb = fminf(tb, fmaxf(t1, t2));
                    } else if (!(base >= lo[i] - 1e-3f && base <= hi[i] + 1e-3f)) {
                        empty = true;
                    }
                }
                if (!empty && ta <= tb + 0.02f) {
                    ha = max((int)ceilf(ta - 0.01f), CUBE ? 0 : -R);
                    const int hb = min((int)floorf(tb + 0.01f), R);
                    cnt = max(hb - ha + 1, 0);
                }
            }
            // inclusive prefix sum of the row counts over the lanes
            int pre = cnt;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(pre, o, 64);
                if (lane >= o) pre += u;
            }
            const int total = __shfl(pre, 63, 64);
            const unsigned vbase = (unsigned)p * (unsigned)((2 * R + 1) * w);
            for (int t0 = 0; t0 < total; t0 += 64) {
                const int t = t0 + lane;
                // row of point t: the number of rows whose inclusive prefix is <= t
                int r = 0;
#pragma unroll
                for (int sft = 32; sft > 0; sft >>= 1)
                    if (__shfl(pre, r + sft - 1, 64) <= t) r += sft;
                const int rcnt = __shfl(cnt, r, 64), rha = __shfl(ha, r, 64), rpre = __shfl(pre, r, 64);
                const int hp = rha + (t - (rpre - rcnt)), kp = ky0 + r;
                // the second candidate, (0, -k), exists on the h = 0 column only
                // (no mirror in the cube: a plane is the half disk hx >= 0 itself)
                const int ncand = !CUBE && __any(t < total && hp == 0) ? 2 : 1;
                for (int cidx = 0; cidx < ncand; ++cidx) {
                    bool ok = t < total && (cidx == 0 || hp == 0);
                    const bool mirror = !CUBE && (cidx == 1 || hp < 0);
                    const int hx = mirror ? -hp : hp, ky = mirror ? -kp : kp;
                    ok = ok && hx >= 0 && hx <= R && ky >= -R && ky <= R;
                    const unsigned vi = vbase + (unsigned)((ky + R) * w + hx);
                    if (ok) ok = hx <= shx[ky + R];   // vals[vi].w != 0
                    int gx = 0, gy = 0, gz = 0;
                    float ax = 0.f, ay = 0.f, az = 0.f;
                    int ox = 0, oy = 0, oz = 0;
                    bool flip = false;
                    if (ok) {
                        if (CUBE) place_cube(m, (float)hx, (float)ky, R, gx, gy, gz, ax, ay, az);
                        else place(m, (float)hx, (float)ky, g, gx, gy, gz, ax, ay, az, flip);
                        ox = gx - bx0;
                        oy = gy - by0;
                        oz = gz - bz0;
                        ok = flip == mirror && ox >= -1 && ox < BX && oy >= -1 && oy < BY && oz >= -1 && oz < BZ;
                    }
                    const unsigned long long mask = __ballot(ok);
                    if (ok) {
                        const int e = qn + __popcll(mask & ((1ull << lane) - 1ull));
                        qpos[wv][e] = (ox + 1) | ((oy + 1) << 5) | ((oz + 1) << 10) | ((int)flip << 15);
                        qval[wv][e] = vi;
                        qax[wv][e] = ax;
                        qay[wv][e] = ay;
                        qaz[wv][e] = az;
                    }
                    qn += __popcll(mask);
                    if (qn >= 64) drain(64);
                }
            }
        }
        __syncthreads();
    }
    if (qn > 0) drain(qn);
    __syncthreads();
    brick_flush<BX, BY, BZ, true>(a, sacc, bx0, by0, bz0, cap);
}

// ---- coset insert: symmetrisation of the cube into the half-volume sums -------
//
// The 12 signed-permutation operators T of the icosahedral group map the
// lattice onto itself and the trilinear weight w(x) = prod (1 - |x_i|) onto
// itself: w(g y) = w(y).  With the 60 operators written g_j c_i (g_j in T, c_i
// the 5 coset representatives), the unfolded sum over the 60 operators
// U60(x) = sum_j U5(g_j^T x), where U5 is the sum over the c_i alone.  The
// reference's half-volume sums are U60(x) + conj U60(-x) for z >= 1; on the
// z = 0 plane they split that sum by the side each sample came from, and only
// the fold's combination of x and -x is ever used, which is unchanged when the
// plane holds half of U60(x) + conj U60(-x) (the origin's real part and weight
// are exactly that half).  One thread per half-volume voxel adds that sum
// (fa, fb, wg; the weights are real) to acc, in float64 from the cube's
// hi + lo, and into hi + lo by TwoSum in the compensated mode.
struct TSym {
    int c[12][3];   // row i of operator j: column c & 3, sign -1 when c & 4
};

__global__ void k_coset_sym(const float4* __restrict__ cube, const float4* __restrict__ cube_lo,
                            float4* __restrict__ acc, float4* __restrict__ lo, int R, int dxy, int dz, TSym t) {
    const long long n2 = (long long)dxy * dxy;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n2 * dz) return;
    const int Z = (int)(id / n2), Y = (int)((id / dxy) % dxy), X = (int)(id % dxy);
    const int x[3] = {X - R, Y - R, Z};
    // the 12 permuted values of x and of -x, summed with TwoSum (s) and their errors and the
    // cube's lo parts (c), in float32
    float fa = 0.f, fac = 0.f, fb = 0.f, fbc = 0.f, wg = 0.f, wgc = 0.f;
    auto add = [&](int u0, int u1, int u2, float sgn_b) {
        if (u0 < -R || u0 > R + 1 || u1 < -R || u1 > R + 1 || u2 < -R || u2 > R + 1) return;
        const long long i = ((long long)(u2 + R) * dxy + (u1 + R)) * dxy + (u0 + R);
        const float4 h = cube[i];
        two_sum_acc(fa, fac, h.x);
        two_sum_acc(fb, fbc, sgn_b * h.y);
        two_sum_acc(wg, wgc, h.z);
        if (cube_lo) {
            const float4 l = cube_lo[i];
            fac = __fadd_rn(fac, l.x);
            fbc = __fadd_rn(fbc, sgn_b * l.y);
            wgc = __fadd_rn(wgc, l.z);
        }
    };
#pragma unroll 1
    for (int j = 0; j < 12; ++j) {
        // u = g_j^T x: u[c_i] = s_i x[i]
        int u[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int c = t.c[j][i];
            u[c & 3] = (c & 4) ? -x[i] : x[i];
        }
        add(u[0], u[1], u[2], 1.f);
        add(-u[0], -u[1], -u[2], -1.f);
    }
    if (Z == 0) {
        fa *= 0.5f; fac *= 0.5f;
        fb *= 0.5f; fbc *= 0.5f;
        wg *= 0.5f; wgc *= 0.5f;
    }
    if (fa == 0.f && fb == 0.f && wg == 0.f && fac == 0.f && fbc == 0.f && wgc == 0.f) return;
    float4 v = acc[id];
    if (lo) {
        float4 l = lo[id];
        two_sum_acc(v.x, l.x, fa);
        two_sum_acc(v.y, l.y, fb);
        two_sum_acc(v.z, l.z, wg);
        l.x = __fadd_rn(l.x, fac);
        l.y = __fadd_rn(l.y, fbc);
        l.z = __fadd_rn(l.z, wgc);
        lo[id] = l;
    } else {
        v.x = __fadd_rn(v.x, __fadd_rn(fa, fac));
        v.y = __fadd_rn(v.y, __fadd_rn(fb, fbc));
        v.z = __fadd_rn(v.z, __fadd_rn(wg, wgc));
    }
    acc[id] = v;
}

// ---- host side -----------------------------------------------------------

// Symmetry operators of icosprocess's local reconstruction (mode 6,
// reconstruction_local.f90:30-55 with icosprocess/SRC/icos2f_sym_mat.f90:250-540):
// z rotations by k * 360/N degrees, each entry evaluated as
// cos(a * 3.1415926 / 180.0) in REAL*4, for c12, c15, c5a, c6, c3, c4, c2,
// c8; the identity for c1.  Any other name (the reference has no branch for
// it there) returns -1.
static int icosprocess_local_ops(const char* sym, std::vector<double>& r) {
    std::string s(sym ? sym : "");
    while (!s.empty() && s.back() == ' ') s.pop_back();
    int N = 0;
    float step = 0.f;
    if (s == "c1") { N = 1; step = 0.f; }
    else if (s == "c2") { N = 2; step = 180.f; }
    else if (s == "c3") { N = 3; step = 120.f; }
    else if (s == "c4") { N = 4; step = 90.f; }
    else if (s == "c5a") { N = 5; step = 72.f; }
    else if (s == "c6") { N = 6; step = 60.f; }
    else if (s == "c7") { N = 7; step = 0.f; }
    else if (s == "c8") { N = 8; step = 45.f; }
    else if (s == "c12") { N = 12; step = 30.f; }
    else if (s == "c15") { N = 15; step = 24.f; }
    else return -1;
    r.assign(9 * (size_t)N, 0.0);
    for (int k = 0; k < N; ++k) {
        // literal angles k*step (icos2f_sym_matc2..c15); c7 computes
        // float(i-1)*360.0/7.0 (icos2f_sym_mat.f90:520-523)
        const float a = step > 0.f ? (float)k * step : (float)k * 360.0f / (float)N;
        const float rad = a * 3.1415926f / 180.0f;
        const float c = cosf(rad), sn = sinf(rad);
        double* m = &r[9 * (size_t)k];
        m[0] = c;
        m[1] = -sn;
        m[3] = sn;
        m[4] = c;
        m[8] = 1.0;
    }
    return N;
}

}  // namespace cim

extern "C" int cim_icosprocess_sym_ops(const char* sym, double* ops, int* nops) {
    std::vector<double> r;
    const int n = cim::icosprocess_local_ops(sym, r);
    if (n < 0) return CIM_E_ARG;
    if (nops) *nops = n;
    if (ops)
        for (size_t i = 0; i < r.size(); ++i) ops[i] = r[i];
    return CIM_OK;
}

namespace cim {


// Brick shapes of the queue kernel (x, y, z voxels; workgroup size): CIM_BRICKQ
// selects one by its index (default 0).
// Measured at box 384 (maxR 190, 64 particles, compensated): 16x16x8 4.44 ms,
// 16^3 4.46 ms, 8^3 5.12 ms per launch.
// Shapes 3 and 4 (16x8x8, 53 KB and 48 KB of LDS) let three workgroups share a
// CU: shape 3 asks the compiler for 6 waves per SIMD (78 VGPRs, no spills;
// shape 0 holds 87 and is held to 4 waves per SIMD by its 77 KB of LDS).
// Measured at box 384 (maxR 190, 512 particles, compensated, same bits):
// 16x16x8 34.8 ms, 16x8x8 at 512 threads 33.8 ms, at 384 threads 48.0 ms.
constexpr int kBrickQ[][4] = {{16, 16, 8, 512}, {16, 16, 16, 1024}, {8, 8, 8, 256}, {16, 8, 8, 512}, {16, 8, 8, 384}};
static int brickq_shape() {
    const char* e = getenv("CIM_BRICKQ");
    const int n = (int)(sizeof(kBrickQ) / sizeof(kBrickQ[0]));
    const int v = e ? atoi(e) : 0;
    return v >= 0 && v < n ? v : 0;
}

int rec_begin(cim_ctx* ctx, int box, int maxR, int layout, float rcut, float wg_init, int nops, const double* ops) {
    if (box < 8 || (box & 1)) return set_error(ctx, CIM_E_ARG, "box %d must be even and >= 8", box);
    if (maxR < 1 || maxR > box / 2 - 2)
        return set_error(ctx, CIM_E_ARG, "maxR %d outside [1, box/2-2] for box %d", maxR, box);
    const bool comp = (layout & CIM_REC_COMPENSATED) != 0;
    layout &= ~CIM_REC_COMPENSATED;
    if (layout != 0 && layout != 1)
        return set_error(ctx, CIM_E_ARG, "layout must be 0 (full) or 1 (compact), optionally | CIM_REC_COMPENSATED");
    if (rcut > (float)maxR) return set_error(ctx, CIM_E_ARG, "rcut %.2f > maxR %d", rcut, maxR);
    if (nops < 1 || nops > 360) return set_error(ctx, CIM_E_ARG, "nops %d outside [1, 360]", nops);
    RecState& r = ctx->rec;
    r.box = box;
    r.maxR = maxR;
    r.layout = layout;
    r.comp = comp;
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
    if (comp) {
        CIM_ALLOC(r.lo, sizeof(float4) * nvox);
        hipLaunchKernelGGL(k_fill_acc, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, ctx->stream,
                           r.lo.as<float4>(), nvox, 0.f);
        CIM_HIP(hipGetLastError());
    }
    std::vector<double> hops(9 * nops);
    if (ops) {
        for (int i = 0; i < 9 * nops; ++i) hops[i] = ops[i];
        // every operator must be a rotation: the accumulator is sized for |S x| = |x| <= maxR
        for (int o = 0; o < nops; ++o) {
            const double* S = hops.data() + 9 * o;
            double dev = 0.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double d = S[i] * S[j] + S[3 + i] * S[3 + j] + S[6 + i] * S[6 + j] - (i == j ? 1.0 : 0.0);
                    dev = std::max(dev, std::fabs(d));
                }
            const double det = S[0] * (S[4] * S[8] - S[5] * S[7]) - S[1] * (S[3] * S[8] - S[5] * S[6]) +
                               S[2] * (S[3] * S[7] - S[4] * S[6]);
            if (!(dev < 1e-5) || !(det > 0.0))
                return set_error(ctx, CIM_E_ARG, "symmetry operator %d is not a proper rotation (|S^T S - I| = %.3g)",
                                 o, dev);
        }
    } else {
        if (nops != 60) return set_error(ctx, CIM_E_ARG, "ops == NULL requires nops == 60 (icosahedral)");
        cim_icos60_ref_order(hops.data());
    }
    CIM_ALLOC(r.ops, sizeof(double) * 9 * nops);
    CIM_HIP(hipMemcpyAsync(r.ops.ptr, hops.data(), sizeof(double) * 9 * nops, hipMemcpyHostToDevice, ctx->stream));
    // half-plane sample list in the loop order of reconstruct.f90:103-106
    std::vector<int2> smp;
    std::vector<int> hxmax(2 * maxR + 1, -1);
    for (int ky = -maxR; ky <= maxR; ++ky)
        for (int hx = 0; hx <= maxR; ++hx) {
            float dr = std::sqrt((float)(hx * hx + ky * ky));
            if (dr <= rcut) {
                smp.push_back(make_int2(hx, ky));
                hxmax[ky + maxR] = hx;
            }
        }
    r.nsamples = (int)smp.size();
    CIM_ALLOC(r.hxmax, sizeof(int) * hxmax.size());
    CIM_HIP(hipMemcpyAsync(r.hxmax.ptr, hxmax.data(), sizeof(int) * hxmax.size(), hipMemcpyHostToDevice, ctx->stream));
    // bricks of the accumulator that a sample of radius <= rcut can reach: 8^3 for the
    // row kernel, the queue kernel's shape for it
    auto brick_list = [&](int bx, int by, int bz, bool central_first, DevBuf& buf, int& nb) -> int {
        std::vector<int> bl;
        std::vector<std::pair<double, int>> order;
        for (int z = 0; z < r.dz; z += bz)
            for (int y = 0; y < r.dxy; y += by)
                for (int x = 0; x < r.dxy; x += bx) {
                    // nearest point of the brick's reach (relative coordinates) to the origin
                    auto nearest = [&](int a0, int n, int shift) {
                        const int lo = a0 - shift - 1, hi = a0 - shift + n;
                        return lo > 0 ? lo : (hi < 0 ? hi : 0);
                    };
                    const double nx = nearest(x, bx, r.maxR), ny = nearest(y, by, r.maxR), nz = nearest(z, bz, 0);
                    const double d = std::sqrt(nx * nx + ny * ny + nz * nz);
                    if (d <= rcut + 2.0) {
                        order.push_back({d, (int)order.size()});
                        bl.push_back(x);
                        bl.push_back(y);
                        bl.push_back(z);
                    }
                }
        // central bricks meet the most planes: starting them first shortens the tail
        // of the launch (the bricks are disjoint, so the order changes no sum)
        if (central_first) std::stable_sort(order.begin(), order.end());
        std::vector<int> ol(bl.size());
        for (size_t i = 0; i < order.size(); ++i)
            for (int j = 0; j < 3; ++j) ol[3 * i + j] = bl[3 * order[i].second + j];
        nb = (int)order.size();
        CIM_ALLOC(buf, sizeof(int) * std::max<size_t>(ol.size(), 3));
        CIM_HIP(hipMemcpyAsync(buf.ptr, ol.data(), sizeof(int) * ol.size(), hipMemcpyHostToDevice, ctx->stream));
        CIM_HIP(hipStreamSynchronize(ctx->stream));
        return CIM_OK;
    };
    int rc = brick_list(BRICK, BRICK, BRICK, false, r.bricks, r.nbricks);
    if (rc != CIM_OK) return rc;
    const int qs = brickq_shape();
    const char* ord = getenv("CIM_BRICK_ORDER");
    const bool central = !(ord && strcmp(ord, "raster") == 0);
    rc = brick_list(kBrickQ[qs][0], kBrickQ[qs][1], kBrickQ[qs][2], central, r.bricksq, r.nbricksq);
    if (rc != CIM_OK) return rc;
    r.qshape = qs;
    // coset insert: the default for the full layout with the 60 icosahedral operators;
    // CIM_INSERT=brick (the 60-plane queue kernel), brickrows or scatter select another
    const char* em = getenv("CIM_INSERT");
    r.cosets = layout == 0 && ops == nullptr && (!em || strcmp(em, "cosets") == 0) && maxR <= 512;
    r.cube_dirty = false;
    if (!r.cosets) {
        r.cube.release();
        r.cube_lo.release();
    }
    if (r.cosets) {
        // T: the operators that are signed permutations; representatives: the first operator of
        // each coset T c, in the table's order
        std::vector<int> tj;
        for (int o = 0; o < 60; ++o) {
            const double* S = hops.data() + 9 * o;
            bool perm = true;
            for (int i = 0; i < 9; ++i) perm = perm && (std::fabs(S[i]) < 1e-9 || std::fabs(std::fabs(S[i]) - 1.0) < 1e-9);
            if (perm) tj.push_back(o);
        }
        std::vector<int> reps;
        std::vector<bool> covered(60, false);
        for (int o = 0; o < 60 && tj.size() == 12; ++o) {
            if (covered[o]) continue;
            reps.push_back(o);
            for (int j : tj) {
                const double* G = hops.data() + 9 * j;
                const double* C = hops.data() + 9 * o;
                double P[9];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) P[3 * a + b] = G[3 * a] * C[b] + G[3 * a + 1] * C[3 + b] + G[3 * a + 2] * C[6 + b];
                for (int q = 0; q < 60; ++q) {
                    double dev = 0.0;
                    for (int i = 0; i < 9; ++i) dev = std::max(dev, std::fabs(P[i] - hops[9 * q + i]));
                    if (dev < 1e-6) covered[q] = true;
                }
            }
        }
        const bool all = std::count(covered.begin(), covered.end(), true) == 60;
        if (tj.size() != 12 || reps.size() != 5 || !all)
            return set_error(ctx, CIM_E_STATE, "icosahedral table: no 12-operator signed-permutation subgroup");
        for (int j = 0; j < 12; ++j)
            for (int i = 0; i < 3; ++i)
                for (int c = 0; c < 3; ++c) {
                    const double v = hops[9 * tj[j] + 3 * i + c];
                    if (std::fabs(v) > 0.5) r.tsym[j][i] = c | (v < 0 ? 4 : 0);
                }
        std::vector<double> cops(45);
        for (int i = 0; i < 5; ++i)
            for (int e = 0; e < 9; ++e) cops[9 * i + e] = hops[9 * reps[i] + e];
        CIM_ALLOC(r.cops, sizeof(double) * 45);
        CIM_HIP(hipMemcpyAsync(r.cops.ptr, cops.data(), sizeof(double) * 45, hipMemcpyHostToDevice, ctx->stream));
        const long long ncube = (long long)r.dxy * r.dxy * r.dxy;
        CIM_ALLOC(r.cube, sizeof(float4) * ncube);
        hipLaunchKernelGGL(k_fill_acc, dim3((unsigned)((ncube + 255) / 256)), dim3(256), 0, ctx->stream,
                           r.cube.as<float4>(), ncube, 0.f);
        CIM_HIP(hipGetLastError());
        if (comp) {
            CIM_ALLOC(r.cube_lo, sizeof(float4) * ncube);
            hipLaunchKernelGGL(k_fill_acc, dim3((unsigned)((ncube + 255) / 256)), dim3(256), 0, ctx->stream,
                               r.cube_lo.as<float4>(), ncube, 0.f);
            CIM_HIP(hipGetLastError());
        }
        // bricks of the cube within reach, central first (16 x 16 x 8)
        std::vector<std::pair<double, int>> order;
        std::vector<int> bl;
        const int bx = 16, by = 16, bz = 8;
        auto nearest = [&](int a0, int n) {
            const int lo = a0 - maxR - 1, hi = a0 - maxR + n;
            return lo > 0 ? lo : (hi < 0 ? hi : 0);
        };
        for (int z = 0; z < r.dxy; z += bz)
            for (int y = 0; y < r.dxy; y += by)
                for (int x = 0; x < r.dxy; x += bx) {
                    const double nx = nearest(x, bx), ny = nearest(y, by), nz = nearest(z, bz);
                    const double d = std::sqrt(nx * nx + ny * ny + nz * nz);
                    if (d <= rcut + 2.0) {
                        order.push_back({d, (int)order.size()});
                        bl.push_back(x);
                        bl.push_back(y);
                        bl.push_back(z);
                    }
                }
        std::stable_sort(order.begin(), order.end());
        std::vector<int> ol(bl.size());
        for (size_t i = 0; i < order.size(); ++i)
            for (int j = 0; j < 3; ++j) ol[3 * i + j] = bl[3 * order[i].second + j];
        r.nbricksc = (int)order.size();
        CIM_ALLOC(r.bricksc, sizeof(int) * std::max<size_t>(ol.size(), 3));
        CIM_HIP(hipMemcpyAsync(r.bricksc.ptr, ol.data(), sizeof(int) * ol.size(), hipMemcpyHostToDevice, ctx->stream));
    }
    CIM_ALLOC(r.samples, sizeof(int2) * smp.size());
    CIM_HIP(hipMemcpyAsync(r.samples.ptr, smp.data(), sizeof(int2) * smp.size(), hipMemcpyHostToDevice, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    r.active = true;
    return CIM_OK;
}

// Insert n particles whose centred, centre-shifted FTs are in ft (device).
int centre_shift(cim_ctx* ctx, float2* ft, int box, int n, const float* parts_dev, int recentre);

// The atomic scatter kernel is selected (CIM_INSERT=scatter, plain sums only).
bool insert_scatter(const cim_ctx* ctx) {
    const char* mode = getenv("CIM_INSERT");
    return mode && strcmp(mode, "scatter") == 0 && !ctx->rec.comp;
}

// shift: ft is not centre-shifted yet (the value staging applies it; the
// scatter kernel, which reads ft itself, gets it applied first).  view: where
// the staging reads the transform (a half-plane crop); null: ft, whole
// box x box transforms.
int rec_insert_ft_dev(cim_ctx* ctx, int n, float2* ft, const float* parts_dev, const float* ctf_img_dev,
                      int apply_ctf, const cim_params* p, int shift, const FtView* view) {
    RecState& r = ctx->rec;
    if (!r.active) return set_error(ctx, CIM_E_STATE, "cim_rec_begin has not been called");
    if (n <= 0) return CIM_OK;
    const FtView whole{ft, (long long)r.box * r.box, r.box, r.box / 2, r.box / 2};
    const FtView fv = view ? *view : whole;
    const int total_vals = (2 * r.maxR + 1) * (r.maxR + 1);
    // the coset kernel indexes the batch's values with 32 bits as the queue kernel does
    const bool cosets = r.cosets && (long long)total_vals * n < (1ll << 32);
    const int nplanes = cosets ? 5 : r.nops;
    CIM_ALLOC(ctx->aux2, sizeof(float) * 9 * (size_t)n * nplanes);
    {
        const int tot = n * nplanes;
        hipLaunchKernelGGL(k_compose, dim3((tot + 255) / 256), dim3(256), 0, ctx->stream, parts_dev, n,
                           cosets ? r.cops.as<double>() : r.ops.as<double>(), nplanes, ctx->aux2.as<float>());
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
    const char* mode = getenv("CIM_INSERT");
    if (insert_scatter(ctx)) {
        if (view) return set_error(ctx, CIM_E_STATE, "the scatter insert reads whole transforms");
        if (shift) {
            const int rc = centre_shift(ctx, ft, r.box, n, parts_dev, 1);
            if (rc) return rc;
        }
        TimedLaunch tl(ctx, "insert");
        hipLaunchKernelGGL(k_insert, dim3((r.nsamples + 255) / 256, n), dim3(256), 0, ctx->stream, r.acc.as<float4>(),
                           ft, r.samples.as<int2>(), r.nsamples, ctx->aux2.as<float>(), parts_dev, ctf_img_dev,
                           apply_ctf, cc, g);
        CIM_HIP(hipGetLastError());
    } else {
        const int total = (2 * r.maxR + 1) * (r.maxR + 1);
        CIM_ALLOC(r.vals, sizeof(float4) * (size_t)total * n);
        CIM_ALLOC(r.maxv, sizeof(unsigned int) * MAXV_STRIDE * ((size_t)n + 1));
        hipLaunchKernelGGL(k_clear_slots, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, r.maxv.as<unsigned int>(), n);
        const int per_block = 256 * VALS_PER_THREAD;
        hipLaunchKernelGGL(k_insert_vals, dim3((total + per_block - 1) / per_block, n), dim3(256), 0, ctx->stream, fv,
                           r.box, r.maxR, r.hxmax.as<int>(), parts_dev, ctf_img_dev, apply_ctf, cc, shift,
                           r.vals.as<float4>(), r.maxv.as<unsigned int>());
        hipLaunchKernelGGL(k_max_slots, dim3(1), dim3(256), 0, ctx->stream, r.maxv.as<unsigned int>(), n);
        CIM_HIP(hipGetLastError());
        BrickArgs ba;
        ba.acc = r.acc.as<float4>();
        ba.lo = r.comp ? r.lo.as<float4>() : nullptr;
        ba.maxbits = r.maxv.as<unsigned int>();
        ba.vals = r.vals.as<float4>();
        ba.M = ctx->aux2.as<float>();
        ba.bricks = r.bricks.as<int>();
        ba.hxmax = r.hxmax.as<int>();
        ba.n = n;
        ba.dz = r.dz;
        ba.g = g;
        // the queue kernel indexes the batch's values with 32 bits
        const bool rows = (mode && strcmp(mode, "brickrows") == 0) || (long long)total * n >= (1ll << 32) ||
                          r.maxR > 512;
        TimedLaunch tl(ctx, "insert");
        if (cosets) {
            // the 5 coset planes per particle into the cube (flushed by rec_flush_cosets)
            ba.acc = r.cube.as<float4>();
            ba.lo = r.comp ? r.cube_lo.as<float4>() : nullptr;
            ba.bricks = r.bricksc.as<int>();
            ba.dz = r.dxy;
            ba.g.nops = 5;
            hipLaunchKernelGGL((k_insert_brickq<16, 16, 8, 512, 1, true>), dim3(r.nbricksc), dim3(512), 0,
                               ctx->stream, ba);
            r.cube_dirty = true;
            ctx->counters["insert_coset_particles"] += n;
        } else if (rows) hipLaunchKernelGGL(k_insert_brick, dim3(r.nbricks), dim3(256), 0, ctx->stream, ba);
        else {
            ba.bricks = r.bricksq.as<int>();
            if (r.qshape == 0)
                hipLaunchKernelGGL((k_insert_brickq<16, 16, 8, 512>), dim3(r.nbricksq), dim3(512), 0, ctx->stream, ba);
            else if (r.qshape == 1)
                hipLaunchKernelGGL((k_insert_brickq<16, 16, 16, 1024>), dim3(r.nbricksq), dim3(1024), 0, ctx->stream, ba);
            else if (r.qshape == 2)
                hipLaunchKernelGGL((k_insert_brickq<8, 8, 8, 256>), dim3(r.nbricksq), dim3(256), 0, ctx->stream, ba);
            else if (r.qshape == 3)
                hipLaunchKernelGGL((k_insert_brickq<16, 8, 8, 512, 6>), dim3(r.nbricksq), dim3(512), 0, ctx->stream, ba);
            else
                hipLaunchKernelGGL((k_insert_brickq<16, 8, 8, 384>), dim3(r.nbricksq), dim3(384), 0, ctx->stream, ba);
        }
        CIM_HIP(hipGetLastError());
    }
    r.inserted += n;
    ctx->counters["insert_samples"] += (long long)n * r.nsamples;
    ctx->counters["insert_particles"] += n;
    return CIM_OK;
}

// Add the coset insert's cube, symmetrised by T, into the half-volume sums
// and clear it; a no-op when nothing was inserted since the last flush.
int rec_flush_cosets(cim_ctx* ctx) {
    RecState& r = ctx->rec;
    if (!r.cosets || !r.cube_dirty) return CIM_OK;
    TSym t;
    for (int j = 0; j < 12; ++j)
        for (int i = 0; i < 3; ++i) t.c[j][i] = r.tsym[j][i];
    const long long nvox = (long long)r.dxy * r.dxy * r.dz;
    {
        TimedLaunch tl(ctx, "coset_sym");
        hipLaunchKernelGGL(k_coset_sym, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, ctx->stream,
                           r.cube.as<float4>(), r.comp ? r.cube_lo.as<float4>() : (const float4*)nullptr,
                           r.acc.as<float4>(), r.comp ? r.lo.as<float4>() : (float4*)nullptr, r.maxR, r.dxy, r.dz, t);
        CIM_HIP(hipGetLastError());
    }
    const long long ncube = (long long)r.dxy * r.dxy * r.dxy;
    hipLaunchKernelGGL(k_fill_acc, dim3((unsigned)((ncube + 255) / 256)), dim3(256), 0, ctx->stream,
                       r.cube.as<float4>(), ncube, 0.f);
    CIM_HIP(hipGetLastError());
    if (r.comp) {
        hipLaunchKernelGGL(k_fill_acc, dim3((unsigned)((ncube + 255) / 256)), dim3(256), 0, ctx->stream,
                           r.cube_lo.as<float4>(), ncube, 0.f);
        CIM_HIP(hipGetLastError());
    }
    r.cube_dirty = false;
    return CIM_OK;
}

int rec_tnf_io(cim_ctx* ctx, float* fa, float* fb, float* wg, int add) {
    RecState& r = ctx->rec;
    if (!r.active) return set_error(ctx, CIM_E_STATE, "cim_rec_begin has not been called");
    if (int rc = rec_flush_cosets(ctx)) return rc;
    const int t = 2 * r.maxR + 1;
    const long long total = (long long)t * t * (r.maxR + 1);
    CIM_ALLOC(ctx->out, sizeof(float) * 3 * total);
    float* d = ctx->out.as<float>();
    if (add) {
        CIM_HIP(hipMemcpyAsync(d, fa, sizeof(float) * total, hipMemcpyHostToDevice, ctx->stream));
        CIM_HIP(hipMemcpyAsync(d + total, fb, sizeof(float) * total, hipMemcpyHostToDevice, ctx->stream));
        CIM_HIP(hipMemcpyAsync(d + 2 * total, wg, sizeof(float) * total, hipMemcpyHostToDevice, ctx->stream));
    }
    hipLaunchKernelGGL(k_tnf_io, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, r.acc.as<float4>(),
                       r.comp ? r.lo.as<float4>() : (const float4*)nullptr, d, d + total, d + 2 * total, r.maxR, r.dxy,
                       add, r.layout == 1 ? 1 : 0, r.wg_init);
    CIM_HIP(hipGetLastError());
    if (!add) {
        CIM_HIP(hipMemcpyAsync(fa, d, sizeof(float) * total, hipMemcpyDeviceToHost, ctx->stream));
        CIM_HIP(hipMemcpyAsync(fb, d + total, sizeof(float) * total, hipMemcpyDeviceToHost, ctx->stream));
        CIM_HIP(hipMemcpyAsync(wg, d + 2 * total, sizeof(float) * total, hipMemcpyDeviceToHost, ctx->stream));
    }
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

// Fold {fa, fb, wg} sums (row pitch dxy, z >= 0) into the full complex
// volume, inverse 3D FFT, real part with the checkerboard (everywhere, or
// only inside radius rmask when rmask >= 0, zero outside).
int fold_to_map_dev(cim_ctx* ctx, const float4* acc, const float4* lo, int dxy, int box, int maxR, int int_radius,
                    float isg, float rmask, float* map_dev) {
    const long long nvox = (long long)box * box * box;
    const unsigned nb = (unsigned)((nvox + 255) / 256);
    CIM_ALLOC(ctx->ft2, sizeof(float2) * nvox);
    float2* vol = ctx->ft2.as<float2>();
    {
        TimedLaunch tl(ctx, "fold");
        hipLaunchKernelGGL(k_fold, dim3(nb), dim3(256), 0, ctx->stream, acc, lo, vol, box, maxR, dxy, int_radius,
                           isg);
        CIM_HIP(hipGetLastError());
    }
    int rc = fft3d_dev(ctx, vol, box, -1, 0);
    if (rc) return rc;
    if (rmask < 0.f)
        hipLaunchKernelGGL(k_checker_real, dim3(nb), dim3(256), 0, ctx->stream, vol, map_dev, box);
    else
        hipLaunchKernelGGL(k_checker_mask, dim3(nb), dim3(256), 0, ctx->stream, vol, map_dev, box, rmask);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

// merge_reconstruct.f90:37-149 on the accumulated sums: integer-radius fold,
// inverse FFT, checkerboard, (box-8)^3 crop.  out_dev holds (box-8)^3 floats.
int rec_merge_dev(cim_ctx* ctx, float* out_dev, int nicosavg, float imgmask) {
    RecState& r = ctx->rec;
    if (!r.active) return set_error(ctx, CIM_E_STATE, "cim_rec_begin has not been called");
    if (int rc = rec_flush_cosets(ctx)) return rc;
    const long long nvox = (long long)r.box * r.box * r.box;
    CIM_ALLOC(ctx->aux, sizeof(float) * nvox);
    int rc = fold_to_map_dev(ctx, r.acc.as<float4>(), r.comp ? r.lo.as<float4>() : nullptr, r.dxy, r.box, r.maxR, 1,
                             -1.f, -1.f, ctx->aux.as<float>());
    if (rc) return rc;
    // merge_reconstruct.f90:142: merge's img3d_icosavg on the full map, before the crop
    if (nicosavg == 1 && (rc = icosavg_dev(ctx, ctx->aux.as<float>(), nvox, r.box, imgmask, CIM_ICOSAVG_MERGE)))
        return rc;
    const long long m3 = (long long)(r.box - 8) * (r.box - 8) * (r.box - 8);
    hipLaunchKernelGGL(k_merge_crop, dim3((unsigned)((m3 + 255) / 256)), dim3(256), 0, ctx->stream,
                       ctx->aux.as<float>(), out_dev, r.box);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

int rec_finish_dev(cim_ctx* ctx, float* map_dev) {
    RecState& r = ctx->rec;
    if (!r.active) return set_error(ctx, CIM_E_STATE, "cim_rec_begin has not been called");
    if (int rc = rec_flush_cosets(ctx)) return rc;
    return fold_to_map_dev(ctx, r.acc.as<float4>(), r.comp ? r.lo.as<float4>() : nullptr, r.dxy, r.box, r.maxR, 0,
                           -1.f, -1.f, map_dev);
}

}  // namespace cim
