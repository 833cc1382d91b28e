// Proposal network of the learned search: six 3x3 convolutions of 64
// channels with stride 2 and ReLU (128^2 -> 2^2), then a linear head of
// CIM_CNN_OUT outputs (per hypothesis: the first two columns of a rotation
// and a confidence logit).  Activations are NHWC float32; every product runs
// on the matrix cores (v_mfma_f32_32x32x16_bf16, bfloat16 operands, float32
// sums) through one tiled GEMM with strided operands:
//   forward  Y = im2col(X) W        (M = pixels, K = 9 Cin, N = 64)
//   dX       col2im(dY W^T)         (gather of the up to 4 outputs a pixel feeds)
//   dW       im2col(X)^T dY         (split over pixels, float32 atomics)
// The reference has no counterpart: its global search scores every grid
// orientation (ort_globesearch.f90:68-106); the network proposes where to look.
#include "cim_device.h"

namespace cim {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int CNN_IMG = 128;
constexpr int CNN_CH = 64;
constexpr int CNN_LAYERS = 6;
constexpr int CNN_FEAT = 2 * 2 * CNN_CH;

// GEMM C[m, n] = sum_k A(m, k) B(k, n) with A(m, k) = A[m sam + k sak] and
// B(k, n) = B[k sbk + n sbn]; 64 x 64 tiles of C per workgroup of 4 waves,
// each 32 x 32; K in steps of 32 staged through LDS as bfloat16.  EPI: 0
// store, 1 add the bias of column n then ReLU, 2 add atomically (split K:
// blockIdx.z takes K range [z kchunk, (z + 1) kchunk)).
struct GemmArgs {
    const float* A;
    const float* B;
    float* C;
    const float* bias;
    long long M, K;
    int N;
    long long sam, sak, sbk, sbn, ldc;
    long long kchunk;
};

template <int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
    constexpr int BM = 64, BN = 64, BK = 32, P = BK + 8;
    __shared__ __attribute__((aligned(16))) __bf16 sA[BM][P];
    __shared__ __attribute__((aligned(16))) __bf16 sB[BN][P];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long long m0 = (long long)blockIdx.x * BM;
    const int n0 = blockIdx.y * BN;
    const long long k0 = (long long)blockIdx.z * g.kchunk;
    const long long k1 = min(g.K, k0 + g.kchunk);
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    f32x16 acc;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.f;
    for (long long kb = k0; kb < k1; kb += BK) {
        // stage A (64 x 32) and B (32 x 64, stored by column) as bfloat16
        for (int e = tid; e < BM * BK; e += 256) {
            const int r = e / BK, c = e - r * BK;
            const long long m = m0 + r, k = kb + c;
            const float v = (m < g.M && k < k1) ? g.A[m * g.sam + k * g.sak] : 0.f;
            sA[r][c] = (__bf16)v;
        }
        for (int e = tid; e < BN * BK; e += 256) {
            const int c = e / BK, kk = e - c * BK;
            const long long k = kb + kk;
            const int n = n0 + c;
            const float v = (n < g.N && k < k1) ? g.B[k * g.sbk + (long long)n * g.sbn] : 0.f;
            sB[c][kk] = (__bf16)v;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            // A: lane l holds row l % 32, k = 8 (l / 32) + e; B: column l % 32, the same k
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[wr + (lane & 31)][ks + 8 * (lane >> 5)]);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sB[wc + (lane & 31)][ks + 8 * (lane >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // register e of lane l: row 8 (e / 4) + 4 (l / 32) + e % 4, column l % 32
    const int n = n0 + wc + (lane & 31);
    if (n >= g.N) return;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const long long m = m0 + wr + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        if (m >= g.M) continue;
        float* c = g.C + m * g.ldc + n;
        if (EPI == 0) *c = acc[e];
        else if (EPI == 1) *c = fmaxf(acc[e] + g.bias[n], 0.f);
        else atomicAdd(c, acc[e]);
    }
}

// 3x3 patches of stride 2 and padding 1 of NHWC x (H x H x C per image):
// cols[(b, y, x), (ky, kx, c)] for the Ho x Ho outputs.
__global__ void k_im2col(const float* __restrict__ x, int B, int H, int C, float* __restrict__ cols) {
    const int Ho = H / 2;
    const long long total = (long long)B * Ho * Ho * 9 * C;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        long long r = i / C;
        const int t = (int)(r % 9);
        r /= 9;
        const int ox = (int)(r % Ho);
        r /= Ho;
        const int oy = (int)(r % Ho);
        const int b = (int)(r / Ho);
        const int iy = 2 * oy + t / 3 - 1, ix = 2 * ox + t % 3 - 1;
        cols[i] = (iy >= 0 && iy < H && ix >= 0 && ix < H) ? x[(((long long)b * H + iy) * H + ix) * C + c] : 0.f;
    }
}

// The adjoint of k_im2col: dx[b, iy, ix, c] = the sum over the taps (ky, kx)
// whose output (y, x) = ((iy + 1 - ky) / 2, (ix + 1 - kx) / 2) exists of
// dcols[(b, y, x), (ky, kx, c)] (a gather, so the sums are fixed in order).
__global__ void k_col2im(const float* __restrict__ dcols, int B, int H, int C, float* __restrict__ dx) {
    const int Ho = H / 2;
    const long long total = (long long)B * H * H * C;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        long long r = i / C;
        const int ix = (int)(r % H);
        r /= H;
        const int iy = (int)(r % H);
        const int b = (int)(r / H);
        float s = 0.f;
        for (int ky = 0; ky < 3; ++ky) {
            const int ty = iy + 1 - ky;
            if (ty < 0 || (ty & 1) || ty / 2 >= Ho) continue;
            for (int kx = 0; kx < 3; ++kx) {
                const int tx = ix + 1 - kx;
                if (tx < 0 || (tx & 1) || tx / 2 >= Ho) continue;
                const long long row = ((long long)b * Ho + ty / 2) * Ho + tx / 2;
                s += dcols[(row * 9 + ky * 3 + kx) * C + c];
            }
        }
        dx[i] = s;
    }
}

// dy *= (y > 0): the ReLU's derivative at its output y.
__global__ void k_relu_bwd(float* __restrict__ dy, const float* __restrict__ y, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        if (!(y[i] > 0.f)) dy[i] = 0.f;
}

// db[c] += the sum over rows of dy[:, c] (rows x C), per block then atomically.
__global__ void k_colsum(const float* __restrict__ dy, long long rows, int C, float* __restrict__ db) {
    __shared__ float s[256];
    const int c = threadIdx.x % C, lanes = blockDim.x / C, sub = threadIdx.x / C;
    float acc = 0.f;
    for (long long r = (long long)blockIdx.x * lanes + sub; r < rows; r += (long long)gridDim.x * lanes)
        acc += dy[r * C + c];
    s[threadIdx.x] = acc;
    __syncthreads();
    if (sub == 0) {
        for (int j = 1; j < lanes; ++j) acc += s[j * C + c];
        atomicAdd(&db[c], acc);
    }
}

__global__ void k_fill(float* __restrict__ p, long long n, float v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// Forward head with the bias row: out[b, j] = sum_i feat[b, i] W[i, j] + bias[j].
__global__ void k_head_bias(float* __restrict__ out, const float* __restrict__ bias, int B, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < B * N) out[i] += bias[i % N];
}

unsigned grid_for(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 65536); }

}  // namespace

struct CnnLayout {
    long long w[CNN_LAYERS + 1], b[CNN_LAYERS + 1], total;
    int cin[CNN_LAYERS], hin[CNN_LAYERS];
};

static CnnLayout cnn_layout() {
    CnnLayout L{};
    long long off = 0;
    int h = CNN_IMG, c = 1;
    for (int l = 0; l < CNN_LAYERS; ++l) {
        L.cin[l] = c;
        L.hin[l] = h;
        L.w[l] = off;
        off += 9LL * c * CNN_CH;
        L.b[l] = off;
        off += CNN_CH;
        c = CNN_CH;
        h /= 2;
    }
    L.w[CNN_LAYERS] = off;
    off += (long long)CNN_FEAT * CIM_CNN_OUT;
    L.b[CNN_LAYERS] = off;
    off += CIM_CNN_OUT;
    L.total = off;
    return L;
}

static int gemm(cim_ctx* ctx, int epi, const float* A, long long sam, long long sak, const float* B, long long sbk,
                long long sbn, float* C, long long ldc, long long M, long long K, int N, const float* bias,
                int splits) {
    GemmArgs g{A, B, C, bias, M, K, N, sam, sak, sbk, sbn, ldc, 0};
    splits = std::max(1, splits);
    g.kchunk = ((K + splits - 1) / splits + 31) / 32 * 32;
    const dim3 grid((unsigned)((M + 63) / 64), (unsigned)((N + 63) / 64), (unsigned)((K + g.kchunk - 1) / g.kchunk));
    if (epi == 0) hipLaunchKernelGGL(k_gemm<0>, grid, dim3(256), 0, ctx->stream, g);
    else if (epi == 1) hipLaunchKernelGGL(k_gemm<1>, grid, dim3(256), 0, ctx->stream, g);
    else hipLaunchKernelGGL(k_gemm<2>, grid, dim3(256), 0, ctx->stream, g);
    CIM_HIP(hipGetLastError());
    return CIM_OK;
}

}  // namespace cim

using namespace cim;

extern "C" {

int cim_cnn_nparams(void) { return (int)cnn_layout().total; }

int cim_cnn_forward(cim_ctx* ctx, int B, const void* images_dev, const float* params, void* out_dev) {
    if (!ctx) return CIM_E_ARG;
    if (B <= 0 || !images_dev || !params || !out_dev) return set_error(ctx, CIM_E_ARG, "cim_cnn_forward: bad arguments");
    CIM_HIP(hipSetDevice(ctx->device));
    const CnnLayout L = cnn_layout();
    CnnState& s = ctx->cnn;
    CIM_ALLOC(s.params, sizeof(float) * (size_t)L.total);
    CIM_HIP(hipMemcpyAsync(s.params.ptr, params, sizeof(float) * (size_t)L.total, hipMemcpyHostToDevice, ctx->stream));
    // activations: the input of layer l at act[l] (layer 0 reads the images), its output at act[l + 1]
    long long aoff[CNN_LAYERS + 2];
    long long tot = 0, colmax = 0;
    for (int l = 0; l <= CNN_LAYERS; ++l) {
        aoff[l] = tot;
        const int h = l == 0 ? CNN_IMG : CNN_IMG >> l;
        tot += (long long)B * h * h * (l == 0 ? 1 : CNN_CH);
    }
    aoff[CNN_LAYERS + 1] = tot;
    for (int l = 0; l < CNN_LAYERS; ++l) {
        const int ho = L.hin[l] / 2;
        colmax = std::max(colmax, (long long)B * ho * ho * 9 * L.cin[l]);
    }
    CIM_ALLOC(s.act, sizeof(float) * (size_t)tot);
    CIM_ALLOC(s.cols, sizeof(float) * (size_t)colmax);
    float* act = s.act.as<float>();
    const float* prm = s.params.as<float>();
    CIM_HIP(hipMemcpyAsync(act, images_dev, sizeof(float) * (size_t)B * CNN_IMG * CNN_IMG, hipMemcpyDeviceToDevice,
                           ctx->stream));
    for (int l = 0; l < CNN_LAYERS; ++l) {
        const int H = L.hin[l], C = L.cin[l], Ho = H / 2;
        const long long M = (long long)B * Ho * Ho, K = 9LL * C;
        hipLaunchKernelGGL(k_im2col, dim3(grid_for(M * K)), dim3(256), 0, ctx->stream, act + aoff[l], B, H, C,
                           s.cols.as<float>());
        CIM_HIP(hipGetLastError());
        int rc = gemm(ctx, 1, s.cols.as<float>(), K, 1, prm + L.w[l], CNN_CH, 1, act + aoff[l + 1], CNN_CH, M, K,
                      CNN_CH, prm + L.b[l], 1);
        if (rc) return rc;
    }
    // head: features (B x 256, the 2 x 2 x 64 NHWC output flattened) times W (256 x OUT)
    float* out = static_cast<float*>(out_dev);
    int rc = gemm(ctx, 0, act + aoff[CNN_LAYERS], CNN_FEAT, 1, prm + L.w[CNN_LAYERS], CIM_CNN_OUT, 1, out, CIM_CNN_OUT,
                  B, CNN_FEAT, CIM_CNN_OUT, nullptr, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_head_bias, dim3((B * CIM_CNN_OUT + 255) / 256), dim3(256), 0, ctx->stream, out,
                       prm + L.b[CNN_LAYERS], B, CIM_CNN_OUT);
    CIM_HIP(hipGetLastError());
    s.batch = B;
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

int cim_cnn_backward(cim_ctx* ctx, const void* dout_dev, float* grads) {
    if (!ctx) return CIM_E_ARG;
    CnnState& s = ctx->cnn;
    if (s.batch <= 0) return set_error(ctx, CIM_E_STATE, "cim_cnn_forward has not been called");
    if (!dout_dev || !grads) return set_error(ctx, CIM_E_ARG, "cim_cnn_backward: NULL argument");
    CIM_HIP(hipSetDevice(ctx->device));
    const CnnLayout L = cnn_layout();
    const int B = s.batch;
    long long aoff[CNN_LAYERS + 2];
    long long tot = 0, dmax = 0;
    for (int l = 0; l <= CNN_LAYERS; ++l) {
        aoff[l] = tot;
        const int h = l == 0 ? CNN_IMG : CNN_IMG >> l;
        const long long sz = (long long)B * h * h * (l == 0 ? 1 : CNN_CH);
        tot += sz;
        dmax = std::max(dmax, sz);
    }
    CIM_ALLOC(s.grads, sizeof(float) * (size_t)L.total);
    CIM_ALLOC(s.dy, sizeof(float) * (size_t)dmax);
    CIM_ALLOC(s.dx, sizeof(float) * (size_t)dmax);
    float* gr = s.grads.as<float>();
    const float* prm = s.params.as<float>();
    const float* act = s.act.as<float>();
    hipLaunchKernelGGL(k_fill, dim3(grid_for(L.total)), dim3(256), 0, ctx->stream, gr, L.total, 0.f);
    CIM_HIP(hipGetLastError());
    const float* dout = static_cast<const float*>(dout_dev);
    // head: dW = feat^T dout, db = colsum(dout), dfeat = dout W^T
    int rc = gemm(ctx, 2, act + aoff[CNN_LAYERS], 1, CNN_FEAT, dout, CIM_CNN_OUT, 1, gr + L.w[CNN_LAYERS],
                  CIM_CNN_OUT, CNN_FEAT, B, CIM_CNN_OUT, nullptr, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_colsum, dim3(1), dim3(CIM_CNN_OUT * (256 / CIM_CNN_OUT)), 0, ctx->stream, dout, (long long)B,
                       CIM_CNN_OUT, gr + L.b[CNN_LAYERS]);
    CIM_HIP(hipGetLastError());
    rc = gemm(ctx, 0, dout, CIM_CNN_OUT, 1, prm + L.w[CNN_LAYERS], 1, CIM_CNN_OUT, s.dy.as<float>(), CNN_FEAT, B,
              CIM_CNN_OUT, CNN_FEAT, nullptr, 1);
    if (rc) return rc;
    for (int l = CNN_LAYERS - 1; l >= 0; --l) {
        const int H = L.hin[l], C = L.cin[l], Ho = H / 2;
        const long long M = (long long)B * Ho * Ho, K = 9LL * C;
        float* dy = s.dy.as<float>();
        // through the ReLU at this layer's output
        hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(M * CNN_CH)), dim3(256), 0, ctx->stream, dy, act + aoff[l + 1],
                           M * CNN_CH);
        CIM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_colsum, dim3((unsigned)std::min<long long>((M + 3) / 4, 1024)), dim3(256), 0, ctx->stream,
                           dy, M, CNN_CH, gr + L.b[l]);
        CIM_HIP(hipGetLastError());
        // dW = cols^T dy, split over the M pixels
        hipLaunchKernelGGL(k_im2col, dim3(grid_for(M * K)), dim3(256), 0, ctx->stream, act + aoff[l], B, H, C,
                           s.cols.as<float>());
        CIM_HIP(hipGetLastError());
        const int splits = (int)std::min<long long>(std::max<long long>(1, M / 2048), 256);
        rc = gemm(ctx, 2, s.cols.as<float>(), 1, K, dy, CNN_CH, 1, gr + L.w[l], CNN_CH, K, M, CNN_CH, nullptr, splits);
        if (rc) return rc;
        if (l == 0) break;
        // dx = col2im(dy W^T)
        rc = gemm(ctx, 0, dy, CNN_CH, 1, prm + L.w[l], 1, CNN_CH, s.cols.as<float>(), K, M, CNN_CH, (int)K, nullptr, 1);
        if (rc) return rc;
        hipLaunchKernelGGL(k_col2im, dim3(grid_for((long long)B * H * H * C)), dim3(256), 0, ctx->stream,
                           s.cols.as<float>(), B, H, C, s.dx.as<float>());
        CIM_HIP(hipGetLastError());
        std::swap(s.dy, s.dx);
    }
    CIM_HIP(hipMemcpyAsync(grads, gr, sizeof(float) * (size_t)L.total, hipMemcpyDeviceToHost, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

}  // extern "C"

extern "C" int cim_debug_cnn_act(cim_ctx* ctx, int layer, float* out) {
    if (!ctx) return CIM_E_ARG;
    CnnState& s = ctx->cnn;
    if (s.batch <= 0) return set_error(ctx, CIM_E_STATE, "cim_cnn_forward has not been called");
    if (layer < 0 || layer > CNN_LAYERS || !out) return set_error(ctx, CIM_E_ARG, "cim_debug_cnn_act: layer %d", layer);
    CIM_HIP(hipSetDevice(ctx->device));
    long long off = 0, sz = 0;
    for (int l = 0; l <= layer; ++l) {
        const int h = l == 0 ? CNN_IMG : CNN_IMG >> l;
        sz = (long long)s.batch * h * h * (l == 0 ? 1 : CNN_CH);
        if (l < layer) off += sz;
    }
    CIM_HIP(hipMemcpyAsync(out, s.act.as<float>() + off, sizeof(float) * (size_t)sz, hipMemcpyDeviceToHost, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}
