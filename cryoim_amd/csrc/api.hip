// C ABI entry points of libcryoim_hip.so: context, FFT and reconstruction.
#include <cmath>
#include <cstring>

#include "cim_device.h"

namespace cim {

int set_error(cim_ctx* c, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

static hipEvent_t take_event(cim_ctx* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

TimedLaunch::TimedLaunch(cim_ctx* c, const char* n) : ctx(c), name(n) {
    if (!ctx->timing) return;
    a = take_event(ctx);
    if (a) (void)hipEventRecord(a, ctx->stream);
}

TimedLaunch::~TimedLaunch() {
    if (!ctx->timing || !a) return;
    hipEvent_t b = take_event(ctx);
    if (!b) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->pending.push_back({name, a, b});
}

void resolve_timing(cim_ctx* c) {
    if (c->pending.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (auto& p : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            KTimer& t = c->timers[p.name];
            t.total_ms += ms;
            t.launches += 1;
        }
        c->event_pool.push_back(p.a);
        c->event_pool.push_back(p.b);
    }
    c->pending.clear();
}

int rec_begin(cim_ctx* ctx, int box, int maxR, int layout, float rcut, float wg_init, int nops, const double* ops);
int rec_insert_ft_dev(cim_ctx* ctx, int n, float2* ft, const float* parts_dev, const float* ctf_img_dev,
                      int apply_ctf, const cim_params* p, int shift,
                      const FtView* view = nullptr);
int rec_tnf_io(cim_ctx* ctx, float* fa, float* fb, float* wg, int add);
int rec_finish_dev(cim_ctx* ctx, float* map_dev);

static int check_params(cim_ctx* ctx, const cim_params* p) {
    if (!p) return set_error(ctx, CIM_E_ARG, "params is NULL");
    if (p->box < 8 || p->box > 1024 || (p->box & 1)) return set_error(ctx, CIM_E_ARG, "box %d unsupported", p->box);
    if (p->maxR < 1 || p->maxR > p->box / 2 - 2)
        return set_error(ctx, CIM_E_ARG, "maxR %d outside [1, box/2-2]", p->maxR);
    if (p->minR < 0 || p->minR > p->maxR) return set_error(ctx, CIM_E_ARG, "minR %d outside [0, maxR]", p->minR);
    if (p->apply_ctf && !(p->apix > 0.f && p->vol > 0.f))
        return set_error(ctx, CIM_E_ARG, "CTF needs apix > 0 and vol > 0");
    if (p->norm_on < 0 || p->norm_on > 2) return set_error(ctx, CIM_E_ARG, "norm_on %d is not 0, 1 or 2", p->norm_on);
    if (p->corr_cycles < 0 || p->corr_cycles > 64)
        return set_error(ctx, CIM_E_ARG, "corr_cycles %d outside [0, 64]", p->corr_cycles);
    if (p->corr_cycles > 0 && !p->apply_ctf)
        return set_error(ctx, CIM_E_ARG, "a defocus series (corr_cycles > 0) needs apply_ctf");
    return CIM_OK;
}

int upload_parts(cim_ctx* ctx, int n, const float* parts) {
    CIM_ALLOC(ctx->parts, sizeof(float) * CIM_PART_STRIDE * (size_t)n);
    CIM_HIP(hipMemcpyAsync(ctx->parts.ptr, parts, sizeof(float) * CIM_PART_STRIDE * (size_t)n, hipMemcpyHostToDevice,
                           ctx->stream));
    return CIM_OK;
}

int check_params_public(cim_ctx* ctx, const cim_params* p) { return check_params(ctx, p); }

}  // namespace cim

using namespace cim;

extern "C" {

int cim_init(cim_ctx** out, int device) {
    if (!out) return CIM_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CIM_E_HIP;
    if (device < 0 || device >= ndev) return CIM_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return CIM_E_HIP;
    cim_ctx* ctx = new cim_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return CIM_E_HIP;
    }
    // this device's copy of the scratch array of cim_local_box_stack_noalloc
    if (lbox_scratch_setup(ctx) != CIM_OK) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return CIM_E_HIP;
    }
    *out = ctx;
    return CIM_OK;
}

static void step_release(cim_ctx* ctx);

int cim_finalize(cim_ctx* ctx) {
    if (!ctx) return CIM_E_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    step_release(ctx);
    ctx->cnn.release();
    for (auto& kv : ctx->plans) (void)hipFree(kv.second.tw);
    ctx->rec.acc.release();
    ctx->rec.lo.release();
    ctx->rec.ops.release();
    ctx->rec.samples.release();
    ctx->rec.bricks.release();
    ctx->rec.bricksq.release();
    ctx->rec.hxmax.release();
    ctx->rec.vals.release();
    ctx->rec.maxv.release();
    ctx->rec.cube.release();
    ctx->rec.cube_lo.release();
    ctx->rec.cops.release();
    ctx->rec.bricksc.release();
    ctx->srch.tft.release();
    ctx->srch.tfcut.release();
    ctx->srch.wk.release();
    ctx->lref.release();
    ctx->cen.release();
    ctx->lbox.release();
    for (DevBuf* b : {&ctx->img, &ctx->ft, &ctx->ft2, &ctx->ftrow, &ctx->ftcrop, &ctx->prepst, &ctx->aux, &ctx->aux2, &ctx->parts,
                      &ctx->out, &ctx->fsc_in, &ctx->fsc_nrm, &ctx->fsc_acc, &ctx->fsc_sums, &ctx->fsc_wt, &ctx->avg_ops, &ctx->avg_acc})
        b->release();
    resolve_timing(ctx);
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CIM_OK;
}

const char* cim_last_error(const cim_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

void* cim_get_stream(cim_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int cim_sync(cim_ctx* ctx) {
    if (!ctx) return CIM_E_ARG;
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

int cim_timing(cim_ctx* ctx, int on) {
    if (!ctx) return CIM_E_ARG;
    resolve_timing(ctx);
    ctx->timing = on != 0;
    ctx->timers.clear();
    return CIM_OK;
}

int cim_timing_get(cim_ctx* ctx, const char* name, double* avg_ms, long long* launches) {
    if (!ctx || !name) return CIM_E_ARG;
    resolve_timing(ctx);
    auto it = ctx->timers.find(name);
    if (it == ctx->timers.end()) {
        if (avg_ms) *avg_ms = 0.0;
        if (launches) *launches = 0;
        return CIM_OK;
    }
    if (avg_ms) *avg_ms = it->second.launches ? it->second.total_ms / it->second.launches : 0.0;
    if (launches) *launches = it->second.launches;
    return CIM_OK;
}

int cim_counter_get(cim_ctx* ctx, const char* name, long long* value) {
    if (!ctx || !name || !value) return CIM_E_ARG;
    auto it = ctx->counters.find(name);
    *value = it == ctx->counters.end() ? 0 : it->second;
    return CIM_OK;
}

int cim_counter_reset(cim_ctx* ctx) {
    if (!ctx) return CIM_E_ARG;
    ctx->counters.clear();
    return CIM_OK;
}

// ---- FFT ------------------------------------------------------------------

static int fft_host(cim_ctx* ctx, float* re, float* im, int n, long long count, int dims, int batch, int dir,
                    int conj_out) {
    if (!ctx) return CIM_E_ARG;
    if (!re || !im || n < 2 || batch < 1 || (dir != 1 && dir != -1)) return set_error(ctx, CIM_E_ARG, "bad FFT args");
    CIM_HIP(hipSetDevice(ctx->device));
    CIM_ALLOC(ctx->img, sizeof(float) * 2 * count);
    CIM_ALLOC(ctx->ft, sizeof(float2) * count);
    float* dre = ctx->img.as<float>();
    float* dim = dre + count;
    CIM_HIP(hipMemcpyAsync(dre, re, sizeof(float) * count, hipMemcpyHostToDevice, ctx->stream));
    CIM_HIP(hipMemcpyAsync(dim, im, sizeof(float) * count, hipMemcpyHostToDevice, ctx->stream));
    int rc = interleave(ctx, dre, dim, ctx->ft.as<float2>(), count);
    if (rc) return rc;
    rc = dims == 2 ? fft2d_dev(ctx, ctx->ft.as<float2>(), n, batch, dir, conj_out)
                   : fft3d_dev(ctx, ctx->ft.as<float2>(), n, dir, conj_out);
    if (rc) return rc;
    rc = deinterleave(ctx, ctx->ft.as<float2>(), dre, dim, count);
    if (rc) return rc;
    CIM_HIP(hipMemcpyAsync(re, dre, sizeof(float) * count, hipMemcpyDeviceToHost, ctx->stream));
    CIM_HIP(hipMemcpyAsync(im, dim, sizeof(float) * count, hipMemcpyDeviceToHost, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

int cim_fft2d(cim_ctx* ctx, float* re, float* im, int n, int batch, int dir, int conj_out) {
    return fft_host(ctx, re, im, n, (long long)n * n * batch, 2, batch, dir, conj_out);
}

int cim_fft3d(cim_ctx* ctx, float* re, float* im, int n, int dir, int conj_out) {
    return fft_host(ctx, re, im, n, (long long)n * n * n, 3, 1, dir, conj_out);
}

int cim_fft2d_dev(cim_ctx* ctx, void* data, int n, int batch, int dir, int conj_out) {
    if (!ctx || !data || batch < 1 || (dir != 1 && dir != -1)) return ctx ? set_error(ctx, CIM_E_ARG, "bad FFT args") : CIM_E_ARG;
    CIM_HIP(hipSetDevice(ctx->device));
    int rc = fft2d_dev(ctx, (float2*)data, n, batch, dir, conj_out);
    if (rc) return rc;
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

int cim_fft3d_dev(cim_ctx* ctx, void* data, int n, int dir, int conj_out) {
    if (!ctx || !data || (dir != 1 && dir != -1)) return ctx ? set_error(ctx, CIM_E_ARG, "bad FFT args") : CIM_E_ARG;
    CIM_HIP(hipSetDevice(ctx->device));
    int rc = fft3d_dev(ctx, (float2*)data, n, dir, conj_out);
    if (rc) return rc;
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

// ---- reconstruction ----------------------------------------------------------

// Preparation, transform and insert of n device images whose records are in
// ctx->parts.  The staging reads only the half plane 0 <= hx <= maxR,
// |ky| <= maxR, so the transform is pruned to it (rows keep those columns, the
// columns pass keeps those rows: the same values as the whole transform);
// the scatter kernel and radii past the box take whole transforms.  The
// centre shift is applied where the values are staged.
static int insert_images_dev(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev) {
    const int box = p->box, R = ctx->rec.maxR, c = box / 2;
    const long long N = (long long)box * box;
    if (R <= c - 1 && !insert_scatter(ctx) && !prep_fused()) {
        // the same with a prepared copy of the images (CIM_PREP_FUSED=0)
        CIM_ALLOC(ctx->ft, sizeof(float2) * N * n);
        int rc = prep_images_dev(ctx, p, n, images_dev, ctx->parts.as<float>(), ctx->ft.as<float2>(), p->norm_on,
                                 p->mask_on);
        if (rc) return rc;
        const int Dx = R + 1, Dy = 2 * R + 1;
        CIM_ALLOC(ctx->ftrow, sizeof(float2) * (size_t)n * box * Dx);
        CIM_ALLOC(ctx->ftcrop, sizeof(float2) * (size_t)n * Dx * Dy);
        rc = fft2d_crop_dev(ctx, ctx->ft.as<float2>(), box, n, c, Dx, c - R, Dy, p->conj_out, nullptr,
                            ctx->ftrow.as<float2>(), ctx->ftcrop.as<float2>());
        if (rc) return rc;
        const FtView v{ctx->ftcrop.as<float2>(), (long long)Dx * Dy, Dx, 0, R};
        return rec_insert_ft_dev(ctx, n, nullptr, ctx->parts.as<float>(), nullptr, p->apply_ctf, p, 1, &v);
    }
    if (R <= c - 1 && !insert_scatter(ctx)) {
        // the images are prepared as the row pass reads them, from their sums
        CIM_ALLOC(ctx->prepst, sizeof(PrepStats) * (size_t)n);
        int rc = prep_stats_dev(ctx, p, n, images_dev, ctx->parts.as<float>(), p->norm_on, p->mask_on,
                                ctx->prepst.as<PrepStats>());
        if (rc) return rc;
        const int Dx = R + 1, Dy = 2 * R + 1;
        CIM_ALLOC(ctx->ftrow, sizeof(float2) * (size_t)n * box * Dx);
        CIM_ALLOC(ctx->ftcrop, sizeof(float2) * (size_t)n * Dx * Dy);
        rc = fft2d_crop_real_dev(ctx, images_dev, ctx->prepst.as<PrepStats>(), p->norm_on, p->mask_on, p->imgmask,
                                 box, n, c, Dx, c - R, Dy, p->conj_out, nullptr, ctx->ftrow.as<float2>(),
                                 ctx->ftcrop.as<float2>());
        if (rc) return rc;
        const FtView v{ctx->ftcrop.as<float2>(), (long long)Dx * Dy, Dx, 0, R};
        return rec_insert_ft_dev(ctx, n, nullptr, ctx->parts.as<float>(), nullptr, p->apply_ctf, p, 1, &v);
    }
    CIM_ALLOC(ctx->ft, sizeof(float2) * N * n);
    int rc = preprocess_images(ctx, p, n, images_dev, ctx->parts.as<float>(), ctx->ft.as<float2>(), p->norm_on,
                               p->mask_on, p->conj_out);
    if (rc) return rc;
    return rec_insert_ft_dev(ctx, n, ctx->ft.as<float2>(), ctx->parts.as<float>(), nullptr, p->apply_ctf, p, 1);
}

int cim_rec_begin(cim_ctx* ctx, int box, int maxR, int layout, float rcut, float wg_init, int nops,
                  const double* ops) {
    if (!ctx) return CIM_E_ARG;
    CIM_HIP(hipSetDevice(ctx->device));
    return rec_begin(ctx, box, maxR, layout, rcut, wg_init, nops, ops);
}

static int rec_insert_images_dev(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev,
                                 const float* parts) {
    int rc = check_params(ctx, p);
    if (rc) return rc;
    if (!ctx->rec.active) return set_error(ctx, CIM_E_STATE, "cim_rec_begin has not been called");
    if (p->box != ctx->rec.box) return set_error(ctx, CIM_E_ARG, "box %d differs from cim_rec_begin box %d", p->box,
                                                 ctx->rec.box);
    if (n <= 0) return CIM_OK;
    // compact the used particles (CIM_P_USE) to the front
    std::vector<float> used;
    std::vector<int> idx;
    for (int i = 0; i < n; ++i)
        if (parts[i * CIM_PART_STRIDE + CIM_P_USE] != 0.f) {
            idx.push_back(i);
            used.insert(used.end(), parts + i * CIM_PART_STRIDE, parts + (i + 1) * CIM_PART_STRIDE);
        }
    const int nu = (int)idx.size();
    if (nu == 0) return CIM_OK;
    const long long N = (long long)p->box * p->box;
    rc = upload_parts(ctx, nu, used.data());
    if (rc) return rc;
    const float* src = images_dev;
    if (nu != n) {
        CIM_ALLOC(ctx->aux, sizeof(float) * N * nu);
        for (int k = 0; k < nu; ++k)
            CIM_HIP(hipMemcpyAsync(ctx->aux.as<float>() + k * N, images_dev + (long long)idx[k] * N, sizeof(float) * N,
                                   hipMemcpyDeviceToDevice, ctx->stream));
        src = ctx->aux.as<float>();
    }
    rc = insert_images_dev(ctx, p, nu, src);
    if (rc) return rc;
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

// The insert of n device records (every record, CIM_P_USE not read), on the
// stream with no host synchronisation.
static int rec_insert_records_dev(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev,
                                  const float* parts_dev) {
    int rc = check_params(ctx, p);
    if (rc) return rc;
    if (!ctx->rec.active) return set_error(ctx, CIM_E_STATE, "cim_rec_begin has not been called");
    if (p->box != ctx->rec.box) return set_error(ctx, CIM_E_ARG, "box %d differs from cim_rec_begin box %d", p->box,
                                                 ctx->rec.box);
    const long long N = (long long)p->box * p->box;
    CIM_ALLOC(ctx->parts, sizeof(float) * CIM_PART_STRIDE * (size_t)n);
    CIM_HIP(hipMemcpyAsync(ctx->parts.ptr, parts_dev, sizeof(float) * CIM_PART_STRIDE * (size_t)n,
                           hipMemcpyDeviceToDevice, ctx->stream));
    return insert_images_dev(ctx, p, n, images_dev);
}

static void step_release(cim_ctx* ctx) {
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->gexec = nullptr;
    ctx->graph = nullptr;
}

int cim_step_capture(cim_ctx* ctx, const cim_params* psearch, const cim_params* pinsert, int n,
                     const void* images_dev, float* parts_dev) {
    if (!ctx) return CIM_E_ARG;
    if (!psearch || !pinsert || !images_dev || !parts_dev || n <= 0)
        return set_error(ctx, CIM_E_ARG, "cim_step_capture: NULL argument or n <= 0");
    CIM_HIP(hipSetDevice(ctx->device));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    step_release(ctx);
    // timing events are not recorded into the graph
    const bool timing = ctx->timing;
    ctx->timing = false;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
        ctx->timing = timing;
        return set_error(ctx, CIM_E_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    int rc = search_records_dev(ctx, psearch, n, (const float*)images_dev, parts_dev);
    if (!rc) rc = rec_insert_records_dev(ctx, pinsert, n, (const float*)images_dev, parts_dev);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(ctx->stream, &g);
    ctx->timing = timing;
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : set_error(ctx, CIM_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        ctx->gexec = nullptr;
        return set_error(ctx, CIM_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    ctx->graph = g;
    return CIM_OK;
}

int cim_search_candidates_dev(cim_ctx* ctx, const cim_params* p, int n, const void* images_dev, float* parts,
                              int ncand, const int* cand, const int* omega) {
    if (!ctx) return CIM_E_ARG;
    if (!images_dev || !parts || !cand || !omega || ncand < 1 || ncand > 256)
        return set_error(ctx, CIM_E_ARG, "cim_search_candidates_dev: NULL array or ncand %d outside 1..256", ncand);
    CIM_HIP(hipSetDevice(ctx->device));
    return search_candidates_dev(ctx, p, n, (const float*)images_dev, parts, ncand, cand, omega);
}

int cim_step_replay(cim_ctx* ctx) {
    if (!ctx) return CIM_E_ARG;
    if (!ctx->gexec) return set_error(ctx, CIM_E_STATE, "cim_step_capture has not been called");
    CIM_HIP(hipSetDevice(ctx->device));
    CIM_HIP(hipGraphLaunch(ctx->gexec, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    // the coset insert's cube now holds samples to fold into the volume
    if (ctx->rec.cosets) ctx->rec.cube_dirty = true;
    return CIM_OK;
}

int cim_step_release(cim_ctx* ctx) {
    if (!ctx) return CIM_E_ARG;
    CIM_HIP(hipSetDevice(ctx->device));
    step_release(ctx);
    return CIM_OK;
}

int cim_rec_insert_dev(cim_ctx* ctx, const cim_params* p, int n, const void* images_dev, const float* parts) {
    if (!ctx) return CIM_E_ARG;
    if (!images_dev || !parts) return set_error(ctx, CIM_E_ARG, "NULL images or parts");
    CIM_HIP(hipSetDevice(ctx->device));
    return rec_insert_images_dev(ctx, p, n, (const float*)images_dev, parts);
}

int cim_rec_insert(cim_ctx* ctx, const cim_params* p, int n, const float* images, const float* parts) {
    if (!ctx) return CIM_E_ARG;
    if (!images || !parts) return set_error(ctx, CIM_E_ARG, "NULL images or parts");
    int rc = check_params(ctx, p);
    if (rc) return rc;
    CIM_HIP(hipSetDevice(ctx->device));
    const long long N = (long long)p->box * p->box;
    CIM_ALLOC(ctx->img, sizeof(float) * N * (n > 0 ? n : 1));
    if (n > 0)
        CIM_HIP(hipMemcpyAsync(ctx->img.ptr, images, sizeof(float) * N * n, hipMemcpyHostToDevice, ctx->stream));
    return rec_insert_images_dev(ctx, p, n, ctx->img.as<float>(), parts);
}

int cim_rec_insert_ft(cim_ctx* ctx, int n, const float* ft_re, const float* ft_im, const float* ctf,
                      const float* euler) {
    if (!ctx) return CIM_E_ARG;
    if (!ft_re || !ft_im || !euler) return set_error(ctx, CIM_E_ARG, "NULL argument");
    if (!ctx->rec.active) return set_error(ctx, CIM_E_STATE, "cim_rec_begin has not been called");
    if (n <= 0) return CIM_OK;
    CIM_HIP(hipSetDevice(ctx->device));
    const long long N = (long long)ctx->rec.box * ctx->rec.box;
    std::vector<float> parts((size_t)n * CIM_PART_STRIDE, 0.f);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) parts[i * CIM_PART_STRIDE + k] = euler[3 * i + k];
    int rc = upload_parts(ctx, n, parts.data());
    if (rc) return rc;
    CIM_ALLOC(ctx->img, sizeof(float) * N * n * (ctf ? 3 : 2));
    float* d = ctx->img.as<float>();
    CIM_HIP(hipMemcpyAsync(d, ft_re, sizeof(float) * N * n, hipMemcpyHostToDevice, ctx->stream));
    CIM_HIP(hipMemcpyAsync(d + N * n, ft_im, sizeof(float) * N * n, hipMemcpyHostToDevice, ctx->stream));
    float* dctf = nullptr;
    if (ctf) {
        dctf = d + 2 * N * n;
        CIM_HIP(hipMemcpyAsync(dctf, ctf, sizeof(float) * N * n, hipMemcpyHostToDevice, ctx->stream));
    }
    CIM_ALLOC(ctx->ft, sizeof(float2) * N * n);
    rc = interleave(ctx, d, d + N * n, ctx->ft.as<float2>(), N * n);
    if (rc) return rc;
    rc = rec_insert_ft_dev(ctx, n, ctx->ft.as<float2>(), ctx->parts.as<float>(), dctf, 0, nullptr, 0);
    if (rc) return rc;
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

int cim_rec_get_tnf(cim_ctx* ctx, float* fa, float* fb, float* wg) {
    if (!ctx) return CIM_E_ARG;
    if (!fa || !fb || !wg) return set_error(ctx, CIM_E_ARG, "NULL argument");
    CIM_HIP(hipSetDevice(ctx->device));
    return rec_tnf_io(ctx, fa, fb, wg, 0);
}

int cim_rec_add_tnf(cim_ctx* ctx, const float* fa, const float* fb, const float* wg) {
    if (!ctx) return CIM_E_ARG;
    if (!fa || !fb || !wg) return set_error(ctx, CIM_E_ARG, "NULL argument");
    CIM_HIP(hipSetDevice(ctx->device));
    return rec_tnf_io(ctx, (float*)fa, (float*)fb, (float*)wg, 1);
}

int cim_rec_finish_dev(cim_ctx* ctx, void* map_dev) {
    if (!ctx) return CIM_E_ARG;
    if (!map_dev) return set_error(ctx, CIM_E_ARG, "NULL map");
    CIM_HIP(hipSetDevice(ctx->device));
    int rc = rec_finish_dev(ctx, (float*)map_dev);
    if (rc) return rc;
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

int cim_rec_finish(cim_ctx* ctx, float* map) {
    if (!ctx) return CIM_E_ARG;
    if (!map) return set_error(ctx, CIM_E_ARG, "NULL map");
    if (!ctx->rec.active) return set_error(ctx, CIM_E_STATE, "cim_rec_begin has not been called");
    CIM_HIP(hipSetDevice(ctx->device));
    const long long nvox = (long long)ctx->rec.box * ctx->rec.box * ctx->rec.box;
    CIM_ALLOC(ctx->out, sizeof(float) * nvox);
    int rc = rec_finish_dev(ctx, ctx->out.as<float>());
    if (rc) return rc;
    CIM_HIP(hipMemcpyAsync(map, ctx->out.ptr, sizeof(float) * nvox, hipMemcpyDeviceToHost, ctx->stream));
    CIM_HIP(hipStreamSynchronize(ctx->stream));
    return CIM_OK;
}

}  // extern "C"
