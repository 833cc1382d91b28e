// Internal declarations shared by the HIP translation units of libcryoim_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/cryoim.h"
#include "cim_symops.h"

namespace cim {

// Growable device buffer owned by a context.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes) return hipSuccess;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&ptr, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Twiddle table and factorisation of one FFT length.
struct FftPlan {
    int n = 0;
    int nfac = 0;
    int fac[16];
    float2* tw = nullptr;   // exp(-2 pi i k / n), k = 0..n-1 (device)
};

// Reconstruction accumulator: float4 {fa, fb, wg, 0} voxels covering
// x, y in [-(maxR), maxR+1] and z in [0, maxR+1] around the origin.
struct RecState {
    bool active = false;
    int box = 0, maxR = 0, layout = 0, nops = 0;
    float rcut = 0.f, wg_init = 0.f;
    int dxy = 0, dz = 0;            // accumulator dims
    DevBuf acc;                     // float4 * dxy*dxy*dz
    bool comp = false;              // compensated: lo holds the TwoSum residues of acc
    DevBuf lo;                      // float4 * dxy*dxy*dz (compensated mode)
    DevBuf ops;                     // double[nops*9]
    DevBuf samples;                 // int2 (hx, ky) per half-plane sample
    int nsamples = 0;
    DevBuf hxmax;                   // largest inserted hx per ky (-1: none)
    DevBuf bricks;                  // int3 origins of the 8^3 bricks a sample can reach
    int nbricks = 0;
    DevBuf bricksq;                 // the same for the queue kernel's brick shape, central first
    int nbricksq = 0, qshape = 0;
    DevBuf vals;                    // per-batch float4 sample values on the half-plane grid
    DevBuf maxv;                    // largest |sample value| of the batch (float bits)
    long long inserted = 0;
    // coset insert (full layout, the 60 icosahedral operators): the 5 coset
    // representatives c_i of the 12 signed-permutation operators T inserted into a
    // full cube [-maxR, maxR+1]^3 without the z fold, symmetrised by T into acc when
    // the sums are read
    bool cosets = false;
    bool cube_dirty = false;
    DevBuf cube, cube_lo;           // float4 * dxy^3 (cube_lo: compensated mode)
    DevBuf cops;                    // double[5*9] coset representatives
    DevBuf bricksc;                 // int3 brick origins in the cube, central first
    int nbricksc = 0;
    int tsym[12][3] = {};           // operator j of T: row i has its +-1 in column tsym[j][i] & 3, sign bit 4
};

// Per-batch working set of the searches.
struct SearchBatch {
    DevBuf pcrop, pcrop4, fcut0, tcrop, tfcut, outv, outi, prfull, ops, rows, pts, pbase, tang, fv, fi, counter, ftrow;
    // tabulated search: polar rays of particles / templates, residual tables, candidates
    DevBuf ppol, tpol, table, cand, kwin, cv, ci;
    // coarse-to-fine gather: coarse point list, 3x3 neighbours, per-particle coarse winners and fine lists
    DevBuf clist, nbr, ccand, ckwin, flist, fcnt;
    // per (AU point, template): the valid common lines' table rows and zero-omega columns, and their count
    DevBuf lines, lcount;
    // coarse E4M3 tables and the D bounds of the particles and templates they are scaled by (sums
    // of the ring maxima rmax); directions of the polar rays
    DevBuf ctab, bound, rmax, rays;
    // fine stage: the two omega windows of each listed point, the coarse second minima and their windows
    DevBuf fwin, fwin2, outi2, ckwin2, outv2;
    // exhaustive fallback: count and list of the particles it scores
    DevBuf fbl;
    void release() {
        for (DevBuf* b : {&pcrop, &pcrop4, &fcut0, &tcrop, &tfcut, &outv, &outi, &prfull, &ops, &rows, &pts, &pbase, &tang,
                          &fv, &fi, &counter, &ppol, &tpol, &table, &cand, &kwin, &cv, &ci, &clist, &nbr, &ccand, &ckwin,
                          &flist, &fcnt, &lines, &lcount, &ctab, &bound, &rmax, &rays, &fwin, &fwin2, &outi2, &ckwin2, &outv2, &fbl, &ftrow})
            b->release();
    }
};

// Template bank for the common-lines search.
struct SearchState {
    bool ready = false;
    int box = 0, T = 0, minR = 0, maxR = 0;
    DevBuf tft;        // float2 [T][box*box] template FTs (masked, FFT, centre shift)
    DevBuf tfcut;      // float [T][nr] Fcut fitted without CTF (load_template.f90:44-47)
    std::vector<float> tang;   // T*3 template angles
    SearchBatch wk;
    long long version = 0;     // template banks set up so far
    std::string step_sig;      // the parameters wk's constant tables were last set up for
};

// Resident model transform and per-batch buffers of local_refine.
struct LrefState {
    bool ready = false;
    int box = 0;
    DevBuf model, model4, acc, disk, symops, mparts, g, tfac;   // model4: tap-paired copy of model
    void release() {
        for (DevBuf* b : {&model, &model4, &acc, &disk, &symops, &mparts, &g, &tfac}) b->release();
    }
};

// Resident model transform and batch buffers of the centre finder.
struct CenterState {
    bool ready = false;
    int box = 0;
    DevBuf model;      // float2 box^3: {Re, -Im} of the checkerboarded model FT
    DevBuf x1, x2;     // float2 [chunk][box^2] working images
    DevBuf acc;        // double[2] shell sums of the model mask
    DevBuf img, parts, cc;
    void release() {
        for (DevBuf* b : {&model, &x1, &x2, &acc, &img, &parts, &cc}) b->release();
    }
};

// Proposal network (cnn.hip): parameters, the activations of the last
// forward pass (kept for the backward pass), patches, gradients.
struct CnnState {
    int batch = 0;
    DevBuf params, act, cols, grads, dy, dx;
    void release() {
        for (DevBuf* b : {&params, &act, &cols, &grads, &dy, &dx}) b->release();
    }
};

// Device buffers of local_box: angles and shifts, operators, per-(particle,
// operator) records, row-shifted complex images, shifted real images.
struct LboxState {
    DevBuf pr, ops, rec, x, r;
    void release() {
        for (DevBuf* b : {&pr, &ops, &rec, &x, &r}) b->release();
    }
};

struct KTimer {
    double total_ms = 0.0;
    long long launches = 0;
};

}  // namespace cim

struct cim_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<int, cim::FftPlan> plans;
    cim::RecState rec;
    cim::SearchState srch;
    cim::LrefState lref;
    cim::CenterState cen;
    cim::LboxState lbox;
    cim::CnnState cnn;
    // scratch buffers reused across calls
    cim::DevBuf img, ft, ft2, aux, aux2, parts, out;
    // the insert's transform pruned to the half plane it reads: row-pass columns, then the crop
    cim::DevBuf ftrow, ftcrop;
    cim::DevBuf prepst;   // per-image preparation sums (PrepStats) of the pruned transforms
    // half-map FSC: input tnf pairs, normalised copies, merged sums, shell sums, weights
    cim::DevBuf fsc_in, fsc_nrm, fsc_acc, fsc_sums, fsc_wt;
    // real-space icosahedral average: operators and the sum volume
    cim::DevBuf avg_ops, avg_acc;
    // kernel timing: event pairs recorded around launches, resolved lazily
    bool timing = false;
    std::map<std::string, cim::KTimer> timers;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    // autotuned launch shapes
    std::map<std::string, int> tuned;
    // work counters (evaluated samples, inserted samples, ...)
    std::map<std::string, long long> counters;
    // captured search + insert step (cim_step_capture)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

namespace cim {

int set_error(cim_ctx* c, int code, const char* fmt, ...);
// Resolve recorded timing events into ctx->timers (synchronises the stream).
void resolve_timing(cim_ctx* c);

#define CIM_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return cim::set_error(ctx, CIM_E_HIP, "%s failed: %s (%s:%d)", #call,      \
                                  hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

#define CIM_ALLOC(buf, n)                                                              \
    do {                                                                               \
        if ((buf).ensure(n) != hipSuccess)                                             \
            return cim::set_error(ctx, CIM_E_NOMEM, "device allocation of %zu bytes failed", \
                                  (size_t)(n));                                        \
    } while (0)

// Kernel timing bracket (enabled with cim_timing).
struct TimedLaunch {
    cim_ctx* ctx;
    const char* name;
    hipEvent_t a = nullptr;
    TimedLaunch(cim_ctx* c, const char* n);
    ~TimedLaunch();
};

// FFT (fft.hip)
int fft_plan(cim_ctx* ctx, int n, FftPlan** out);
// Tabulated global search of n device records on the stream, with no host
// synchronisation (search.hip); needs a host-records call with the same
// parameters and count first.
int search_records_dev(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev, float* parts_dev);
// The exact score of given (AU point, omega window) candidates per particle (search.hip).
int search_candidates_dev(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev, float* parts,
                          int ncand, const int* cand, const int* omega);
// In-place transforms of interleaved complex data on ctx->stream.
int fft2d_dev(cim_ctx* ctx, float2* data, int n, int batch, int dir, int conj_out);
// A batch of centred transforms (or crops of them): image i's value at centred
// frequency (hx, ky) is p[i * stride + (ky + oy) * ld + (hx + ox)].
struct FtView {
    const float2* p;
    long long stride;
    int ld, ox, oy;
};
bool insert_scatter(const cim_ctx* ctx);
int fft2d_crop_dev(cim_ctx* ctx, const float2* data, int n, int batch, int ox0, int Dx, int oy0, int Dy,
                   int conj_out, const float* parts, float2* tmp, float2* crop);
// The same transform of batch real images prepared as it reads them, from
// their sums (prep_stats_dev; the same values as prep_images_dev's).
struct PrepStats;
// The pruned transforms prepare the images as the row pass reads them unless
// CIM_PREP_FUSED=0 (then from a prepared copy, the same bits).
bool prep_fused();
int fft2d_crop_real_dev(cim_ctx* ctx, const float* imgs, const PrepStats* stats, int do_norm, int do_mask,
                        float mask_r, int n, int batch, int ox0, int Dx, int oy0, int Dy, int conj_out,
                        const float* parts, float2* tmp, float2* crop);
int fft3d_dev(cim_ctx* ctx, float2* data, int n, int dir, int conj_out);
// Fold {fa, fb, wg} (pitch dxy, z >= 0) -> complex volume -> inverse FFT ->
// real map with checkerboard (inside radius rmask only when rmask >= 0).
int lbox_scratch_setup(cim_ctx* ctx);
int fft_plan_host(cim_ctx* ctx, int n, FftPlan& pl, std::vector<float2>& tw);
int fold_to_map_dev(cim_ctx* ctx, const float4* acc, const float4* lo, int dxy, int box, int maxR, int int_radius,
                    float isg, float rmask, float* map_dev);

// Real-space icosahedral average (icosavg.hip) of the first n^3 values of
// map_dev (len values readable), variant CIM_ICOSAVG_FULL / _MERGE.
// ops: nops row-major float64 operators, or NULL for the 60 icosahedral ones.
int icosavg_dev(cim_ctx* ctx, float* map_dev, long long len, int n, float r0, int variant, int nops = 0,
                const double* ops = nullptr);

// split <-> interleaved conversion (prep.hip)
int interleave(cim_ctx* ctx, const float* re, const float* im, float2* out, long long count);
int deinterleave(cim_ctx* ctx, const float2* in, float* re, float* im, long long count);

// Per-particle preprocessing to a centred FT (prep.hip).
// mode_norm: normalise first; mask centre from parts[CIM_P_MASKX/Y].
int prep_images_dev(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev, const float* parts_dev,
                    float2* ft_out, int do_norm, int do_mask);
int prep_stats_dev(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev, const float* parts_dev,
                   int do_norm, int do_mask, PrepStats* stats);
int preprocess_images(cim_ctx* ctx, const cim_params* p, int n, const float* images_dev,
                      const float* parts_dev, float2* ft_out, int do_norm, int do_mask,
                      int conj_out);
// Multiply centred FTs by exp(-i phase) of img2d_FT_centshift (recentre=1).
int centre_shift(cim_ctx* ctx, float2* ft, int box, int n, const float* parts_dev, int recentre);
// Fcut of FT_amp_expfit for n centred FTs (optionally multiplied by the CTF
// of parts_dev record (p % nparts_ctf)).
int amp_expfit(cim_ctx* ctx, const float2* ft, int box, int n, int minR, int maxR, float* fcut_dev);
// General form: item i at ft + i*stride, rows of ld, zero frequency at (cidx, cidx).
int amp_expfit_general(cim_ctx* ctx, const float2* ft, long long stride, int ld, int cidx, int nitems, int minR,
                       int maxR, float* fcut_dev);

}  // namespace cim
