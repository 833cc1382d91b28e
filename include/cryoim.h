/*
 * cryoim.h -- C ABI of libcryoim_hip.so, the MI355X (gfx950) implementation
 * of cryoIM's icosahedral search and direct Fourier inversion hot path.
 *
 * Conventions (all kept from the reference):
 *   - images are box*box row-major f32 with x fastest: the memory order of
 *     the reference's flat a(i), i = ky*box + hx + 1;
 *   - Fourier data cross the boundary as split re/im arrays (a, b), centred
 *     (DC at index box/2 on each axis), forward transforms scaled 1/N;
 *   - angles in degrees, ZYZ (euler_mat.f90:3-21); shifts in pixels.
 * Every call is synchronous at the boundary, returns 0 on success or a
 * negative CIM_E* code, and never aborts; cim_last_error() explains the
 * failure.  The caller owns every host array; the library owns every
 * device buffer it allocates inside the context.  Pointers marked "dev" in
 * the *_dev variants are device pointers (hipMalloc memory of the device
 * the context was created on).
 *
 * Reference interfaces replaced (citations into chenwenyuan2009/cryoIM):
 *   FFT layer ........ icosprocess/SRC/FFTW.f90:179-219, img2d_FFT.f90:4-32,
 *                      local_refine/SRC/img3d_FFT.f90:2-31
 *   insert ........... icosprocess/SRC/reconstruct.f90:53-149 (and
 *                      reconstruct_substep.f90, reconstruct_half.f90)
 *   fold + inversion . icosprocess/SRC/reconstruct.f90:151-247
 *   global search .... icosprocess/SRC/ort_globesearch.f90:34-134
 *   local refinement . icosprocess/SRC/ort_refinement.f90:31-124
 *   slice + PR ....... local_refine/SRC/refine_PR.f90:69-245, FFT3d_slice.f90,
 *                      cal_PR_FT.f90, cal_template_FT.f90, refine_PR_Zhigh.f90,
 *                      refine_particle_df.f90
 *   half maps, FSC ... icosprocess/SRC/reconstruct_half.f90, reconstruct_FSC.f90
 *   tnf merges ....... merge_icosprocess_subrec/SRC/merge_reconstruct.f90,
 *                      merge_reconstruct_FSC.f90
 *   local insert ..... icosprocess/icosprocess/reconstruction_local*.f90
 *   centre finding ... icosprocess/SRC/calc_center.f90,
 *                      calc_center_selfrot180.f90
 *   real-space avg ... icosprocess/SRC/img3d_icosavg.f90,
 *                      merge_icosprocess_subrec/SRC/img3d_icosavg.f90
 *   local_reconstruct  local_reconstruct/SRC/reconstruct.f90,
 *                      reconstruct_substep.f90, C_matrix.f90, icos_matrix_i*.f90
 *   local_box ........ local_box/SRC/loc_box_coordinate.f90, loc_box.f90,
 *                      equalort.f90, img2d_centshift.f90
 */
#ifndef CRYOIM_H
#define CRYOIM_H

#ifdef __cplusplus
extern "C" {
#endif

#define CIM_OK 0
#define CIM_E_ARG -1      /* invalid argument or size */
#define CIM_E_HIP -2      /* HIP runtime error */
#define CIM_E_STATE -3    /* call out of order (e.g. insert before begin) */
#define CIM_E_NOMEM -4    /* device allocation failed */

typedef struct cim_ctx cim_ctx;

/* Scalar parameters shared by the entry points.  Field meaning follows the
 * reference's key=value parameters after calc_initialvalues
 * (icosprocess/SRC/0calc_initialvalues.f90:1-21). */
typedef struct cim_params {
    int box;          /* FFTsize (image edge, pixels) */
    int minR;         /* inner search radius (Fourier pixels) */
    int maxR;         /* outer search / insert radius (Fourier pixels) */
    int apply_ctf;    /* 1: multiply by getCTF2d (applyCTF=y) */
    int mask_on;      /* 1: soft circular mask (imgmask given/default y) */
    int norm_on;      /* 1: zero-mean unit-RMS normalisation before masking;
                         2: after masking (local_reconstruct's substep) */
    int ncycle;       /* refinement cycles (Ncycle) */
    int conj_out;     /* FFT stores b = -Im (1, icosprocess) or +Im (0, local_refine) */
    int search_method;/* global search: CIM_SEARCH_PLAIN (0) or CIM_SEARCH_TABLES (1) */
    int search_topk;  /* CIM_SEARCH_TABLES: AU points re-scored exactly per particle (0 = 32) */
    int corr_cycles;  /* insert with a defocus series (reconstruct_corrCTF): C defocus steps on
                         each side of df; 0 = the plain CTF */
    int search_second;/* CIM_SEARCH_TABLES: second windows re-scored per particle besides the
                         search_topk points (0 = none): the best other minima of the table
                         scores over omega, at least a window from their point's first */
    int reserved_i[4];
    float apix;       /* pixel size (A) */
    float imgmask;    /* mask radius in pixels */
    float vol;        /* accelerating voltage (kV) */
    float cs;         /* spherical aberration (mm) */
    float ampwg;      /* amplitude contrast */
    float df_scale;   /* defocus unit to A: 1e4 (um, icosprocess) or 1 (A) */
    float searchstep; /* global search grid step (deg) */
    float delta_angle;/* refinement angle step (deg) */
    float delta_trans;/* refinement shift step (px) */
    float corr_step;  /* defocus step of the series, in the df unit (corrCTF=step,cycles) */
    float reserved_f[6];
} cim_params;

/* Global search methods (cim_search_global):
 * CIM_SEARCH_PLAIN   the reference's exhaustive search: every AU point x
 *                    omega scored by cal_crosslin_PR_omega's phase residual.
 * CIM_SEARCH_TABLES  the same residual tabulated per (particle, template)
 *                    over all pairs of rays on a polar grid of step
 *                    searchstep; (AU point, omega) pairs are scored from the
 *                    tables (common-line angles rounded to the grid) coarse
 *                    to fine: first the grid of twice the step in all three
 *                    angles (the AU points of even theta and phi index plus
 *                    the AU border points, every other omega) from E4M3
 *                    tables of the polar grid of twice the step, then every
 *                    omega of the 3x3 grid neighbourhoods of each particle's
 *                    64 best coarse points (CIM_TABLE_COARSE), so a point
 *                    that is neither coarse nor next to a good coarse point
 *                    is never scored at the step.  The search_topk best AU
 *                    points of each particle are scored again exactly as
 *                    CIM_SEARCH_PLAIN does, but only over a window of 16
 *                    omega columns (plus their Friedel pairs;
 *                    CIM_RESCORE_CPW, CIM_TABLE_WINDOW) centred on the omega
 *                    of the table minimum; with search_second = S, the S
 *                    best second minima of the table scores over omega (at
 *                    least a window from their point's first) get windows
 *                    too; the first minimum among them all is returned.  A
 *                    particle whose exact minimum ranks in the last quarter
 *                    of the search_topk points, or lies within 2 columns of
 *                    its window's edge, is searched again exhaustively (the
 *                    plain search).  Equality with the plain result is
 *                    therefore measured, not guaranteed: it holds when the
 *                    plain minimum lies in a re-scored window.  Measured at
 *                    box 384 (0.5 deg, maxR 72, T = 4, defaults): 20,477 of
 *                    20,480 distinct particles bit for bit; at box 128
 *                    (1 deg, maxR 24, T = 4) 1,987 of 2,000 without and
 *                    1,995 with search_second = 8. */
#define CIM_SEARCH_PLAIN 0
#define CIM_SEARCH_TABLES 1

/* Particle records: CIM_PART_STRIDE floats per particle. */
#define CIM_PART_STRIDE 12
#define CIM_P_THETA 0
#define CIM_P_PHI 1
#define CIM_P_OMEGA 2
#define CIM_P_X 3
#define CIM_P_Y 4
#define CIM_P_PR 5
#define CIM_P_DF1 6
#define CIM_P_DF2 7
#define CIM_P_ASTIG 8
#define CIM_P_MASKX 9     /* mask centre x (see cim_rec_insert) */
#define CIM_P_MASKY 10
#define CIM_P_USE 11      /* 1.0 = insert, 0.0 = skip (PR/boundX filter) */
/* local_refine records reuse slots 9, 10 for the sub-particle box centre in
 * the full particle image (local_hx0, local_ky0 of the local ortfile). */
#define CIM_P_BOXX 9
#define CIM_P_BOXY 10
#define CIM_P_Z 11        /* local_refine: local_z0 (A) in; cim_lref_defocus writes the chosen z */

/* ---- context ---------------------------------------------------------- */
int cim_init(cim_ctx** ctx, int device);
int cim_finalize(cim_ctx* ctx);
const char* cim_last_error(const cim_ctx* ctx);
/* hipStream_t every kernel of this context runs on. */
void* cim_get_stream(cim_ctx* ctx);
int cim_sync(cim_ctx* ctx);
/* Average duration (ms) of the named kernel family over the launches timed
 * since the last reset; timing is off until cim_timing(ctx, 1). */
int cim_timing(cim_ctx* ctx, int on);
int cim_timing_get(cim_ctx* ctx, const char* name, double* avg_ms, long long* launches);
/* Work counters accumulated since the last reset: "cl_search_samples"
 * (bilinear samples scored by the global search), "cl_search_particles",
 * "insert_samples" (half-plane samples inserted, each into nops*8 voxels),
 * "insert_particles". */
int cim_counter_get(cim_ctx* ctx, const char* name, long long* value);
int cim_counter_reset(cim_ctx* ctx);

/* ---- FFT: replaces FFTcmplx2d_f/B, FFTcmplx3d_f/B and img2d_FFT ------- */
/* dir = 1 forward (scaled 1/N), -1 backward (unscaled).  conj_out = 1
 * negates the imaginary part of the result (img2d_FFT.f90:19).  Sizes:
 * any n whose prime factors are 2, 3, 5 and n <= 1024. */
int cim_fft2d(cim_ctx* ctx, float* re, float* im, int n, int batch, int dir, int conj_out);
int cim_fft3d(cim_ctx* ctx, float* re, float* im, int n, int dir, int conj_out);
/* Same on interleaved complex device data (n*n*batch or n^3 float2). */
int cim_fft2d_dev(cim_ctx* ctx, void* data, int n, int batch, int dir, int conj_out);
int cim_fft3d_dev(cim_ctx* ctx, void* data, int n, int dir, int conj_out);

/* ---- direct Fourier inversion ---------------------------------------- */
/* OR'd into cim_rec_begin's layout: accumulate compensated FP32 sums.  Each
 * batch's contributions are summed in FP32 per 8^3 brick, then added to a
 * hi/lo pair per voxel with TwoSum; fold, tnf readout and the map use hi+lo.
 * Costs a second accumulator (the same size) and its read-modify-write. */
#define CIM_REC_COMPENSATED 16
/* layout: 0 = full box accumulator as reconstruct.f90 (offset box/2),
 *         1 = compact tnf accumulator as reconstruct_substep.f90 (offset maxR).
 * rcut: radius cut of inserted samples (maxR, maxR-1 or maxR-2 in the
 * reference variants; the half maps of reconstruct_half.f90:98 use maxR with
 * the compact layout, whose edge corners wrap as that array's flat index
 * does -- see cim_rec_get_tnf).  wg_init: initial weight (0.01 reconstruct.f90:28,
 * 0 substep).  ops: nops row-major 3x3 float64 symmetry operators, or NULL
 * with nops = 60 for the icosahedral group (icos2f_sym_mat order).
 * Full layout with ops = NULL (the icosahedral group) selects the coset
 * insert: the 12 operators that are signed permutations map the lattice and
 * the trilinear weights onto themselves, so each particle is inserted only
 * under the 5 representatives of their cosets, into a full cube without the
 * z fold, and that cube is symmetrised by the 12 into the sums when they
 * are read (cim_rec_get_tnf, cim_rec_finish, cim_merge_map): 12x fewer
 * corner updates for the same sums up to float rounding of the weights.
 * CIM_INSERT=brick keeps the 60-plane brick gather. */
int cim_rec_begin(cim_ctx* ctx, int box, int maxR, int layout, float rcut, float wg_init,
                  int nops, const double* ops);
/* Full per-particle pipeline of reconstruct.f90:77-147: optional norm, CTF,
 * mask (centre CIM_P_MASKX/Y, which the caller sets to the first particle's
 * shift to reproduce reconstruct.f90:92), checkerboard, FFT, centre shift
 * by (x, y), 60-fold insert.  images: n*box*box real.  parts: n records. */
int cim_rec_insert(cim_ctx* ctx, const cim_params* p, int n, const float* images, const float* parts);
int cim_rec_insert_dev(cim_ctx* ctx, const cim_params* p, int n, const void* images_dev,
                       const float* parts);
/* Insert already-transformed particles: centred FT split arrays (n*box*box),
 * optional CTF images (n*box*box, NULL = 1) and angles (n*3). */
int cim_rec_insert_ft(cim_ctx* ctx, int n, const float* ft_re, const float* ft_im,
                      const float* ctf, const float* euler);
/* Copy the accumulated sums out in tnf order ((2maxR+1)^2 (maxR+1) each,
 * lz slowest, hx fastest; reconstruct_substep.f90:148-155).  With the
 * compact layout, corners at x or y = maxR+1 land where the reference's flat
 * index puts them (next row / next plane, reconstruct_half.f90:127-130) and
 * corners past the last plane are dropped.  With the full layout and the 60
 * icosahedral operators (the coset insert, see cim_rec_begin) the lz = 0
 * plane holds half of each voxel's Friedel combination F(x) + conj F(-x)
 * instead of the reference's split by the side each sample came from; the
 * fold (cim_rec_finish, merges) only ever reads that combination, so every
 * map is the reference's.  Planes lz >= 1 are the reference's sums. */
int cim_rec_get_tnf(cim_ctx* ctx, float* fa, float* fb, float* wg);
/* Add tnf-order partial sums into the accumulator (merge of several
 * reconstruct_substep outputs, merge_reconstruct.f90:23-34). */
int cim_rec_add_tnf(cim_ctx* ctx, const float* fa, const float* fb, const float* wg);
/* Fold of the lz=0 plane, division by the weights (float64), (-1)^(h+k+l),
 * Hermitian completion, inverse 3D FFT and checkerboard: the real map,
 * box^3 f32 (reconstruct.f90:151-247). */
int cim_rec_finish(cim_ctx* ctx, float* map);
int cim_rec_finish_dev(cim_ctx* ctx, void* map_dev);
/* merge_icosprocess_subrec without FSC (merge_reconstruct.f90:14-157): the
 * caller opens a compact accumulator with weights 0.01
 * (cim_rec_begin(ctx, box, maxR, 1, maxR, 0.01f, 60, NULL)), adds each
 * partial tnf with cim_rec_add_tnf, then this folds with the integer shell
 * test of :42 (int(r) <= maxR), inverse-transforms and writes the
 * (box-8)^3 block of :143-149, including its plane index i/(box-8)^2+4 on
 * the 1-based i (the last voxel of each plane comes from the next plane).
 * nicosavg = 1 (the tnf header's flag, written by realspaceavg=y) runs
 * merge's img3d_icosavg (CIM_ICOSAVG_MERGE, radius imgmask in pixels) on the
 * full map before the crop (:142). */
int cim_merge_map(cim_ctx* ctx, int nicosavg, float imgmask, float* map);
/* icosprocess's cyclic groups (generators icos2f_sym_mat.f90:250-540), used
 * by mode 1 with sym != icos (reconstruct_C1.f90:33-59) and mode 6
 * (reconstruction_local.f90:30-55): sym = "c1", "c2", "c3", "c4", "c5a",
 * "c6", "c7", "c8", "c12" or "c15", z rotations by k*360/N degrees
 * evaluated in REAL*4 as the reference (mode 6 has no "c7" branch; the
 * reference program faults there).  ops (may be NULL): nops row-major 3x3
 * float64, ready for cim_rec_begin.  Returns CIM_E_ARG for other names. */
int cim_icosprocess_sym_ops(const char* sym, double* ops, int* nops);

/* ---- real-space icosahedral average (realspaceavg=y) ---------------------- */
#define CIM_ICOSAVG_FULL 0    /* icosprocess/SRC/img3d_icosavg.f90:2-126 */
#define CIM_ICOSAVG_MERGE 1   /* merge_icosprocess_subrec/SRC/img3d_icosavg.f90:1-91 */
/* Replaces the first n^3 values of map (taken as an n^3 cube, centre n/2)
 * by the mean over the operators of the trilinearly interpolated map at S x,
 * within radius r0 (pixels) of the centre and 0 beyond.  ops: nops (<= 60)
 * row-major float64 operators -- icosprocess's img3d_icosavg averages over
 * the group of its sym (cim_icosprocess_sym_ops for the cyclic ones) -- or
 * NULL for the 60 icosahedral operators.  CIM_ICOSAVG_MERGE (ops NULL)
 * averages the quarter lz, ky >= 0 and fills the rest with the copies of the
 * reference's second loop, in its order and flat indexing.  Taps read the
 * flat index the reference reads; len values of map are readable (reads
 * outside them give 0).  Host array. */
int cim_icosavg(cim_ctx* ctx, float* map, long long len, int n, float r0, int variant, int nops,
                const double* ops);
/* Same on a device array. */
int cim_icosavg_dev(cim_ctx* ctx, void* map_dev, long long len, int n, float r0, int variant, int nops,
                    const double* ops);

/* ---- half-map FSC ------------------------------------------------------ */
#define CIM_FSC_RECONSTRUCT 0   /* icosprocess reconstruct_FSC.f90:10-197 */
#define CIM_FSC_MERGE 1         /* merge_reconstruct_FSC.f90:9-210 */
/* odd/even tnf sums (as written by reconstruct_half.f90:143-151, summed over
 * files for the merge) -> fsc[0..maxR] = the FSC column of FSC.txt
 * (shells ir = int(r); reconstruct_FSC.f90:37-63), and the FSC-weighted map:
 *   CIM_FSC_RECONSTRUCT: box^3 floats, cmplx(fa, fb) inverse-transformed
 *     (reconstruct_FSC.f90:173-179);
 *   CIM_FSC_MERGE: the (box-8)^3 floats merge_reconstruct_FSC writes, i.e.
 *     the first (box-8)^3 values of the flat box^3 map masked to radius
 *     imgmask (:188-206).
 * fsc may be NULL.  The weighting keeps the reference's stale-index zeroing
 * (reconstruct_FSC.f90:82-84).  nicosavg = 1 (realspaceavg=y) averages the
 * map in real space: CIM_ICOSAVG_FULL on the box^3 map for
 * CIM_FSC_RECONSTRUCT (:185), CIM_ICOSAVG_MERGE on the first (box-8)^3
 * values taken as a (box-8)^3 cube for CIM_FSC_MERGE
 * (merge_reconstruct_FSC.f90:202, after maphead has set FFTsize = box-8);
 * radius imgmask. */
int cim_fsc_map(cim_ctx* ctx, int box, int maxR, float imgmask, int variant, int nicosavg, const float* odd_fa,
                const float* odd_fb, const float* odd_wg, const float* eve_fa, const float* eve_fb,
                const float* eve_wg, float* fsc, float* map);

/* ---- icosprocess global search (mode 4) and refinement (mode 5) ------- */
/* Template bank of load_template.f90:18-48: T real images (templatestck)
 * with their ortfile angles and shifts (tort: T*5 theta phi omega x y);
 * masked, checkerboarded, transformed and centre-shifted on the device. */
int cim_search_templates(cim_ctx* ctx, const cim_params* p, int T, const float* images,
                         const float* tort);
/* The same bank with the templates projected from a model instead
 * (load_template.f90:32-33 when no templatestck is given):
 * project_template2d.f90:3-80, the real-space projection of the box^3 map
 * model at each template's angles (tort: T*5 theta phi omega x y) --
 * per pixel the trilinear model values along its rotated voxel column,
 * voxels strictly inside radius box/2 - 1 -- then the template-stack path. */
int cim_search_templates_model(cim_ctx* ctx, const cim_params* p, int T, const float* model, const float* tort);
/* The projections alone (icosprocess mode 0, project2d.f90:2-86, which
 * writes them as the template stack; the same arithmetic as
 * project_template2d.f90): out = T box^2 real images for angles T*3. */
int cim_project_templates(cim_ctx* ctx, int box, int T, const float* model, const float* angles, float* out);
/* Exhaustive common-lines phase-residual search over the AU grid
 * (ort_globesearch.f90:34-134).  parts in/out: theta, phi, omega and PR are
 * replaced by the search result; x, y are used as given. */
int cim_search_global(cim_ctx* ctx, const cim_params* p, int n, const float* images, float* parts);
int cim_search_global_dev(cim_ctx* ctx, const cim_params* p, int n, const void* images_dev,
                          float* parts);
/* One search + insert step captured as a HIP graph and replayed (no
 * counterpart in the reference, whose loops are host programs; for callers
 * that run the same batch shape step after step and want the step as one
 * launch with no host work inside it).  n device records parts_dev
 * (n x CIM_PART_STRIDE floats) and n box^2 device images: the step runs the
 * tabulated global search (psearch, CIM_SEARCH_TABLES, one batch) with its
 * fallback on the device, writes theta, phi, omega and PR into parts_dev,
 * then inserts every record (CIM_P_USE is not read) with pinsert into the
 * volume of cim_rec_begin.  Needs, before capture: cim_search_templates, one
 * cim_search_global / _dev call with the same psearch and n (it sets up the
 * search's constant tables) and cim_rec_begin.  cim_step_replay launches the
 * graph and waits for it; refill images_dev and parts_dev between replays.
 * The graph holds the context's buffers: release it before calls that use
 * other sizes on the context.  Replays do not update the work counters. */
int cim_step_capture(cim_ctx* ctx, const cim_params* psearch, const cim_params* pinsert, int n,
                     const void* images_dev, float* parts_dev);
int cim_step_replay(cim_ctx* ctx);
int cim_step_release(cim_ctx* ctx);
/* ---- learned proposals (no counterpart in the reference) --------------
 * The proposal network of the learned search: six 3x3 convolutions of 64
 * channels, stride 2, ReLU (128^2 -> 2^2), then a linear head of
 * CIM_CNN_OUT outputs: per hypothesis h < CIM_CNN_HYP, outputs 7h .. 7h+5
 * are the first two columns of a rotation (column-major) and 7h+6 its
 * confidence logit.  Products on the matrix cores in bfloat16 with float32
 * sums.  params: cim_cnn_nparams() floats, per layer W [ky][kx][cin][cout]
 * then the 64 biases, then the head W [256][CIM_CNN_OUT] (features in
 * (y, x, channel) order) and its biases.  cim_cnn_forward takes B device
 * images of 128 x 128 float32 and writes B x CIM_CNN_OUT device outputs,
 * keeping the activations for cim_cnn_backward, which takes the device
 * gradient of the loss with respect to those outputs and writes the
 * parameter gradient to the host. */
/* The exact score (the plain search's arithmetic, bit for bit) of ncand
 * given candidates per particle: cand (n x ncand) AU point indices in the
 * search grid's order (theta index outer; per phi index ip >= 0 the point
 * at +ip, then at -ip when ip > 0) and omega (n x ncand) the omega column
 * (0 .. Nomega-1) each window of 16 columns (and their Friedel columns) is
 * centred on.  parts in/out as cim_search_global: the first minimum over
 * the candidates.  For searches whose candidates come from elsewhere
 * (learned proposals); ncand <= 256. */
int cim_search_candidates_dev(cim_ctx* ctx, const cim_params* p, int n, const void* images_dev, float* parts,
                              int ncand, const int* cand, const int* omega);
#define CIM_CNN_HYP 4
#define CIM_CNN_OUT 28
int cim_cnn_nparams(void);
int cim_cnn_forward(cim_ctx* ctx, int B, const void* images_dev, const float* params, void* out_dev);
int cim_cnn_backward(cim_ctx* ctx, const void* dout_dev, float* grads);

/* Full phase-residual map of one particle (Nomega x (2Nphi+1) x (Ntheta+1),
 * omega fastest, 1000 outside the AU) -- the PR_array of ort_globesearch. */
int cim_search_pr_array(cim_ctx* ctx, const cim_params* p, const float* image, const float* part,
                        float* pr_array);
/* Coordinate-descent refinement of orientation and centre
 * (ort_refinement.f90:31-124, ort_cent_11.f90, return_icos_asymmetry_unit.f90). */
int cim_refine_local(cim_ctx* ctx, const cim_params* p, int n, const float* images, float* parts);

/* ---- local_refine: central slices + full-plane phase residual ---------- */
/* Symmetry operators of the local programs (local_refine/SRC/icos2f_sym_matc.f90):
 * "c1".."c360" (rotations by i*360/N, i = 1..N), "asym", "i2", "i5", and
 * "icos" (the icos2f_sym_mat table).  ops may be NULL to query nops. */
int cim_sym_ops(const char* sym, double* ops, int* nops);
/* Symmetry operators of local_reconstruct (0calc_initialvalues.f90:33-62):
 * "i2"/"I2" (icos_matrix_i2.f90), "i3"/"I3" and "i5"/"I5" (I2 conjugated onto
 * the 3-fold and 5-fold frames), or "c1".."c15" ("C1".."C9", "C_10".."C_15")
 * by C_matrix.f90, whose first operator is the rotation by 360/N and whose
 * last is the identity -- generated in float32 in the reference's order.
 * ops may be NULL to query nops; CIM_E_ARG for other names. */
int cim_local_reconstruct_sym_ops(const char* sym, double* ops, int* nops);
/* Resident model transform (refine_PR.f90:21-54): soft mask of radius
 * p->imgmask (img3d_softmask.f90), checkerboard, forward 3D FFT (1/N, +Im),
 * checkerboard.  vol: box^3 real map, box = p->box. */
int cim_lref_model(cim_ctx* ctx, const cim_params* p, const float* vol);
/* Refine n sub-particles (refine_PR.f90:69-233).  images: n full particle
 * images of box0^2; each record gives the local box centre (CIM_P_BOXX/Y),
 * the starting local theta, phi, omega, x, y (replaced by the result, PR in
 * CIM_P_PR) and df1, df2 (A, p->df_scale = 1), astig.  p->ncycle cycles of
 * ort_cent_11 with steps p->delta_angle, p->delta_trans shrunk by convF;
 * sym != "c1" runs the icos_equalort stage first.  maxres: the resolution
 * limit in Fourier pixels (apix*box/maxRes) that bounds the residual and sets
 * the Gaussian low-pass exp(-r^2 / (2 maxres^2)). */
int cim_lref_refine(cim_ctx* ctx, const cim_params* p, int n, const float* images, int box0, float* parts,
                    const char* sym, float convF, float maxres);
/* refine_PR_Zhigh (local_refine/SRC/refine_PR_Zhigh.f90:3-283, Zhigh_df=y):
 * cim_lref_refine with each record's df1, df2 replaced by df -/+ local_z0
 * (the sign key) and apply_ctf = 1 -- the variant differs from refine_PR only
 * in that CTF and in the output fields it writes.
 * refine_particle_df (refine_particle_df.f90:3-177, particle_df=y): no
 * orientation refinement; at each record's local angles and shift the
 * central slice (phase-shifted, checkerboarded, low-passed) is compared by
 * cal_PR_FT with the sub-particle under the CTF of 401 defoci
 * df -/+ z_i, z_i = z0 + 4 (i - 200) A, z0 = CIM_P_Z (sign = +1 / -1 for the
 * sign key '+' / '-'); the first minimum gives CIM_P_Z and CIM_P_PR.  Images,
 * box0 and maxres as cim_lref_refine; needs cim_lref_model. */
int cim_lref_defocus(cim_ctx* ctx, const cim_params* p, int n, const float* images, int box0, float* parts,
                     float sign, float maxres);

/* ---- centre finding (icosprocess mode 3) -------------------------------- */
#define CIM_CENTER_MODEL 0        /* calc_center.f90:2-113 (model3d given) */
#define CIM_CENTER_SELFROT180 1   /* calc_center_selfrot180.f90:1-95 */
/* Resident model of calc_center.f90:16-41: box^3 real map (the model3d
 * data, box = p->box), img3d_mask of radius p->imgmask when p->mask_on,
 * checkerboard, forward 3D FFT (1/N), checkerboard, imaginary part negated. */
int cim_center_model(cim_ctx* ctx, const cim_params* p, const float* model);
/* Centre n particles (the loops of calc_center.f90:44-102 and
 * calc_center_selfrot180.f90:18-92).  images: n*box*box real.  parts: n
 * records; reads theta, phi, omega, df1, df2, astig (CTF when
 * p->apply_ctf, CIM_CENTER_MODEL only) and x, y (the mask centre of the
 * particle image, CIM_CENTER_MODEL); writes x, y (the fitted centre, 0 where
 * |shift| >= box/16) and CIM_P_PR = acos(cc) in degrees.  Low-pass radius
 * p->maxR; p->mask_on / p->imgmask as imgmask.  cc (may be NULL): the n
 * fitted peak values before clipping to [-1, 1]. */
int cim_center(cim_ctx* ctx, const cim_params* p, int method, int n, const float* images, float* parts, float* cc);
/* Same with the images already in device memory. */
int cim_center_dev(cim_ctx* ctx, const cim_params* p, int method, int n, const void* images_dev, float* parts,
                   float* cc);

/* ---- local_box: sub-particle coordinates and boxes per operator ---------- */
/* Symmetry operators of local_box (local_box/SRC/0main.f90:80-122):
 * "i2"/"I2", "i3"/"I3", "i5"/"I5" as cim_local_reconstruct_sym_ops, and
 * "c1".."c15"/"C1".."C15" by C_matrix.f90; any other name keeps local_box's
 * default Nsym = 60 and gives C_matrix with 60 operators (float32).  ops may
 * be NULL to query nops.  Never fails. */
int cim_local_box_sym_ops(const char* sym, double* ops, int* nops);
/* One record of CIM_LBOX_REC floats per (particle i, operator k), at
 * rec[(i*nops + k)*CIM_LBOX_REC]:
 *   0..2  theta, phi, omega of S_k R(particle) (equalort.f90:10-21)
 *   3, 4  hx0, ky0 = int(c1), int(c2) (truncation toward zero)
 *   5, 6  c1 - hx0, c2 - ky0
 *   7     coordinates: (c3 - box/2) * apix; stack: c3 * apix
 * with c = R_k^T x0 (R_k the matrix of the local angles, x0 = local_x,
 * local_y, local_z), plus box/2 + (x, y, 0) in coordinate mode. */
#define CIM_LBOX_REC 8
/* loc_box_coordinate.f90:16-41: the records of n particles (parts: records
 * of CIM_PART_STRIDE; theta, phi, omega, x, y read) for nops operators (ops
 * row-major float64, values of float32 operators).  box = FFTsize of the
 * particle images, apix in A/pixel. */
int cim_local_box_coords(cim_ctx* ctx, int n, const float* parts, int nops, const double* ops, const float* x0,
                         int box, float apix, float* rec);
/* loc_box.f90:50-91: n images of box^2 (the particle images); each is
 * shifted by -(x, y) through its transform (img2d_centshift.f90), then for
 * each operator k the sub x sub box with corner (hx0, ky0) + box/2 - sub/2
 * (periodic) is cut and normalised to mean 0, rms 1 into
 * boxes[(i*nops + k)*sub^2].  rec (may be NULL): the stack-mode records.
 * Requires sub <= box and |x0| <= box - 2 (so every index stays within one
 * periodic wrap).  box with prime factors 2, 3, 5 only, <= 1024. */
int cim_local_box_stack(cim_ctx* ctx, int n, const float* images, const float* parts, int box, int nops,
                        const double* ops, const float* x0, int sub, float apix, float* boxes, float* rec);
/* Same with images and boxes in device memory. */
int cim_local_box_stack_dev(cim_ctx* ctx, int n, const void* images_dev, const float* parts, int box, int nops,
                            const double* ops, const float* x0, int sub, float apix, void* boxes_dev, float* rec);
/* cim_local_box_stack_dev without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * working memory (twiddles, operators, per pass of particles their angles,
 * records, complex rows and shifted images) lives in a 64 MiB array in device
 * memory that every GPU the library runs on holds until the process ends;
 * particles are taken in passes that fit it.  CIM_E_ARG when one particle's
 * working set (12 box^2 + 32 nops + 20 bytes, plus the twiddles and
 * operators) does not fit.  Nothing in the array needs clearing between
 * calls: every pass writes what it reads.  Calls that overlap on one GPU
 * (several contexts or threads) share the array and race.  Same results as
 * cim_local_box_stack_dev (the same kernels). */
int cim_local_box_stack_noalloc(cim_ctx* ctx, int n, const void* images_dev, const float* parts, int box, int nops,
                                const double* ops, const float* x0, int sub, float apix, void* boxes_dev, float* rec);

#ifdef __cplusplus
}
#endif
#endif
