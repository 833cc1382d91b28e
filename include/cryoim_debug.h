/*
 * cryoim_debug.h -- kernel-level entry points of libcryoim_hip.so used by
 * the tests to check single kernels against float64 restatements.  They are
 * not part of the drop-in boundary (include/cryoim.h) and have no
 * counterpart in the reference.
 */
#ifndef CRYOIM_DEBUG_H
#define CRYOIM_DEBUG_H

#include "cryoim.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The residual tables of the tabulated search (CIM_SEARCH_TABLES) for np
 * particles x T templates from polar rays given on the host:
 * ppol np x nr x npsi and tpol (tpol_per_particle ? np T : T) x nr x npsi
 * pairs {phase (radians), |F|/2 or -1e30 when gated}; table receives
 * np T x (npsi/2) x npsi entries {N (low 16 bits), D (high 16 bits)} as
 * bfloat16.  nh = 0: the exact
 * folded phase residual on the VALU; nh = 1, 2, 4, 8: its cosine series cut
 * after nh odd harmonics, on the matrix cores. */
int cim_debug_cl_table(cim_ctx* ctx, int nh, int np, int T, int tpol_per_particle, int nr, int npsi, const float* ppol,
                       const float* tpol, unsigned int* table);

/* The activations the last cim_cnn_forward kept: layer 0 the B input images
 * (128 x 128), layer l = 1..6 the output of convolution l (NHWC, B x
 * (128 >> l)^2 x 64 floats). */
int cim_debug_cnn_act(cim_ctx* ctx, int layer, float* out);

#ifdef __cplusplus
}
#endif
#endif
