// The PyramidProcessor handle (the per-frame state of
// MetalPipeline/pyramid_processor.hpp:22-49), shared by the single-frame and
// the burst entry points.
#pragma once

#include "gcam_internal.hpp"

struct gcam_pyramid {
    int width = 0, height = 0, levels = 0, radius = 10, precision = GCAM_PRECISION_FP32, bm_rows = 0, device = 0;
    bool bilateral = false;  // usePatchSimiliarity == false
    int W[GCAM_MAX_LEVELS] = {}, H[GCAM_MAX_LEVELS] = {};
    char* arena = nullptr;        // one allocation for every fixed-size buffer
    float2* grad_arena = nullptr; // gradient pyramid, allocated on first use
    float4* layer[GCAM_MAX_LEVELS] = {};  // level-i noisy input, i >= 1 (imagePyramid[i-1])
    float2* gradl[GCAM_MAX_LEVELS] = {};  // gradient pyramid, i >= 1
    float4* sub[GCAM_MAX_LEVELS] = {};    // subtractedImagePyramid
    float4* den[GCAM_MAX_LEVELS] = {};    // denoisedImagePyramid
    float* feat = nullptr;                // PCA features of one level (reused per level)
    double* moment_partials = nullptr;
    double* nlf_partials = nullptr;
    gcam::NlfFitState* nlf_state = nullptr;
    gcam::NlfDev* nlf_dev = nullptr;
    float* basis = nullptr;    // levels x 25 x 8
    double* eigvals = nullptr; // levels x 25 (9..25 NaN until gcam_pyramid_basis completes them)
    double* tri = nullptr;     // levels x 50: each level's tridiagonal form (diagonal, off-diagonal, scale)
    int feature_level = -1;
    // differentiation form (allocated by the first gcam_pyramid_denoise_fwd)
    char* adj_arena = nullptr;
    double* eig_full = nullptr;       // levels x kEigStride
    float4* adj_coef = nullptr;       // 3 float4 per level-0 pixel
    float4* adj_gS = nullptr;         // level-0 sized
    float4* adj_gF = nullptr;         // 2 float4 per level-0 pixel
    float4* adj_gD[2] = {};           // ping-pong dL/d(denoised), level-1 sized
    float* adj_gthr_part = nullptr;
    float* adj_gthr = nullptr;        // levels x 3
    double* adj_gB_part = nullptr;
    double* adj_gC = nullptr;
    bool fwd_ready = false;
    gcam_image fwd_image = {};
    float fwd_thr[GCAM_MAX_LEVELS][3] = {};
    float fwd_cb[GCAM_MAX_LEVELS] = {};
    bool profiling = false;
    bool ran[GCAM_MAX_LEVELS][GCAM_STAGE_COUNT] = {};
    hipEvent_t ev[GCAM_MAX_LEVELS][GCAM_STAGE_COUNT][2] = {};
};
