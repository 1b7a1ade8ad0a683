// CPU oracle for the pyramid PCA block-matching denoise.
//
// TEST INFRASTRUCTURE ONLY.  This is a restatement, in plain C++, of the
// reference's denoise path, written to check the HIP implementation.  Only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark load
// it; the product library never links, loads or calls it.
//
// Every function cites the reference file:line it restates.  Semantics the
// reference leaves undefined are fixed here, once, and documented at the
// function that fixes them:
//   * out-of-bounds texture.read() (collectPatches, pcaProjection,
//     blockMatchingDenoiseImage, basicNoiseStatistics) -> clamp-to-edge;
//   * bilinear sample positions are computed exactly in integer arithmetic
//     (the reference's sampler uses fixed-point sub-texel positions, which are
//     exact for the power-of-two pyramids used here);
//   * patch slots the collect grid never writes hold zeros;
//   * the PCA basis and features stay in the working precision (no fp16
//     storage) in the FP32 and FP64 forms;
//   * gls::max(NaN, floor) in the noise-model fit returns the floor.
#pragma once
#include <cstdint>

extern "C" {

// Per-level denoise parameters, mirroring DenoiseParameters
// (MetalPipeline/demosaic.hpp:34-41).
typedef struct {
    float luma, chroma, chromaBoost, gradientBoost, gradientThreshold, sharpening;
} oracle_denoise_params;

void oracle_set_threads(int n);
int oracle_get_threads(void);

// MetalPipeline/demosaic.metal:886-920 (downsampleImageXYZ / downsampleImageXY).
// `ch` is the number of interleaved channels of both images (4 or 2); the
// fourth channel of a 4-channel output is written as 0, like the reference.
void oracle_downsample_f32(const float* in, int win, int hin, int ch, float* out, int wout, int hout);
void oracle_downsample_f64(const double* in, int win, int hin, int ch, double* out, int wout, int hout);

// MetalPipeline/demosaic.metal:922-959 (subtractNoiseImage).
void oracle_subtract_noise_f32(const float* in, int w, int h, const float* in1, const float* den1, int w1, int h1,
                               const float* grad, float luma_weight, float sharpening, float nlf_x, float nlf_y,
                               float* out);
void oracle_subtract_noise_f64(const double* in, int w, int h, const double* in1, const double* den1, int w1, int h1,
                               const double* grad, double luma_weight, double sharpening, double nlf_x, double nlf_y,
                               double* out);

// MetalPipeline/demosaic.metal:755-772 (collectPatches), grid from
// MetalPipeline/demosaic_kernels.hpp:377-391 and sample size from
// MetalPipeline/pyramid_processor.cpp:142.  Writes n = w*h/64 slots of 25
// values (zero where the grid never writes) and returns n.
int oracle_collect_patches_f32(const float* in, int w, int h, float* patches);
int oracle_collect_patches_f64(const double* in, int w, int h, double* patches);

// MetalPipeline/PCA.cpp:23-56 (build_pca_space<25,8>), FP64.
// Covariance of the centred patches divided by (n-1), eigen-decomposition by
// cyclic Jacobi, the 8 largest eigenvectors (descending) as basis[r*8+c].
// Sign convention: each eigenvector is flipped so that its dot product with
// the fixed vector gcam_sign_ref (oracle and device use the same table) is
// non-negative.  eigvals (nullable) receives all 25 eigenvalues, descending.
void oracle_patch_covariance_f32(const float* patches, int n, double* cov);
void oracle_patch_covariance_f64(const double* patches, int n, double* cov);
void oracle_eigh_jacobi(const double* cov, int n, double* eigvals, double* eigvecs /* column j = vector j */);
void oracle_build_pca_space_f32(const float* patches, int n, double* basis, double* eigvals);
void oracle_build_pca_space_f64(const double* patches, int n, double* basis, double* eigvals);

// MetalPipeline/demosaic.metal:774-796 (pcaProjection).  feat is w*h*8.
void oracle_pca_project_f32(const float* in, int w, int h, const float* basis, float* feat);
void oracle_pca_project_f64(const double* in, int w, int h, const double* basis, double* feat);

// MetalPipeline/demosaic.metal:806-884 (blockMatchingDenoiseImage), window
// (2*radius+1)^2; the reference hard-codes radius 10 (:853).
void oracle_block_match_f32(const float* in, const float* grad, const float* feat, int w, int h,
                            const float var_a[3], const float var_b[3], const float thr[3], float chromaBoost,
                            float gradientBoost, float gradientThreshold, float lensShading, int radius, float* out);
void oracle_block_match_f64(const double* in, const double* grad, const double* feat, int w, int h,
                            const double var_a[3], const double var_b[3], const double thr[3], double chromaBoost,
                            double gradientBoost, double gradientThreshold, double lensShading, int radius,
                            double* out);

// MetalPipeline/demosaic.metal:664-719 (tunnel + denoiseImage), the non-PCA
// bilateral used when usePatchSimiliarity is false
// (MetalPipeline/pyramid_processor.hpp:27, .cpp:162-168).
void oracle_denoise_image_f32(const float* in, const float* grad, int w, int h, const float var_a[3],
                              const float var_b[3], const float thr[3], float chromaBoost, float gradientBoost,
                              float gradientThreshold, float* out);
void oracle_denoise_image_f64(const double* in, const double* grad, int w, int h, const double var_a[3],
                              const double var_b[3], const double thr[3], double chromaBoost, double gradientBoost,
                              double gradientThreshold, double* out);

// MetalPipeline/demosaic.metal:1138-1159 (basicNoiseStatistics).
void oracle_noise_stats_f32(const float* in, int w, int h, float* stats);
void oracle_noise_stats_f64(const double* in, int w, int h, double* stats);

// MetalPipeline/pyramid_processor.cpp:174-324 (MeasureYCbCrNLF after the
// statistics kernel), FP64.  Returns 1 if the second (filtered) regression was
// accepted, 0 if the first-pass fit was kept.  diag (nullable, 16 doubles):
// histogram[6], N1, N2, err2[3], newErr2[3], discarded, varianceMaxFirst.
int oracle_fit_ycbcr_nlf(const float* stats, int w, int h, double exposure_multiplier, double A[3], double B[3],
                         double* diag);

// MetalPipeline/pyramid_processor.cpp:96-172 (PyramidProcessor::denoise).
//   image:    w*h*4 YCbCr(+w) level-0 input
//   gradient: w*h*2 or NULL (treated as zeros)
//   params:   levels entries
//   nlf:      levels*6 doubles {A0,A1,A2,B0,B1,B2}; written when calibrate != 0
//   denoised: sum_i w_i*h_i*4, level 0 first (the per-level denoised images)
//   subtracted (nullable): same layout; level levels-1 holds the undenoised
//             coarsest pyramid level (the layer denoised there)
//   bases (nullable): levels*200 doubles, basis of each level
//   eigvals (nullable): levels*25 doubles
// Returns 0 on success, -1 if a level is too small to sample two patches.
int oracle_pyramid_denoise_f32(const float* image, const float* gradient, int w, int h, int levels, int radius,
                               int patch_similarity,
                               const oracle_denoise_params* params, double* nlf, float exposure_multiplier,
                               float lensShading, int calibrate, float* denoised, float* subtracted, double* bases,
                               double* eigvals);
int oracle_pyramid_denoise_f64(const double* image, const double* gradient, int w, int h, int levels, int radius,
                               int patch_similarity,
                               const oracle_denoise_params* params, double* nlf, double exposure_multiplier,
                               double lensShading, int calibrate, double* denoised, double* subtracted,
                               double* bases, double* eigvals);

// Differentiation form, FP64: the whole multi-level denoise with independent
// per-level, per-channel thresholds thr[levels*3] (Y, Cb, Cr; the reference
// ties Cb and Cr to DenoiseParameters::chroma), chromaBoost per level, no
// gradient image, gradientBoost 0, sharpening 1, no lens shading (the
// differentiation workload's setting).  Loss = 1/2 sum over pixels and Y, Cb,
// Cr of (denoised0 - clean)^2.  Returns the loss; writes dL/dthr (nullable)
// and denoised level 0 (nullable, w*h*4).  The backward pass differentiates
// every stage, the PCA basis included (eigenvector derivative
// dv_k = sum_{j!=k} v_j (v_j^T dC v_k) / (lambda_k - lambda_j)), by scatter.
double oracle_pyramid_loss_grad_f64(const double* image, const double* clean, int w, int h, int levels, int radius,
                                    const double* thr, const double* chroma_boost, const double* nlf,
                                    double* grad_thr, double* denoised0);

}  // extern "C"
