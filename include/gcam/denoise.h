/*
 * gcam denoise -- C ABI of the MI355X (gfx950) pyramid PCA block-matching
 * denoise.
 *
 * Every entry point is plain C: POD structs, device pointers, sizes and a
 * hipStream_t passed as void*.  Calls are stream-ordered and asynchronous
 * unless their comment says they synchronize.  No exception crosses the ABI;
 * failures return a gcam_status and gcam_last_error() describes the last one
 * on the calling thread.
 *
 * The reference interface each entry point replaces is cited beside it
 * (paths relative to the GlassCameraPublic repository).
 */
#ifndef GCAM_DENOISE_H
#define GCAM_DENOISE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GCAM_ABI_VERSION 3
#define GCAM_MAX_LEVELS 8
#define GCAM_MAX_SEARCH_RADIUS 10
#define GCAM_MAX_BURST 64

typedef enum gcam_status {
    GCAM_OK = 0,
    GCAM_ERR_INVALID_ARGUMENT = 1,
    GCAM_ERR_HIP = 2,
    GCAM_ERR_OUT_OF_MEMORY = 3,
    GCAM_ERR_UNSUPPORTED = 4,
    GCAM_ERR_NO_DEVICE = 5
} gcam_status;

/* Pixel formats of gls::mtl_image_2d<T> on this path
 * (MetalPipeline/gls_mtl_image.hpp:35-83). */
typedef enum gcam_pixel_format {
    GCAM_PIXEL_RGBA32F = 0,  /* gls::pixel_float4: YCbCr + w                   */
    GCAM_PIXEL_RG32F = 1,    /* gls::pixel_float2: gradient                    */
    GCAM_PIXEL_R32F = 2,     /* gls::pixel_float                               */
    GCAM_PIXEL_PCA8_F32 = 3, /* 8 x float PCA features (32 B per pixel)        */
    GCAM_PIXEL_PCA8_F16 = 4, /* 8 x half in a uint4 texel, the reference's
                                pcaImage format (demosaic.metal:721-741)       */
    GCAM_PIXEL_R16U = 5      /* gls::luma_pixel_16: 16-bit raw Bayer data      */
} gcam_pixel_format;

/* A device image: gls::mtl_image_2d<T> (MetalPipeline/gls_mtl_image.hpp:116-208).
 * data is a device pointer; stride is the row pitch in pixels (>= width). */
typedef struct gcam_image {
    void* data;
    int32_t width;
    int32_t height;
    int32_t stride;
    int32_t format; /* gcam_pixel_format */
} gcam_image;

/* DenoiseParameters (MetalPipeline/demosaic.hpp:34-41); defaults {1,1,1,0,1,1}. */
typedef struct gcam_denoise_params {
    float luma;
    float chroma;
    float chromaBoost;
    float gradientBoost;
    float gradientThreshold;
    float sharpening;
} gcam_denoise_params;

/* YCbCrNLF = std::pair<Vector<3>, Vector<3>> (MetalPipeline/demosaic.hpp:54):
 * per-channel noise variance var = a + b * Y. */
typedef struct gcam_ycbcr_nlf {
    float a[3];
    float b[3];
} gcam_ycbcr_nlf;

typedef enum gcam_precision {
    GCAM_PRECISION_FP32 = 0, /* plain FP32 form                                */
    GCAM_PRECISION_FP16 = 1, /* FP16 MFMA contractions, FP32 accumulation       */
    GCAM_PRECISION_BF16 = 2  /* BF16 MFMA contractions, FP32 accumulation       */
} gcam_precision;

/* Build-time knobs the reference hard-codes: the search window radius
 * (demosaic.metal:853, size = 10) and the arithmetic precision. */
typedef struct gcam_config {
    int32_t search_radius; /* 1..GCAM_MAX_SEARCH_RADIUS; window (2r+1)^2      */
    int32_t precision;     /* gcam_precision                                   */
    int32_t bm_rows_per_thread; /* block-matching tile shape: output rows per
                                   thread, 1, 2 or 4; 0 = tuned default         */
    int32_t bilateral;     /* 0: PCA block matching (usePatchSimiliarity =
                              true, pyramid_processor.hpp:27); 1: the non-PCA
                              denoiseImage bilateral (demosaic.metal:669-719) */
    int32_t reserved[4];   /* must be zero                                     */
} gcam_config;

typedef struct gcam_pyramid gcam_pyramid;

/* Which per-level image gcam_pyramid_level() returns. */
typedef enum gcam_level_image {
    GCAM_LEVEL_DENOISED = 0,   /* denoisedImagePyramid[i]                      */
    GCAM_LEVEL_SUBTRACTED = 1, /* subtractedImagePyramid[i] (i < levels-1)    */
    GCAM_LEVEL_NOISY = 2,      /* the level-i input: imagePyramid[i-1]        */
    GCAM_LEVEL_FEATURES = 3    /* PCA features of the last denoised level     */
} gcam_level_image;

int32_t gcam_abi_version(void);
const char* gcam_status_string(gcam_status status);
const char* gcam_last_error(void);

/* PyramidProcessor<levels>::PyramidProcessor(MetalContext*, int width, int height)
 * (MetalPipeline/pyramid_processor.hpp:57, .cpp:25-50).  Allocates every
 * per-level buffer on the current HIP device.  config may be NULL (radius 10,
 * FP32).  Fails if a level is too small to sample two PCA patches. */
gcam_status gcam_pyramid_create(int32_t width, int32_t height, int32_t levels, const gcam_config* config,
                                gcam_pyramid** out);
void gcam_pyramid_destroy(gcam_pyramid* pyramid);

/* PyramidProcessor<levels>::denoise(...) (MetalPipeline/pyramid_processor.hpp:59-62,
 * .cpp:96-172).
 *   params:   `levels` entries (read only)
 *   image:    level-0 YCbCr, GCAM_PIXEL_RGBA32F, width x height of the handle
 *   gradient: GCAM_PIXEL_RG32F or NULL (then the gradient is zero)
 *   nlf:      `levels` entries, in/out: written when calibrate_from_image != 0
 *   out:      NULL, or receives a view of denoised level 0 (owned by the
 *             handle, valid until the next call or destroy), like the
 *             reference's returned pointer.
 * Stream-ordered; with calibrate_from_image it synchronizes the stream once at
 * the end to return the fitted noise model. */
gcam_status gcam_pyramid_denoise(gcam_pyramid* pyramid, const gcam_denoise_params* params, const gcam_image* image,
                                 const gcam_image* gradient, gcam_ycbcr_nlf* nlf, float exposure_multiplier,
                                 float lens_shading_correction, int32_t calibrate_from_image, void* stream,
                                 gcam_image* out);

/* Direct access to denoisedImagePyramid[level] and the other per-level images
 * (read by RawConverter::denoise, MetalPipeline/raw_converter.cpp:88,107-109). */
gcam_status gcam_pyramid_level(const gcam_pyramid* pyramid, int32_t level, int32_t which, gcam_image* out);

/* The FP32 PCA basis (25 x 8, row-major [row][component]) and the 25
 * eigenvalues (descending) of `level` from the last denoise call; host
 * pointers, either may be NULL.  Synchronizes the device. */
gcam_status gcam_pyramid_basis(const gcam_pyramid* pyramid, int32_t level, float* basis, double* eigenvalues);

/* Stage timing with HIP events recorded on the call's stream; replaces the
 * reference's std::chrono prints (PipelineTest/pipelineTest.cpp:161-169,
 * MetalPipeline/PCA.cpp:44-46).  With profiling on, every denoise call records
 * one event pair per (level, stage); gcam_pyramid_stage_times() synchronizes
 * on them and writes ms[level * GCAM_STAGE_COUNT + stage] of the last call
 * (0 for stages that did not run). */
typedef enum gcam_stage {
    GCAM_STAGE_DOWNSAMPLE = 0,  /* building level i+1 from level i          */
    GCAM_STAGE_NLF = 1,         /* noise statistics + fit (calibrating)     */
    GCAM_STAGE_SUBTRACT = 2,    /* subtractNoiseImage                       */
    GCAM_STAGE_PCA = 3,         /* patch moments + eigensolve + projection  */
    GCAM_STAGE_BLOCK_MATCH = 4, /* blockMatchingDenoiseImage                */
    GCAM_STAGE_MOMENTS = 5,     /* the patch-moment kernel alone (inside GCAM_STAGE_PCA) */
    GCAM_STAGE_COUNT = 6
} gcam_stage;

gcam_status gcam_pyramid_set_profiling(gcam_pyramid* pyramid, int32_t enable);
gcam_status gcam_pyramid_stage_times(const gcam_pyramid* pyramid, float* ms);

/* ---- differentiation form (no counterpart in the reference) ----
 *
 * gcam_pyramid_denoise_fwd: PyramidProcessor::denoise
 * (MetalPipeline/pyramid_processor.cpp:96-172) with independent thresholds per
 * level and channel, thresholds[level * 3 + c] for c = Y, Cb, Cr (the
 * reference ties Cb and Cr to DenoiseParameters::chroma).  The handle keeps
 * what the reverse pass needs (each level's layer, output, and full PCA
 * eigen-decomposition); `image` must stay valid until the backward call.
 * Requires an FP32 PCA handle, gradientBoost == 0 and sharpening <= 1 on every
 * level, no gradient image and no lens shading (the differentiation workload's
 * setting); other cases return GCAM_ERR_UNSUPPORTED.
 *
 * gcam_pyramid_denoise_bwd: reverse pass of the last forward on the same
 * handle.  grad_output = dL/d(denoised level 0) (RGBA32F device image, .w
 * ignored); writes dL/dthresholds (levels * 3 floats, host).  Match weights
 * are recomputed, not stored; the PCA basis is differentiated too.
 * Synchronizes the stream. */
gcam_status gcam_pyramid_denoise_fwd(gcam_pyramid* pyramid, const gcam_denoise_params* params,
                                     const float* thresholds, const gcam_image* image, const gcam_ycbcr_nlf* nlf,
                                     void* stream, gcam_image* out);
gcam_status gcam_pyramid_denoise_bwd(gcam_pyramid* pyramid, const gcam_image* grad_output, float* grad_thresholds,
                                     void* stream);

/* ---- batched form: a burst of frames of uneven sizes (no counterpart in the reference) ----
 *
 * A burst handle holds one pyramid per frame (frame sizes fixed at creation,
 * in any mix) and denoises all frames with the PyramidProcessor::denoise
 * algorithm (MetalPipeline/pyramid_processor.cpp:96-172) in one graph launch:
 * every stage of a level covers all frames in one kernel launch, the block
 * matching runs as one persistent kernel over the tiles of every frame, and
 * the whole sequence is a hipGraph captured once per handle and replayed.
 * Frames, parameters and noise models reach the captured kernels through a
 * device-side table that each call refreshes (one host-to-device copy).
 *
 * The block matching launch (output rows per thread, persistent grid size,
 * tile order) is picked per level by an autotuner that times the candidates
 * on the first call's frames (or on gcam_burst_tune); gcam_burst_set_tuning
 * replaces it with a fixed choice. */
typedef struct gcam_burst gcam_burst;

typedef struct gcam_burst_tuning {
    int32_t rows_per_thread[GCAM_MAX_LEVELS]; /* 1, 2 or 4 (radius > 6: 1 or 2)          */
    int32_t grid[GCAM_MAX_LEVELS];            /* persistent workgroups; 0 = one per tile */
    int32_t schedule[GCAM_MAX_LEVELS];        /* 0: tiles in frame order, strided over the
                                                 grid; 1: contiguous chunks per XCD     */
    int32_t tile_width[GCAM_MAX_LEVELS];      /* output tile width: 32, or 64 with 2 rows
                                                 per thread (radius <= 6)               */
    float level_ms[GCAM_MAX_LEVELS];          /* measured block-matching time (output)  */
    int32_t tuned;                            /* 1 once chosen (output)                  */
} gcam_burst_tuning;

/* n_frames in 1..GCAM_MAX_BURST; widths/heights: host arrays of n_frames.
 * config->precision FP16 or BF16 (the precision form); bilateral must be 0. */
gcam_status gcam_burst_create(int32_t n_frames, const int32_t* widths, const int32_t* heights, int32_t levels,
                              const gcam_config* config, gcam_burst** out);
void gcam_burst_destroy(gcam_burst* burst);

/* params: n_frames * levels (frame-major); frames: n_frames RGBA32F device
 * images of the creation sizes; gradients: NULL or n_frames images (RG32F;
 * an entry with data == NULL means no gradient for that frame); nlf:
 * n_frames * levels noise models (frame-major).  The first call without a
 * tuning autotunes (and synchronizes the stream); later calls are one
 * host-to-device table copy plus one graph launch, asynchronous. */
gcam_status gcam_burst_denoise(gcam_burst* burst, const gcam_denoise_params* params, const gcam_image* frames,
                               const gcam_image* gradients, const gcam_ycbcr_nlf* nlf, float lens_shading_correction,
                               void* stream);
/* Denoised level `level` of frame `frame` (handle-owned, valid until the next call). */
gcam_status gcam_burst_output(const gcam_burst* burst, int32_t frame, int32_t level, gcam_image* out);
/* Run the autotuner on these inputs now (outputs are left valid); synchronizes. */
gcam_status gcam_burst_tune(gcam_burst* burst, const gcam_denoise_params* params, const gcam_image* frames,
                            const gcam_image* gradients, const gcam_ycbcr_nlf* nlf, float lens_shading_correction,
                            void* stream);
gcam_status gcam_burst_get_tuning(const gcam_burst* burst, gcam_burst_tuning* tuning);
/* Fix the launch choice (validated); the graph is re-captured on the next call. */
gcam_status gcam_burst_set_tuning(gcam_burst* burst, const gcam_burst_tuning* tuning);

/* ---- per-operation entry points (MetalPipeline/demosaic_kernels.hpp wrappers) ---- */

/* resampleImageKernel (demosaic_kernels.hpp:470-483) running downsampleImageXYZ
 * (RGBA32F) or downsampleImageXY (RG32F), demosaic.metal:886-920. */
gcam_status gcam_downsample_image(const gcam_image* input, gcam_image* output, void* stream);

/* subtractNoiseImageKernel (demosaic_kernels.hpp:410-436), demosaic.metal:922-959.
 * gradient may be NULL (zero gradient). */
gcam_status gcam_subtract_noise_image(const gcam_image* input, const gcam_image* input1,
                                      const gcam_image* input_denoised1, const gcam_image* gradient,
                                      float luma_weight, float sharpening, float nlf_x, float nlf_y,
                                      gcam_image* output, void* stream);

/* collectPatchesKernel (demosaic_kernels.hpp:377-391), demosaic.metal:755-772.
 * patches: device buffer of `capacity` x 25 floats; the first
 * width*height/64 slots are written (slots the grid never reaches are set to
 * zero).  *count (host, may be NULL) receives width*height/64. */
gcam_status gcam_collect_patches(const gcam_image* input, float* patches, int64_t capacity, int64_t* count,
                                 void* stream);

/* build_pca_space<25,8> (MetalPipeline/PCA.hpp:21-23, PCA.cpp:23-56) on the
 * device, in FP64: covariance of the n device-resident patches (n x 25
 * floats), Jacobi eigen-decomposition, the 8 leading eigenvectors.  basis:
 * device 25 x 8 floats; eigenvalues: device 25 doubles or NULL. */
gcam_status gcam_build_pca_space(const float* patches, int64_t n, float* basis, double* eigenvalues, void* stream);

/* No counterpart in the reference: gcam_build_pca_space for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results; the
 * moment partials live in a fixed device array of the library (2.87 MB per
 * GPU, held until the process ends) instead of a per-call allocation, and no
 * call allocates.  n must be 2 .. GCAM_PCA_NOALLOC_MAX_PATCHES.  Every call on
 * one GPU shares that array: calls that overlap on one GPU (other streams or
 * host threads) race on it. */
#define GCAM_PCA_NOALLOC_MAX_PATCHES ((int64_t)1 << 24)
#define GCAM_PCA_NOALLOC_SCRATCH_BYTES ((size_t)1025 * 350 * 8)
gcam_status gcam_build_pca_space_noalloc(const float* patches, int64_t n, float* basis, double* eigenvalues,
                                         void* stream);

/* pcaProjectionKernel (demosaic_kernels.hpp:393-408), demosaic.metal:774-796.
 * basis: device 25 x 8 floats.  features: GCAM_PIXEL_PCA8_F32. */
gcam_status gcam_pca_projection(const gcam_image* input, const float* basis, gcam_image* features, void* stream);

/* blockMatchingDenoiseImageKernel (demosaic_kernels.hpp:345-375),
 * demosaic.metal:806-884, with a (2*search_radius+1)^2 window.
 * gradient may be NULL. */
gcam_status gcam_block_matching_denoise_image(const gcam_image* input, const gcam_image* gradient,
                                              const gcam_image* features, const float var_a[3],
                                              const float var_b[3], const float threshold_multipliers[3],
                                              float chroma_boost, float gradient_boost, float gradient_threshold,
                                              float lens_shading_correction, int32_t search_radius,
                                              gcam_image* output, void* stream);

/* Precision form of pcaProjection + blockMatchingDenoiseImage (demosaic.metal:774-796,
 * 806-884): the features of each tile are projected in-kernel on MFMA in
 * FP16 or BF16 (precision = GCAM_PRECISION_FP16 / _BF16) with FP32
 * accumulation and never written to memory; patch distances use FP16/BF16
 * products with FP32 accumulation; weights and sums are FP32.
 * basis: device 25 x 8 floats. */
gcam_status gcam_block_matching_denoise_image_lp(const gcam_image* input, const gcam_image* gradient,
                                                 const float* basis, const float var_a[3], const float var_b[3],
                                                 const float threshold_multipliers[3], float chroma_boost,
                                                 float gradient_boost, float gradient_threshold,
                                                 float lens_shading_correction, int32_t search_radius,
                                                 int32_t precision, gcam_image* output, void* stream);

/* denoiseImageKernel (demosaic_kernels.hpp:317-343), demosaic.metal:664-719:
 * the non-PCA bilateral (5x5 window, 9x9 when gradient_boost > 0).
 * gradient may be NULL.  Output .w = gradient magnitude, as the reference. */
gcam_status gcam_denoise_image(const gcam_image* input, const gcam_image* gradient, const float var_a[3],
                               const float var_b[3], const float threshold_multipliers[3], float chroma_boost,
                               float gradient_boost, float gradient_threshold, gcam_image* output, void* stream);

/* basicNoiseStatisticsKernel (demosaic_kernels.hpp:438-450), demosaic.metal:1138-1159. */
gcam_status gcam_basic_noise_statistics(const gcam_image* input, gcam_image* statistics, void* stream);

/* PyramidProcessor::MeasureYCbCrNLF (MetalPipeline/pyramid_processor.hpp:72-75,
 * .cpp:174-324): statistics into `scratch` (RGBA32F, same size as input), the
 * FP64 regression on the device, result to the host `nlf`.  Synchronizes the
 * stream. */
gcam_status gcam_measure_ycbcr_nlf(const gcam_image* input, gcam_image* scratch, float exposure_multiplier,
                                   gcam_ycbcr_nlf* nlf, void* stream);

/* ---- synthetic input (no counterpart in the reference, which reads DNG files) ---- */

typedef struct gcam_synth_params {
    uint64_t scene_seed; /* per-channel scene seeds are scene_seed + c           */
    uint64_t noise_seed;
    float shot_gain;     /* a: Poisson gain, var = a*x + b^2                      */
    float read_sigma;    /* b                                                     */
    float lo, hi;        /* scene range in linear units, e.g. 0.02 .. 0.9         */
    float period;        /* base period of the gradient-noise scene, in pixels    */
    int32_t octaves;     /* octaves (persistence 0.5, lacunarity 2)               */
} gcam_synth_params;

/* Seeded scene -> RGGB 12-bit Poisson-Gaussian raw (uint16 DN, width x height,
 * dense) -> bilinear demosaic -> sRGB YCbCr (RGBA32F).  clean_ycc (may be
 * NULL) receives the noiseless scene in YCbCr.  raw may be NULL. */
gcam_status gcam_synth_frame(int32_t width, int32_t height, const gcam_synth_params* params, uint16_t* raw,
                             gcam_image* ycc, gcam_image* clean_ycc, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GCAM_DENOISE_H */
