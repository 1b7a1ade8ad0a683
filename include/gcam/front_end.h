/*
 * gcam front end -- C ABI of the raw front end that produces the inputs of
 * the pyramid denoise (gcam/denoise.h) on MI355X (gfx950): the steps of
 * RawConverter::demosaic up to the YCbCr image and the gradient image
 * (MetalPipeline/raw_converter.cpp:148-214), the raw noise-model fit
 * (MeasureRawNLF, :339-490), and the despeckle that RawConverter::denoise runs
 * before PyramidProcessor::denoise (raw_converter.cpp:70-78).
 *
 * Same conventions as gcam/denoise.h: device images, stream-ordered
 * asynchronous calls, status codes, gcam_last_error() for the message.
 * Images of one call share one size; raw frames have even width and height.
 * Paths are relative to the GlassCameraPublic repository.
 */
#ifndef GCAM_FRONT_END_H
#define GCAM_FRONT_END_H

#include "gcam/denoise.h"

#ifdef __cplusplus
extern "C" {
#endif

/* BayerPattern (MetalPipeline/demosaic.hpp, demosaic.metal:24-29). */
typedef enum gcam_bayer_pattern {
    GCAM_BAYER_GRBG = 0,
    GCAM_BAYER_GBRG = 1,
    GCAM_BAYER_RGGB = 2,
    GCAM_BAYER_BGGR = 3
} gcam_bayer_pattern;

/* scaleRawDataKernel (MetalPipeline/demosaic_kernels.hpp:30-49, demosaic.metal:95-117):
 * scaled = max(lens * scale_mul[c] * (raw / 65535 - black_level) * 0.9 + 0.1, 0)
 * per quad colour c (red, green, blue, green2), lens = 1 + lens_shading * d^2
 * at the quad's distance d from the centre.  black_level is normalised
 * (DemosaicParameters.black_level / 0xffff, raw_converter.cpp:182).
 * raw: GCAM_PIXEL_R16U; scaled: GCAM_PIXEL_R32F. */
gcam_status gcam_scale_raw_data(const gcam_image* raw, gcam_image* scaled, int32_t bayer_pattern,
                                const float scale_mul[4], float black_level, float lens_shading, void* stream);

/* rawImageSobelKernel (demosaic_kernels.hpp:51-64, demosaic.metal:120-139):
 * sobel = (gx, gy, |gx|, |gy|) of the scaled raw (RGBA32F). */
gcam_status gcam_raw_image_sobel(const gcam_image* scaled, gcam_image* sobel, void* stream);

/* gaussianKernelBilinearWeights (MetalPipeline/demosaic_utils.cpp:987-1006):
 * writes the bilinear taps {weight, x offset, y offset} (3 floats each) when
 * capacity >= the count; returns the count (-1 for a bad radius).  Host only. */
int32_t gcam_gaussian_kernel_bilinear_weights(float radius, float* taps, int32_t capacity);

/* gaussianBlurSobelImageKernel (demosaic_kernels.hpp:66-97, sampledConvolutionSobel
 * demosaic.metal:188-210): blur the Sobel image with the radius1 kernel; where
 * its gradient is below 4 sigma of the raw noise (raw_variance = {a, b},
 * var = a + b * scaled), with the radius2 kernel.  gradient (RG32F) = the
 * blurred (|gx|, |gy|) with the signs of the blurred (gx, gy).  sobel may be
 * NULL: the Sobel filter of `scaled` is then computed inside the blur and
 * never stored.  radius1 <= 1.5, radius2 <= 5.5 (the reference uses 1.5 and
 * 4.5, raw_converter.hpp:134). */
gcam_status gcam_gaussian_blur_sobel_image(const gcam_image* scaled, const gcam_image* sobel, float radius1,
                                           float radius2, const float raw_variance[2], gcam_image* gradient,
                                           void* stream);

/* getRawVariance (MetalPipeline/demosaic.hpp:186-193): RawNLF (a[4], b[4]
 * per quad colour) -> {red a, b, green a, b, blue a, b}.  Host only. */
void gcam_raw_variance(const float raw_nlf_a[4], const float raw_nlf_b[4], float out[6]);

/* demosaicImageKernel (demosaic_kernels.hpp:124-181): interpolateGreen ->
 * green (R32F), interpolateRedBlue -> rgb_tmp, interpolateRedBlueAtGreen ->
 * rgb_out (RGBA32F, linear camera RGB, w = 0) (demosaic.metal:214-521).
 * raw_variance as returned by gcam_raw_variance. */
gcam_status gcam_demosaic_image(const gcam_image* scaled, const gcam_image* gradient, gcam_image* green,
                                gcam_image* rgb_tmp, gcam_image* rgb_out, int32_t bayer_pattern,
                                const float raw_variance[6], void* stream);

/* blendHighlightsImageKernel (demosaic_kernels.hpp:255-269, demosaic.metal:536-564). */
gcam_status gcam_blend_highlights_image(const gcam_image* in, float clip, gcam_image* out, void* stream);

/* transformImageKernel (demosaic_kernels.hpp:271-295, demosaic.metal:649-662):
 * out.xyz = m (row-major 3x3) * in.xyz, out.w = 0. */
gcam_status gcam_transform_image(const gcam_image* in, const float m[9], gcam_image* out, void* stream);

/* despeckleImageKernel (demosaic_kernels.hpp:297-315, despeckleLumaMedianChromaImage
 * demosaic.metal:1242-1283): chroma = 3x3 median, luma clamped between its
 * neighbourhood's extremes by the noise model var = var_a + var_b * Y
 * (channel 0 used).  in and out must differ. */
gcam_status gcam_despeckle_image(const gcam_image* in, const float var_a[3], const float var_b[3], gcam_image* out,
                                 void* stream);

/* basicRawNoiseStatistics (demosaic.metal:1096-1136, driven at
 * raw_converter.cpp:340): per 2x2 quad, the mean and variance of the red,
 * green, blue and green2 samples over a 9x9 window of quad reads (the window
 * steps one pixel at a time, as in the reference).  mean, var: RGBA32F of half
 * the scaled raw's size. */
gcam_status gcam_raw_noise_statistics(const gcam_image* scaled, int32_t bayer_pattern, gcam_image* mean,
                                      gcam_image* var, void* stream);

/* MeasureRawNLF (MetalPipeline/raw_converter.cpp:339-490): the statistics and
 * the reference's FP64 least-squares fit var = a + b * mean per quad colour
 * (histogram-chosen variance ceiling, refit on the quads within half the mean
 * squared error), on the device; synchronizes the stream and writes the
 * RawNLF (a[4], b[4], scaled by exposure_multiplier^2) to the host arrays. */
gcam_status gcam_measure_raw_nlf(const gcam_image* scaled, int32_t bayer_pattern, float exposure_multiplier,
                                 float raw_nlf_a[4], float raw_nlf_b[4], void* stream);

/* histogramImageKernel::histogram_data (MetalPipeline/demosaic_kernels.hpp:494-503):
 * the 256-bin luma histogram RawConverter::denoise takes of denoisedImagePyramid[3]
 * (raw_converter.cpp:87-89) and its statistics.  Lives in device memory. */
typedef struct gcam_histogram_data {
    uint32_t histogram[256];
    uint32_t bands[8];
    float black_level;
    float white_level;
    float shadows;
    float highlights;
    float mean;
    float median;
} gcam_histogram_data;

/* histogramImageKernel::reset (demosaic_kernels.hpp:507-509), on the stream. */
gcam_status gcam_histogram_reset(gcam_histogram_data* data, void* stream);

/* histogramImage (demosaic.metal:1373-1383): adds the image's luma, bin
 * clamp((int)(255 sqrt(Y / 2)), 0, 255), to data->histogram (RGBA32F image). */
gcam_status gcam_histogram_image(const gcam_image* image, gcam_histogram_data* data, void* stream);

/* histogramStatistics (demosaic.metal:1385-1453): bands, black / white level,
 * median, mean, highlights and shadows of data->histogram for an image of
 * width x height pixels (bands accumulate, as in the reference). */
gcam_status gcam_histogram_statistics(gcam_histogram_data* data, int32_t width, int32_t height, void* stream);

/* LTMParameters (MetalPipeline/demosaic.hpp:95-100); defaults {0.01, 0.8, 1.05,
 * {1.1, 1.2, 1.3}}.  The mask reads shadows and highlights from the histogram. */
typedef struct gcam_ltm_params {
    float eps;
    float shadows;
    float highlights;
    float detail[3];
} gcam_ltm_params;

/* GuidedFilterABImage (demosaic.metal:1295-1323): per guide pixel the guided
 * filter's (a, b) = (var / (var + eps), mean (1 - a)) of the 5x5 luma box.
 * guide RGBA32F, ab RG32F, one size. */
gcam_status gcam_guided_filter_ab_image(const gcam_image* guide, float eps, gcam_image* ab, void* stream);

/* BoxFilterGFImage (demosaic.metal:1325-1355): 5x5 box mean of an RG32F image. */
gcam_status gcam_box_filter_gf_image(const gcam_image* in, gcam_image* out, void* stream);

/* localToneMappingMaskImage (demosaic.metal:1481-1535), as driven by
 * LocalToneMapping::createMask (raw_converter.hpp:61-75, demosaic_kernels.hpp:
 * 565-600): the R32F mask for `image` (denoisedImagePyramid[0]) from the box
 * means of the guided-filter (a, b) of denoised levels 4, 2 and 0 (ab_mean_lf
 * at W/16, ab_mean_mf at W/4, ab_mean_hf at W), the gradient, the level-0
 * luma noise model (nlf_a, nlf_b = pyramidNlf[0].first[0], .second[0]) and the
 * shadows / highlights of a histogram after gcam_histogram_statistics
 * (device memory). */
gcam_status gcam_local_tone_mapping_mask_image(const gcam_image* image, const gcam_image* gradient,
                                               const gcam_image* ab_mean_lf, const gcam_image* ab_mean_mf,
                                               const gcam_image* ab_mean_hf, const gcam_ltm_params* params,
                                               float nlf_a, float nlf_b, const gcam_histogram_data* histogram,
                                               gcam_image* mask, void* stream);

/* cam_ycbcr(rgb_cam, xyz_sRGB) (MetalPipeline/demosaic_utils.cpp:703-720,
 * raw_converter.cpp:166): the camera-RGB -> YCbCr matrix, row-major.  Host only. */
void gcam_cam_ycbcr(const float rgb_cam[9], float out[9]);

/* inverse(cam_to_ycbcr) (raw_converter.cpp:169): 3x3 inverse, row-major.  Host only. */
gcam_status gcam_matrix3_inverse(const float m[9], float out[9]);

/* The front end as one call: RawConverter::demosaic from the raw frame to the
 * YCbCr image (raw_converter.cpp:179-214; the high-noise raw denoise of
 * :195-202 is not part of it), then RawConverter::denoise's despeckle
 * (:70-78).  Three passes: scaleRawData; the gradient blur with the Sobel
 * filter fused in; interpolateGreen through the YCbCr transform and the
 * despeckle as one tiled pass (intermediates in on-chip memory), equal bit for
 * bit to the per-operation entry points in sequence. */
typedef struct gcam_front_end_params {
    int32_t bayer_pattern;       /* gcam_bayer_pattern                               */
    float black_level;           /* normalised, as for gcam_scale_raw_data           */
    float lens_shading;          /* DemosaicParameters.lensShadingCorrection         */
    float scale_mul[4];          /* DemosaicParameters.scale_mul                     */
    float raw_nlf_a[4];          /* NoiseModel.rawNlf.first                          */
    float raw_nlf_b[4];          /* NoiseModel.rawNlf.second                         */
    float rgb_cam[9];            /* DemosaicParameters.rgb_cam, row-major            */
    float highlight_clip;        /* blendHighlightsImage clip; 1.0 in the reference  */
    int32_t despeckle;           /* 1: run the despeckle with despeckle_nlf          */
    gcam_ycbcr_nlf despeckle_nlf; /* NoiseModel.pyramidNlf[0]                        */
    int32_t calibrate_raw_nlf;   /* 1: fit the raw noise model from the frame
                                    (MeasureRawNLF, raw_converter.cpp:187-189) and
                                    use it instead of raw_nlf_a / raw_nlf_b; the
                                    call then synchronizes once, after the fit      */
    float exposure_multiplier;   /* MeasureRawNLF's exposure_multiplier              */
} gcam_front_end_params;

typedef struct gcam_front_end gcam_front_end;

/* Allocates the one intermediate image (the scaled raw) for one frame size on
 * the current device. */
gcam_status gcam_front_end_create(int32_t width, int32_t height, gcam_front_end** out);
void gcam_front_end_destroy(gcam_front_end* front_end);

/* raw (R16U) -> ycbcr (RGBA32F, the image for gcam_pyramid_denoise) and
 * gradient (RG32F, its gradient image).  One handle per stream. */
gcam_status gcam_front_end_run(gcam_front_end* front_end, const gcam_image* raw, const gcam_front_end_params* params,
                               gcam_image* ycbcr, gcam_image* gradient, void* stream);

/* The raw noise model the last gcam_front_end_run used (fitted when it
 * calibrated; NoiseModel.rawNlf, written back by the reference at
 * raw_converter.cpp:187-189).  Host only. */
gcam_status gcam_front_end_raw_nlf(const gcam_front_end* front_end, float raw_nlf_a[4], float raw_nlf_b[4]);

#ifdef __cplusplus
}
#endif

#endif /* GCAM_FRONT_END_H */
