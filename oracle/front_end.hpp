// CPU oracle for the raw front end (scale, gradient, demosaic, highlights,
// YCbCr transform, despeckle) that produces the denoise's inputs, and for the
// raw noise-model fit.
//
// TEST INFRASTRUCTURE ONLY: restates MetalPipeline/demosaic.metal and
// MetalPipeline/demosaic_utils.cpp in FP32 to check the HIP front end; the
// product library never links, loads or calls it.  Images are row-major,
// tightly packed, with 1 (raw, green), 2 (gradient) or 4 (Sobel, RGB, YCbCr)
// interleaved float channels; raw input is 16-bit.
#pragma once
#include <cstdint>

extern "C" {

void oracle_scale_raw(const uint16_t* raw, int w, int h, int bayer, const float* scale_mul, float black_level,
                      float lens_correction, float* out);
void oracle_raw_sobel(const float* in, int w, int h, float* out4);
int oracle_gaussian_bilinear_weights(float radius, float* out, int cap);
void oracle_blur_sobel(const float* raw, const float* sobel4, int w, int h, const float* w1, int n1,
                       const float* w2, int n2, float var_a, float var_b, float* out2);
void oracle_interpolate_green(const float* raw, const float* grad2, int w, int h, int bayer, float gvar_a,
                              float gvar_b, float* green);
void oracle_interpolate_red_blue(const float* raw, const float* green, const float* grad2, int w, int h, int bayer,
                                 const float* rvar, const float* bvar, float* rgb4);
void oracle_interpolate_red_blue_at_green(const float* rgb4_in, const float* grad2, int w, int h, int bayer,
                                          const float* rvar, const float* bvar, float* rgb4_out);
void oracle_blend_highlights(const float* in4, int w, int h, float clip, float* out4);
void oracle_transform(const float* in4, int w, int h, const float* m, float* out4);
void oracle_despeckle(const float* in4, int w, int h, const float* var_a, const float* var_b, float* out4);
void oracle_cam_ycbcr(const float* rgb_cam, float* out);
void oracle_histogram(const float* in4, int w, int h, uint32_t* hist);
void oracle_histogram_statistics(const uint32_t* hist, uint32_t* bands, int w, int h, float* stats);
void oracle_guided_filter_ab(const float* guide4, int w, int h, float eps, float* ab2);
void oracle_box_filter_gf(const float* in2, int w, int h, float* out2);
void oracle_ltm_mask(const float* in4, const float* grad2, int w, int h, const float* lf2, int lw, int lh,
                     const float* mf2, int mw, int mh, const float* hf2, int hw, int hh, const float* params,
                     float nlf_a, float nlf_b, float hist_shadows, float hist_highlights, float* mask);
void oracle_raw_noise_stats(const float* raw, int w, int h, int bayer, float* mean4, float* var4);
void oracle_fit_raw_nlf(const float* mean4, const float* var4, int qw, int qh, double exposure_multiplier,
                        double npx, double* a4, double* b4, double* diag);

}  // extern "C"
