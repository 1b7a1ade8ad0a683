// C++ facade over the gcam C ABI with the shape of the reference's operator
// surface: gcam::PyramidProcessor<levels> mirrors PyramidProcessor<levels>
// (MetalPipeline/pyramid_processor.hpp:22-76), DenoiseParameters mirrors
// MetalPipeline/demosaic.hpp:34-41 and YCbCrNLF mirrors demosaic.hpp:54.
//
// Header-only; link against libgcam_denoise.so.  Errors become
// gcam::Error exceptions on the C++ side of the boundary only.
#pragma once

#include <array>
#include <stdexcept>
#include <string>
#include <utility>

#include "gcam/denoise.h"

namespace gcam {

struct Error : std::runtime_error {
    gcam_status status;
    Error(gcam_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(gcam_status s) {
    if (s != GCAM_OK) throw Error(s, std::string(gcam_status_string(s)) + ": " + gcam_last_error());
}

using DenoiseParameters = gcam_denoise_params;
inline constexpr DenoiseParameters defaultDenoiseParameters{1.0f, 1.0f, 1.0f, 0.0f, 1.0f, 1.0f};

// std::pair<Vector<3>, Vector<3>>: (A, B) of var = A + B * Y.
using YCbCrNLF = std::pair<std::array<float, 3>, std::array<float, 3>>;

template <size_t levels>
class PyramidProcessor {
   public:
    const int width, height;

    PyramidProcessor(int width_, int height_, int searchRadius = 10,
                     gcam_precision precision = GCAM_PRECISION_FP32, bool usePatchSimiliarity = true)
        : width(width_), height(height_) {
        gcam_config cfg{searchRadius, precision, 0, usePatchSimiliarity ? 0 : 1, {0, 0, 0, 0}};
        check(gcam_pyramid_create(width, height, (int32_t)levels, &cfg, &handle_));
    }
    ~PyramidProcessor() { gcam_pyramid_destroy(handle_); }
    PyramidProcessor(const PyramidProcessor&) = delete;
    PyramidProcessor& operator=(const PyramidProcessor&) = delete;

    // PyramidProcessor::denoise (MetalPipeline/pyramid_processor.hpp:59-62).
    // image / gradientImage are device images; gradientImage may be null.
    // Returns the processor-owned level-0 denoised image.
    gcam_image denoise(std::array<DenoiseParameters, levels>* denoiseParameters, const gcam_image& image,
                       const gcam_image* gradientImage, std::array<YCbCrNLF, levels>* nlfParameters,
                       float exposure_multiplier, float lensShadingCorrection, bool calibrateFromImage = false,
                       void* stream = nullptr) {
        std::array<gcam_ycbcr_nlf, levels> nlf;
        for (size_t i = 0; i < levels; i++)
            for (int c = 0; c < 3; c++) {
                nlf[i].a[c] = (*nlfParameters)[i].first[c];
                nlf[i].b[c] = (*nlfParameters)[i].second[c];
            }
        gcam_image out;
        check(gcam_pyramid_denoise(handle_, denoiseParameters->data(), &image, gradientImage, nlf.data(),
                                   exposure_multiplier, lensShadingCorrection, calibrateFromImage ? 1 : 0, stream,
                                   &out));
        if (calibrateFromImage)
            for (size_t i = 0; i < levels; i++)
                for (int c = 0; c < 3; c++) {
                    (*nlfParameters)[i].first[c] = nlf[i].a[c];
                    (*nlfParameters)[i].second[c] = nlf[i].b[c];
                }
        return out;
    }

    // denoisedImagePyramid[i] (read by RawConverter::denoise, raw_converter.cpp:88,107-109).
    gcam_image denoisedImagePyramid(int level) const {
        gcam_image out;
        check(gcam_pyramid_level(handle_, level, GCAM_LEVEL_DENOISED, &out));
        return out;
    }

    // PyramidProcessor::MeasureYCbCrNLF (MetalPipeline/pyramid_processor.hpp:72-75).
    static YCbCrNLF MeasureYCbCrNLF(const gcam_image& inputImage, gcam_image* noiseStats,
                                    float exposure_multiplier, void* stream = nullptr) {
        gcam_ycbcr_nlf n;
        check(gcam_measure_ycbcr_nlf(&inputImage, noiseStats, exposure_multiplier, &n, stream));
        return {{n.a[0], n.a[1], n.a[2]}, {n.b[0], n.b[1], n.b[2]}};
    }

    gcam_pyramid* handle() const { return handle_; }

   private:
    gcam_pyramid* handle_ = nullptr;
};

}  // namespace gcam
