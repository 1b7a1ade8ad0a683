// Runs the reference-shaped C++ surface, gcam::PyramidProcessor<5>
// (include/gcam/pyramid_processor.hpp, mirroring
// MetalPipeline/pyramid_processor.hpp:57-75), on the device the way
// RawConverter::denoise drives PyramidProcessor<5> (MetalPipeline/raw_converter.cpp:80-90):
// denoise with calibrateFromImage, then read denoisedImagePyramid[3] (the
// histogram's level, raw_converter.cpp:88), then MeasureYCbCrNLF on the input.
//
//   facade_denoise W H RADIUS input.f32 params.f32 outdir
//
// input.f32: H x W x 4 float32 YCbCr; params.f32: 5 x 6 float32
// DenoiseParameters.  Writes outdir/level0.f32, outdir/level3.f32 (H/8 x W/8
// x 4), outdir/nlf.f32 (the fitted 5 x (A[3], B[3])) and outdir/measured.f32
// (MeasureYCbCrNLF of the input: A[3], B[3]).  Exit status 0 on success.
#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gcam/pyramid_processor.hpp"

namespace {

bool read_file(const std::string& path, void* dst, size_t bytes) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    const size_t n = std::fread(dst, 1, bytes, f);
    std::fclose(f);
    return n == bytes;
}

bool write_file(const std::string& path, const void* src, size_t bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const size_t n = std::fwrite(src, 1, bytes, f);
    std::fclose(f);
    return n == bytes;
}

#define HIP_OK(expr)                                                               \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s: %s\n", #expr, hipGetErrorString(e_));       \
            return 2;                                                              \
        }                                                                          \
    } while (0)

int copy_out(const gcam_image& im, const std::string& path) {
    std::vector<float> host((size_t)im.width * im.height * 4);
    HIP_OK(hipMemcpy2D(host.data(), (size_t)im.width * 16, im.data, (size_t)im.stride * 16, (size_t)im.width * 16,
                       im.height, hipMemcpyDeviceToHost));
    return write_file(path, host.data(), host.size() * sizeof(float)) ? 0 : 3;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s W H RADIUS input.f32 params.f32 outdir\n", argv[0]);
        return 1;
    }
    const int w = std::atoi(argv[1]), h = std::atoi(argv[2]), radius = std::atoi(argv[3]);
    const std::string out = argv[6];
    std::vector<float> img((size_t)w * h * 4);
    std::array<gcam::DenoiseParameters, 5> params;
    if (!read_file(argv[4], img.data(), img.size() * sizeof(float)) ||
        !read_file(argv[5], params.data(), sizeof(params))) {
        std::fprintf(stderr, "cannot read the inputs\n");
        return 1;
    }
    void* dimg = nullptr;
    void* dstats = nullptr;
    HIP_OK(hipMalloc(&dimg, img.size() * sizeof(float)));
    HIP_OK(hipMalloc(&dstats, img.size() * sizeof(float)));
    HIP_OK(hipMemcpy(dimg, img.data(), img.size() * sizeof(float), hipMemcpyHostToDevice));
    const gcam_image image{dimg, w, h, w, GCAM_PIXEL_RGBA32F};
    gcam_image stats{dstats, w, h, w, GCAM_PIXEL_RGBA32F};
    try {
        gcam::PyramidProcessor<5> processor(w, h, radius);
        std::array<gcam::YCbCrNLF, 5> nlf{};
        // RawConverter::denoise: denoise(context, &denoiseParameters, image, gradientImage, &nlfParameters,
        // exposure_multiplier, lensShadingCorrection, calibrateFromImage)
        const gcam_image level0 = processor.denoise(&params, image, nullptr, &nlf, 1.0f, 0.0f, true);
        HIP_OK(hipDeviceSynchronize());
        if (copy_out(level0, out + "/level0.f32") != 0) return 3;
        if (copy_out(processor.denoisedImagePyramid(3), out + "/level3.f32") != 0) return 3;
        std::array<float, 30> fitted;
        for (int i = 0; i < 5; i++)
            for (int c = 0; c < 3; c++) {
                fitted[i * 6 + c] = nlf[i].first[c];
                fitted[i * 6 + 3 + c] = nlf[i].second[c];
            }
        if (!write_file(out + "/nlf.f32", fitted.data(), sizeof(fitted))) return 3;
        const gcam::YCbCrNLF m = gcam::PyramidProcessor<5>::MeasureYCbCrNLF(image, &stats, 1.0f);
        HIP_OK(hipDeviceSynchronize());
        const std::array<float, 6> measured{m.first[0], m.first[1], m.first[2], m.second[0], m.second[1], m.second[2]};
        if (!write_file(out + "/measured.f32", measured.data(), sizeof(measured))) return 3;
    } catch (const gcam::Error& e) {
        std::fprintf(stderr, "gcam::Error %d: %s\n", (int)e.status, e.what());
        return 4;
    }
    HIP_OK(hipFree(dimg));
    HIP_OK(hipFree(dstats));
    std::printf("facade_denoise %dx%d ok\n", w, h);
    return 0;
}
