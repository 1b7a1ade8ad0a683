// Internal declarations shared by the gcam HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "gcam/denoise.h"

namespace gcam {

// Patch geometry of the PCA path (MetalPipeline/pyramid_processor.hpp:28-29).
constexpr int kPatch = 25;      // 5 x 5 luma patch
constexpr int kComponents = 8;  // PCA components kept
constexpr int kMoments = kPatch + kPatch * (kPatch + 1) / 2;  // sum x (25) + upper triangle of sum x x^T (325)
constexpr int kMomentBlocks = 1024;  // most workgroups of the moment reduction (one partial each)
constexpr int kMomentSlots = kMomentBlocks + 1;  // partials of each workgroup, then their sum
constexpr int kBasisGradBlocks = 240;  // fixed grid of the basis-gradient reduction (reverse pass)
constexpr int kNlfBlocks = 256;     // fixed grid of the noise-model reductions

// Device view of an image: base pointer, size and row pitch in pixels.
template <class P>
struct View {
    P* ptr;
    int w, h, stride;
    __device__ __forceinline__ P& at(int x, int y) const { return ptr[(size_t)y * stride + x]; }
};

template <class P>
inline View<P> view(const gcam_image& im) {
    return View<P>{reinterpret_cast<P*>(im.data), im.width, im.height, im.stride};
}
template <class P>
inline View<P> view(P* p, int w, int h) {
    return View<P>{p, w, h, w};
}

// XCD-contiguous tile order for a 1-D grid of 8 * per_xcd workgroups over
// tiles_x x (tiles / tiles_x) tiles.  The dispatcher deals workgroups to the
// 8 XCDs in turn (workgroup b runs on XCD b % 8), so tile (b % 8) * per_xcd +
// b / 8 gives each XCD a contiguous run of tiles in row order: the aprons that
// neighbouring tiles share, left-right and up-down, then meet in that XCD's
// L2 instead of being fetched once per XCD.
struct XcdTiles {
    int tiles_x, tiles, per_xcd;
    // The grid size of the launch.
    int grid() const { return 8 * per_xcd; }
    // This workgroup's tile (column tx, row ty); false for the few workgroups
    // past the last tile, which must return before any barrier.
    __device__ __forceinline__ bool tile(int& tx, int& ty) const {
        const int t = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
        if (t >= tiles) return false;
        tx = t % tiles_x;
        ty = t / tiles_x;
        return true;
    }
};
inline XcdTiles xcd_tiles(int w, int h, int tile_w, int tile_h) {
    const int tx = (w + tile_w - 1) / tile_w, n = tx * ((h + tile_h - 1) / tile_h);
    return XcdTiles{tx, n, n > 0 ? (n + 7) / 8 : 1};
}

// Per-level noise model on the device: var = a + b * Y per channel.
struct NlfDev {
    float a[3];
    float b[3];
};

// Host-side parameters of one block-matching launch (by value).
struct BmParams {
    float thr[3];
    float chromaBoost, gradientBoost, gradientThreshold, lensShading;
};

// State of the on-device noise-model fit (MetalPipeline/pyramid_processor.cpp:186-323).
struct NlfFitState {
    double varianceMax[3];
    double A[3], B[3];
    double err2[3];
    double hist[6];
    double n1, n2, discarded;
    double newErr2[3];
    int accepted;
};

// ---- error plumbing ----
gcam_status set_error(gcam_status st, const std::string& msg);
#define GCAM_HIP_TRY(expr)                                                                   \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            return ::gcam::set_error(_e == hipErrorOutOfMemory ? GCAM_ERR_OUT_OF_MEMORY      \
                                                               : GCAM_ERR_HIP,               \
                                     std::string(#expr) + ": " + hipGetErrorString(_e));     \
    } while (0)
#define GCAM_LAUNCH_CHECK(what)                                                              \
    do {                                                                                     \
        hipError_t _e = hipGetLastError();                                                   \
        if (_e != hipSuccess)                                                                \
            return ::gcam::set_error(GCAM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(_e)); \
    } while (0)

// ---- launchers (stream-ordered; all return GCAM_OK or set the error) ----
gcam_status launch_downsample4(View<const float4> in, View<float4> out, hipStream_t s);
gcam_status launch_downsample2(View<const float2> in, View<float2> out, hipStream_t s);
// nlf_level: device noise model of this level (or nullptr to use nlf_x/nlf_y as given);
// when set, nlf = (a.x * thr_y, b.x * thr_y).
gcam_status launch_subtract_noise(View<const float4> in, View<const float4> in1, View<const float4> den1,
                                  View<const float2> grad, float luma_weight, float sharpening, float nlf_x,
                                  float nlf_y, const NlfDev* nlf_level, float thr_y, View<float4> out,
                                  hipStream_t s);
// The subtract of launch_subtract_noise and the PCA moments of its output in
// one pass: writes out and *nblocks per-workgroup moment partials of the
// (w / 8) x (h / 8) patches (the 5 x 5 luma around pixel (8 px, 8 py),
// clamp-to-edge), the sample collectPatches would take; the sample itself is
// never materialized.
gcam_status launch_subtract_moments(View<const float4> in, View<const float4> in1, View<const float4> den1,
                                    View<const float2> grad, float luma_weight, float sharpening, float nlf_x,
                                    float nlf_y, const NlfDev* nlf_level, float thr_y, View<float4> out,
                                    double* partials, int* nblocks, hipStream_t s);
gcam_status launch_noise_stats(View<const float4> in, View<float4> out, hipStream_t s);
gcam_status launch_nlf_fit(View<const float4> stats, float exposure_multiplier, double* partials,
                           NlfFitState* state, NlfDev* nlf_out, hipStream_t s);
gcam_status launch_set_nlf(const gcam_ycbcr_nlf* host_levels, int levels, NlfDev* dev, hipStream_t s);

gcam_status launch_collect_patches(View<const float4> in, float* patches, int64_t n, hipStream_t s);
// Jacobi stops when the off-diagonal energy is below this fraction of the
// total: off-diagonal entries ~1e-11 of the matrix norm, so the leading
// eigenvectors are accurate to ~1e-11 / (relative eigenvalue gap), far below
// the FP32 precision of the features they produce.
constexpr double kJacobiTol = 1e-22;
// Moments of the patches sampled from an image (never materialized); *nblocks
// receives the number of per-workgroup partials written.
gcam_status launch_image_moments(View<const float4> in, double* partials, int* nblocks, hipStream_t s);
// Moments of a device patch buffer (n x 25).
gcam_status launch_buffer_moments(const float* patches, int64_t n, double* partials, int* nblocks, hipStream_t s);
// Reduce the partials, form the covariance, eigensolve, write basis (25 x 8 f32) and eigenvalues.
// tri (nullable, 2 x 25 doubles): bracket only the 8 leading eigenvalues,
// leave eigvals[8..24] NaN and keep the tridiagonal form there for
// launch_tridiag_eigvals (the basis needs only the 8 leading eigenpairs).
gcam_status launch_pca_solve(double* partials, int nblocks, int64_t n, float* basis, double* eigvals,
                             double* eig_full, double* tri, hipStream_t s);
// Completes eigvals[8..24] of a solve that kept its tridiagonal form in tri
// (no-op when they are already set).
gcam_status launch_tridiag_eigvals(const double* tri, double* eigvals, hipStream_t s);
gcam_status launch_pca_project(View<const float4> in, const float* basis, View<float> feat8, hipStream_t s);

// Default output rows per thread of block matching (rows_per_thread == 0):
// two (32 x 16 output tiles) while that still gives at least 1024 workgroups,
// else one (32 x 8 tiles: twice the workgroups for the small coarse levels);
// one for radii above 6.
inline int default_bm_rows(int radius, int w, int h) {
    if (radius > 6) return 1;
    const long tiles2 = (long)((w + 31) / 32) * ((h + 15) / 16);
    return tiles2 >= 1024 ? 2 : 1;
}

gcam_status launch_block_match(View<const float4> in, View<const float2> grad, View<const float> feat8,
                               const NlfDev* nlf_level, const BmParams& p, int radius, int rows_per_thread,
                               View<float4> out, hipStream_t s);

// Precision form: projection fused in (features never materialized), basis = 25 x 8 f32 on the device.
gcam_status launch_block_match_lp(View<const float4> in, View<const float2> grad, const float* basis,
                                  const NlfDev* nlf_level, const BmParams& p, int radius, int rows_per_thread,
                                  int precision, View<float4> out, hipStream_t s);

// denoiseImage, the non-PCA bilateral (usePatchSimiliarity == false).
gcam_status launch_denoise_image(View<const float4> in, View<const float2> grad, const NlfDev* nlf_level,
                                 const BmParams& p, View<float4> out, hipStream_t s);

gcam_status launch_synth(int w, int h, const gcam_synth_params& p, uint16_t* raw, View<float4> ycc,
                         View<float4> clean, hipStream_t s);

// ---- reverse pass (differentiation form) ----
// Per level, the full eigen-decomposition kept by the forward: V (25 x 25,
// column = eigenvector, descending) at 0, eigenvalues at 625, the 8 basis
// orientations at 650, the mean patch at 658.
constexpr int kEigStride = 704;

struct AdjointLevel {
    View<const float4> S, D, gD;  // layer, denoised, dL/d(denoised) of this level
    const float* feat;            // features of S (8 floats per pixel, packed)
    const float* basis;           // 25 x 8
    const double* eig;            // kEigStride doubles
    const NlfDev* nlf;
    float thr[3], chromaBoost;
    int radius;
    double n;                     // sampled patch slots, w*h/64
    float4* coef;                 // 3 float4 per pixel (scratch)
    float4* gS;                   // dL/dS (Y, Cb, Cr)
    float4* gF;                   // dL/dF (2 float4 per pixel)
    float* gthr_part;             // adjoint_thr_blocks(w, h) x 3
    float* gthr;                  // 3 floats: dL/dthr of the level
    double* gB_part;              // kBasisGradBlocks x 200
    double* gC;                   // 625
    float4* gD1;                  // dL/d(denoised) of level i+1, or nullptr at the coarsest level
    int w1, h1;
};

int adjoint_thr_blocks(int w, int h);
gcam_status launch_level_backward(const AdjointLevel& level, hipStream_t s);

}  // namespace gcam
