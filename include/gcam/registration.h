/*
 * gcam registration -- C ABI of burst registration and fusion on MI355X
 * (gfx950): the homography warps of ImageRegistration/SURF.metal:489-559
 * behind their kernel wrappers (MetalPipeline/demosaic_kernels.hpp:690-749),
 * and the robust homography fit of ImageRegistration/Homography.cpp:72-116
 * (RANSAC, RTL/RANSAC.hpp:53-118, then a least-squares refit on the inliers,
 * LeastSquaresHomography.cpp:356-421), as used by the burst driver
 * PyramidFusion/pyramidFusion.cpp:287-353.
 *
 * Same conventions as gcam/denoise.h: device images, status codes,
 * gcam_last_error() for the message.  A homography is a row-major 3x3 float
 * matrix (gls::Matrix<3, 3>) that maps output coordinates to input
 * coordinates.  Matches are device arrays of count x 4 floats
 * (x1, y1, x2, y2), the layout of std::pair<Point2f, Point2f>, 16-byte
 * aligned; a fitted homography maps (x1, y1) to (x2, y2).
 * Paths are relative to the GlassCameraPublic repository.
 */
#ifndef GCAM_REGISTRATION_H
#define GCAM_REGISTRATION_H

#include "gcam/denoise.h"

#ifdef __cplusplus
extern "C" {
#endif

/* RegisterAndFuseKernel (demosaic_kernels.hpp:690-710, registerAndFuse
 * SURF.metal:493-516): out(x, y) = ((count - 1) fused(x, y) + input(H(x, y))) / count
 * over fused's size; input is sampled bilinearly (clamp to edge) at the
 * warped position, its normalised coordinate taken with out's size as the
 * reference does.  RGBA32F; out may be fused (in place); count >= 1.
 * Stream-ordered. */
gcam_status gcam_register_and_fuse_image(const gcam_image* fused, const gcam_image* input, gcam_image* out,
                                         const float homography[9], int32_t count, void* stream);

/* RegisterImageKernel<T> (demosaic_kernels.hpp:712-730, registerImage
 * SURF.metal:518-538): out(x, y) = input(H(x, y)), bilinear, clamp to edge.
 * R32F or RGBA32F, input and out of one size and format.  Stream-ordered. */
gcam_status gcam_register_image(const gcam_image* input, gcam_image* out, const float homography[9], void* stream);

/* RegisterBayerImageKernel (demosaic_kernels.hpp:732-749, registerBayerImage
 * SURF.metal:540-559): nearest-neighbour warp of each colour plane of a raw
 * frame, out(x, y) = input(2 round(H(x / 2, y / 2)) + (x % 2, y % 2)); reads
 * outside the frame clamp to its edge.  R16U, one size.  Stream-ordered. */
gcam_status gcam_register_bayer_image(const gcam_image* input, gcam_image* out, const float homography[9],
                                      void* stream);

/* FindLeastSquaresHomography (LeastSquaresHomography.cpp:356-421): the
 * normalised DLT fit of all count >= 4 matches, computed on the device
 * (FP64).  Fails with GCAM_ERR_INVALID_ARGUMENT where the reference throws
 * std::range_error (a point set with no spread in x or y).  Writes the host
 * array homography and synchronizes the stream. */
gcam_status gcam_find_least_squares_homography(const float* matches, int32_t count, float homography[9],
                                               void* stream);

/* FindHomography (Homography.cpp:72-116): max_iterations RANSAC hypotheses
 * (minimal samples of 4 distinct matches, generated by
 * gcam_ransac_samples from seed), each scored by the number of matches with
 * squared transfer error > threshold; the first hypothesis with the fewest
 * such matches wins; its inliers (error < threshold, ascending indices) are
 * refitted by least squares when there are at least 4 (the RANSAC model is
 * kept where that fit fails).  All of it runs on the device.  Writes
 * homography (host), *inlier_count (host) and, when inlier_indices is not
 * NULL, the inlier indices (host array of count entries), and synchronizes
 * the stream.  count >= 4 (the reference's sampler never ends below 4). */
gcam_status gcam_find_homography(const float* matches, int32_t count, float threshold, int32_t max_iterations,
                                 uint64_t seed, float homography[9], int32_t* inlier_indices, int32_t* inlier_count,
                                 void* stream);

/* The minimal samples gcam_find_homography draws: for hypotheses
 * first .. first + n - 1, 4 distinct ascending indices in [0, count) each
 * (Floyd's subset sampling driven by SplitMix64 of seed, the hypothesis and
 * the draw), written to out (n x 4).  Host only: the reference draws with
 * std::mt19937 (RTL/RANSAC.hpp:94-98), whose distribution code differs
 * between C++ libraries, so this generator is the build's own and exported
 * for reproducibility. */
gcam_status gcam_ransac_samples(uint64_t seed, int32_t count, int32_t first, int32_t n, int32_t* out);

/* ScaleHomography (Homography.hpp:29-37): S H S^-1 with S = diag(scale, scale, 1).
 * The reference builds S with 2 whatever scale is; its one call passes 2.
 * Host only. */
void gcam_scale_homography(const float homography[9], float scale, float out[9]);

#ifdef __cplusplus
}
#endif

#endif /* GCAM_REGISTRATION_H */
