// CPU oracle for burst registration and fusion: the homography warps of
// ImageRegistration/SURF.metal:489-559, FindLeastSquaresHomography
// (ImageRegistration/LeastSquaresHomography.cpp:189-421) and FindHomography
// (ImageRegistration/Homography.cpp:72-116 over RTL/RANSAC.hpp:53-118).
//
// TEST INFRASTRUCTURE ONLY: checks the HIP registration path; the product
// library never links, loads or calls it.  Images are row-major, tightly
// packed; matches are n x 4 floats (x1, y1, x2, y2); homographies are
// row-major 3 x 3 floats.
#pragma once
#include <cstdint>

extern "C" {

// registerAndFuse / registerImage (channels 1 or 4): output (ow x oh), input (iw x ih).
void oracle_register_and_fuse(const float* fused4, const float* in4, int ow, int oh, int iw, int ih, const float* H,
                              int count, float* out4);
void oracle_register_image(const float* in, int channels, int ow, int oh, int iw, int ih, const float* H, float* out);
void oracle_register_bayer(const uint16_t* in, int w, int h, const float* H, uint16_t* out);

// Symmetric 9 x 9 eigensolve of LeastSquaresHomography.cpp:216-349:
// eigenvalues descending in w, eigenvectors as the rows of v.
void oracle_jacobi9(const double* a, double* w, double* v);
// FindLeastSquaresHomography on n matches; returns 0, or 1 where the reference
// throws std::range_error (H untouched).
int oracle_least_squares_homography(const float* matches, int n, float* H);
// The minimal samples of gcam_ransac_samples.
void oracle_ransac_samples(uint64_t seed, int count, int first, int n, int* out);
// FindHomography with those samples; writes H, the inliers (ascending) and
// their count; best_out receives the winning hypothesis and its outlier count.
void oracle_find_homography(const float* matches, int n, float threshold, int iters, uint64_t seed, float* H,
                            int* inliers, int* inlier_count, int* best_out);
}
