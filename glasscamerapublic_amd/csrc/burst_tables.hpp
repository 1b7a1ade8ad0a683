// Device-side frame tables of the batched (burst) denoise.
//
// Every stage of the pyramid denoise runs once per level for all frames of a
// burst: a launch's workgroups are numbered across the frames, and a
// workgroup finds its frame in the stage's table (`block0`/`tile0` hold the
// first workgroup or tile of each frame, ascending).  The tables live in one
// device buffer that a burst call refreshes before it launches its graph, so
// the captured kernels never change: frames, noise models and parameters
// reach them through the tables.
#pragma once

#include "gcam_internal.hpp"

namespace gcam {

constexpr int kMaxBurst = GCAM_MAX_BURST;

// 64 x 4-pixel workgroups over an output image of w x h: bx = ceil(w / 64).
struct BurstDownFrame {
    View<const float4> in;
    View<float4> out;
    View<const float2> gin;  // gradient (ptr == nullptr: none)
    View<float2> gout;
    int bx, block0;
};

struct BurstSubFrame {
    View<const float4> in, in1, den1;
    View<const float2> grad;
    const NlfDev* nlf;  // this frame's noise model of the level
    float sharpening, thr_y;
    View<float4> out;
    double* partials;  // this frame's moment partials (nblk x kMoments)
    int nblk, block0;  // workgroups of the fused subtract-and-moments pass; block0 a multiple of 8
};

struct BurstMomFrame {
    View<const float4> in;  // the level the moment kernel samples (top level)
    double* partials;  // kMomentSlots x kMoments of this frame
    int nblocks, block0;
    double n;          // patch slots, w * h / 64
    float* basis;      // 25 x 8
    double* eigvals;   // 25
    double* tri;       // 50: the tridiagonal form (eigenvalues 9..25 on demand)
};

struct BurstBmFrame {
    View<const float4> in;
    View<const float2> grad;
    const float* basis;
    const NlfDev* nlf;
    BmParams p;
    View<float4> out;
    int tiles_x, tile0;
};

// Direction weights of the precision-form block matching, indexed by the
// squared offset |q - p|^2 (0 .. 2 R^2): w = the weight divided by
// sqrt(2 log2 e) (the luma weight enters pre-scaled by that factor), lw =
// log2 of the weight.  The burst reads it from its device table buffer.
struct BmDirTable {
    float w[2 * GCAM_MAX_SEARCH_RADIUS * GCAM_MAX_SEARCH_RADIUS + 1];
    float lw[2 * GCAM_MAX_SEARCH_RADIUS * GCAM_MAX_SEARCH_RADIUS + 1];
};
BmDirTable make_bm_dir_table();

template <class F>
struct BurstTable {
    int nframes;
    int total;  // workgroups (or tiles) over all frames
    F frame[kMaxBurst];
};

// Frame of workgroup (or tile) `idx`: the last frame whose first index is <= idx.
template <class F>
__device__ __forceinline__ int burst_frame_of(const BurstTable<F>* t, int idx) {
    int lo = 0, hi = t->nframes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t->frame[mid].block0 <= idx) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return lo;
}
__device__ __forceinline__ int burst_tile_frame(const BurstTable<BurstBmFrame>* t, int idx) {
    int lo = 0, hi = t->nframes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t->frame[mid].tile0 <= idx) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return lo;
}

// Launch configuration of the persistent block-matching kernel.
struct BurstBmConfig {
    int rows;      // output rows per thread: 1, 2 or 4
    int grid;      // workgroups; <= 0: one workgroup per tile
    int schedule;  // 0: tiles in frame order, strided over the grid; 1: XCD-chunked
    int tile_w;    // output tile width: 32, or 64 (two rows per thread, radius <= 6)
};

gcam_status launch_burst_downsample(const BurstTable<BurstDownFrame>* dev, int total_blocks, hipStream_t s);
// The fused subtract-and-moments pass of every frame of a level.
gcam_status launch_burst_subtract_moments(const BurstTable<BurstSubFrame>* dev, int total_blocks, hipStream_t s);
// Moments, their reduction and the eigensolve of every frame of a level.
gcam_status launch_burst_pca(const BurstTable<BurstMomFrame>* dev, int total_blocks, int nframes, hipStream_t s);
// The reduction and eigensolve only, of partials written by the fused subtract-and-moments pass.
gcam_status launch_burst_pca_solve(const BurstTable<BurstMomFrame>* dev, int nframes, hipStream_t s);
// Workgroups of the moment reduction for npatch patches (1 .. kMomentBlocks).
int moment_block_count(int64_t npatch);
// Output tile of the persistent kernel for `rows` rows per thread: tw x th pixels.
void burst_bm_tile(int radius, int rows, int tile_w, int precision, int* tw, int* th);
// dir: the direction table in device memory (make_bm_dir_table).
gcam_status launch_burst_block_match_lp(const BurstTable<BurstBmFrame>* dev, const BmDirTable* dir, int total_tiles,
                                        int radius, int precision, const BurstBmConfig& cfg, hipStream_t s);

}  // namespace gcam
