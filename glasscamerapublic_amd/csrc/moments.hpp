// FP64 patch-moment helpers shared by the moment kernels and the fused
// subtract-and-moments pass.
//
// Moments of a set of 5 x 5 luma patches x (PCA.cpp:28-32 needs sum x and
// sum x x^T): kMoments = 350 entries, 0..24 the sum x_r, then the upper
// triangle r <= s of sum x_r x_s row by row.
#pragma once

#include "gcam_internal.hpp"

namespace gcam {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Row/column of moment entry e: entries 0..24 are sum x_r (r = e, s = -1),
// entries 25.. enumerate the upper triangle r <= s row by row.
// Row r of triangle entry t = e - 25 is the largest r with start(r) <= t,
// start(r) = 25 r - r (r - 1) / 2 (the entries of rows 0 .. r - 1): the root
// of the quadratic, corrected by one step either way.
__device__ __forceinline__ void moment_rs(int e, int& r, int& s) {
    if (e < kPatch) {
        r = e;
        s = -1;
        return;
    }
    const int t = e - kPatch;
    auto start = [](int q) { return kPatch * q - q * (q - 1) / 2; };
    int q = (int)((51.0f - sqrtf(2601.0f - 8.0f * (float)t)) * 0.5f);
    q = q < 0 ? 0 : (q > kPatch - 1 ? kPatch - 1 : q);
    if (q < kPatch - 1 && start(q + 1) <= t) q++;
    if (start(q) > t) q--;
    r = q;
    s = q + (t - start(q));
}

// One group of 4 patches into the Gram accumulators: X is the patch x 32
// matrix (25 luma values, 7 zero columns); lane l holds lo = X[4 g + l / 16]
// [l % 16] and hi = X[4 g + l / 16][16 + l % 16], both the A (16 x 4) and the
// B (4 x 16) operand of v_mfma_f64_16x16x4f64, in the three 16 x 16 tiles
// (00, 01, 11; 10 is the transpose of 01) of X^T X; sum x accumulates on the
// vector ALU.
__device__ __forceinline__ void moments_group(double lo, double hi, f64x4& c00, f64x4& c01, f64x4& c11, double& slo,
                                              double& shi) {
    slo += lo;
    shi += hi;
    c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(lo, lo, c00, 0, 0, 0);
    c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(lo, hi, c01, 0, 0, 0);
    c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(hi, hi, c11, 0, 0, 0);
}

// The accumulators of NW wavefronts into one partial (kMoments doubles): the
// result tile holds C[l / 16 + 4 r][l % 16] in register r of lane l (measured
// on gfx950; unlike the FP32 16 x 16 forms, rows interleave by 4).  Wavefront
// w < NW stores its tiles, then every thread of the workgroup (nthreads)
// adds the wavefronts' entries in index order (deterministic).  gram and sums
// are LDS arrays of the caller.
template <int NW>
__device__ __forceinline__ void moments_store(int tid, int nthreads, const f64x4& c00, const f64x4& c01,
                                              const f64x4& c11, double slo, double shi, double (*gram)[32][33],
                                              double (*sums)[64][2], double* __restrict__ partial) {
    const int lane = tid & 63, wave = tid >> 6;
    const int f = lane & 15, kq = lane >> 4;
    if (wave < NW) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = kq + 4 * r;
            gram[wave][i][f] = c00[r];
            gram[wave][i][16 + f] = c01[r];
            gram[wave][16 + i][16 + f] = c11[r];
        }
        sums[wave][lane][0] = slo;
        sums[wave][lane][1] = shi;
    }
    __syncthreads();
    for (int e = tid; e < kMoments; e += nthreads) {
        int r, s;
        moment_rs(e, r, s);
        double acc = 0.0;
        if (s < 0) {
            for (int w = 0; w < NW; w++)
                for (int q = 0; q < 4; q++) acc += sums[w][16 * q + (r & 15)][r >> 4];
        } else {
            for (int w = 0; w < NW; w++) acc += gram[w][r][s];
        }
        partial[e] = acc;
    }
}

// Fused subtract-and-moments pass: 64 x 8-pixel tiles anchored at
// (64 tx - 2, 8 ty - 2), so the 5 x 5 patches (8 tx + p, ty), p = 0..7, lie
// inside their tile (rows 0..4, columns 8 p .. 8 p + 4).
__host__ __device__ inline int submom_tiles_x(int w) { return (w + 2 + 63) / 64; }
__host__ __device__ inline int submom_tiles(int w, int h) { return submom_tiles_x(w) * ((h + 2 + 7) / 8); }
// Workgroups (= moment partials) of the fused pass over a w x h level: one
// per tile up to kMomentBlocks, and at least 8 (one per XCD tile range).
inline int submom_blocks(int w, int h) {
    const int t = submom_tiles(w, h);
    return t < 8 ? 8 : (t > kMomentBlocks ? kMomentBlocks : t);
}

}  // namespace gcam
