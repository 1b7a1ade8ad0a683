// PCA-space block-matching denoise (FP32 form) for gfx950.
//
//   blockMatchingDenoiseImage  MetalPipeline/demosaic.metal:806-884
//   gaussian / lensShading     MetalPipeline/demosaic.metal:87-93
//
// For every pixel p and every neighbour q of its (2R+1)^2 window:
//   d2 = |f_q - f_p|^2 (8 PCA features)
//   wl = exp(-2 d2 (dm.x / (1 + gB edge))^2)                 lumaWeight
//   wc = exp(-2 (wl^2 + |(C_q - C_p) dm.yz / cB|^2))          chromaWeight
//   wd = exp(-2 (0.1 |q - p|)^2)                              directionWeight
// and the output is sum wd (wl, wc, wc) YCC_q / sum wd (wl, wc, wc), with
// sum wd wl in .w.  The reference takes sqrt(d2) and squares it again; here
// the square root is skipped, and the exponentials run as single v_exp_f32
// (base 2) with log2(e) folded into the per-pixel coefficients.
//
// The work is VALU/transcendental-bound (~34 VALU + 2 v_exp per pixel pair),
// so the kernel stages the tile and its R-pixel apron in LDS once (YCC and the
// two halves of the features as three float4 planes, conflict-free 16-byte
// reads) and runs the window fully unrolled with the direction weights as
// kernel-argument constants (scalar operands).
#include "gcam_internal.hpp"

#include <cmath>

namespace gcam {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// directionWeight per squared offset |q - p|^2 (0 .. 2 R^2).
struct DirTable {
    float w[2 * GCAM_MAX_SEARCH_RADIUS * GCAM_MAX_SEARCH_RADIUS + 1];
};

constexpr float kLog2e = 1.4426950408889634f;

template <int R>
struct BmTile {
    static constexpr int TW = 32;
    static constexpr int TH = R > 6 ? 4 : 8;
    static constexpr int LW = TW + 2 * R;
    static constexpr int LH = TH + 2 * R;
    static constexpr int LN = LW * LH;
};

template <int R>
__global__ __launch_bounds__(256) void k_block_match(View<const float4> in, View<const float2> grad,
                                                     const float* __restrict__ feat, int feat_stride,
                                                     const NlfDev* __restrict__ nlf, BmParams p, DirTable dir,
                                                     View<float4> out) {
    using T = BmTile<R>;
    __shared__ float4 sY[T::LN];   // Y, Cb, Cr, (w)
    __shared__ float4 sF0[T::LN];  // features 0..3
    __shared__ float4 sF1[T::LN];  // features 4..7
    const int x0 = blockIdx.x * T::TW, y0 = blockIdx.y * T::TH;
    const int tid = threadIdx.y * T::TW + threadIdx.x;
    for (int e = tid; e < T::LN; e += T::TW * T::TH) {
        const int ly = e / T::LW, lx = e % T::LW;
        const int gx = clampi(x0 + lx - R, 0, in.w - 1), gy = clampi(y0 + ly - R, 0, in.h - 1);
        sY[e] = in.at(gx, gy);
        const float4* f = reinterpret_cast<const float4*>(feat + ((size_t)gy * feat_stride + gx) * 8);
        sF0[e] = f[0];
        sF1[e] = f[1];
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= out.w || y >= out.h) return;

    const int c = (threadIdx.y + R) * T::LW + threadIdx.x + R;
    const float4 ycc = sY[c];
    const float4 fa = sF0[c], fb = sF1[c];

    float lens = 1.0f;
    if (p.lensShading > 0.0f) {
        const float cx = (float)(in.w / 2), cy = (float)(in.h / 2);
        const float dx = (float)x - cx, dy = (float)y - cy;
        const float d = sqrtf(dx * dx + dy * dy) / sqrtf(cx * cx + cy * cy);
        lens = 1.0f + p.lensShading * d * d;
    }
    const float s0 = sqrtf(nlf->a[0] + nlf->b[0] * lens * ycc.x);
    const float s1 = sqrtf(nlf->a[1] + nlf->b[1] * lens * ycc.x);
    const float s2 = sqrtf(nlf->a[2] + nlf->b[2] * lens * ycc.x);
    const float dm0 = 1.0f / (p.thr[0] * s0), dm1 = 1.0f / (p.thr[1] * s1), dm2 = 1.0f / (p.thr[2] * s2);
    float mag = 0.0f;
    if (grad.ptr) {
        const float2 g = grad.at(x, y);
        mag = sqrtf(g.x * g.x + g.y * g.y);
    }
    const float edge = smoothstepf(2.0f, 16.0f, p.gradientThreshold * mag / s0);
    const float lm = dm0 / (1.0f + p.gradientBoost * edge);
    const float kl = -2.0f * kLog2e * lm * lm;  // wl = 2^(kl d2)
    const float cb1 = dm1 / p.chromaBoost, cb2 = dm2 / p.chromaBoost;
    const float kb = -2.0f * kLog2e * cb1 * cb1, kr = -2.0f * kLog2e * cb2 * cb2, kw = -2.0f * kLog2e;

    float fY = 0.0f, fCb = 0.0f, fCr = 0.0f, nY = 0.0f, nC = 0.0f;
#pragma unroll
    for (int dy = -R; dy <= R; dy++) {
#pragma unroll
        for (int dx = -R; dx <= R; dx++) {
            const int q = c + dy * T::LW + dx;
            const float4 qa = sF0[q], qb = sF1[q], qy = sY[q];
            float t, d2;
            t = qa.x - fa.x; d2 = t * t;
            t = qa.y - fa.y; d2 = fmaf(t, t, d2);
            t = qa.z - fa.z; d2 = fmaf(t, t, d2);
            t = qa.w - fa.w; d2 = fmaf(t, t, d2);
            t = qb.x - fb.x; d2 = fmaf(t, t, d2);
            t = qb.y - fb.y; d2 = fmaf(t, t, d2);
            t = qb.z - fb.z; d2 = fmaf(t, t, d2);
            t = qb.w - fb.w; d2 = fmaf(t, t, d2);
            const float wl = __builtin_amdgcn_exp2f(kl * d2);
            const float db = qy.y - ycc.y, dr = qy.z - ycc.z;
            const float e = fmaf(db * kb, db, fmaf(dr * kr, dr, kw * wl * wl));
            const float wc = __builtin_amdgcn_exp2f(e);
            const float wd = dir.w[dx * dx + dy * dy];
            const float wY = wd * wl, wC = wd * wc;
            fY = fmaf(wY, qy.x, fY);
            fCb = fmaf(wC, qy.y, fCb);
            fCr = fmaf(wC, qy.z, fCr);
            nY += wY;
            nC += wC;
        }
    }
    out.at(x, y) = make_float4(fY / nY, fCb / nC, fCr / nC, nY);
}

template <int R>
gcam_status launch_r(View<const float4> in, View<const float2> grad, View<const float> feat8, const NlfDev* nlf,
                     const BmParams& p, const DirTable& dir, View<float4> out, hipStream_t s) {
    using T = BmTile<R>;
    dim3 grid((out.w + T::TW - 1) / T::TW, (out.h + T::TH - 1) / T::TH);
    k_block_match<R><<<grid, dim3(T::TW, T::TH), 0, s>>>(in, grad, feat8.ptr, feat8.stride, nlf, p, dir, out);
    GCAM_LAUNCH_CHECK("blockMatchingDenoiseImage");
    return GCAM_OK;
}

DirTable make_dir_table() {
    // demosaic.metal:869: gaussian(0.1 * length(float2(x, y))), FP32.
    DirTable t{};
    for (int d2 = 0; d2 < (int)(sizeof(t.w) / sizeof(float)); d2++) {
        const float a = 0.1f * std::sqrt((float)d2);
        t.w[d2] = std::exp(-2.0f * a * a);
    }
    return t;
}

}  // namespace

gcam_status launch_block_match(View<const float4> in, View<const float2> grad, View<const float> feat8,
                               const NlfDev* nlf_level, const BmParams& p, int radius, View<float4> out,
                               hipStream_t s) {
    static const DirTable dir = make_dir_table();
    switch (radius) {
        case 1: return launch_r<1>(in, grad, feat8, nlf_level, p, dir, out, s);
        case 2: return launch_r<2>(in, grad, feat8, nlf_level, p, dir, out, s);
        case 3: return launch_r<3>(in, grad, feat8, nlf_level, p, dir, out, s);
        case 4: return launch_r<4>(in, grad, feat8, nlf_level, p, dir, out, s);
        case 5: return launch_r<5>(in, grad, feat8, nlf_level, p, dir, out, s);
        case 6: return launch_r<6>(in, grad, feat8, nlf_level, p, dir, out, s);
        case 7: return launch_r<7>(in, grad, feat8, nlf_level, p, dir, out, s);
        case 8: return launch_r<8>(in, grad, feat8, nlf_level, p, dir, out, s);
        case 9: return launch_r<9>(in, grad, feat8, nlf_level, p, dir, out, s);
        case 10: return launch_r<10>(in, grad, feat8, nlf_level, p, dir, out, s);
        default: return set_error(GCAM_ERR_INVALID_ARGUMENT, "search radius must be 1..10");
    }
}

}  // namespace gcam
