// denoiseImage: the non-PCA bilateral of the reference (demosaic.metal:664-719),
// used by PyramidProcessor::denoise when usePatchSimiliarity is false
// (pyramid_processor.hpp:27, pyramid_processor.cpp:162-168).
//
// Its luma and chroma weights are step functions of the scaled differences,
// so one rounding difference can flip a weight between 0 and 1.  Contraction
// into FMA is therefore off in this file, and the comparison inputs (sigma,
// the scaled differences, their lengths, the edge term) are formed with the
// same IEEE operations, in the same order, as the reference's expressions.
#include "gcam_internal.hpp"

#pragma clang fp contract(off)

namespace gcam {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return t * t * (3.0f - 2.0f * t);
}

template <int SIZE>
__device__ __forceinline__ void bilateral_window(View<const float4> in, View<const float2> grad, int px, int py,
                                                 const float4 ip, float gx, float gy, const float dm[3], float sigma0,
                                                 float ca, float sa, float edge, float lumaLimit, float chromaBoost,
                                                 float fil[3], float norm[3]) {
    for (int y = -SIZE; y <= SIZE; y++) {
        const int sy = clampi(py + y, 0, in.h - 1);
        for (int x = -SIZE; x <= SIZE; x++) {
            const int sx = clampi(px + x, 0, in.w - 1);
            const float4 s = in.at(sx, sy);
            float sgx = 0.0f, sgy = 0.0f;
            if (grad.ptr) {
                const float2 g = grad.at(sx, sy);
                sgx = g.x;
                sgy = g.y;
            }
            const float d0 = (s.x - ip.x) * dm[0], d1 = (s.y - ip.y) * dm[1], d2 = (s.z - ip.z) * dm[2];
            const float gdx = (sgx - gx) / sigma0, gdy = (sgy - gy) / sigma0;
            const float a = (float)x * ca + (float)y * sa;
            const float tunnel = expf(-(a * a) / 0.25f);
            const float directionWeight = 1.0f + (tunnel - 1.0f) * edge;
            const float gradientWeight = 1.0f - smoothstepf(2.0f, 8.0f, sqrtf(gdx * gdx + gdy * gdy));
            const float lumaWeight = fabsf(d0) < lumaLimit ? 1.0f : 0.0f;
            const float chromaWeight = sqrtf(d0 * d0 + d1 * d1 + d2 * d2) < chromaBoost ? 1.0f : 0.0f;
            const float wY = directionWeight * gradientWeight * lumaWeight;
            fil[0] += wY * s.x;
            fil[1] += chromaWeight * s.y;
            fil[2] += chromaWeight * s.z;
            norm[0] += wY;
            norm[1] += chromaWeight;
            norm[2] += chromaWeight;
        }
    }
}

__global__ __launch_bounds__(256) void k_denoise_image(View<const float4> in, View<const float2> grad,
                                                       const NlfDev* __restrict__ nlf, BmParams p, View<float4> out) {
    const int px = blockIdx.x * 64 + threadIdx.x, py = blockIdx.y * 4 + threadIdx.y;
    if (px >= out.w || py >= out.h) return;
    const float4 ip = in.at(px, py);
    float sigma[3], dm[3];
    for (int c = 0; c < 3; c++) {
        sigma[c] = sqrtf(nlf->a[c] + nlf->b[c] * ip.x);
        dm[c] = 1.0f / (p.thr[c] * sigma[c]);
    }
    float gx = 0.0f, gy = 0.0f;
    if (grad.ptr) {
        const float2 g = grad.at(px, py);
        gx = g.x;
        gy = g.y;
    }
    const float angle = atan2f(gy, gx);
    const float magnitude = sqrtf(gx * gx + gy * gy);
    const float edge = smoothstepf(4.0f, 16.0f, p.gradientThreshold * magnitude / sigma[0]);
    const float ca = cosf(angle), sa = sinf(angle);
    const float lumaLimit = 1.0f + p.gradientBoost * edge;
    float fil[3] = {0.0f, 0.0f, 0.0f}, norm[3] = {0.0f, 0.0f, 0.0f};
    if (p.gradientBoost > 0.0f)
        bilateral_window<4>(in, grad, px, py, ip, gx, gy, dm, sigma[0], ca, sa, edge, lumaLimit, p.chromaBoost, fil,
                            norm);
    else
        bilateral_window<2>(in, grad, px, py, ip, gx, gy, dm, sigma[0], ca, sa, edge, lumaLimit, p.chromaBoost, fil,
                            norm);
    out.at(px, py) = make_float4(fil[0] / norm[0], fil[1] / norm[1], fil[2] / norm[2], magnitude);
}

}  // namespace

gcam_status launch_denoise_image(View<const float4> in, View<const float2> grad, const NlfDev* nlf_level,
                                 const BmParams& p, View<float4> out, hipStream_t s) {
    k_denoise_image<<<dim3((out.w + 63) / 64, (out.h + 3) / 4), dim3(64, 4), 0, s>>>(in, grad, nlf_level, p, out);
    GCAM_LAUNCH_CHECK("denoiseImage");
    return GCAM_OK;
}

}  // namespace gcam
