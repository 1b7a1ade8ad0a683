// Raw front end for gfx950: the steps of RawConverter::demosaic that produce
// the denoise's inputs, and the despeckle that RawConverter::denoise runs
// before the pyramid.
//
//   scaleRawData                  MetalPipeline/demosaic.metal:95-117
//   rawImageSobel                 MetalPipeline/demosaic.metal:120-139
//   sampledConvolutionSobel       MetalPipeline/demosaic.metal:170-210
//   interpolateGreen              MetalPipeline/demosaic.metal:214-300
//   interpolateRedBlue            MetalPipeline/demosaic.metal:309-416
//   interpolateRedBlueAtGreen     MetalPipeline/demosaic.metal:420-521
//   blendHighlightsImage          MetalPipeline/demosaic.metal:523-564
//   transformImage                MetalPipeline/demosaic.metal:649-662
//   despeckleLumaMedianChromaImage MetalPipeline/demosaic.metal:1160-1283
//   basicRawNoiseStatistics       MetalPipeline/demosaic.metal:1096-1136
//   MeasureRawNLF                 MetalPipeline/raw_converter.cpp:339-490
//   histogramImage, histogramStatistics MetalPipeline/demosaic.metal:1357-1453
//   GuidedFilterABImage, BoxFilterGFImage, localToneMappingMaskImage
//                                 MetalPipeline/demosaic.metal:1293-1355, 1456-1535
//   gaussianKernelBilinearWeights MetalPipeline/demosaic_utils.cpp:939-1006
//   cam_ycbcr                     MetalPipeline/demosaic_utils.cpp:703-720
//   orchestration                 MetalPipeline/raw_converter.cpp:148-214, 70-78
//
// The kernels are stencils over FP32 images, HBM-bound except the gradient
// blur (a 4-tap and, below the noise threshold, a 9 x 9 Gaussian of a
// 4-channel image per pixel: vector-ALU and LDS-bound).  The Sobel filter is
// fused into that blur (the Sobel tile of a workgroup lives in LDS and never
// reaches HBM), and the last demosaic pass, the highlight blend and the YCbCr
// transform run as one pass.  Arithmetic is FP32 evaluated in source order (no
// contraction): out-of-bounds reads clamp to the edge, the linear sampler
// interpolates at exact sub-texel positions, so every stage except the
// gradient direction (atan2f) and the wide blur reproduces a plain FP32
// evaluation of the formulas bit for bit.  The wide blur runs separably
// (SepKernel), equal to the reference's bilinear taps up to FP32 rounding.
#pragma clang fp contract(off)

#include "gcam_internal.hpp"
#include "gcam/front_end.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace gcam {

namespace {

// MetalPipeline/demosaic.metal:24-38: offsets of the red, green, blue and
// second green sample of a 2x2 quad, per BayerPattern (grbg, gbrg, rggb, bggr).
const int kBayerOffsets[4][4][2] = {
    {{1, 0}, {0, 0}, {0, 1}, {1, 1}},
    {{0, 1}, {0, 0}, {1, 0}, {1, 1}},
    {{0, 0}, {0, 1}, {1, 1}, {1, 0}},
    {{1, 1}, {0, 1}, {0, 0}, {1, 0}},
};
enum { kRed = 0, kGreen = 1, kBlue = 2, kGreen2 = 3 };

struct Quad {
    int2 o[4];  // red, green, blue, green2
};

Quad quad_of(int pattern) {
    Quad q;
    for (int c = 0; c < 4; c++) q.o[c] = make_int2(kBayerOffsets[pattern][c][0], kBayerOffsets[pattern][c][1]);
    return q;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <class P>
__device__ __forceinline__ P fetch(const View<const P>& v, int x, int y) {
    return v.at(clampi(x, 0, v.w - 1), clampi(y, 0, v.h - 1));
}

// An image region held in LDS: cell (x, y) of the image plane at
// p[(y - oy) * pitch + (x - ox)].  A cell outside the image holds the value at
// the clamped position, so fetch needs no clamp.
template <class P>
struct Tile {
    P* p;
    int ox, oy, pitch;
    __device__ __forceinline__ P& at(int x, int y) const { return p[(y - oy) * pitch + (x - ox)]; }
};
template <class P>
__device__ __forceinline__ P fetch(const Tile<P>& t, int x, int y) {
    return t.at(x, y);
}

// Metal's smoothstep / mix / clamp.
__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
__device__ __forceinline__ float mixf(float a, float b, float t) { return a + (b - a) * t; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

constexpr float kPi = 3.14159265358979323846f;

// ---- scaleRawData: one thread per 2x2 quad -------------------------------

__global__ __launch_bounds__(256) void k_scale_raw(View<const uint16_t> raw, View<float> out, Quad q, float4 sm,
                                                   float black, float lens_corr) {
    const int qx = blockIdx.x * blockDim.x + threadIdx.x, qy = blockIdx.y * blockDim.y + threadIdx.y;
    if (qx >= out.w / 2 || qy >= out.h / 2) return;
    const int x = 2 * qx, y = 2 * qy;
    float lens = 1.0f;
    if (lens_corr > 0.0f) {
        const float cx = (float)(out.w / 2), cy = (float)(out.h / 2);
        const float dx = (float)x - cx, dy = (float)y - cy;
        const float d = sqrtf(dx * dx + dy * dy) / sqrtf(cx * cx + cy * cy);
        lens = 1.0f + lens_corr * d * d;
    }
    const float s[4] = {sm.x, sm.y, sm.z, sm.w};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int px = x + q.o[c].x, py = y + q.o[c].y;
        const float v = (float)raw.at(px, py) / 65535.0f;
        out.at(px, py) = fmaxf(lens * s[c] * (v - black) * 0.9f + 0.1f, 0.0f);
    }
}

// ---- rawImageSobel -------------------------------------------------------

// (gx, gy, |gx|, |gy|) at (x, y) from a clamped reader of the scaled raw:
// gx = left - right and gy = top - bottom, 1-2-1 weighted, / sqrt(4.5),
// accumulated row by row, left to right (demosaic.metal:120-129).
template <class Raw>
__device__ __forceinline__ float4 sobel_at(const Raw& r, int x, int y) {
    float gx = 0.0f, gy = 0.0f;
#pragma unroll
    for (int j = -1; j <= 1; j++) {
#pragma unroll
        for (int i = -1; i <= 1; i++) {
            const float kx = (float)(-i * (j == 0 ? 2 : 1));
            const float ky = (float)(-j * (i == 0 ? 2 : 1));
            const float s = r(x + i, y + j);
            gx = gx + kx * s;
            gy = gy + ky * s;
        }
    }
    const float norm = sqrtf(4.5f);
    gx = gx / norm;
    gy = gy / norm;
    return make_float4(gx, gy, fabsf(gx), fabsf(gy));
}

__global__ __launch_bounds__(256) void k_raw_sobel(View<const float> in, View<float4> out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    out.at(x, y) = sobel_at([&](int u, int v) { return fetch(in, u, v); }, x, y);
}

// ---- gaussianBlurSobelImage (sampledConvolutionSobel) --------------------

// One bilinear tap: weight, integer offset and exact fractional part of the
// sample position relative to the output pixel.
struct Tap {
    float w, fx, fy;
    int ix, iy;
};
constexpr int kMaxTaps1 = 4, kMaxTaps2 = 36;  // radius <= 1.5 and <= 5.5
struct BlurTaps {
    int n1, n2;
    Tap t1[kMaxTaps1];
    Tap t2[kMaxTaps2];
};

// The wide kernel evaluated separably.  The bilinear taps of a k x k Gaussian
// (k = 2h + 1) cover the texel kernel G(dx, dy) = g(dx) g(dy), |dx|, |dy| <= h,
// except the corner tap, which the reference places at (h + 1/2, h + 1/2), so
// that G(h, h) is spread as G(h, h) / 4 over the texels {h, h + 1}^2.  Hence
// sum = sum_dy g(dy) sum_dx g(dx) S(dx, dy)
//     + G(h, h) (S(h+1, h) + S(h, h+1) + S(h+1, h+1) - 3 S(h, h)) / 4,
// divided by the taps' weight sum: 2h + 2 horizontal and 2h + 2 vertical
// weights per output instead of (h + 1)^2 bilinear taps of four texels.
// Equal to the tap evaluation up to FP32 rounding of the folded taps (~1e-7
// relative).
constexpr int kMaxSepHalf = 5;
struct SepKernel {
    int h;
    float g[kMaxSepHalf + 1];  // g(d), d = 0 .. h
    float gc;                  // G(h, h)
    float norm;                // sum of the bilinear taps' weights, in tap order
};

// Output tile of a workgroup and its apron: taps reach texels x - 5 .. x + 6
// for offsets in [-5, 5.5].  64 x 32 tiles of 512 threads (four rows each)
// measured fastest against 32-64 x 8-64 tiles of 1-8 rows per thread on the
// 12 MP frame: taller tiles recompute less Sobel apron.
constexpr int kBTW = 64, kBTH = 32, kBApL = 5, kBApR = 6;
constexpr int kBAW = kBTW + kBApL + kBApR, kBAH = kBTH + kBApL + kBApR;

// Rows of output a thread blurs: consecutive, so that the horizontal
// interpolation of a Sobel row under one tap is computed once and serves as
// the lower row of one output and the upper row of the next.
constexpr int kBRows = 4;

// Weighted sums over the taps of the bilinear samples of the LDS Sobel tile
// (demosaic.metal:170-186) for kBRows consecutive outputs starting at tile
// row ly: s = (T00 (1 - fx) + T10 fx) (1 - fy) + (T01 (1 - fx) + T11 fx) fy
// per channel, summed in tap order; the caller divides by `norm`.
__device__ __forceinline__ void blur_taps(const float4 (*sb)[kBAW], int lx, int ly, const Tap* taps, int n,
                                          float4 (&sum)[kBRows], float& norm) {
#pragma unroll
    for (int k = 0; k < kBRows; k++) sum[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    norm = 0.0f;
    for (int i = 0; i < n; i++) {
        const Tap t = taps[i];
        const float gx = 1.0f - t.fx, gy = 1.0f - t.fy;
        const int ax = lx + t.ix, ay = ly + t.iy;
        auto hrow = [&](int r) {
            const float4 a = sb[r][ax], b = sb[r][ax + 1];
            return make_float4(a.x * gx + b.x * t.fx, a.y * gx + b.y * t.fx, a.z * gx + b.z * t.fx,
                               a.w * gx + b.w * t.fx);
        };
        float4 h0 = hrow(ay);
#pragma unroll
        for (int k = 0; k < kBRows; k++) {
            const float4 h1 = hrow(ay + k + 1);
            sum[k].x = sum[k].x + t.w * (h0.x * gy + h1.x * t.fy);
            sum[k].y = sum[k].y + t.w * (h0.y * gy + h1.y * t.fy);
            sum[k].z = sum[k].z + t.w * (h0.z * gy + h1.z * t.fy);
            sum[k].w = sum[k].w + t.w * (h0.w * gy + h1.w * t.fy);
            h0 = h1;
        }
        norm = norm + t.w;
    }
}

// The separable wide kernel for kBRows consecutive outputs starting at tile
// row ly: each Sobel row's horizontal sum is formed once (in registers) and
// added into every output whose window holds the row.
template <int HW>
__device__ __forceinline__ void blur_sep(const float4 (*sb)[kBAW], int lx, int ly, const SepKernel& k,
                                         float4 (&sum)[kBRows]) {
#pragma unroll
    for (int r = 0; r < kBRows; r++) sum[r] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int j = -HW; j <= kBRows - 1 + HW; j++) {
        float4 hr = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int dx = -HW; dx <= HW; dx++) {
            const float g = k.g[dx < 0 ? -dx : dx];
            const float4 v = sb[ly + j][lx + dx];
            hr = make_float4(fmaf(g, v.x, hr.x), fmaf(g, v.y, hr.y), fmaf(g, v.z, hr.z), fmaf(g, v.w, hr.w));
        }
#pragma unroll
        for (int r = 0; r < kBRows; r++) {
            const int d = j - r;
            if (d < -HW || d > HW) continue;
            const float g = k.g[d < 0 ? -d : d];
            sum[r] = make_float4(fmaf(g, hr.x, sum[r].x), fmaf(g, hr.y, sum[r].y), fmaf(g, hr.z, sum[r].z),
                                 fmaf(g, hr.w, sum[r].w));
        }
    }
    const float c4 = 0.25f * k.gc;
#pragma unroll
    for (int r = 0; r < kBRows; r++) {
        const float4 a = sb[ly + r + HW][lx + HW], b = sb[ly + r + HW][lx + HW + 1];
        const float4 c = sb[ly + r + HW + 1][lx + HW], d = sb[ly + r + HW + 1][lx + HW + 1];
        sum[r].x = fmaf(c4, (b.x + c.x) + (d.x - 3.0f * a.x), sum[r].x);
        sum[r].y = fmaf(c4, (b.y + c.y) + (d.y - 3.0f * a.y), sum[r].y);
        sum[r].z = fmaf(c4, (b.z + c.z) + (d.z - 3.0f * a.z), sum[r].z);
        sum[r].w = fmaf(c4, (b.w + c.w) + (d.w - 3.0f * a.w), sum[r].w);
    }
}

// kFromRaw: the Sobel tile is computed here from the scaled raw (the fused
// front end); otherwise it is read from `sobel` (the per-operation entry
// point, like the reference's two-texture kernel).  A tile entry outside the
// image holds the Sobel value of the nearest image pixel, which is what the
// clamp-to-edge sampler reads there.  The narrow kernel, which decides
// between the two radii per pixel, is evaluated tap by tap in the reference's
// order; the wide one separably (HW = its half width).
constexpr int kBThreads = kBTW * (kBTH / kBRows);
static_assert(kBTH % kBRows == 0 && kBThreads <= 1024, "whole rows per thread, one workgroup");

template <bool kFromRaw, int HW>
__global__ __launch_bounds__(kBThreads) void k_blur_sobel(View<const float> raw, View<const float4> sobel,
                                                          BlurTaps taps, SepKernel wide_k, float var_a, float var_b,
                                                          View<float2> out) {
    __shared__ float4 sb[kBAH][kBAW];
    constexpr int RW = kBAW + 2, RH = kBAH + 2;
    __shared__ float rw[kFromRaw ? RH : 1][kFromRaw ? RW : 1];
    const int x0 = blockIdx.x * kBTW, y0 = blockIdx.y * kBTH;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x, nth = blockDim.x * blockDim.y;
    const int W = out.w, H = out.h;
    if (kFromRaw) {
        // raw over [x0 - 6, x0 + TW + 6] (clamped positions)
        for (int e = tid; e < RW * RH; e += nth) {
            const int ry = e / RW, rx = e % RW;
            rw[ry][rx] = fetch(raw, x0 - kBApL - 1 + rx, y0 - kBApL - 1 + ry);
        }
        __syncthreads();
    }
    for (int e = tid; e < kBAW * kBAH; e += nth) {
        const int ay = e / kBAW, ax = e % kBAW;
        const int cx = clampi(x0 - kBApL + ax, 0, W - 1), cy = clampi(y0 - kBApL + ay, 0, H - 1);
        if (kFromRaw) {
            // raw at the neighbours of (cx, cy): cell (u, v) of rw holds the raw at
            // the clamped position, and (cx +- 1, cy +- 1) lie in rw's region.
            sb[ay][ax] = sobel_at([&](int u, int v) { return rw[v - (y0 - kBApL - 1)][u - (x0 - kBApL - 1)]; },
                                  cx, cy);
        } else {
            sb[ay][ax] = sobel.at(cx, cy);
        }
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x >= W) return;
    const int r0 = threadIdx.y * kBRows, lx = threadIdx.x + kBApL;
    float4 nar[kBRows], wid[kBRows];
    float nn;
    blur_taps(sb, lx, r0 + kBApL, taps.t1, taps.n1, nar, nn);
    bool wide[kBRows], any = false;
#pragma unroll
    for (int k = 0; k < kBRows; k++) {
        const int y = min(y0 + r0 + k, H - 1);
        nar[k] = make_float4(nar[k].x / nn, nar[k].y / nn, nar[k].z / nn, nar[k].w / nn);
        const float sigma = sqrtf(var_a + var_b * raw.at(x, y));
        wide[k] = sqrtf(nar[k].x * nar[k].x + nar[k].y * nar[k].y) < 4.0f * sigma;
        any = any || wide[k];
    }
    if (any) blur_sep<HW>(sb, lx, r0 + kBApL, wide_k, wid);
    // The separable wide sums already differ from the tap-order sums by FP32
    // rounding; they are normalized by the reciprocal of the weight sum.
    const float rnw = 1.0f / wide_k.norm;
#pragma unroll
    for (int k = 0; k < kBRows; k++) {
        const int y = y0 + r0 + k;
        if (y >= H) break;
        const float4 res = wide[k] ? make_float4(wid[k].x * rnw, wid[k].y * rnw, wid[k].z * rnw, wid[k].w * rnw) : nar[k];
        out.at(x, y) = make_float2(copysignf(res.z, res.x), copysignf(res.w, res.y));
    }
}

// ---- demosaic ------------------------------------------------------------

struct Var2 {
    float a, b;
};

// The demosaic passes (interpolateGreen, interpolateRedBlue,
// interpolateRedBlueAtGreen) are held to the oracle within FP32 tolerance, not
// bit for bit: their gradient direction comes from the device atan2f.  Their
// quotients and square roots use the hardware reciprocal and square root
// (1 ulp) instead of the correctly rounded sequences.  Divisions by 2, 4 and 8
// stay exact products.
__device__ __forceinline__ float dm_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float dm_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float dm_smoothstep(float e0, float e1, float x) {
    float t = dm_div(x - e0, e1 - e0);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// The gradient as the demosaic passes use it: t = 2 atan2(|gy|, |gx|) / pi
// (0: horizontal gradient, 1: vertical) and its length.
struct GradDir {
    float t, len;
};
__device__ __forceinline__ GradDir grad_dir(float2 g) {
    const float ax = fabsf(g.x), ay = fabsf(g.y);
    return GradDir{dm_div(2 * atan2f(ay, ax), kPi), dm_sqrt(ax * ax + ay * ay)};
}

__device__ __forceinline__ bool is_site(const Quad& q, int c, int x, int y) {
    return (q.o[c].x & 1) == (x & 1) && (q.o[c].y & 1) == (y & 1);
}

// interpolateGreen at a red or blue site (demosaic.metal:214-300).  R reads the
// scaled raw (an image or an LDS tile), gdir is the site's gradient.
template <class RawR>
__device__ __forceinline__ float green_at_site(const RawR& R, GradDir gdir, Var2 gv, int x, int y) {
    const float low_noise = 1.0f - dm_smoothstep(3.5e-4f, 2e-3f, gv.b);
    const float gl = fetch(R, x - 1, y), gr = fetch(R, x + 1, y), gu = fetch(R, x, y - 1), gd = fetch(R, x, y + 1);
    const float c = R.at(x, y);
    const float cl = fetch(R, x - 2, y), cr = fetch(R, x + 2, y), cu = fetch(R, x, y - 2), cd = fetch(R, x, y + 2);
    const float d_ave = (fetch(R, x - 1, y - 1) + fetch(R, x + 1, y - 1) + fetch(R, x - 1, y + 1) + fetch(R, x + 1, y + 1)) / 4;
    const float g_ave = (gl + gr + gu + gd) / 4;
    const float lf_h = (gl + gr) / 2, lf_v = (gu + gd) / 2;
    const float hf_h = ((cl + cr) - 2 * c) / 4, hf_v = ((cu + cd) - 2 * c) / 4;
    const float sd = dm_sqrt(gv.a + gv.b * g_ave);
    const float glen = gdir.len;
    const float thr = dm_smoothstep(sd, 4 * sd, glen);
    const float low_thr = 1.0f - dm_smoothstep(2 * sd, 8 * sd, glen);
    const float sharpening = (0.5f + 0.5f * low_noise) * thr * (1 + low_noise * low_thr);
    const float highlights = 1.0f - dm_smoothstep(0.25f, 1.0f, fmaxf(c, fmaxf(fmaxf(cl, cr), fmaxf(cu, cd))));
    float dir = gdir.t;
    if (gv.b < 1e-3f) {
        dir = dir < 0.5f ? mixf(dir, 0.0f, 1.0f - dm_smoothstep(0.3f, 0.45f, dir))
                         : mixf(dir, 1.0f, dm_smoothstep(1.0f - 0.45f, 1.0f - 0.3f, dir));
    }
    const float cmin = fminf(c, fminf(g_ave, d_ave)), cmax = fmaxf(c, fmaxf(g_ave, d_ave));
    const float white = dm_div(cmin, cmax);
    const float k = highlights * white * sharpening;
    const float est_h = lf_h - k * hf_h, est_v = lf_v - k * hf_v;
    const float g = mixf(est_v, est_h, dir);
    const float gmax = fmaxf(fmaxf(gl, gr), fmaxf(gu, gd)), gmin = fminf(fminf(gl, gr), fminf(gu, gd));
    return clampf(g, mixf(1.0f, 0.5f, white) * gmin, mixf(1.0f, 1.5f, white) * gmax);
}

// interpolateGreen at any pixel: green sites keep their sample.
template <class RawR>
__device__ __forceinline__ float green_at(const RawR& R, const View<const float2>& Gr, Quad q, Var2 gv, int x, int y) {
    if (!(is_site(q, kRed, x, y) || is_site(q, kBlue, x, y))) return R.at(x, y);
    return green_at_site(R, grad_dir(Gr.at(x, y)), gv, x, y);
}

__global__ __launch_bounds__(256) void k_interpolate_green(View<const float> raw, View<const float2> grad,
                                                           View<float> green, Quad q, Var2 gv) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= green.w || y >= green.h) return;
    green.at(x, y) = green_at(raw, grad, q, gv, x, y);
}

// interpolateRedBluePixel (demosaic.metal:309-380): the missing colour at a
// red (red_site) or blue site from the colour differences of its diagonal
// neighbours, steered by the gradient direction (gdir: the site's gradient).
template <class RawR, class GreenR>
__device__ __forceinline__ float4 red_blue_site(const RawR& R, const GreenR& Gn, GradDir gdir, int x, int y,
                                                bool red_site, Var2 rv, Var2 bv) {
    const float g = Gn.at(x, y), c1 = R.at(x, y);
    const float g_tl = fetch(Gn, x - 1, y - 1), g_tr = fetch(Gn, x + 1, y - 1);
    const float g_bl = fetch(Gn, x - 1, y + 1), g_br = fetch(Gn, x + 1, y + 1);
    const float c_tl = fetch(R, x - 1, y - 1), c_tr = fetch(R, x + 1, y - 1);
    const float c_bl = fetch(R, x - 1, y + 1), c_br = fetch(R, x + 1, y + 1);
    const float c2_ave = (c_tl + c_tr + c_bl + c_br) / 4;
    const float d_tl = g_tl - c_tl, d_tr = g_tr - c_tr, d_bl = g_bl - c_bl, d_br = g_br - c_br;
    const float g2_tl = fetch(Gn, x - 2, y - 2), g2_tr = fetch(Gn, x + 2, y - 2);
    const float g2_bl = fetch(Gn, x - 2, y + 2), g2_br = fetch(Gn, x + 2, y + 2);
    const float d2_tl = g2_tl - fetch(R, x - 2, y - 2), d2_tr = g2_tr - fetch(R, x + 2, y - 2);
    const float d2_bl = g2_bl - fetch(R, x - 2, y + 2), d2_br = g2_br - fetch(R, x + 2, y + 2);
    const Var2 v = red_site ? rv : bv;
    const float sd = dm_sqrt(v.a + v.b * c2_ave);
    const float dir = 1.0f - gdir.t;
    const float thr = dm_smoothstep(sd, 4 * sd, gdir.len);
    const float alpha = mixf(0.5f, dir, thr);
    const float highlights = 1.0f - dm_smoothstep(0.25f, 1.0f, fmaxf(g, fmaxf(fmaxf(g2_tr, g2_bl), fmaxf(g2_tl, g2_br))));
    const float centre = 2 * (g - c1);
    const float anti = (d_tr + d_bl) / 2 + highlights * (d2_tr + d2_bl - centre) / 8;
    const float diag = (d_tl + d_br) / 2 + highlights * (d2_tl + d2_br - centre) / 8;
    float c2 = g - mixf(anti, diag, alpha);
    const float c2max = fmaxf(fmaxf(c_tl, c_tr), fmaxf(c_bl, c_br));
    const float c2min = fminf(fminf(c_tl, c_tr), fminf(c_bl, c_br));
    c2 = clampf(c2, c2min, c2max);
    return red_site ? make_float4(c1, g, c2, 0.0f) : make_float4(c2, g, c1, 0.0f);
}

__global__ __launch_bounds__(256) void k_interpolate_red_blue(View<const float> raw, View<const float> green,
                                                              View<const float2> grad, View<float4> rgb, Quad q,
                                                              Var2 rv, Var2 bv) {
    const int qx = blockIdx.x * blockDim.x + threadIdx.x, qy = blockIdx.y * blockDim.y + threadIdx.y;
    if (qx >= rgb.w / 2 || qy >= rgb.h / 2) return;
    const int x = 2 * qx, y = 2 * qy;
    const int rx = x + q.o[kRed].x, ry = y + q.o[kRed].y, bx = x + q.o[kBlue].x, by = y + q.o[kBlue].y;
    rgb.at(rx, ry) = red_blue_site(raw, green, grad_dir(grad.at(rx, ry)), rx, ry, true, rv, bv);
    rgb.at(bx, by) = red_blue_site(raw, green, grad_dir(grad.at(bx, by)), bx, by, false, rv, bv);
    for (int c : {kGreen, kGreen2}) {
        const int px = x + q.o[c].x, py = y + q.o[c].y;
        rgb.at(px, py) = make_float4(0.0f, green.at(px, py), 0.0f, 0.0f);
    }
}

// interpolateRedBlueAtGreenPixel (demosaic.metal:420-502).  C reads the
// interpolateRedBlue output as (red, green, blue, 0); gdir is the site's
// gradient.
template <class RgbR>
__device__ __forceinline__ float4 red_blue_at_green_site(const RgbR& C, GradDir gdir, int x, int y, Var2 rv, Var2 bv) {
    const float gc = C.at(x, y).y;
    const float4 l = fetch(C, x - 1, y), r = fetch(C, x + 1, y), u = fetch(C, x, y - 1), d = fetch(C, x, y + 1);
    const float4 l3 = fetch(C, x - 3, y), r3 = fetch(C, x + 3, y), u3 = fetch(C, x, y - 3), d3 = fetch(C, x, y + 3);
    const float red_ave = (l.x + r.x + u.x + d.x) / 4, blue_ave = (l.z + r.z + u.z + d.z) / 4;
    const float dr_l = l.y - l.x, dr_r = r.y - r.x, dr_u = u.y - u.x, dr_d = d.y - d.x;
    const float db_l = l.y - l.z, db_r = r.y - r.z, db_u = u.y - u.z, db_d = d.y - d.z;
    const float dr_l3 = l3.y - l3.x, dr_r3 = r3.y - r3.x, dr_u3 = u3.y - u3.x, dr_d3 = d3.y - d3.x;
    const float db_l3 = l3.y - l3.z, db_r3 = r3.y - r3.z, db_u3 = u3.y - u3.z, db_d3 = d3.y - d3.z;
    const float dir = gdir.t;
    const float glen = gdir.len;
    const float rsd = dm_sqrt(rv.a + rv.b * red_ave), bsd = dm_sqrt(bv.a + bv.b * blue_ave);
    const float ralpha = mixf(0.5f, dir, dm_smoothstep(rsd, 4 * rsd, glen));
    const float balpha = mixf(0.5f, dir, dm_smoothstep(bsd, 4 * bsd, glen));
    const float highlights = 1.0f - dm_smoothstep(0.25f, 1.0f, fmaxf(gc, fmaxf(fmaxf(r.y, l.y), fmaxf(d.y, u.y))));
    const float red = gc - mixf((dr_u + dr_d) / 2 - highlights * ((dr_d3 - dr_u) - (dr_d - dr_u3)) / 8,
                                (dr_l + dr_r) / 2 - highlights * ((dr_r3 - dr_l) - (dr_r - dr_l3)) / 8, ralpha);
    const float blue = gc - mixf((db_u + db_d) / 2 - highlights * ((db_d3 - db_u) - (db_d - db_u3)) / 8,
                                 (db_l + db_r) / 2 - highlights * ((db_r3 - db_l) - (db_r - db_l3)) / 8, balpha);
    return make_float4(clampf(red, fminf(fminf(l.x, r.x), fminf(u.x, d.x)), fmaxf(fmaxf(l.x, r.x), fmaxf(u.x, d.x))),
                       gc,
                       clampf(blue, fminf(fminf(l.z, r.z), fminf(u.z, d.z)), fmaxf(fmaxf(l.z, r.z), fmaxf(u.z, d.z))),
                       0.0f);
}

// blendHighlightsImage at one pixel (demosaic.metal:523-564): above `clip`,
// keep the sum of channels and scale the opponent chroma to the length of
// the clipped pixel's.
__device__ __forceinline__ float4 blend_highlight(float4 p, float clip) {
    if (!(p.x > clip || p.y > clip || p.z > clip)) return make_float4(p.x, p.y, p.z, 0.0f);
    const float s3 = 1.7320508f;
    const float cx = fminf(p.x, clip), cy = fminf(p.y, clip), cz = fminf(p.z, clip);
    // trans rows: (1, 1, 1), (s3, -s3, 0), (-1, -1, 2)
    const float a0 = 1 * p.x + 1 * p.y + 1 * p.z;
    float a1 = s3 * p.x + -s3 * p.y + 0 * p.z;
    float a2 = -1 * p.x + -1 * p.y + 2 * p.z;
    const float b1 = s3 * cx + -s3 * cy + 0 * cz;
    const float b2 = -1 * cx + -1 * cy + 2 * cz;
    const float sum0 = a1 * a1 + a2 * a2, sum1 = b1 * b1 + b2 * b2;
    const float ratio = sum0 > 0 ? sqrtf(sum1 / sum0) : 1.0f;
    a1 *= ratio;
    a2 *= ratio;
    // itrans rows: (1, s3/2, -0.5), (1, -s3/2, -0.5), (1, 0, 1)
    return make_float4((1 * a0 + s3 / 2 * a1 + -0.5f * a2) / 3, (1 * a0 + -s3 / 2 * a1 + -0.5f * a2) / 3,
                       (1 * a0 + 0 * a1 + 1 * a2) / 3, 0.0f);
}

struct Mat3 {
    float m[9];
};

__device__ __forceinline__ float4 transform3(const Mat3& t, float4 p) {
    return make_float4(t.m[0] * p.x + t.m[1] * p.y + t.m[2] * p.z, t.m[3] * p.x + t.m[4] * p.y + t.m[5] * p.z,
                       t.m[6] * p.x + t.m[7] * p.y + t.m[8] * p.z, 0.0f);
}

// interpolateRedBlueAtGreen, one quad per thread.
__global__ __launch_bounds__(256) void k_interpolate_red_blue_at_green(View<const float4> in, View<const float2> grad,
                                                                       View<float4> out, Quad q, Var2 rv, Var2 bv) {
    const int qx = blockIdx.x * blockDim.x + threadIdx.x, qy = blockIdx.y * blockDim.y + threadIdx.y;
    if (qx >= out.w / 2 || qy >= out.h / 2) return;
    const int x = 2 * qx, y = 2 * qy;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int px = x + q.o[c].x, py = y + q.o[c].y;
        out.at(px, py) = (c == kGreen || c == kGreen2)
                             ? red_blue_at_green_site(in, grad_dir(grad.at(px, py)), px, py, rv, bv)
                             : in.at(px, py);
    }
}

__global__ __launch_bounds__(256) void k_blend_highlights(View<const float4> in, View<float4> out, float clip) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    out.at(x, y) = blend_highlight(in.at(x, y), clip);
}

__global__ __launch_bounds__(256) void k_transform(View<const float4> in, View<float4> out, Mat3 t) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    out.at(x, y) = transform3(t, in.at(x, y));
}

// ---- despeckleLumaMedianChromaImage --------------------------------------

__device__ __forceinline__ void sort2(float& a, float& b) {
    const float lo = fminf(a, b), hi = fmaxf(a, b);
    a = lo;
    b = hi;
}

// Median of nine by the forgetful selection of demosaic.metal:1162-1192
// (each pass drops the minimum and maximum of what is left).
__device__ __forceinline__ float median9(float v[9]) {
    float a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3], a4 = v[4], a5 = v[5];
    sort2(a0, a1); sort2(a2, a3); sort2(a4, a5); sort2(a0, a2); sort2(a1, a3); sort2(a0, a4); sort2(a3, a5);
    a0 = v[6];
    sort2(a0, a1); sort2(a2, a3); sort2(a0, a2); sort2(a1, a3); sort2(a0, a4); sort2(a3, a4);
    a0 = v[7];
    sort2(a0, a1); sort2(a2, a3); sort2(a0, a2); sort2(a1, a3);
    a0 = v[8];
    sort2(a0, a1); sort2(a0, a2); sort2(a1, a2);
    return a1;
}

// despeckleLumaMedianChromaImage at one pixel (demosaic.metal:1262-1283): the
// luma clamped into its neighbourhood's second extremes (softly, by how far
// the extreme stands out of the noise), the chroma replaced by its median.
// `in` reads the YCbCr image (an image or an LDS tile).
template <class YccR>
__device__ __forceinline__ float4 despeckle_at(const YccR& in, int x, int y, Var2 yv) {
    // The reference's reading order: the right two columns top to bottom,
    // then (-1, 1), (-1, 0), (-1, -1).
    const int ox[9] = {0, 1, 0, 1, 0, 1, -1, -1, -1}, oy[9] = {-1, -1, 0, 0, 1, 1, 1, 0, -1};
    float cb[9], cr[9];
    // The reference's running extremes: v > max1 shifts max1 into max2, else
    // v > max2 replaces max2 (and the mirror image for the minima), from the
    // starts 0 and 65535.  As max2 <= max1 always, that is max2 = max(max2,
    // min(v, max1)), max1 = max(max1, v): a NaN neighbour changes nothing in
    // either form (the IEEE minimum / maximum propagate it into fmaxf / fminf,
    // which drop it), and the two forms can differ only in the sign of a zero
    // extreme, which the mix and smoothstep below turn into the same result.
    float max1 = 0.0f, max2 = 0.0f, min1 = 65535.0f, min2 = 65535.0f;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const float4 p = fetch(in, x + ox[k], y + oy[k]);
        cb[k] = p.y;
        cr[k] = p.z;
        const float v = p.x;
        max2 = fmaxf(max2, __builtin_elementwise_minimum(v, max1));
        max1 = fmaxf(max1, v);
        min2 = fminf(min2, __builtin_elementwise_maximum(v, min1));
        min1 = fminf(min1, v);
    }
    const float centre = in.at(x, y).x;
    const float sigma = sqrtf(yv.a + yv.b * centre);
    const float lo = mixf(min2, min1, smoothstepf(sigma, 4 * sigma, min2 - min1));
    const float hi = mixf(max2, max1, smoothstepf(sigma, 4 * sigma, max1 - max2));
    return make_float4(clampf(centre, lo, hi), median9(cb), median9(cr), 0.0f);
}

__global__ __launch_bounds__(256) void k_despeckle(View<const float4> in, View<float4> out, Var2 yv) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    out.at(x, y) = despeckle_at(in, x, y, yv);
}

// ---- the demosaic chain in one pass ----------------------------------------
//
// k_demosaic_tile runs interpolateGreen, interpolateRedBlue,
// interpolateRedBlueAtGreen with blendHighlightsImage and transformImage, and
// (kDespeckle) the despeckle for one TW x TH tile of the output, with every
// intermediate image in LDS: the scaled raw and the gradient are read once
// and the YCbCr image written once.  Each stage covers the tile grown by the
// reach of the stages after it (the despeckle reads +-1, red/blue at green
// +-3, red/blue +-2, green +-2), and a cell outside the image holds the value
// at the clamped position, which is what the per-stage kernels' clamped reads
// return there: each value is the same FP32 evaluation as theirs, bit for bit.

// interpolateRedBlue's output read as (red, green, blue, 0) from the fused
// kernel's green plane and (red, blue) plane; green sites hold (0, 0).
struct RgbTile {
    Tile<float> g;
    Tile<float2> rb;
    __device__ __forceinline__ float4 at(int x, int y) const {
        const float2 v = rb.at(x, y);
        return make_float4(v.x, g.at(x, y), v.y, 0.0f);
    }
};
__device__ __forceinline__ float4 fetch(const RgbTile& t, int x, int y) { return t.at(x, y); }

// The YCbCr image as three LDS planes of one geometry, `plane` floats apart.
struct YccTile {
    float* p;
    int ox, oy, pitch, plane;
    __device__ __forceinline__ int index(int x, int y) const { return (y - oy) * pitch + (x - ox); }
    __device__ __forceinline__ float4 at(int x, int y) const {
        const int i = index(x, y);
        return make_float4(p[i], p[i + plane], p[i + 2 * plane], 0.0f);
    }
    __device__ __forceinline__ void store(int x, int y, float4 v) const {
        const int i = index(x, y);
        p[i] = v.x;
        p[i + plane] = v.y;
        p[i + 2 * plane] = v.z;
    }
};
__device__ __forceinline__ float4 fetch(const YccTile& t, int x, int y) { return t.at(x, y); }

// Calls f(x, y) for every position of the w x h region at (ox, oy), w and h
// even, one 2x2 block per thread and pass: in each of the four calls of a
// pass the threads of a wave hold positions of one Bayer colour (away from
// the image border), so the colour branches do not diverge.
template <int NT, class F>
__device__ __forceinline__ void for_each_quad(int ox, int oy, int w, int h, F&& f) {
    const int bw = w / 2, n = bw * (h / 2);
    for (int b = threadIdx.x; b < n; b += NT) {
        const int x = ox + 2 * (b % bw), y = oy + 2 * (b / bw);
        f(x, y);
        f(x + 1, y);
        f(x, y + 1);
        f(x + 1, y + 1);
    }
}

template <int TW, int TH, bool kDespeckle>
struct TileGeometry {
    static constexpr int kH3 = kDespeckle ? 1 : 0;  // YCbCr
    static constexpr int kH2 = kH3 + 3;             // red / blue
    static constexpr int kH1 = kH2 + 2;             // green
    static constexpr int kH0 = kH1 + 2;             // scaled raw
    static constexpr int kW0 = TW + 2 * kH0, kR0 = TH + 2 * kH0;
    static constexpr int kW1 = TW + 2 * kH1, kR1 = TH + 2 * kH1;
    static constexpr int kW2 = TW + 2 * kH2, kR2 = TH + 2 * kH2;
    static constexpr int kW3 = TW + 2 * kH3, kR3 = TH + 2 * kH3;
    static constexpr int kGreenFloats = kW1 * kR1, kRbFloats = 2 * kW2 * kR2;
    // The gradient of the red and blue sites of the red / blue region (half its
    // cells, a GradDir each), computed by the green pass and read by the
    // red / blue pass.  It and the raw plane are dead once red / blue is done;
    // the YCbCr planes reuse their space.
    static constexpr int kRawFloats = kW0 * kR0, kDirFloats = kW2 * kR2, kYccFloats = 3 * kW3 * kR3;
    static constexpr int kReusedFloats =
        kRawFloats + kDirFloats > kYccFloats ? kRawFloats + kDirFloats : kYccFloats;
    static constexpr int kFloats = kGreenFloats + kRbFloats + kReusedFloats;
    static_assert(TW % 2 == 0 && TH % 2 == 0, "tiles hold whole 2x2 quads");
    static_assert(kGreenFloats % 2 == 0 && kRawFloats % 2 == 0, "float2 planes are 8-byte aligned");
    static_assert(sizeof(GradDir) == 2 * sizeof(float), "a GradDir is two floats");
};

// One workgroup per tile, tiles in XCD-contiguous order (their halos meet in
// one XCD's L2).
template <int TW, int TH, int NT, bool kDespeckle>
__global__ __launch_bounds__(NT) void k_demosaic_tile(View<const float> raw, View<const float2> grad,
                                                                View<float4> out, Quad q, Var2 rv, Var2 gv, Var2 bv,
                                                                float clip, Mat3 t, Var2 yv, XcdTiles order) {
    using G = TileGeometry<TW, TH, kDespeckle>;
    __shared__ float lds[G::kFloats];
    int tx, ty;
    if (!order.tile(tx, ty)) return;
    const int x0 = tx * TW, y0 = ty * TH;
    const int W = out.w, H = out.h;
    const Tile<float> green{lds, x0 - G::kH1, y0 - G::kH1, G::kW1};
    const Tile<float2> rb{reinterpret_cast<float2*>(lds + G::kGreenFloats), x0 - G::kH2, y0 - G::kH2, G::kW2};
    float* reused = lds + G::kGreenFloats + G::kRbFloats;
    const Tile<float> sraw{reused, x0 - G::kH0, y0 - G::kH0, G::kW0};
    // dirs: one entry per red / blue site of the red / blue region; the sites
    // of a row share the parity of x, so x / 2 indexes them.
    GradDir* dirs = reinterpret_cast<GradDir*>(reused + G::kRawFloats);
    auto dir_at = [&](int x, int y) -> GradDir& { return dirs[(y - rb.oy) * (G::kW2 / 2) + ((x - rb.ox) >> 1)]; };
    const YccTile ycc{reused, x0 - G::kH3, y0 - G::kH3, G::kW3, G::kW3 * G::kR3};

    for (int i = threadIdx.x; i < G::kW0 * G::kR0; i += NT) {
        const int x = sraw.ox + i % G::kW0, y = sraw.oy + i / G::kW0;
        sraw.p[i] = raw.at(clampi(x, 0, W - 1), clampi(y, 0, H - 1));
    }
    __syncthreads();
    for_each_quad<NT>(green.ox, green.oy, G::kW1, G::kR1, [&](int x, int y) {
        const int cx = clampi(x, 0, W - 1), cy = clampi(y, 0, H - 1);
        float g;
        if (is_site(q, kRed, cx, cy) || is_site(q, kBlue, cx, cy)) {
            const GradDir d = grad_dir(grad.at(cx, cy));
            g = green_at_site(sraw, d, gv, cx, cy);
            // the red / blue pass reads the gradient of its in-image sites
            if (x == cx && y == cy && x - rb.ox >= 0 && x - rb.ox < G::kW2 && y - rb.oy >= 0 && y - rb.oy < G::kR2)
                dir_at(x, y) = d;
        } else {
            g = sraw.at(cx, cy);
        }
        green.at(x, y) = g;
    });
    __syncthreads();
    for_each_quad<NT>(rb.ox, rb.oy, G::kW2, G::kR2, [&](int x, int y) {
        const int cx = clampi(x, 0, W - 1), cy = clampi(y, 0, H - 1);
        const bool red = is_site(q, kRed, cx, cy);
        float2 v = make_float2(0.0f, 0.0f);
        if (red || is_site(q, kBlue, cx, cy)) {
            const float4 c = red_blue_site(sraw, green, dir_at(cx, cy), cx, cy, red, rv, bv);
            v = make_float2(c.x, c.z);
        }
        rb.at(x, y) = v;
    });
    __syncthreads();
    const RgbTile rgb{green, rb};
    for_each_quad<NT>(ycc.ox, ycc.oy, G::kW3, G::kR3, [&](int x, int y) {
        const int cx = clampi(x, 0, W - 1), cy = clampi(y, 0, H - 1);
        const bool site = is_site(q, kRed, cx, cy) || is_site(q, kBlue, cx, cy);
        const float4 v = site ? rgb.at(cx, cy) : red_blue_at_green_site(rgb, grad_dir(grad.at(cx, cy)), cx, cy, rv, bv);
        ycc.store(x, y, transform3(t, blend_highlight(v, clip)));
    });
    __syncthreads();
    for (int i = threadIdx.x; i < TW * TH; i += NT) {
        const int x = x0 + i % TW, y = y0 + i / TW;
        if (x < W && y < H) out.at(x, y) = kDespeckle ? despeckle_at(ycc, x, y, yv) : ycc.at(x, y);
    }
}

// ---- raw noise model: basicRawNoiseStatistics + MeasureRawNLF ------------

// Per quad: mean and population variance of the quad samples read at
// 2 (qx, qy) + (x, y) + offset[c], x, y in [-4, 4] (the window steps by one
// pixel, as the reference's readRAWQuad is called), summed row by row.
__global__ __launch_bounds__(256) void k_raw_noise_stats(View<const float> raw, Quad q, View<float4> mean,
                                                         View<float4> var) {
    const int qx = blockIdx.x * blockDim.x + threadIdx.x, qy = blockIdx.y * blockDim.y + threadIdx.y;
    if (qx >= mean.w || qy >= mean.h) return;
    float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f}, sq[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int y = -4; y <= 4; y++) {
        for (int x = -4; x <= 4; x++) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const float v = fetch(raw, 2 * qx + x + q.o[c].x, 2 * qy + y + q.o[c].y);
                sum[c] = sum[c] + v;
                sq[c] = sq[c] + v * v;
            }
        }
    }
    const float count = 81.0f;
    mean.at(qx, qy) = make_float4(sum[0] / count, sum[1] / count, sum[2] / count, sum[3] / count);
    var.at(qx, qy) = make_float4((sq[0] - (sum[0] * sum[0]) / count) / count, (sq[1] - (sum[1] * sum[1]) / count) / count,
                                 (sq[2] - (sum[2] * sum[2]) / count) / count, (sq[3] - (sum[3] * sum[3]) / count) / count);
}

// Fit state on the device (FP64), advanced by the finalize kernels.
struct RawNlfState {
    double vmax[4], A[4], B[4], err2[4];
    double n1, n2, accepted, vmax1;
};
constexpr int kRawNlfBlocks = 512, kRawNlfValues = 21;

template <int PASS>
struct RawPassWidth;
template <>
struct RawPassWidth<0> { static constexpr int value = 6; };   // histogram
template <>
struct RawPassWidth<1> { static constexpr int value = 17; };  // sx, sy, sxx, sxy (x4), N
template <>
struct RawPassWidth<2> { static constexpr int value = 4; };   // squared errors
template <>
struct RawPassWidth<3> { static constexpr int value = 21; };  // pass-1 sums over the fitting quads, errors

__device__ __forceinline__ bool nan4(float4 v) { return isnan(v.x) || isnan(v.y) || isnan(v.z) || isnan(v.w); }

// Fixed-order block sum of NV per-thread doubles; thread 0 writes them.
template <int NV>
__device__ void block_sum(double (&acc)[NV], double* dst) {
    __shared__ double red[256];
    for (int v = 0; v < NV; v++) {
        red[threadIdx.x] = acc[v];
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) dst[v] = red[0];
        __syncthreads();
    }
}

// One pass over the statistics images (raw_converter.cpp:353-363, 379-394,
// 402-414, 425-446); per-block partials.
template <int PASS>
__global__ __launch_bounds__(256) void k_raw_nlf_pass(View<const float4> mean, View<const float4> var,
                                                      const RawNlfState* st, double* partials) {
    constexpr int NV = RawPassWidth<PASS>::value;
    double acc[NV];
    for (int v = 0; v < NV; v++) acc[v] = 0.0;
    double vmax[4], A[4], B[4], e2[4];
    for (int c = 0; c < 4; c++) {
        vmax[c] = st->vmax[c];
        A[c] = st->A[c];
        B[c] = st->B[c];
        e2[c] = st->err2[c];
    }
    const size_t n = (size_t)mean.w * mean.h;
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) {
        const int x = (int)(k % mean.w), y = (int)(k / mean.w);
        const float4 vv = var.at(x, y);
        const double v[4] = {vv.x, vv.y, vv.z, vv.w};
        if (PASS == 0) {
            if (nan4(vv) || !(v[0] > 0 && v[1] > 0 && v[2] > 0 && v[3] > 0)) continue;
            double mx = fmax(fmax(log10(v[0]), log10(v[1])), fmax(log10(v[2]), log10(v[3])));
            mx = fmin(fmax(mx, -6.0), -1.0);
            const int entry = 6 + (int)mx;
            for (int e = 0; e < 6; e++) acc[e] += (e == entry) ? 1.0 : 0.0;
            continue;
        }
        const float4 mm = mean.at(x, y);
        if (nan4(mm) || nan4(vv)) continue;
        const double m[4] = {mm.x, mm.y, mm.z, mm.w};
        bool sel = true;
        for (int c = 0; c < 4; c++) sel = sel && m[c] >= 0.001 && m[c] <= 0.9 && v[c] <= vmax[c];
        if (!sel) continue;
        if (PASS == 2) {
            for (int c = 0; c < 4; c++) {
                const double d = A[c] + B[c] * m[c] - v[c];
                acc[c] += d * d;
            }
            continue;
        }
        double dsq[4];
        if (PASS == 3) {
            bool fits = true;
            for (int c = 0; c < 4; c++) {
                const double d = fabs(A[c] + B[c] * m[c] - v[c]);
                dsq[c] = d * d;
                fits = fits && dsq[c] <= 0.5 * e2[c];
            }
            if (!fits) continue;
        }
        for (int c = 0; c < 4; c++) {
            acc[c] += m[c];
            acc[4 + c] += v[c];
            acc[8 + c] += m[c] * m[c];
            acc[12 + c] += m[c] * v[c];
        }
        acc[16] += 1.0;
        if (PASS == 3)
            for (int c = 0; c < 4; c++) acc[17 + c] += dsq[c];
    }
    block_sum<NV>(acc, partials + (size_t)blockIdx.x * kRawNlfValues);
}

__device__ __forceinline__ double floor_max(double v, double fl) { return isnan(v) ? fl : fmax(v, fl); }

// Clears the fit state with a kernel (no memset node if a caller captures the stream).
__global__ void k_raw_nlf_init(RawNlfState* st) {
    if (threadIdx.x < 20) reinterpret_cast<double*>(st)[threadIdx.x] = 0.0;
}
static_assert(sizeof(RawNlfState) == 20 * sizeof(double), "RawNlfState layout");

// Sum the partials (fixed order) and advance the fit (raw_converter.cpp:366-372,
// 397-398, 416-421, 448-484).  out: a[4], b[4] after the last pass.
template <int PASS>
__global__ __launch_bounds__(256) void k_raw_nlf_finalize(const double* partials, int nblocks, RawNlfState* st,
                                                          double exposure_multiplier, double npx, double* out) {
    constexpr int NV = RawPassWidth<PASS>::value;
    __shared__ double tot[kRawNlfValues];
    __shared__ double red[256];
    for (int v = 0; v < NV; v++) {
        double a = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += 256) a += partials[(size_t)b * kRawNlfValues + v];
        red[threadIdx.x] = a;
        __syncthreads();
        for (int s2 = 128; s2 > 0; s2 >>= 1) {
            if ((int)threadIdx.x < s2) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s2];
            __syncthreads();
        }
        if (threadIdx.x == 0) tot[v] = red[0];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (PASS == 0) {
        const double* h = tot;
        const double vm = h[0] > 1e3 ? 1.0e-5 : h[1] > 1e3 ? 1.0e-4 : h[2] > 1e3 ? 1.0e-3
                        : h[3] > 1e3 ? 1.0e-2 : h[4] > 1e3 ? 1.0e-1 : 1.0;
        for (int c = 0; c < 4; c++) st->vmax[c] = vm;
        st->vmax1 = vm;
    } else if (PASS == 1 || PASS == 3) {
        const double N = tot[16];
        bool take = true;
        if (PASS == 3) {
            take = N > 0.001 * npx;
            for (int c = 0; c < 4; c++) {
                const double e = tot[17 + c] / N;
                take = take && !isnan(e) && e < st->err2[c];
            }
            st->n2 = N;
            st->accepted = take ? 1.0 : 0.0;
        } else {
            st->n1 = N;
        }
        if (take) {
            for (int c = 0; c < 4; c++) {
                const double sx = tot[c], sy = tot[4 + c], sxx = tot[8 + c], sxy = tot[12 + c];
                const double B = floor_max((N * sxy - sx * sy) / (N * sxx - sx * sx), 1e-8);
                st->B[c] = B;
                st->A[c] = floor_max((sy - B * sx) / N, 1e-8);
                if (PASS == 3) st->err2[c] = tot[17 + c] / N;
            }
        }
        if (PASS == 3) {
            const double adj = exposure_multiplier * exposure_multiplier;
            for (int c = 0; c < 4; c++) {
                out[c] = st->A[c] * adj;
                out[4 + c] = st->B[c] * adj;
            }
        }
    } else {  // PASS 2
        for (int c = 0; c < 4; c++) {
            st->err2[c] = tot[c] / st->n1;
            st->vmax[c] = st->B[c];
        }
    }
}


// ---- histogramImage + histogramStatistics ---------------------------------

// Each workgroup counts its pixels in an LDS histogram, then adds its nonzero
// bins to the global one (integer counts: exact in any order).
__global__ __launch_bounds__(256) void k_histogram(View<const float4> in, uint32_t* hist) {
    __shared__ uint32_t lh[256];
    const int tid = threadIdx.x;
    lh[tid] = 0;
    __syncthreads();
    const size_t n = (size_t)in.w * in.h;
    for (size_t k = (size_t)blockIdx.x * 256 + tid; k < n; k += (size_t)gridDim.x * 256) {
        const float y = in.at((int)(k % in.w), (int)(k / in.w)).x;
        const float v = sqrtf(y / 2.0f) * 255.0f;
        const int idx = v >= 0.0f ? min((int)v, 255) : 0;  // NaN (negative luma): bin 0
        atomicAdd(&lh[idx], 1u);
    }
    __syncthreads();
    if (lh[tid]) atomicAdd(&hist[tid], lh[tid]);
}

// The reference's single-thread pass over the bins (demosaic.metal:1386-1453).
__global__ void k_histogram_statistics(gcam_histogram_data* d, uint32_t image_size) {
    if (threadIdx.x != 0) return;
    float mean = 0.0f;
    bool found_median = false, found_black = false, found_white = false;
    int median_index = 0, black_index = 0, white_index = 0;
    uint32_t sum = 0;
    for (int band = 0; band < 8; band++) {
        for (int i = 0; i < 32; i++) {
            const int index = band * 32 + i;
            const uint32_t entry = d->histogram[index];
            d->bands[band] += entry;
            mean += (float)(entry * (uint32_t)(index + 1)) / 256.0f;
            sum += entry;
            if (!found_median && sum >= (image_size + 1) / 2) {
                median_index = index;
                found_median = true;
            }
            if (!found_black && sum >= 0.001 * image_size) {
                black_index = index;
                found_black = true;
            }
            if (!found_white && sum >= 0.999 * image_size) {
                white_index = index;
                found_white = true;
            }
        }
    }
    mean /= (float)image_size;
    const float vb = (float)(black_index - 1) / 255.0f, vw = (float)white_index / 255.0f;
    d->black_level = 2.0f * vb * vb;
    d->white_level = 2.0f * vw * vw;
    d->median = (float)median_index / 255.0f - d->black_level;
    d->mean = mean - d->black_level;
    d->highlights = 1.0f + 2.0f * smoothstepf(0.01f, 0.1f, (float)(d->bands[5] + d->bands[6] + d->bands[7]) /
                                                             (float)image_size);
    d->shadows = 0.8f - 0.5f * smoothstepf(0.25f, 0.5f, (float)(d->bands[0] + d->bands[1]) / (float)image_size);
}

// Zeroes a histogram_data with a kernel (histogramImageKernel::reset).
__global__ void k_histogram_reset(gcam_histogram_data* d) {
    uint32_t* w = reinterpret_cast<uint32_t*>(d);
    for (int i = threadIdx.x; i < (int)(sizeof(gcam_histogram_data) / 4); i += blockDim.x) w[i] = 0u;
}


// ---- local tone-mapping mask ----------------------------------------------

// GuidedFilterABImage: the sampler reads the guide at texel centres, so the
// 5x5 window is a plain clamped box of channel 0.
__global__ __launch_bounds__(256) void k_guided_filter_ab(View<const float4> g, float eps, View<float2> ab) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= ab.w || y >= ab.h) return;
    float sum = 0.0f, sq = 0.0f;
    for (int j = -2; j <= 2; j++)
        for (int i = -2; i <= 2; i++) {
            const float v = fetch(g, x + i, y + j).x;
            sum = sum + v;
            sq = sq + v * v;
        }
    const float mean = sum / 25, var = (sq - sum * sum / 25) / 25;
    const float a = var / (var + eps);
    ab.at(x, y) = make_float2(a, mean * (1 - a));
}

// Bilinear sample of a 2-channel image at texel-space x0 + fx, y0 + fy.
__device__ __forceinline__ float2 sample2(const View<const float2>& im, int x0, float fx, int y0, float fy) {
    const float2 a = fetch(im, x0, y0), b = fetch(im, x0 + 1, y0), c = fetch(im, x0, y0 + 1), d = fetch(im, x0 + 1, y0 + 1);
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    return make_float2((a.x * gx + b.x * fx) * gy + (c.x * gx + d.x * fx) * fy,
                       (a.y * gx + b.y * fx) * gy + (c.y * gx + d.y * fx) * fy);
}

// BoxFilterGFImage: the 5x5 box as 9 bilinear taps (demosaic.metal:1329-1339).
__global__ __launch_bounds__(256) void k_box_filter_gf(View<const float2> in, View<float2> out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    const float w[9] = {0.16f, 0.16f, 0.08f, 0.16f, 0.16f, 0.08f, 0.08f, 0.08f, 0.04f};
    const int ix[9] = {-2, 0, 2, -2, 0, 2, -2, 0, 2}, iy[9] = {-2, -2, -2, 0, 0, 0, 2, 2, 2};
    const float fx[9] = {0.5f, 0.5f, 0.0f, 0.5f, 0.5f, 0.0f, 0.5f, 0.5f, 0.0f};
    const float fy[9] = {0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.0f, 0.0f, 0.0f};
    float2 acc = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int t = 0; t < 9; t++) {
        const float2 s2 = sample2(in, x + ix[t], fx[t], y + iy[t], fy[t]);
        acc.x = acc.x + w[t] * s2.x;
        acc.y = acc.y + w[t] * s2.y;
    }
    out.at(x, y) = acc;
}

// Centre of output pixel o (of `out` pixels) in the texel space of an image of
// `in` pixels: ((2 o + 1) in - out) / (2 out) = i0 + frac, in integers.
__device__ __forceinline__ void centre_map(int o, int in, int out, int& i0, float& frac) {
    const long long num = (long long)(2 * o + 1) * in - out, den = 2LL * out;
    long long q = num / den;
    if (num % den != 0 && num < 0) q -= 1;
    i0 = (int)q;
    frac = (float)(num - q * den) / (float)den;
}

// computeLtmMultiplier (demosaic.metal:1456-1472).
__device__ __forceinline__ float ltm_multiplier(float yv, float ill, float shadows, float highlights, float detail) {
    const float mid = 0.22f;
    const float reflectance = yv / ill;
    const float adj = ill <= mid ? mid * powf(ill / mid, shadows)
                                 : (1 - mid) * powf((ill - mid) / (1 - mid), highlights) + mid;
    return adj * powf(reflectance, detail) / yv;
}

struct LtmBands {
    View<const float2> b[3];  // box-filtered (a, b): low, mid, high frequency
    float detail[3];
};

// localToneMappingMaskImage (demosaic.metal:1481-1535).
__global__ __launch_bounds__(256) void k_ltm_mask(View<const float4> in, View<const float2> grad, LtmBands bands,
                                                  float nlf_a, float nlf_b, const gcam_histogram_data* hist,
                                                  View<float> mask) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= mask.w || y >= mask.h) return;
    const float yv = in.at(x, y).x;
    const float shadows = hist->shadows, highlights = hist->highlights;
    float m = 1.0f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int x0, y0;
        float fx, fy;
        centre_map(x, bands.b[k].w, mask.w, x0, fx);
        centre_map(y, bands.b[k].h, mask.h, y0, fy);
        const float2 ab = sample2(bands.b[k], x0, fx, y0, fy);
        const float ill = ab.x * yv + ab.y;
        float detail = bands.detail[k];
        if (detail > 1.0f) {
            float d = (1.0f - smoothstepf(0.75f, 0.95f, ill)) * smoothstepf(0.05f, 0.1f, ill);
            if (k == 2) {
                const float2 gg = grad.at(x, y);
                const float g = sqrtf(gg.x * gg.x + gg.y * gg.y);
                const float sigma = sqrtf(nlf_a + nlf_b * ill);
                d = smoothstepf(2 * sigma, 8 * sigma, g) * (1 - smoothstepf(32 * sigma, 128 * sigma, g)) *
                    (1.0f - smoothstepf(0.75f, 0.95f, ill)) * smoothstepf(0.05f, 0.1f, ill);
            }
            detail = 1 + (detail - 1) * d;
        }
        const float mk = ltm_multiplier(yv, ill, shadows, k == 0 ? highlights : 1.0f, detail);
        m = k == 0 ? mk : m * mk;
    }
    mask.at(x, y) = m;
}

// ---- host side -----------------------------------------------------------

// gaussianKernelBilinearWeights (demosaic_utils.cpp:987-1006) with
// KernelOptimizeBilinear2d (:939-985): the k x k Gaussian (k = max(ceil(2r),
// 2) made odd) folded into ((k / 2) + 1)^2 bilinear taps {weight, x, y}; the
// corner tap sits at (k / 2.0, k / 2.0) as the reference places it.
std::vector<float> bilinear_gaussian(float radius) {
    int k = std::max((int)std::ceil(2 * radius), 2);
    if (k % 2 == 0) k++;
    const int half = k / 2, ow = half + 1;
    std::vector<float> g((size_t)k * k), taps((size_t)ow * ow * 3);
    for (int y = 0; y < k; y++)
        for (int x = 0; x < k; x++) {
            const int dx = x - half, dy = y - half;
            g[(size_t)y * k + x] = std::exp(-((float)(dx * dx + dy * dy) / (2 * radius * radius)));
        }
    auto G = [&](int r, int c) { return g[(size_t)r * k + c]; };
    auto set = [&](int r, int c, float wt, float ox, float oy) {
        float* t = &taps[((size_t)(r / 2) * ow + c / 2) * 3];
        t[0] = wt;
        t[1] = ox;
        t[2] = oy;
    };
    // The last row / column index after the pair loops is k - 1.
    for (int r = 0; r < k; r += 2) {
        for (int c = 0; c < k; c += 2) {
            const bool rp = r + 1 < k, cp = c + 1 < k;  // a pair of rows / columns
            const float x1 = (float)(c - half), y1 = (float)(r - half);
            if (rp && cp) {
                const float w1 = G(r, c), w2 = G(r, c + 1), w3 = G(r + 1, c), w4 = G(r + 1, c + 1);
                set(r, c, w1 + w2 + w3 + w4, (x1 * w1 + (x1 + 1.0f) * w2) / (w1 + w2),
                    (y1 * w1 + (y1 + 1.0f) * w3) / (w1 + w3));
            } else if (rp) {
                const float w1 = G(r, c), w2 = G(r + 1, c);
                set(r, c, w1 + w2, x1, (y1 * w1 + (y1 + 1.0f) * w2) / (w1 + w2));
            } else if (cp) {
                const float w1 = G(r, c), w2 = G(r, c + 1);
                set(r, c, w1 + w2, (x1 * w1 + (x1 + 1.0f) * w2) / (w1 + w2), y1);
            } else {
                set(r, c, G(r, c), k / 2.0f, k / 2.0f);
            }
        }
    }
    return taps;
}

// Taps of one kernel as integer offset + exact fraction; false if they do not
// fit the device table or reach beyond the tile apron.
bool make_taps(float radius, Tap* out, int cap, int* n) {
    if (!(radius > 0.0f) || !std::isfinite(radius)) return false;
    const std::vector<float> t = bilinear_gaussian(radius);
    const int m = (int)t.size() / 3;
    if (m > cap) return false;
    for (int i = 0; i < m; i++) {
        const float fx0 = std::floor(t[3 * i + 1]), fy0 = std::floor(t[3 * i + 2]);
        Tap& d = out[i];
        d.w = t[3 * i];
        d.ix = (int)fx0;
        d.iy = (int)fy0;
        d.fx = t[3 * i + 1] - fx0;
        d.fy = t[3 * i + 2] - fy0;
        if (d.ix < -kBApL || d.ix + 1 > kBApR || d.iy < -kBApL || d.iy + 1 > kBApR) return false;
    }
    *n = m;
    return true;
}

// The wide kernel's separable form (see SepKernel), from the same FP32
// Gaussian the taps are folded from; norm = the taps' weights summed in tap
// order, as the tap-by-tap evaluation divides by.
SepKernel make_sep_kernel(float radius, const Tap* t, int n) {
    SepKernel k{};
    int ks = std::max((int)std::ceil(2 * radius), 2);
    if (ks % 2 == 0) ks++;
    k.h = ks / 2;
    for (int d = 0; d <= k.h && d <= kMaxSepHalf; d++)
        k.g[d] = std::exp(-((float)(d * d) / (2 * radius * radius)));
    k.gc = std::exp(-((float)(2 * k.h * k.h) / (2 * radius * radius)));
    float norm = 0.0f;
    for (int i = 0; i < n; i++) norm = norm + t[i].w;
    k.norm = norm;
    return k;
}

dim3 grid_for(int w, int h, dim3 b) { return dim3((w + b.x - 1) / b.x, (h + b.y - 1) / b.y); }

bool valid_image(const gcam_image* im, int fmt) {
    return im && im->data && im->width > 0 && im->height > 0 && im->stride >= im->width && im->format == fmt;
}
bool same_size(const gcam_image* a, const gcam_image* b) { return a->width == b->width && a->height == b->height; }

gcam_status bad(const char* what) { return set_error(GCAM_ERR_INVALID_ARGUMENT, what); }

gcam_status launch_status(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(GCAM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return GCAM_OK;
}


// Statistics and the four fit passes on stream s; scratch: kRawNlfBlocks *
// kRawNlfValues doubles, a RawNlfState, 8 doubles of output (a[4], b[4]).
gcam_status raw_nlf_device(View<const float> scaled, const Quad& q, View<float4> mean, View<float4> var,
                           double* partials, RawNlfState* state, double* out, double exposure, hipStream_t s) {
    const dim3 b(64, 4);
    k_raw_noise_stats<<<grid_for(mean.w, mean.h, b), b, 0, s>>>(scaled, q, mean, var);
    k_raw_nlf_init<<<1, 64, 0, s>>>(state);
    const View<const float4> m{mean.ptr, mean.w, mean.h, mean.stride}, v{var.ptr, var.w, var.h, var.stride};
    const double npx = 4.0 * mean.w * mean.h;
    k_raw_nlf_pass<0><<<kRawNlfBlocks, 256, 0, s>>>(m, v, state, partials);
    k_raw_nlf_finalize<0><<<1, 256, 0, s>>>(partials, kRawNlfBlocks, state, exposure, npx, out);
    k_raw_nlf_pass<1><<<kRawNlfBlocks, 256, 0, s>>>(m, v, state, partials);
    k_raw_nlf_finalize<1><<<1, 256, 0, s>>>(partials, kRawNlfBlocks, state, exposure, npx, out);
    k_raw_nlf_pass<2><<<kRawNlfBlocks, 256, 0, s>>>(m, v, state, partials);
    k_raw_nlf_finalize<2><<<1, 256, 0, s>>>(partials, kRawNlfBlocks, state, exposure, npx, out);
    k_raw_nlf_pass<3><<<kRawNlfBlocks, 256, 0, s>>>(m, v, state, partials);
    k_raw_nlf_finalize<3><<<1, 256, 0, s>>>(partials, kRawNlfBlocks, state, exposure, npx, out);
    return launch_status("MeasureRawNLF");
}

// The demosaic chain of the front end (interpolateGreen through the YCbCr
// transform, then the despeckle when asked) as one k_demosaic_tile pass.
template <bool kDespeckle>
gcam_status demosaic_tiles(View<const float> scaled, View<const float2> grad, View<float4> out, const Quad& q,
                           const float rv[6], float clip, const Mat3& t, Var2 yv, hipStream_t s) {
    // 48 x 32 tiles, 512 threads: two workgroups (32 wavefronts) per CU; measured
    // against 32-64 x 16-64 tiles of 256-1024 threads on the 12 MP frame.
    constexpr int TW = 48, TH = 32, NT = 512;
    const XcdTiles order = xcd_tiles(out.w, out.h, TW, TH);
    k_demosaic_tile<TW, TH, NT, kDespeckle><<<order.grid(), NT, 0, s>>>(
        scaled, grad, out, q, Var2{rv[0], rv[1]}, Var2{rv[2], rv[3]}, Var2{rv[4], rv[5]}, clip, t, yv, order);
    return launch_status(kDespeckle ? "demosaic + despeckle" : "demosaic");
}
}  // namespace

}  // namespace gcam

using namespace gcam;

namespace {
const dim3 kBlock(64, 4);
}

extern "C" {

int32_t gcam_gaussian_kernel_bilinear_weights(float radius, float* taps, int32_t capacity) {
    if (!(radius > 0.0f) || !std::isfinite(radius) || radius > 64.0f) return -1;
    const std::vector<float> t = bilinear_gaussian(radius);
    const int n = (int)t.size() / 3;
    if (taps && capacity >= n) std::copy(t.begin(), t.end(), taps);
    return n;
}

void gcam_raw_variance(const float raw_nlf_a[4], const float raw_nlf_b[4], float out[6]) {
    out[0] = raw_nlf_a[0];
    out[1] = raw_nlf_b[0];
    out[2] = (raw_nlf_a[1] + raw_nlf_a[3]) / 2;
    out[3] = (raw_nlf_b[1] + raw_nlf_b[3]) / 2;
    out[4] = raw_nlf_a[2];
    out[5] = raw_nlf_b[2];
}

void gcam_cam_ycbcr(const float rgb_cam[9], float out[9]) {
    const float yr[3] = {0.2126f, 0.7152f, 0.0722f};  // row 1 of xyz_sRGB (demosaic.hpp:138-142)
    float K[3];
    for (int j = 0; j < 3; j++) K[j] = yr[0] * rgb_cam[j] + yr[1] * rgb_cam[3 + j] + yr[2] * rgb_cam[6 + j];
    const float kr = K[0], kg = K[1], kb = K[2];
    const float m[9] = {kr, kg, kb, -0.5f * kr / (1.0f - kb), -0.5f * kg / (1 - kb), 0.5f,
                        0.5f, -0.5f * kg / (1 - kr), -0.5f * kb / (1 - kr)};
    std::copy(m, m + 9, out);
}

gcam_status gcam_matrix3_inverse(const float m[9], float out[9]) {
    if (!m || !out) return bad("gcam_matrix3_inverse: null matrix");
    double a[9];
    for (int i = 0; i < 9; i++) a[i] = m[i];
    const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (det == 0.0 || !std::isfinite(det)) return bad("gcam_matrix3_inverse: singular matrix");
    const double r = 1.0 / det;
    const double inv[9] = {c00 * r, (a[2] * a[7] - a[1] * a[8]) * r, (a[1] * a[5] - a[2] * a[4]) * r,
                           c01 * r, (a[0] * a[8] - a[2] * a[6]) * r, (a[2] * a[3] - a[0] * a[5]) * r,
                           c02 * r, (a[1] * a[6] - a[0] * a[7]) * r, (a[0] * a[4] - a[1] * a[3]) * r};
    for (int i = 0; i < 9; i++) out[i] = (float)inv[i];
    return GCAM_OK;
}

gcam_status gcam_scale_raw_data(const gcam_image* raw, gcam_image* scaled, int32_t bayer, const float scale_mul[4],
                                float black_level, float lens_shading, void* stream) {
    if (!valid_image(raw, GCAM_PIXEL_R16U) || !valid_image(scaled, GCAM_PIXEL_R32F) || !same_size(raw, scaled))
        return bad("gcam_scale_raw_data: raw must be R16U and scaled R32F of the same size");
    if ((raw->width | raw->height) & 1) return bad("gcam_scale_raw_data: width and height must be even");
    if (bayer < 0 || bayer > 3 || !scale_mul) return bad("gcam_scale_raw_data: bad Bayer pattern or scale_mul");
    const dim3 g = grid_for(raw->width / 2, raw->height / 2, kBlock);
    k_scale_raw<<<g, kBlock, 0, (hipStream_t)stream>>>(view<const uint16_t>(*raw), view<float>(*scaled),
                                                        quad_of(bayer),
                                                        make_float4(scale_mul[0], scale_mul[1], scale_mul[2], scale_mul[3]),
                                                        black_level, lens_shading);
    return launch_status("scaleRawData");
}

gcam_status gcam_raw_image_sobel(const gcam_image* scaled, gcam_image* sobel, void* stream) {
    if (!valid_image(scaled, GCAM_PIXEL_R32F) || !valid_image(sobel, GCAM_PIXEL_RGBA32F) || !same_size(scaled, sobel))
        return bad("gcam_raw_image_sobel: scaled must be R32F and sobel RGBA32F of the same size");
    k_raw_sobel<<<grid_for(sobel->width, sobel->height, kBlock), kBlock, 0, (hipStream_t)stream>>>(
        view<const float>(*scaled), view<float4>(*sobel));
    return launch_status("rawImageSobel");
}

}  // extern "C"

template <int HW>
static void launch_blur_sobel(const gcam_image* scaled, const gcam_image* sobel, const BlurTaps& taps,
                              const SepKernel& wk, const float raw_variance[2], gcam_image* gradient, hipStream_t s) {
    const dim3 g((gradient->width + kBTW - 1) / kBTW, (gradient->height + kBTH - 1) / kBTH);
    const dim3 b(kBTW, kBTH / kBRows);
    if (sobel) {
        k_blur_sobel<false, HW><<<g, b, 0, s>>>(view<const float>(*scaled), view<const float4>(*sobel), taps, wk,
                                                raw_variance[0], raw_variance[1], view<float2>(*gradient));
    } else {
        k_blur_sobel<true, HW><<<g, b, 0, s>>>(view<const float>(*scaled), View<const float4>{nullptr, 0, 0, 0}, taps,
                                               wk, raw_variance[0], raw_variance[1], view<float2>(*gradient));
    }
}

extern "C" {

static gcam_status blur_sobel(const gcam_image* scaled, const gcam_image* sobel, float radius1, float radius2,
                              const float raw_variance[2], gcam_image* gradient, hipStream_t s) {
    BlurTaps taps{};
    if (!make_taps(radius1, taps.t1, kMaxTaps1, &taps.n1) || !make_taps(radius2, taps.t2, kMaxTaps2, &taps.n2))
        return bad("gaussianBlurSobelImage: radius1 must be in (0, 1.5] and radius2 in (0, 5.5]");
    const SepKernel wk = make_sep_kernel(radius2, taps.t2, taps.n2);
    switch (wk.h) {
        case 1: launch_blur_sobel<1>(scaled, sobel, taps, wk, raw_variance, gradient, s); break;
        case 2: launch_blur_sobel<2>(scaled, sobel, taps, wk, raw_variance, gradient, s); break;
        case 3: launch_blur_sobel<3>(scaled, sobel, taps, wk, raw_variance, gradient, s); break;
        case 4: launch_blur_sobel<4>(scaled, sobel, taps, wk, raw_variance, gradient, s); break;
        case 5: launch_blur_sobel<5>(scaled, sobel, taps, wk, raw_variance, gradient, s); break;
        default: return bad("gaussianBlurSobelImage: radius2 must be in (0, 5.5]");
    }
    return launch_status("gaussianBlurSobelImage");
}

gcam_status gcam_gaussian_blur_sobel_image(const gcam_image* scaled, const gcam_image* sobel, float radius1,
                                           float radius2, const float raw_variance[2], gcam_image* gradient,
                                           void* stream) {
    if (!valid_image(scaled, GCAM_PIXEL_R32F) || !valid_image(gradient, GCAM_PIXEL_RG32F) ||
        !same_size(scaled, gradient) || !raw_variance)
        return bad("gcam_gaussian_blur_sobel_image: scaled R32F, gradient RG32F of the same size, raw_variance");
    if (sobel && (!valid_image(sobel, GCAM_PIXEL_RGBA32F) || !same_size(sobel, scaled)))
        return bad("gcam_gaussian_blur_sobel_image: sobel must be RGBA32F of the image size (or NULL)");
    return blur_sobel(scaled, sobel, radius1, radius2, raw_variance, gradient, (hipStream_t)stream);
}

static gcam_status demosaic(const gcam_image* scaled, const gcam_image* gradient, gcam_image* green,
                            gcam_image* rgb_tmp, gcam_image* rgb_out, int32_t bayer, const float rv[6], hipStream_t s) {
    const Quad q = quad_of(bayer);
    const Var2 red{rv[0], rv[1]}, gv{rv[2], rv[3]}, blue{rv[4], rv[5]};
    const int w = scaled->width, h = scaled->height;
    k_interpolate_green<<<grid_for(w, h, kBlock), kBlock, 0, s>>>(view<const float>(*scaled),
                                                                  view<const float2>(*gradient), view<float>(*green),
                                                                  q, gv);
    gcam_status st = launch_status("interpolateGreen");
    if (st != GCAM_OK) return st;
    const dim3 gq = grid_for(w / 2, h / 2, kBlock);
    k_interpolate_red_blue<<<gq, kBlock, 0, s>>>(view<const float>(*scaled), view<const float>(*green),
                                                 view<const float2>(*gradient), view<float4>(*rgb_tmp), q, red, blue);
    if ((st = launch_status("interpolateRedBlue")) != GCAM_OK) return st;
    k_interpolate_red_blue_at_green<<<gq, kBlock, 0, s>>>(view<const float4>(*rgb_tmp), view<const float2>(*gradient),
                                                          view<float4>(*rgb_out), q, red, blue);
    return launch_status("interpolateRedBlueAtGreen");
}

gcam_status gcam_demosaic_image(const gcam_image* scaled, const gcam_image* gradient, gcam_image* green,
                                gcam_image* rgb_tmp, gcam_image* rgb_out, int32_t bayer, const float raw_variance[6],
                                void* stream) {
    if (!valid_image(scaled, GCAM_PIXEL_R32F) || !valid_image(gradient, GCAM_PIXEL_RG32F) ||
        !valid_image(green, GCAM_PIXEL_R32F) || !valid_image(rgb_tmp, GCAM_PIXEL_RGBA32F) ||
        !valid_image(rgb_out, GCAM_PIXEL_RGBA32F) || !same_size(scaled, gradient) || !same_size(scaled, green) ||
        !same_size(scaled, rgb_tmp) || !same_size(scaled, rgb_out) || !raw_variance)
        return bad("gcam_demosaic_image: scaled R32F, gradient RG32F, green R32F, rgb_tmp / rgb_out RGBA32F, one size");
    if (rgb_tmp->data == rgb_out->data) return bad("gcam_demosaic_image: rgb_tmp and rgb_out must differ");
    if ((scaled->width | scaled->height) & 1) return bad("gcam_demosaic_image: width and height must be even");
    if (bayer < 0 || bayer > 3) return bad("gcam_demosaic_image: bad Bayer pattern");
    return demosaic(scaled, gradient, green, rgb_tmp, rgb_out, bayer, raw_variance, (hipStream_t)stream);
}

gcam_status gcam_blend_highlights_image(const gcam_image* in, float clip, gcam_image* out, void* stream) {
    if (!valid_image(in, GCAM_PIXEL_RGBA32F) || !valid_image(out, GCAM_PIXEL_RGBA32F) || !same_size(in, out))
        return bad("gcam_blend_highlights_image: RGBA32F images of one size");
    k_blend_highlights<<<grid_for(out->width, out->height, kBlock), kBlock, 0, (hipStream_t)stream>>>(
        view<const float4>(*in), view<float4>(*out), clip);
    return launch_status("blendHighlightsImage");
}

gcam_status gcam_transform_image(const gcam_image* in, const float m[9], gcam_image* out, void* stream) {
    if (!valid_image(in, GCAM_PIXEL_RGBA32F) || !valid_image(out, GCAM_PIXEL_RGBA32F) || !same_size(in, out) || !m)
        return bad("gcam_transform_image: RGBA32F images of one size and a 3x3 matrix");
    Mat3 t;
    std::copy(m, m + 9, t.m);
    k_transform<<<grid_for(out->width, out->height, kBlock), kBlock, 0, (hipStream_t)stream>>>(
        view<const float4>(*in), view<float4>(*out), t);
    return launch_status("transformImage");
}

gcam_status gcam_despeckle_image(const gcam_image* in, const float var_a[3], const float var_b[3], gcam_image* out,
                                 void* stream) {
    if (!valid_image(in, GCAM_PIXEL_RGBA32F) || !valid_image(out, GCAM_PIXEL_RGBA32F) || !same_size(in, out) ||
        !var_a || !var_b)
        return bad("gcam_despeckle_image: RGBA32F images of one size and the noise model");
    if (in->data == out->data) return bad("gcam_despeckle_image: in place is not supported");
    k_despeckle<<<grid_for(out->width, out->height, kBlock), kBlock, 0, (hipStream_t)stream>>>(
        view<const float4>(*in), view<float4>(*out), Var2{var_a[0], var_b[0]});
    return launch_status("despeckleLumaMedianChromaImage");
}


gcam_status gcam_raw_noise_statistics(const gcam_image* scaled, int32_t bayer, gcam_image* mean, gcam_image* var,
                                      void* stream) {
    if (!valid_image(scaled, GCAM_PIXEL_R32F) || !valid_image(mean, GCAM_PIXEL_RGBA32F) ||
        !valid_image(var, GCAM_PIXEL_RGBA32F) || !same_size(mean, var) || mean->width != scaled->width / 2 ||
        mean->height != scaled->height / 2)
        return bad("gcam_raw_noise_statistics: scaled R32F, mean and var RGBA32F of half its size");
    if (bayer < 0 || bayer > 3) return bad("gcam_raw_noise_statistics: bad Bayer pattern");
    const dim3 b(64, 4);
    k_raw_noise_stats<<<grid_for(mean->width, mean->height, b), b, 0, (hipStream_t)stream>>>(
        view<const float>(*scaled), quad_of(bayer), view<float4>(*mean), view<float4>(*var));
    return launch_status("basicRawNoiseStatistics");
}

gcam_status gcam_measure_raw_nlf(const gcam_image* scaled, int32_t bayer, float exposure_multiplier,
                                 float raw_nlf_a[4], float raw_nlf_b[4], void* stream) {
    if (!valid_image(scaled, GCAM_PIXEL_R32F) || ((scaled->width | scaled->height) & 1) || !raw_nlf_a || !raw_nlf_b)
        return bad("gcam_measure_raw_nlf: scaled must be R32F with even sizes, outputs non-null");
    if (bayer < 0 || bayer > 3) return bad("gcam_measure_raw_nlf: bad Bayer pattern");
    const hipStream_t s = (hipStream_t)stream;
    const int qw = scaled->width / 2, qh = scaled->height / 2;
    const size_t img = (size_t)qw * qh * sizeof(float4);
    const size_t part = (size_t)kRawNlfBlocks * kRawNlfValues * sizeof(double);
    char* buf = nullptr;
    if (hipMalloc(&buf, 2 * img + part + sizeof(RawNlfState) + 8 * sizeof(double)) != hipSuccess)
        return set_error(GCAM_ERR_OUT_OF_MEMORY, "gcam_measure_raw_nlf: device allocation failed");
    View<float4> mean{reinterpret_cast<float4*>(buf), qw, qh, qw}, var{reinterpret_cast<float4*>(buf + img), qw, qh, qw};
    double* partials = reinterpret_cast<double*>(buf + 2 * img);
    RawNlfState* state = reinterpret_cast<RawNlfState*>(buf + 2 * img + part);
    double* out = reinterpret_cast<double*>(buf + 2 * img + part + sizeof(RawNlfState));
    gcam_status st = raw_nlf_device(view<const float>(*scaled), quad_of(bayer), mean, var, partials, state, out,
                                    exposure_multiplier, s);
    double host[8];
    if (st == GCAM_OK) {
        const hipError_t e1 = hipMemcpyAsync(host, out, sizeof(host), hipMemcpyDeviceToHost, s);
        const hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(s) : e1;
        if (e2 != hipSuccess) st = set_error(GCAM_ERR_HIP, std::string("gcam_measure_raw_nlf: ") + hipGetErrorString(e2));
    }
    hipFree(buf);
    if (st != GCAM_OK) return st;
    for (int c = 0; c < 4; c++) {
        raw_nlf_a[c] = (float)host[c];
        raw_nlf_b[c] = (float)host[4 + c];
    }
    return GCAM_OK;
}


gcam_status gcam_histogram_reset(gcam_histogram_data* data, void* stream) {
    if (!data) return bad("gcam_histogram_reset: null histogram");
    k_histogram_reset<<<1, 256, 0, (hipStream_t)stream>>>(data);
    return launch_status("histogramImageKernel::reset");
}

gcam_status gcam_histogram_image(const gcam_image* in, gcam_histogram_data* data, void* stream) {
    if (!valid_image(in, GCAM_PIXEL_RGBA32F) || !data) return bad("gcam_histogram_image: RGBA32F image and a histogram");
    const size_t n = (size_t)in->width * in->height;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 1024);
    k_histogram<<<blocks, 256, 0, (hipStream_t)stream>>>(view<const float4>(*in), data->histogram);
    return launch_status("histogramImage");
}

gcam_status gcam_histogram_statistics(gcam_histogram_data* data, int32_t width, int32_t height, void* stream) {
    if (!data || width <= 0 || height <= 0) return bad("gcam_histogram_statistics: histogram and image size");
    k_histogram_statistics<<<1, 64, 0, (hipStream_t)stream>>>(data, (uint32_t)width * (uint32_t)height);
    return launch_status("histogramStatistics");
}


gcam_status gcam_guided_filter_ab_image(const gcam_image* guide, float eps, gcam_image* ab, void* stream) {
    if (!valid_image(guide, GCAM_PIXEL_RGBA32F) || !valid_image(ab, GCAM_PIXEL_RG32F) || !same_size(guide, ab))
        return bad("gcam_guided_filter_ab_image: RGBA32F guide and RG32F output of one size");
    k_guided_filter_ab<<<grid_for(ab->width, ab->height, kBlock), kBlock, 0, (hipStream_t)stream>>>(
        view<const float4>(*guide), eps, view<float2>(*ab));
    return launch_status("GuidedFilterABImage");
}

gcam_status gcam_box_filter_gf_image(const gcam_image* in, gcam_image* out, void* stream) {
    if (!valid_image(in, GCAM_PIXEL_RG32F) || !valid_image(out, GCAM_PIXEL_RG32F) || !same_size(in, out) ||
        in->data == out->data)
        return bad("gcam_box_filter_gf_image: two distinct RG32F images of one size");
    k_box_filter_gf<<<grid_for(out->width, out->height, kBlock), kBlock, 0, (hipStream_t)stream>>>(
        view<const float2>(*in), view<float2>(*out));
    return launch_status("BoxFilterGFImage");
}

gcam_status gcam_local_tone_mapping_mask_image(const gcam_image* image, const gcam_image* gradient,
                                               const gcam_image* ab_mean_lf, const gcam_image* ab_mean_mf,
                                               const gcam_image* ab_mean_hf, const gcam_ltm_params* params,
                                               float nlf_a, float nlf_b, const gcam_histogram_data* histogram,
                                               gcam_image* mask, void* stream) {
    if (!valid_image(image, GCAM_PIXEL_RGBA32F) || !valid_image(gradient, GCAM_PIXEL_RG32F) ||
        !valid_image(mask, GCAM_PIXEL_R32F) || !same_size(image, gradient) || !same_size(image, mask) || !params ||
        !histogram)
        return bad("gcam_local_tone_mapping_mask_image: RGBA32F image, RG32F gradient, R32F mask of one size");
    const gcam_image* b[3] = {ab_mean_lf, ab_mean_mf, ab_mean_hf};
    LtmBands bands;
    for (int k = 0; k < 3; k++) {
        if (!valid_image(b[k], GCAM_PIXEL_RG32F)) return bad("gcam_local_tone_mapping_mask_image: RG32F band images");
        bands.b[k] = view<const float2>(*b[k]);
        bands.detail[k] = params->detail[k];
    }
    k_ltm_mask<<<grid_for(mask->width, mask->height, kBlock), kBlock, 0, (hipStream_t)stream>>>(
        view<const float4>(*image), view<const float2>(*gradient), bands, nlf_a, nlf_b, histogram, view<float>(*mask));
    return launch_status("localToneMappingMaskImage");
}

// ---- the front end as one call -------------------------------------------

struct gcam_front_end {
    int w, h, device;
    float* scaled;
    char* nlf_scratch;      // raw statistics images, fit partials, state, output
    float raw_a[4], raw_b[4];  // the raw noise model of the last run
};

gcam_status gcam_front_end_create(int32_t width, int32_t height, gcam_front_end** out) {
    if (!out) return bad("gcam_front_end_create: null output");
    *out = nullptr;
    if (width < 2 || height < 2 || ((width | height) & 1) || (int64_t)width * height > (int64_t)1 << 30)
        return bad("gcam_front_end_create: width and height must be even, >= 2");
    auto* f = new gcam_front_end{};
    f->w = width;
    f->h = height;
    hipGetDevice(&f->device);
    const size_t n = (size_t)width * height;
    if (hipMalloc(&f->scaled, n * sizeof(float)) != hipSuccess) {
        gcam_front_end_destroy(f);
        return set_error(GCAM_ERR_OUT_OF_MEMORY, "gcam_front_end_create: device allocation failed");
    }
    *out = f;
    return GCAM_OK;
}

void gcam_front_end_destroy(gcam_front_end* f) {
    if (!f) return;
    hipFree(f->scaled);
    hipFree(f->nlf_scratch);
    delete f;
}

gcam_status gcam_front_end_run(gcam_front_end* f, const gcam_image* raw, const gcam_front_end_params* p,
                               gcam_image* ycbcr, gcam_image* gradient, void* stream) {
    if (!f || !p) return bad("gcam_front_end_run: null handle or parameters");
    if (!valid_image(raw, GCAM_PIXEL_R16U) || raw->width != f->w || raw->height != f->h)
        return bad("gcam_front_end_run: raw must be R16U of the handle's size");
    if (!valid_image(ycbcr, GCAM_PIXEL_RGBA32F) || !same_size(raw, ycbcr) || !valid_image(gradient, GCAM_PIXEL_RG32F) ||
        !same_size(raw, gradient))
        return bad("gcam_front_end_run: ycbcr (RGBA32F) and gradient (RG32F) must have the raw's size");
    if (p->bayer_pattern < 0 || p->bayer_pattern > 3) return bad("gcam_front_end_run: bad Bayer pattern");
    const hipStream_t s = (hipStream_t)stream;
    const int w = f->w, h = f->h;
    gcam_image sc{f->scaled, w, h, w, GCAM_PIXEL_R32F};
    gcam_status st = gcam_scale_raw_data(raw, &sc, p->bayer_pattern, p->scale_mul, p->black_level, p->lens_shading, s);
    if (st != GCAM_OK) return st;
    for (int c = 0; c < 4; c++) {
        f->raw_a[c] = p->raw_nlf_a[c];
        f->raw_b[c] = p->raw_nlf_b[c];
    }
    if (p->calibrate_raw_nlf) {
        // MeasureRawNLF on the scaled raw (raw_converter.cpp:187-189): statistics
        // and fit on the device, then one synchronization for the fitted model.
        const int qw = w / 2, qh = h / 2;
        const size_t img = (size_t)qw * qh * sizeof(float4), part = (size_t)kRawNlfBlocks * kRawNlfValues * sizeof(double);
        if (!f->nlf_scratch && hipMalloc(&f->nlf_scratch, 2 * img + part + sizeof(RawNlfState) + 8 * sizeof(double)) !=
                                   hipSuccess) {
            f->nlf_scratch = nullptr;
            return set_error(GCAM_ERR_OUT_OF_MEMORY, "gcam_front_end_run: device allocation failed");
        }
        char* buf = f->nlf_scratch;
        View<float4> mean{reinterpret_cast<float4*>(buf), qw, qh, qw}, var{reinterpret_cast<float4*>(buf + img), qw, qh, qw};
        double* out = reinterpret_cast<double*>(buf + 2 * img + part + sizeof(RawNlfState));
        if ((st = raw_nlf_device(View<const float>{f->scaled, w, h, w}, quad_of(p->bayer_pattern), mean, var,
                                 reinterpret_cast<double*>(buf + 2 * img),
                                 reinterpret_cast<RawNlfState*>(buf + 2 * img + part), out, p->exposure_multiplier, s)) !=
            GCAM_OK)
            return st;
        double host[8];
        hipError_t e = hipMemcpyAsync(host, out, sizeof(host), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return set_error(GCAM_ERR_HIP, std::string("gcam_front_end_run: ") + hipGetErrorString(e));
        for (int c = 0; c < 4; c++) {
            f->raw_a[c] = (float)host[c];
            f->raw_b[c] = (float)host[4 + c];
        }
    }
    float rv[6];
    gcam_raw_variance(f->raw_a, f->raw_b, rv);
    if ((st = blur_sobel(&sc, nullptr, 1.5f, 4.5f, rv + 2, gradient, s)) != GCAM_OK) return st;
    Mat3 t;
    gcam_cam_ycbcr(p->rgb_cam, t.m);
    const View<const float> scaled_v{f->scaled, w, h, w};
    const Quad q = quad_of(p->bayer_pattern);
    if (p->despeckle)
        return demosaic_tiles<true>(scaled_v, view<const float2>(*gradient), view<float4>(*ycbcr), q, rv,
                                    p->highlight_clip, t, Var2{p->despeckle_nlf.a[0], p->despeckle_nlf.b[0]}, s);
    return demosaic_tiles<false>(scaled_v, view<const float2>(*gradient), view<float4>(*ycbcr), q, rv,
                                 p->highlight_clip, t, Var2{}, s);
}

gcam_status gcam_front_end_raw_nlf(const gcam_front_end* f, float raw_nlf_a[4], float raw_nlf_b[4]) {
    if (!f || !raw_nlf_a || !raw_nlf_b) return bad("gcam_front_end_raw_nlf: null argument");
    for (int c = 0; c < 4; c++) {
        raw_nlf_a[c] = f->raw_a[c];
        raw_nlf_b[c] = f->raw_b[c];
    }
    return GCAM_OK;
}

}  // extern "C"
