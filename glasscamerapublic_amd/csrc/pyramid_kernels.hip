// Pyramid resampling, Laplacian reconstruct and the noise-model fit for gfx950.
//
//   downsample      MetalPipeline/demosaic.metal:886-920
//   subtract_noise  MetalPipeline/demosaic.metal:922-959
//   noise_stats     MetalPipeline/demosaic.metal:1138-1159
//   nlf_fit         MetalPipeline/pyramid_processor.cpp:186-323
//
// The resampling kernels are HBM-bound: one thread per output pixel, 64-wide
// rows per wavefront so every load and store is a coalesced 16-byte (float4)
// or 8-byte (float2) access; the neighbouring taps hit L1/L2.
#include "gcam_internal.hpp"
#include "burst_tables.hpp"
#include "moments.hpp"

namespace gcam {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// One axis of a linear-filtered, clamp-to-edge sample at texel position
// num/den: u = num/den - 1/2 = (2 num - den) / (2 den), taps floor(u) and
// floor(u)+1 clamped to the image, weight of the second tap frac(u).  All in
// exact integer arithmetic, so power-of-two pyramids get exact 1/2, 1/4, 3/4.
struct Tap {
    int i0, i1;
    float f;
};
__device__ __forceinline__ Tap lin_tap(int num, int den, int size) {
    int n2 = 2 * num - den, d2 = 2 * den;
    int fl = n2 >= 0 ? n2 / d2 : -((-n2 + d2 - 1) / d2);
    int rem = n2 - fl * d2;
    return Tap{clampi(fl, 0, size - 1), clampi(fl + 1, 0, size - 1), (float)rem / (float)d2};
}

template <class P>
__device__ __forceinline__ float3 bilerp3(const View<const P>& im, const Tap& tx, const Tap& ty);

template <>
__device__ __forceinline__ float3 bilerp3<float4>(const View<const float4>& im, const Tap& tx, const Tap& ty) {
    float4 t00 = im.at(tx.i0, ty.i0), t01 = im.at(tx.i1, ty.i0);
    float4 t10 = im.at(tx.i0, ty.i1), t11 = im.at(tx.i1, ty.i1);
    const float a = tx.f, b = ty.f;
    float3 top = make_float3((1.0f - a) * t00.x + a * t01.x, (1.0f - a) * t00.y + a * t01.y,
                             (1.0f - a) * t00.z + a * t01.z);
    float3 bot = make_float3((1.0f - a) * t10.x + a * t11.x, (1.0f - a) * t10.y + a * t11.y,
                             (1.0f - a) * t10.z + a * t11.z);
    return make_float3((1.0f - b) * top.x + b * bot.x, (1.0f - b) * top.y + b * bot.y,
                       (1.0f - b) * top.z + b * bot.z);
}

__device__ __forceinline__ float2 bilerp2(const View<const float2>& im, const Tap& tx, const Tap& ty) {
    float2 t00 = im.at(tx.i0, ty.i0), t01 = im.at(tx.i1, ty.i0);
    float2 t10 = im.at(tx.i0, ty.i1), t11 = im.at(tx.i1, ty.i1);
    const float a = tx.f, b = ty.f;
    float2 top = make_float2((1.0f - a) * t00.x + a * t01.x, (1.0f - a) * t00.y + a * t01.y);
    float2 bot = make_float2((1.0f - a) * t10.x + a * t11.x, (1.0f - a) * t10.y + a * t11.y);
    return make_float2((1.0f - b) * top.x + b * bot.x, (1.0f - b) * top.y + b * bot.y);
}

// downsampleImageXYZ: output o averages 4 linear samples at texel positions
// (o + k) * in/out, k in {0, 1}, per axis (the reference's +-1/2 output-texel
// offsets around the output centre).
__device__ __forceinline__ void downsample4_px(const View<const float4>& in, const View<float4>& out, int ox, int oy) {
    const Tap ty0 = lin_tap(oy * in.h, out.h, in.h), ty1 = lin_tap((oy + 1) * in.h, out.h, in.h);
    const Tap tx0 = lin_tap(ox * in.w, out.w, in.w), tx1 = lin_tap((ox + 1) * in.w, out.w, in.w);
    float3 s = bilerp3(in, tx0, ty0);
    float3 acc = s;
    s = bilerp3(in, tx1, ty0);
    acc.x += s.x, acc.y += s.y, acc.z += s.z;
    s = bilerp3(in, tx0, ty1);
    acc.x += s.x, acc.y += s.y, acc.z += s.z;
    s = bilerp3(in, tx1, ty1);
    acc.x += s.x, acc.y += s.y, acc.z += s.z;
    out.at(ox, oy) = make_float4(0.25f * acc.x, 0.25f * acc.y, 0.25f * acc.z, 0.0f);
}

// RO vertically adjacent outputs oy0 .. oy0 + RO - 1 of column ox.  The
// second vertical sample of output oy is the first of output oy + 1 (both at
// texel row (oy + 1) * in.h / out.h), so RO outputs take RO + 1 sample rows,
// 8 (RO + 1) texel loads instead of 16 RO; each output adds its four samples
// in the order of downsample4_px (identical results).
template <int RO>
__device__ __forceinline__ void downsample4_col(const View<const float4>& in, const View<float4>& out, int ox,
                                                int oy0) {
    const Tap tx0 = lin_tap(ox * in.w, out.w, in.w), tx1 = lin_tap((ox + 1) * in.w, out.w, in.w);
    float3 s0[RO + 1], s1[RO + 1];
#pragma unroll
    for (int b = 0; b <= RO; b++) {
        const Tap ty = lin_tap((oy0 + b) * in.h, out.h, in.h);
        s0[b] = bilerp3(in, tx0, ty);
        s1[b] = bilerp3(in, tx1, ty);
    }
#pragma unroll
    for (int r = 0; r < RO; r++) {
        if (oy0 + r >= out.h) break;
        float3 acc = s0[r];
        acc.x += s1[r].x, acc.y += s1[r].y, acc.z += s1[r].z;
        acc.x += s0[r + 1].x, acc.y += s0[r + 1].y, acc.z += s0[r + 1].z;
        acc.x += s1[r + 1].x, acc.y += s1[r + 1].y, acc.z += s1[r + 1].z;
        out.at(ox, oy0 + r) = make_float4(0.25f * acc.x, 0.25f * acc.y, 0.25f * acc.z, 0.0f);
    }
}

// The per-pixel kernels run 64 x 4 blocks in XCD-contiguous order: the input
// rows that vertically adjacent blocks share stay in one XCD's L2.  The
// downsample's threads take kDownRows outputs of a column each (blocks of
// 64 x 4 kDownRows outputs).
constexpr int kDownRows = 4;
__global__ __launch_bounds__(256) void k_downsample4(View<const float4> in, View<float4> out, XcdTiles order) {
    int tx, ty;
    if (!order.tile(tx, ty)) return;
    const int ox = tx * 64 + threadIdx.x, oy = (ty * 4 + threadIdx.y) * kDownRows;
    if (ox < out.w && oy < out.h) downsample4_col<kDownRows>(in, out, ox, oy);
}

// downsampleImageXY, the 2-channel gradient version.
__device__ __forceinline__ void downsample2_px(const View<const float2>& in, const View<float2>& out, int ox, int oy) {
    const Tap ty0 = lin_tap(oy * in.h, out.h, in.h), ty1 = lin_tap((oy + 1) * in.h, out.h, in.h);
    const Tap tx0 = lin_tap(ox * in.w, out.w, in.w), tx1 = lin_tap((ox + 1) * in.w, out.w, in.w);
    float2 acc = bilerp2(in, tx0, ty0);
    float2 s = bilerp2(in, tx1, ty0);
    acc.x += s.x, acc.y += s.y;
    s = bilerp2(in, tx0, ty1);
    acc.x += s.x, acc.y += s.y;
    s = bilerp2(in, tx1, ty1);
    acc.x += s.x, acc.y += s.y;
    out.at(ox, oy) = make_float2(0.25f * acc.x, 0.25f * acc.y);
}

__global__ __launch_bounds__(256) void k_downsample2(View<const float2> in, View<float2> out, XcdTiles order) {
    int tx, ty;
    if (!order.tile(tx, ty)) return;
    const int ox = tx * 64 + threadIdx.x, oy = ty * 4 + threadIdx.y;
    if (ox < out.w && oy < out.h) downsample2_px(in, out, ox, oy);
}

// subtractNoiseImage: Laplacian-style reconstruct of level i from the noisy
// and the denoised level i+1 (both upsampled bilinearly), with optional
// gradient-driven sharpening.
__device__ __forceinline__ float4 subtract_noise_value(const View<const float4>& in, const View<const float4>& in1,
                                                      const View<const float4>& den1, const View<const float2>& grad,
                                                      float luma_weight, float sharpening, float nlf_x, float nlf_y,
                                                      const NlfDev* nlf_level, float thr_y, const View<float4>& out,
                                                      int ox, int oy) {
    const Tap tyy = lin_tap((2 * oy + 1) * in1.h, 2 * out.h, in1.h);
    const Tap txx = lin_tap((2 * ox + 1) * in1.w, 2 * out.w, in1.w);
    const float4 ip = in.at(ox, oy);
    const float3 p1 = bilerp3(in1, txx, tyy);
    const float3 d1 = bilerp3(den1, txx, tyy);
    float3 dp = make_float3(ip.x - luma_weight * (p1.x - d1.x), ip.y - (p1.y - d1.y), ip.z - (p1.z - d1.z));
    float alpha = sharpening;
    if (alpha > 1.0f) {
        if (nlf_level) {
            nlf_x = nlf_level->a[0] * thr_y;
            nlf_y = nlf_level->b[0] * thr_y;
        }
        float gradient = 0.0f;
        if (grad.ptr) {
            const float2 g = grad.at(ox, oy);
            gradient = sqrtf(g.x * g.x + g.y * g.y);
        }
        const float sigma = sqrtf(nlf_x + nlf_y * d1.x);
        const float detail = smoothstepf(sigma, 4.0f * sigma, gradient) *
                             (1.0f - smoothstepf(0.75f, 0.95f, d1.x)) * smoothstepf(0.05f, 0.1f, d1.x);
        alpha = 1.0f + (alpha - 1.0f) * detail;
    }
    float4 o;
    o.x = fmaxf(d1.x + (dp.x - d1.x) * alpha, 0.0f);
    o.y = d1.y + (dp.y - d1.y) * alpha;
    o.z = d1.z + (dp.z - d1.z) * alpha;
    o.w = ip.w;
    return o;
}

__device__ __forceinline__ float subtract_noise_px(const View<const float4>& in, const View<const float4>& in1,
                                                  const View<const float4>& den1, const View<const float2>& grad,
                                                  float luma_weight, float sharpening, float nlf_x, float nlf_y,
                                                  const NlfDev* nlf_level, float thr_y, const View<float4>& out,
                                                  int ox, int oy) {
    const float4 o = subtract_noise_value(in, in1, den1, grad, luma_weight, sharpening, nlf_x, nlf_y, nlf_level,
                                          thr_y, out, ox, oy);
    out.at(ox, oy) = o;
    return o.x;
}

// Subtract and PCA moments in one pass over a level (the sample is never
// written): 64 x 8-pixel tiles anchored at (64 tx - 2, 8 ty - 2) hold whole
// patches -- patch (8 tx + p, ty), p = 0..7, is tile rows 0..4, columns
// 8 p .. 8 p + 4 -- and a fixed grid of nblk workgroups (one moment partial
// each) walks the tiles.  Workgroup b belongs to the tile range of XCD b % 8
// (a contiguous eighth of the tiles, so neighbouring tiles, which share rows
// of the coarse levels, meet in that XCD's L2) and takes every nw-th tile of
// it, nw the number of workgroups of that XCD.  A thread computes two pixels
// of a tile (rows y and y + 4); pixels outside the image are computed at the
// clamped position (which is what the clamp-to-edge patch sampling reads) and
// not written.  The luma of tile rows 0..4 goes to LDS, in two buffers: the
// barrier that publishes a tile's rows also retires the reads of the tile
// before.  Wavefronts 0 and 1 then add patches 0-3 and 4-7 of the tile to
// their FP64 Gram accumulators.  The tiles of a workgroup are taken in
// increasing order and the wavefronts' sums added in a fixed order, so the
// partial of a workgroup is deterministic, and the burst form, which gives a
// frame the same workgroups and walk, produces the same partials.
struct SubMomArgs {
    View<const float4> in, in1, den1;
    View<const float2> grad;
    float luma_weight, sharpening, nlf_x, nlf_y;
    const NlfDev* nlf_level;
    float thr_y;
    View<float4> out;
};

__device__ __forceinline__ void subtract_moments_body(const SubMomArgs& a, int b, int nblk,
                                                      double* __restrict__ partial) {
    __shared__ float lum[2][5][64];
    __shared__ double gram[2][32][33];
    __shared__ double sums[2][64][2];
    const int tid = threadIdx.y * 64 + threadIdx.x, lane = tid & 63, wave = threadIdx.y;
    const int W = a.out.w, H = a.out.h, gw = W / 8, gh = H / 8;
    const int tiles_x = submom_tiles_x(W), tiles = submom_tiles(W, H);
    const int v = b & 7, chunk = (tiles + 7) >> 3, nw = (nblk - v + 7) >> 3;
    const int end = min(tiles, (v + 1) * chunk);
    // this lane's patch values: feature f (row f / 5, column f % 5) and f + 16
    const int f = lane & 15, kq = lane >> 4;
    const int r0 = f / 5, c0 = f - 5 * r0, r1 = (f + 16) / 5, c1 = (f + 16) - 5 * r1;
    f64x4 c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
    double slo = 0.0, shi = 0.0;
    int buf = 0;
    for (int t = v * chunk + (b >> 3); t < end; t += nw) {
        const int tx = t % tiles_x, ty = t / tiles_x;
        const int cx = 64 * tx - 2 + (int)threadIdx.x, sx = clampi(cx, 0, W - 1);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int r = (int)threadIdx.y + 4 * h;
            const int cy = 8 * ty - 2 + r, sy = clampi(cy, 0, H - 1);
            const float4 o = subtract_noise_value(a.in, a.in1, a.den1, a.grad, a.luma_weight, a.sharpening, a.nlf_x,
                                                  a.nlf_y, a.nlf_level, a.thr_y, a.out, sx, sy);
            if (cx == sx && cy == sy) a.out.at(sx, sy) = o;
            if (r < 5) lum[buf][r][threadIdx.x] = o.x;
        }
        __syncthreads();
        if (wave < 2 && ty < gh && 8 * tx + 4 * wave < gw) {  // the group holds a patch of the sample
            const int p = 4 * wave + kq;  // patch of this lane's row of the group
            const bool ok = 8 * tx + p < gw;
            const double lo = ok ? (double)lum[buf][r0][8 * p + c0] : 0.0;
            const double hi = ok && f < kPatch - 16 ? (double)lum[buf][r1][8 * p + c1] : 0.0;
            moments_group(lo, hi, c00, c01, c11, slo, shi);
        }
        buf ^= 1;
    }
    moments_store<2>(tid, 256, c00, c01, c11, slo, shi, gram, sums, partial);
}

__global__ __launch_bounds__(256) void k_subtract_moments(SubMomArgs a, double* partials) {
    subtract_moments_body(a, blockIdx.x, gridDim.x, partials + (size_t)blockIdx.x * kMoments);
}

__global__ __launch_bounds__(256) void k_subtract_noise(View<const float4> in, View<const float4> in1,
                                                        View<const float4> den1, View<const float2> grad,
                                                        float luma_weight, float sharpening, float nlf_x,
                                                        float nlf_y, const NlfDev* nlf_level, float thr_y,
                                                        View<float4> out, XcdTiles order) {
    int tx, ty;
    if (!order.tile(tx, ty)) return;
    const int ox = tx * 64 + threadIdx.x, oy = ty * 4 + threadIdx.y;
    if (ox < out.w && oy < out.h)
        subtract_noise_px(in, in1, den1, grad, luma_weight, sharpening, nlf_x, nlf_y, nlf_level, thr_y, out, ox, oy);
}

// Burst forms: one launch for the level of every frame; block b covers the
// 64 x 4 block b - block0 of its frame (row-major, bx blocks a row).  Blocks
// run in XCD-contiguous order (`order` spans all frames' blocks), so the rows
// that vertically adjacent blocks share stay in one XCD's L2.
__global__ __launch_bounds__(256) void k_burst_downsample(const BurstTable<BurstDownFrame>* __restrict__ tab,
                                                          XcdTiles order) {
    int b, unused;
    if (!order.tile(b, unused)) return;
    const BurstDownFrame& f = tab->frame[burst_frame_of(tab, b)];
    const int lb = b - f.block0;
    const int ox = (lb % f.bx) * 64 + threadIdx.x, oy = (lb / f.bx) * 4 + threadIdx.y;
    if (ox < f.out.w && oy < f.out.h) {
        downsample4_px(f.in, f.out, ox, oy);
        if (f.gin.ptr) downsample2_px(f.gin, f.gout, ox, oy);
    }
}

// Burst form of the fused pass: frame f owns workgroups block0 .. block0 +
// nblk - 1 (block0 a multiple of 8, so a workgroup's XCD is that of the
// single-frame pass's workgroup of the same index); workgroups between
// frames have no frame and return.
__global__ __launch_bounds__(256) void k_burst_subtract_moments(const BurstTable<BurstSubFrame>* __restrict__ tab) {
    const BurstSubFrame& f = tab->frame[burst_frame_of(tab, blockIdx.x)];
    const int lb = (int)blockIdx.x - f.block0;
    if (lb >= f.nblk) return;
    const SubMomArgs a{f.in, f.in1, f.den1, f.grad, 1.0f, f.sharpening, 0.0f, 0.0f, f.nlf, f.thr_y, f.out};
    subtract_moments_body(a, lb, f.nblk, f.partials + (size_t)lb * kMoments);
}

// basicNoiseStatistics: 5x5 mean of Y and population variances of Y, Cb, Cr.
__global__ __launch_bounds__(256) void k_noise_stats(View<const float4> in, View<float4> out) {
    const int px = blockIdx.x * 64 + threadIdx.x, py = blockIdx.y * 4 + threadIdx.y;
    if (px >= out.w || py >= out.h) return;
    float3 sum = make_float3(0, 0, 0), sq = make_float3(0, 0, 0);
    for (int y = -2; y <= 2; y++) {
        const int sy = clampi(py + y, 0, in.h - 1);
        for (int x = -2; x <= 2; x++) {
            const float4 s = in.at(clampi(px + x, 0, in.w - 1), sy);
            sum.x += s.x, sum.y += s.y, sum.z += s.z;
            sq.x += s.x * s.x, sq.y += s.y * s.y, sq.z += s.z * s.z;
        }
    }
    float4 o;
    o.x = sum.x / 25.0f;
    o.y = (sq.x - (sum.x * sum.x) / 25.0f) / 25.0f;
    o.z = (sq.y - (sum.y * sum.y) / 25.0f) / 25.0f;
    o.w = (sq.z - (sum.z * sum.z) / 25.0f) / 25.0f;
    out.at(px, py) = o;
}

// ---------------- noise-model fit (FP64) ----------------

constexpr int kNlfMaxValues = 13;

__device__ __forceinline__ double floor_max(double v, double fl) { return isnan(v) ? fl : fmax(v, fl); }

__device__ __forceinline__ bool nlf_selected(const float4 ns, const double vmax[3], double& m, double v[3]) {
    m = ns.x;
    v[0] = ns.y;
    v[1] = ns.z;
    v[2] = ns.w;
    const bool valid = !(isnan(m) || isnan(v[0]) || isnan(v[1]) || isnan(v[2]));
    return valid && m >= 0.001 && m <= 0.9 && v[0] <= vmax[0] && v[1] <= vmax[1] && v[2] <= vmax[2];
}

// Block-wide deterministic sum of `nv` per-thread doubles; thread 0 writes them.
template <int NV>
__device__ void block_sum_write(double (&acc)[NV], double* dst) {
    __shared__ double red[256];
    for (int v = 0; v < NV; v++) {
        red[threadIdx.x] = acc[v];
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) dst[v] = red[0];
        __syncthreads();
    }
}

template <int PASS>
struct NlfPassWidth;
template <>
struct NlfPassWidth<0> { static constexpr int value = 6; };
template <>
struct NlfPassWidth<1> { static constexpr int value = 9; };
template <>
struct NlfPassWidth<2> { static constexpr int value = 3; };
template <>
struct NlfPassWidth<3> { static constexpr int value = 13; };

// One statistics pass over every pixel; per-block partial sums.
template <int PASS>
__global__ __launch_bounds__(256) void k_nlf_pass(View<const float4> stats, const NlfFitState* st,
                                                  double* partials) {
    constexpr int NV = NlfPassWidth<PASS>::value;
    double acc[NV];
    for (int v = 0; v < NV; v++) acc[v] = 0.0;
    double vmax[3], A[3], B[3], e2[3];
    for (int c = 0; c < 3; c++) {
        vmax[c] = st->varianceMax[c];
        A[c] = st->A[c];
        B[c] = st->B[c];
        e2[c] = st->err2[c];
    }
    const size_t npx = (size_t)stats.w * stats.h;
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < npx; k += (size_t)gridDim.x * 256) {
        const int x = (int)(k % stats.w), y = (int)(k / stats.w);
        const float4 ns = stats.at(x, y);
        if (PASS == 0) {  // pyramid_processor.cpp:190-201
            const double v0 = ns.y, v1 = ns.z, v2 = ns.w;
            const bool valid = !isnan(v0) && !isnan(v1) && !isnan(v2) && v0 > 0.0 && v1 > 0.0 && v2 > 0.0;
            if (valid) {
                double mx = fmax(fmax(log10(v0), log10(v1)), log10(v2));
                mx = fmin(fmax(mx, -6.0), -1.0);
                const int entry = 6 + (int)mx;
                for (int e = 0; e < 6; e++) acc[e] += (e == entry) ? 1.0 : 0.0;
            }
        } else {
            double m, v[3];
            if (!nlf_selected(ns, vmax, m, v)) continue;
            if (PASS == 1) {  // :223-236
                acc[0] += m;
                acc[1] += m * m;
                for (int c = 0; c < 3; c++) {
                    acc[2 + c] += v[c];
                    acc[5 + c] += m * v[c];
                }
                acc[8] += 1.0;
            } else if (PASS == 2) {  // :243-256
                for (int c = 0; c < 3; c++) {
                    const double diff = A[c] + B[c] * m - v[c];
                    acc[c] += diff * diff;
                }
            } else {  // :269-291
                double dsq[3];
                bool fits = true;
                for (int c = 0; c < 3; c++) {
                    const double diff = fabs(A[c] + B[c] * m - v[c]);
                    dsq[c] = diff * diff;
                    fits = fits && dsq[c] <= e2[c];
                }
                if (fits) {
                    acc[0] += m;
                    acc[1] += m * m;
                    for (int c = 0; c < 3; c++) {
                        acc[2 + c] += v[c];
                        acc[5 + c] += m * v[c];
                        acc[9 + c] += dsq[c];
                    }
                    acc[8] += 1.0;
                } else {
                    acc[12] += 1.0;
                }
            }
        }
    }
    block_sum_write<NV>(acc, partials + (size_t)blockIdx.x * kNlfMaxValues);
}

// Sum the per-block partials (fixed order) and advance the fit state.
template <int PASS>
__global__ __launch_bounds__(256) void k_nlf_finalize(const double* partials, int nblocks, NlfFitState* st,
                                                      float exposure_multiplier, double npx, NlfDev* nlf_out) {
    constexpr int NV = NlfPassWidth<PASS>::value;
    __shared__ double tot[kNlfMaxValues];
    __shared__ double red[256];
    for (int v = 0; v < NV; v++) {
        double a = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += 256) a += partials[(size_t)b * kNlfMaxValues + v];
        red[threadIdx.x] = a;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) tot[v] = red[0];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (PASS == 0) {  // :204-208
        for (int e = 0; e < 6; e++) st->hist[e] = tot[e];
        const double* h = tot;
        const double vm = h[0] > 1e3 ? 1.0e-5 : h[1] > 1e3 ? 1.0e-4 : h[2] > 1e3 ? 1.0e-3
                        : h[3] > 1e3 ? 1.0e-2 : h[4] > 1e3 ? 1.0e-1 : 1.0;
        for (int c = 0; c < 3; c++) st->varianceMax[c] = vm;
    } else if (PASS == 1) {  // :239-240
        const double N = tot[8], s_x = tot[0], s_xx = tot[1];
        for (int c = 0; c < 3; c++) {
            const double B = floor_max((N * tot[5 + c] - s_x * tot[2 + c]) / (N * s_xx - s_x * s_x), 1e-8);
            st->B[c] = B;
            st->A[c] = floor_max((tot[2 + c] - B * s_x) / N, 1e-8);
        }
        st->n1 = N;
    } else if (PASS == 2) {  // :256, :259
        for (int c = 0; c < 3; c++) {
            st->err2[c] = tot[c] / st->n1;
            st->varianceMax[c] = st->B[c];
        }
    } else {  // :292-319
        const double N = tot[8], s_x = tot[0], s_xx = tot[1];
        bool better = true;
        for (int c = 0; c < 3; c++) {
            st->newErr2[c] = tot[9 + c] / N;
            better = better && st->newErr2[c] <= st->err2[c];
        }
        st->n2 = N;
        st->discarded = tot[12];
        st->accepted = 0;
        if (better && N / npx > 0.01) {
            for (int c = 0; c < 3; c++) {
                const double B = floor_max((N * tot[5 + c] - s_x * tot[2 + c]) / (N * s_xx - s_x * s_x), 1e-8);
                st->B[c] = B;
                st->A[c] = floor_max((tot[2 + c] - B * s_x) / N, 1e-8);
            }
            st->accepted = 1;
        }
        const double adj = (double)exposure_multiplier * (double)exposure_multiplier;
        for (int c = 0; c < 3; c++) {
            nlf_out->a[c] = (float)(st->A[c] * adj);
            nlf_out->b[c] = (float)(st->B[c] * adj);
        }
    }
}

__global__ void k_nlf_init(NlfFitState* st) {
    if (threadIdx.x == 0) {
        for (int c = 0; c < 3; c++) st->varianceMax[c] = st->A[c] = st->B[c] = st->err2[c] = 0.0;
    }
}

struct NlfLevels {
    NlfDev lv[GCAM_MAX_LEVELS];
};

__global__ void k_set_nlf(NlfLevels v, int levels, NlfDev* dev) {
    const int i = threadIdx.x;
    if (i < levels) dev[i] = v.lv[i];
}

inline dim3 grid64x4(int w, int h) { return dim3((w + 63) / 64, (h + 3) / 4); }

}  // namespace

gcam_status launch_downsample4(View<const float4> in, View<float4> out, hipStream_t s) {
    const XcdTiles order = xcd_tiles(out.w, out.h, 64, 4 * kDownRows);
    k_downsample4<<<order.grid(), dim3(64, 4), 0, s>>>(in, out, order);
    GCAM_LAUNCH_CHECK("downsampleImageXYZ");
    return GCAM_OK;
}

gcam_status launch_downsample2(View<const float2> in, View<float2> out, hipStream_t s) {
    const XcdTiles order = xcd_tiles(out.w, out.h, 64, 4);
    k_downsample2<<<order.grid(), dim3(64, 4), 0, s>>>(in, out, order);
    GCAM_LAUNCH_CHECK("downsampleImageXY");
    return GCAM_OK;
}

gcam_status launch_subtract_noise(View<const float4> in, View<const float4> in1, View<const float4> den1,
                                  View<const float2> grad, float luma_weight, float sharpening, float nlf_x,
                                  float nlf_y, const NlfDev* nlf_level, float thr_y, View<float4> out,
                                  hipStream_t s) {
    const XcdTiles order = xcd_tiles(out.w, out.h, 64, 4);
    k_subtract_noise<<<order.grid(), dim3(64, 4), 0, s>>>(in, in1, den1, grad, luma_weight, sharpening, nlf_x, nlf_y,
                                                          nlf_level, thr_y, out, order);
    GCAM_LAUNCH_CHECK("subtractNoiseImage");
    return GCAM_OK;
}

gcam_status launch_noise_stats(View<const float4> in, View<float4> out, hipStream_t s) {
    k_noise_stats<<<grid64x4(out.w, out.h), dim3(64, 4), 0, s>>>(in, out);
    GCAM_LAUNCH_CHECK("basicNoiseStatistics");
    return GCAM_OK;
}

// partials: kNlfBlocks * 13 doubles of scratch.
gcam_status launch_nlf_fit(View<const float4> stats, float exposure_multiplier, double* partials,
                           NlfFitState* state, NlfDev* nlf_out, hipStream_t s) {
    const double npx = (double)stats.w * stats.h;
    k_nlf_init<<<1, 64, 0, s>>>(state);
    k_nlf_pass<0><<<kNlfBlocks, 256, 0, s>>>(stats, state, partials);
    k_nlf_finalize<0><<<1, 256, 0, s>>>(partials, kNlfBlocks, state, exposure_multiplier, npx, nlf_out);
    k_nlf_pass<1><<<kNlfBlocks, 256, 0, s>>>(stats, state, partials);
    k_nlf_finalize<1><<<1, 256, 0, s>>>(partials, kNlfBlocks, state, exposure_multiplier, npx, nlf_out);
    k_nlf_pass<2><<<kNlfBlocks, 256, 0, s>>>(stats, state, partials);
    k_nlf_finalize<2><<<1, 256, 0, s>>>(partials, kNlfBlocks, state, exposure_multiplier, npx, nlf_out);
    k_nlf_pass<3><<<kNlfBlocks, 256, 0, s>>>(stats, state, partials);
    k_nlf_finalize<3><<<1, 256, 0, s>>>(partials, kNlfBlocks, state, exposure_multiplier, npx, nlf_out);
    GCAM_LAUNCH_CHECK("MeasureYCbCrNLF");
    return GCAM_OK;
}

gcam_status launch_set_nlf(const gcam_ycbcr_nlf* host_levels, int levels, NlfDev* dev, hipStream_t s) {
    NlfLevels v{};
    for (int i = 0; i < levels && i < GCAM_MAX_LEVELS; i++)
        for (int c = 0; c < 3; c++) {
            v.lv[i].a[c] = host_levels[i].a[c];
            v.lv[i].b[c] = host_levels[i].b[c];
        }
    k_set_nlf<<<1, 64, 0, s>>>(v, levels, dev);
    GCAM_LAUNCH_CHECK("set noise model");
    return GCAM_OK;
}

gcam_status launch_burst_downsample(const BurstTable<BurstDownFrame>* dev, int total_blocks, hipStream_t s) {
    if (total_blocks < 1) return GCAM_OK;
    const XcdTiles order = xcd_tiles(total_blocks, 1, 1, 1);
    k_burst_downsample<<<order.grid(), dim3(64, 4), 0, s>>>(dev, order);
    GCAM_LAUNCH_CHECK("burst downsampleImageXYZ");
    return GCAM_OK;
}

gcam_status launch_subtract_moments(View<const float4> in, View<const float4> in1, View<const float4> den1,
                                    View<const float2> grad, float luma_weight, float sharpening, float nlf_x,
                                    float nlf_y, const NlfDev* nlf_level, float thr_y, View<float4> out,
                                    double* partials, int* nblocks, hipStream_t s) {
    *nblocks = submom_blocks(out.w, out.h);
    const SubMomArgs a{in, in1, den1, grad, luma_weight, sharpening, nlf_x, nlf_y, nlf_level, thr_y, out};
    k_subtract_moments<<<*nblocks, dim3(64, 4), 0, s>>>(a, partials);
    GCAM_LAUNCH_CHECK("subtractNoiseImage with PCA moments");
    return GCAM_OK;
}

gcam_status launch_burst_subtract_moments(const BurstTable<BurstSubFrame>* dev, int total_blocks, hipStream_t s) {
    if (total_blocks < 1) return GCAM_OK;
    k_burst_subtract_moments<<<total_blocks, dim3(64, 4), 0, s>>>(dev);
    GCAM_LAUNCH_CHECK("burst subtractNoiseImage with PCA moments");
    return GCAM_OK;
}


}  // namespace gcam
