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
#include <type_traits>

namespace gcam {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// f(std::integral_constant<int, i>) for i = I .. E (nothing when E < I).
template <int I, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I <= E) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, E>(f);
    }
}

__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// directionWeight per squared offset |q - p|^2 (0 .. 2 R^2), and its log2.
struct DirTable {
    float w[2 * GCAM_MAX_SEARCH_RADIUS * GCAM_MAX_SEARCH_RADIUS + 1];
    float lw[2 * GCAM_MAX_SEARCH_RADIUS * GCAM_MAX_SEARCH_RADIUS + 1];
};

constexpr float kLog2e = 1.4426950408889634f;

// Tile of one workgroup: TW x TH threads, each owning PV vertically adjacent
// output pixels, so the tile is TW x (TH * PV) pixels plus an R-pixel apron.
template <int R, int PV, int TW, int TH>
struct BmTile {
    static constexpr int LW = TW + 2 * R;
    static constexpr int LH = TH * PV + 2 * R;
    static constexpr int LN = LW * LH;
    static constexpr int THREADS = TW * TH;
    static_assert(LN * 48 <= 160 * 1024, "LDS tile larger than a CU's");
};

template <int R, int PV, int TW, int TH>
__global__ __launch_bounds__(TW* TH) void k_block_match(View<const float4> in, View<const float2> grad,
                                                        const float* __restrict__ feat, int feat_stride,
                                                        const NlfDev* __restrict__ nlf, BmParams p, DirTable dir,
                                                        View<float4> out, XcdTiles order) {
    using T = BmTile<R, PV, TW, TH>;
    __shared__ float4 sY[T::LN];   // Y, Cb, Cr, (w)
    __shared__ float4 sF0[T::LN];  // features 0..3
    __shared__ float4 sF1[T::LN];  // features 4..7
    int tx, ty;
    if (!order.tile(tx, ty)) return;  // tiles in XCD-contiguous order
    const int x0 = tx * TW, y0 = ty * TH * PV;
    const int tid = threadIdx.y * TW + threadIdx.x;
    for (int e = tid; e < T::LN; e += T::THREADS) {
        const int ly = e / T::LW, lx = e % T::LW;
        const int gx = clampi(x0 + lx - R, 0, in.w - 1), gy = clampi(y0 + ly - R, 0, in.h - 1);
        sY[e] = in.at(gx, gy);
        const float4* f = reinterpret_cast<const float4*>(feat + ((size_t)gy * feat_stride + gx) * 8);
        sF0[e] = f[0];
        sF1[e] = f[1];
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, yb = y0 + threadIdx.y * PV;
    if (x >= out.w || yb >= out.h) return;

    // Per-pixel coefficients (demosaic.metal:820-838), log2(e) folded in:
    //   wl = 2^(kl d2),  wc = 2^(-(sb dCb)^2 - (sr dCr)^2 + kw wl^2 + log2 wd)
    float4 fa[PV], fb[PV];
    float kl[PV], sb[PV], sr[PV], sbc[PV], src[PV];
    float fY[PV], fCb[PV], fCr[PV], nY[PV], nC[PV];
    const float kw = -2.0f * kLog2e, kroot = sqrtf(2.0f * kLog2e);
#pragma unroll
    for (int k = 0; k < PV; k++) {
        const int y = min(yb + k, out.h - 1);
        const int c = (threadIdx.y * PV + k + R) * T::LW + threadIdx.x + R;
        const float4 ycc = sY[c];
        fa[k] = sF0[c];
        fb[k] = sF1[c];
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
        kl[k] = -2.0f * kLog2e * lm * lm;
        sb[k] = kroot * (dm1 / p.chromaBoost);
        sr[k] = kroot * (dm2 / p.chromaBoost);
        sbc[k] = sb[k] * ycc.y;
        src[k] = sr[k] * ycc.z;
        fY[k] = fCb[k] = fCr[k] = nY[k] = nC[k] = 0.0f;
    }

    // Neighbour rows dy = -R .. R+PV-1 relative to the first owned row: each
    // staged neighbour is loaded once and used by every owned pixel whose
    // window contains it.  Summation order per pixel: rows ascending, columns
    // ascending, as the reference's loops (demosaic.metal:858-859).
    // One window row dy against the owned outputs k = KLO .. KHI (those whose
    // window holds the row).
    auto window_row = [&](auto klo, auto khi, int dy, auto check) {
        constexpr int KLO = decltype(klo)::value, KHI = decltype(khi)::value;
        const int rowq = (threadIdx.y * PV + R + dy) * T::LW + threadIdx.x + R;
        // Operands of position dx are loaded while position dx - 1 is computed;
        // the scheduling barrier keeps the compiler from hoisting the loads of
        // the whole row (registers, and so occupancy).
        float4 na = sF0[rowq - R], nb = sF1[rowq - R], ny = sY[rowq - R];
#pragma unroll
        for (int dx = -R; dx <= R; dx++) {
            const float4 qa = na, qb = nb, qy = ny;
            if (dx < R) {
                na = sF0[rowq + dx + 1];
                nb = sF1[rowq + dx + 1];
                ny = sY[rowq + dx + 1];
            }
#pragma unroll
            for (int k = KLO; k <= KHI; k++) {
                const int ddy = dy - k;
                if (decltype(check)::value && (ddy < -R || ddy > R)) continue;  // wave-uniform
                const int o2 = dx * dx + ddy * ddy;
                const float wd = dir.w[o2], lwd = dir.lw[o2];
                float t, d2;
                t = qa.x - fa[k].x; d2 = t * t;
                t = qa.y - fa[k].y; d2 = fmaf(t, t, d2);
                t = qa.z - fa[k].z; d2 = fmaf(t, t, d2);
                t = qa.w - fa[k].w; d2 = fmaf(t, t, d2);
                t = qb.x - fb[k].x; d2 = fmaf(t, t, d2);
                t = qb.y - fb[k].y; d2 = fmaf(t, t, d2);
                t = qb.z - fb[k].z; d2 = fmaf(t, t, d2);
                t = qb.w - fb[k].w; d2 = fmaf(t, t, d2);
                const float wl = __builtin_amdgcn_exp2f(kl[k] * d2);
                const float wY = wd * wl;
                const float tb = fmaf(sb[k], qy.y, -sbc[k]), tr = fmaf(sr[k], qy.z, -src[k]);
                float e = fmaf(-tb, tb, lwd);
                e = fmaf(-tr, tr, e);
                e = fmaf(kw * wl, wl, e);
                const float wC = __builtin_amdgcn_exp2f(e);
                fY[k] = fmaf(wY, qy.x, fY[k]);
                nY[k] += wY;
                fCb[k] = fmaf(wC, qy.y, fCb[k]);
                fCr[k] = fmaf(wC, qy.z, fCr[k]);
                nC[k] += wC;
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // Window rows dy = -R .. R + PV - 1 of the PV outputs; output k takes the
    // rows with dy - k in [-R, R].  With one or two outputs per thread the
    // first and last PV - 1 rows (where only some outputs take part) are
    // unrolled with their outputs fixed at compile time and the band between
    // them, where all take part, is a loop with no per-output test.  With
    // four, the unrolled edge rows would hold too many registers: every row
    // runs through the loop and tests its outputs (a wave-uniform test).
    if constexpr (PV <= 2) {
        constexpr int kMidLo = -R + PV - 1;
        constexpr int kBotLo = R + 1 > kMidLo ? R + 1 : kMidLo;
        static_for<-R, (kMidLo - 1 < R ? kMidLo - 1 : R)>([&](auto d) {
            constexpr int dy = decltype(d)::value;
            window_row(std::integral_constant<int, (dy - R > 0 ? dy - R : 0)>{},
                       std::integral_constant<int, (dy + R < PV - 1 ? dy + R : PV - 1)>{}, dy,
                           std::false_type{});
        });
#pragma unroll 1
        for (int dy = kMidLo; dy <= R; dy++)
            window_row(std::integral_constant<int, 0>{}, std::integral_constant<int, PV - 1>{}, dy, std::false_type{});
        static_for<kBotLo, R + PV - 1>([&](auto d) {
            constexpr int dy = decltype(d)::value;
            window_row(std::integral_constant<int, (dy - R > 0 ? dy - R : 0)>{},
                       std::integral_constant<int, (dy + R < PV - 1 ? dy + R : PV - 1)>{}, dy,
                           std::false_type{});
        });
    } else {
#pragma unroll 1
        for (int dy = -R; dy <= R + PV - 1; dy++)
            window_row(std::integral_constant<int, 0>{}, std::integral_constant<int, PV - 1>{}, dy, std::true_type{});
    }
#pragma unroll
    for (int k = 0; k < PV; k++) {
        if (yb + k < out.h) out.at(x, yb + k) = make_float4(fY[k] / nY[k], fCb[k] / nC[k], fCr[k] / nC[k], nY[k]);
    }
}

template <int R, int PV, int TW, int TH>
gcam_status launch_shape(View<const float4> in, View<const float2> grad, View<const float> feat8, const NlfDev* nlf,
                         const BmParams& p, const DirTable& dir, View<float4> out, hipStream_t s) {
    const XcdTiles order = xcd_tiles(out.w, out.h, TW, TH * PV);
    k_block_match<R, PV, TW, TH><<<order.grid(), dim3(TW, TH), 0, s>>>(in, grad, feat8.ptr, feat8.stride, nlf, p, dir,
                                                                       out, order);
    GCAM_LAUNCH_CHECK("blockMatchingDenoiseImage");
    return GCAM_OK;
}

// Tile shapes: 1 row per thread uses a 32 x 8 workgroup; 2 and 4 rows per
// thread (R <= 6) use 32 x 8 and 32 x 4.  Above radius 6 every setting takes
// one row on 32 x 8 (70 KB of LDS at radius 10, two workgroups per CU; 12 MP
// level 0 at radius 10: 4.55 ms, against 7.41 ms on 32 x 4 workgroups and
// 4.63 ms on 32 x 16).
template <int R>
gcam_status launch_r(View<const float4> in, View<const float2> grad, View<const float> feat8, const NlfDev* nlf,
                     const BmParams& p, const DirTable& dir, int rows, View<float4> out, hipStream_t s) {
    if constexpr (R <= 6) {
        if (rows == 4) return launch_shape<R, 4, 32, 4>(in, grad, feat8, nlf, p, dir, out, s);
        if (rows == 2) return launch_shape<R, 2, 32, 8>(in, grad, feat8, nlf, p, dir, out, s);
    }
    return launch_shape<R, 1, 32, 8>(in, grad, feat8, nlf, p, dir, out, s);
}

DirTable make_dir_table() {
    // demosaic.metal:869: gaussian(0.1 * length(float2(x, y))), FP32.
    DirTable t{};
    for (int d2 = 0; d2 < (int)(sizeof(t.w) / sizeof(float)); d2++) {
        const float a = 0.1f * std::sqrt((float)d2);
        t.w[d2] = std::exp(-2.0f * a * a);
        t.lw[d2] = std::log2(t.w[d2]);
    }
    return t;
}

}  // namespace

gcam_status launch_block_match(View<const float4> in, View<const float2> grad, View<const float> feat8,
                               const NlfDev* nlf_level, const BmParams& p, int radius, int rows_per_thread,
                               View<float4> out, hipStream_t s) {
    static const DirTable dir = make_dir_table();
    const int rows = rows_per_thread == 0 ? default_bm_rows(radius, out.w, out.h) : rows_per_thread;
    switch (radius) {
        case 1: return launch_r<1>(in, grad, feat8, nlf_level, p, dir, rows, out, s);
        case 2: return launch_r<2>(in, grad, feat8, nlf_level, p, dir, rows, out, s);
        case 3: return launch_r<3>(in, grad, feat8, nlf_level, p, dir, rows, out, s);
        case 4: return launch_r<4>(in, grad, feat8, nlf_level, p, dir, rows, out, s);
        case 5: return launch_r<5>(in, grad, feat8, nlf_level, p, dir, rows, out, s);
        case 6: return launch_r<6>(in, grad, feat8, nlf_level, p, dir, rows, out, s);
        case 7: return launch_r<7>(in, grad, feat8, nlf_level, p, dir, rows, out, s);
        case 8: return launch_r<8>(in, grad, feat8, nlf_level, p, dir, rows, out, s);
        case 9: return launch_r<9>(in, grad, feat8, nlf_level, p, dir, rows, out, s);
        case 10: return launch_r<10>(in, grad, feat8, nlf_level, p, dir, rows, out, s);
        default: return set_error(GCAM_ERR_INVALID_ARGUMENT, "search radius must be 1..10");
    }
}

}  // namespace gcam
