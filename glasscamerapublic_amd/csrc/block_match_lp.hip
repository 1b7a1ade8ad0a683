// Block-matching denoise, precision form (FP16 / BF16 contractions, FP32
// accumulation) for gfx950.
//
//   pcaProjection              MetalPipeline/demosaic.metal:774-796
//   blockMatchingDenoiseImage  MetalPipeline/demosaic.metal:806-884
//
// One kernel per level computes the PCA features of its tile in-workgroup
// (they never touch HBM) and then the weighted window average:
//
// 1. Luma of the tile plus an (R+2)-pixel apron is staged in LDS as 16-bit
//    values (BF16: high part and rounding remainder), centred on the tile's
//    first pixel and scaled by the level scale s (~1/sigma of the level), in
//    pair planes (word e = the values at positions e and e + 1 of a row);
//    the same pass stages the YCbCr of the tile plus its R-pixel apron.
//    Feature distances are invariant to the centring; the scale divides the
//    luma coefficient by s^2.
// 2. Projection on MFMA, transposed (F^T = B^T P^T): two
//    v_mfma_f32_32x32x16_{f16,bf16}, (8 components padded to 32) x (25 patch
//    values as 15 column pairs and a zero pair, K = 2 x 16) x 32 pixels, FP32
//    accumulation, so every lane ends with 4 consecutive components of one
//    pixel (with 16 x 16 x 32 products, 16 pixels per product, half the lanes
//    would hold only the padding rows); a lane's im2col is four LDS reads per
//    product.  With BF16 (8-bit mantissa) each luma value enters as a high
//    part plus its rounding remainder (twice the products).
// 3. Component 0 (the largest-variance one, in practice the patch mean) spans
//    the widest range across a tile, so its rounding error would dominate the
//    distances: it stays FP32.  Components 1..7 are rounded to FP16/BF16 and
//    their distance is taken as a Gram form |f_p|^2 + |f_q|^2 - 2 f_p.f_q, the
//    dot product as 4 packed dot2 (v_dot2c_f32_{f16,bf16}: 16-bit products,
//    FP32 accumulation) started from -|f_q|^2 / 2 (kept in LDS), so the luma
//    exponent is k2 (f_p.f_q - |f_q|^2/2) + kl |f_p|^2 + kl t0^2 -- exactly the
//    squared distance of the rounded features up to FP32 rounding.  Weights
//    and sums are FP32; reciprocals are the hardware v_rcp_f32 (1 ulp).
// 4. The inner loop is bound by vector-ALU issue (a transcendental costs
//    about four FMA slots, so exp2 stays at the two weights): component 0
//    enters the exponent as -(r f0q - r f0p)^2 with r = sqrt(-kl), one FMA for
//    the difference instead of a subtraction and a product.  The luma weight
//    is produced pre-scaled, wl' = sqrt(2 log2 e) wl (its exponent carries
//    log2 sqrt(2 log2 e)), so the chroma exponent's -2 log2(e) wl^2 is one
//    FMA -wl' wl' and the direction table holds w / sqrt(2 log2 e).
#include "gcam_internal.hpp"
#include "burst_tables.hpp"

#include <cmath>
#include <type_traits>

namespace gcam {

namespace {

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <class H>
struct Prec;

template <>
struct Prec<_Float16> {
    using h2 = f16x2;
    using h8 = f16x8;
    static __device__ __forceinline__ f32x16 mfma(h8 a, h8 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ float dot2(h2 a, h2 b, float c) { return __builtin_amdgcn_fdot2(a, b, c, false); }
};

template <>
struct Prec<__bf16> {
    using h2 = bf16x2;
    using h8 = bf16x8;
    static __device__ __forceinline__ f32x16 mfma(h8 a, h8 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ float dot2(h2 a, h2 b, float c) {
        return __builtin_amdgcn_fdot2_f32_bf16(a, b, c, false);
    }
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// f(std::integral_constant<int, i>) for i = I .. E (nothing when E < I).
template <int I, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I <= E) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, E>(f);
    }
}

// Two 16-bit values in one word, the first in the low half.
template <class H>
__device__ __forceinline__ uint32_t pack2(H a, H b) {
    return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}

__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

using DirTable = BmDirTable;

constexpr float kLog2e = 1.4426950408889634f;

// Tile TW x (TH*PV) outputs; feature region with an R apron; luma region
// with a further 2-pixel apron (the 5x5 patch).
template <int R, int PV, int TW, int TH>
struct LpTile {
    static constexpr int FW = TW + 2 * R, FH = TH * PV + 2 * R, FN = FW * FH;
    static constexpr int LW = FW + 4, LH = FH + 4, LN = LW * LH;
    static constexpr int THREADS = TW * TH, WAVES = THREADS / 64, GROUPS = (FN + 31) / 32;
    static_assert(THREADS % 64 == 0, "whole wavefronts");
    static_assert(LN * 8 + FN * 36 <= 160 * 1024, "LDS tile larger than a CU's");
};

// One output tile with origin (x0, y0); every thread of the workgroup calls
// it (the LDS staging has workgroup barriers).
template <int R, int PV, int TW, int TH, class H>
__device__ __forceinline__ void bm_lp_tile(const View<const float4>& in, const View<const float2>& grad,
                                           const float* __restrict__ basis, const NlfDev* __restrict__ nlf,
                                           const BmParams& p, const DirTable& dir, const View<float4>& out, int x0,
                                           int y0, int tx = threadIdx.x, int ty = threadIdx.y) {
    using T = LpTile<R, PV, TW, TH>;
    using P = Prec<H>;
    using h2 = typename P::h2;
    using h8 = typename P::h8;
    constexpr bool kSplit = !std::is_same<H, _Float16>::value;
    // Luma pair planes: word e holds the centred, scaled, rounded luma at
    // luma-region positions e and e + 1 of its row (low half first), so an
    // im2col pair of neighbouring patch values is one 4-byte LDS read.
    __shared__ uint32_t lum_pr[T::LN];
    __shared__ uint32_t lum_pr_lo[kSplit ? T::LN : 1];  // BF16: the rounding remainders
    __shared__ float4 ycn[T::FN];  // Y, Cb, Cr, -|f_1..7|^2 / 2 (scaled, rounded)
    __shared__ float f0s[T::FN];   // component 0, FP32
    __shared__ h8 fh[T::FN];       // components 1..7 rounded, 0
    const int tid = ty * TW + tx;
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    // Level scale ~ 1/sigma of mid-grey luma; centre = the tile's first pixel.
    const float sref = sqrtf(nlf->a[0] + nlf->b[0] * 0.5f);
    const float scale = sref > 0.0f ? 1.0f / sref : 1.0f;
    const float mu = in.at(clampi(x0, 0, in.w - 1), clampi(y0, 0, in.h - 1)).x;

    // One pass over the luma region (tile + (R+2)-pixel apron) stages the
    // luma pairs of the projection and, inside the R-pixel apron, the YCbCr.
    for (int e = tid; e < T::LN; e += T::THREADS) {
        const int ly = e / T::LW, lx = e % T::LW;
        const int gy = clampi(y0 - R - 2 + ly, 0, in.h - 1);
        const float4 v = in.at(clampi(x0 - R - 2 + lx, 0, in.w - 1), gy);
        const float v1 = reinterpret_cast<const float*>(&in.at(clampi(x0 - R - 1 + lx, 0, in.w - 1), gy))[0];
        const float yv = (v.x - mu) * scale, yv1 = (v1 - mu) * scale;
        const H h0 = (H)yv, h1 = (H)yv1;
        lum_pr[e] = pack2(h0, h1);
        if (kSplit) lum_pr_lo[e] = pack2((H)(yv - (float)h0), (H)(yv1 - (float)h1));
        const int fx = lx - 2, fy = ly - 2;
        if (fx >= 0 && fx < T::FW && fy >= 0 && fy < T::FH) ycn[fy * T::FW + fx] = make_float4(v.x, v.y, v.z, 0.0f);
    }
    // A operands of the projection (constant): the transposed basis in the
    // pair order of K, split over two v_mfma_f32_32x32x16 (K = 16 each).  K
    // slice kg (of 4, 8 slots each) = pairs d = 0 .. 3: pairs 0-2 are patch
    // row kg, columns (0, 1), (2, 3), (4, 5), pair 3 is row 4, columns (2 kg,
    // 2 kg + 1) for kg < 3 and zero for kg = 3 (column 5 has a zero basis row).
    // Lane half h = lane / 32 holds slice kg = h in the first product and
    // kg = 2 + h in the second; lane holds basis[k][i] for component
    // i = lane % 32 (zero outside 25 x 8).  ro / xo: LDS word offsets of the
    // slice's row and of its fourth pair from the patch's top-left corner.
    h8 aop[2];
    int ro[2], xo[2];
    {
        const int i = lane & 31, h = lane >> 5;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int kg = 2 * q + h;
#pragma unroll
            for (int v = 0; v < 8; v++) {
                const int d = v / 2, r = d < 3 ? kg : 4, c = (d < 3 ? 2 * d : 2 * kg) + (v & 1);
                const bool valid = (d < 3 || kg < 3) && c < 5 && i < kComponents;
                aop[q][v] = (H)(valid ? basis[(r * 5 + c) * kComponents + i] : 0.0f);
            }
            ro[q] = kg * T::LW;
            xo[q] = kg < 3 ? 4 * T::LW + 2 * kg : kg * T::LW;  // kg = 3: a zero basis pair, any word
        }
    }
    __syncthreads();

    // Projection C[component][pixel], 32 feature-region pixels per pair of
    // v_mfma_f32_32x32x16.  B operands: lane holds the K slice of its half (as
    // in the A operands) of pixel lane % 32 of the group; the patch is centred
    // on the pixel's clamped image position (clamp-to-edge, like the FP32
    // projection).  Interior tiles (feature region inside the image) step the
    // pixel position incrementally.  BF16 adds the rounding remainder of each
    // luma value as further products.
    const bool interior =
        x0 - R >= 0 && y0 - R >= 0 && x0 + TW + R <= in.w && y0 + TH * PV + R <= in.h;
    constexpr int kStep = 32 * T::WAVES, kStepY = kStep / T::FW, kStepX = kStep % T::FW;
    int ex = (32 * wave + (lane & 31)) % T::FW, ey = (32 * wave + (lane & 31)) / T::FW;
    for (int g = wave; g < T::GROUPS; g += T::WAVES) {
        const int j = lane & 31, h = lane >> 5;
        const int er = g * 32 + j;
        int base;
        if (interior) {
            base = er < T::FN ? ey * T::LW + ex : 0;
        } else {
            const int e = min(er, T::FN - 1);
            const int cx = clampi(x0 - R + e % T::FW, 0, in.w - 1) - (x0 - R);
            const int cy = clampi(y0 - R + e / T::FW, 0, in.h - 1) - (y0 - R);
            base = cy * T::LW + cx;
        }
        ex += kStepX;
        ey += kStepY;
        if (ex >= T::FW) {
            ex -= T::FW;
            ey++;
        }
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        f32x16 c = {};
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const uint32_t* ph = lum_pr + base + ro[q];
            const u32x4 whi = u32x4{ph[0], ph[2], ph[4], lum_pr[base + xo[q]]};
            c = P::mfma(aop[q], __builtin_bit_cast(h8, whi), c);
            if (kSplit) {
                const uint32_t* pl = lum_pr_lo + base + ro[q];
                const u32x4 wlo = u32x4{pl[0], pl[2], pl[4], lum_pr_lo[base + xo[q]]};
                c = P::mfma(aop[q], __builtin_bit_cast(h8, wlo), c);
            }
        }
        // C: registers 0-3 of lane l hold components 4 (l / 32) + r of pixel
        // l % 32 (rows 8..31 are the zero basis rows).  Lanes 0-31 keep
        // component 0 in FP32 and write halves 0-3 = (0, c1, c2, c3); lanes
        // 32-63 write halves 4-7 = (c4 .. c7).  The squared norm of the rounded
        // components: two dot2 per lane, then lane l + (l ^ 32).
        const h2 p01 = {(h == 0) ? (H)0.0f : (H)c[0], (H)c[1]}, p23 = {(H)c[2], (H)c[3]};
        float n = P::dot2(p01, p01, P::dot2(p23, p23, 0.0f));
        const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(n), __float_as_uint(n), false, false);
        n = __uint_as_float(sw[0]) + __uint_as_float(sw[1]);
        if (er < T::FN) {
            reinterpret_cast<uint2*>(&fh[er])[h] =
                make_uint2(__builtin_bit_cast(uint32_t, p01), __builtin_bit_cast(uint32_t, p23));
            if (h == 0) {
                f0s[er] = c[0];
                ycn[er].w = -0.5f * n;
            }
        }
    }
    __syncthreads();

    const int x = x0 + tx, yb = y0 + ty * PV;
    if (x < out.w && yb < out.h) {
        h8 fp[PV];
        float r1[PV], pr[PV], kl[PV], k2[PV], cp[PV], sb[PV], sr[PV], sbc[PV], src[PV];
        float fY[PV], fCb[PV], fCr[PV], nY[PV], nC[PV];
        const float kroot = sqrtf(2.0f * kLog2e), lkroot = 0.5f * log2f(2.0f * kLog2e);
        const float rcb = __builtin_amdgcn_rcpf(p.chromaBoost), rscale = sref > 0.0f ? sref : 1.0f;
        const float rthr0 = __builtin_amdgcn_rcpf(p.thr[0]), rthr1 = __builtin_amdgcn_rcpf(p.thr[1]);
        const float rthr2 = __builtin_amdgcn_rcpf(p.thr[2]);
#pragma unroll
        for (int k = 0; k < PV; k++) {
            const int y = min(yb + k, out.h - 1);
            const int c = (ty * PV + k + R) * T::FW + tx + R;
            const float4 v = ycn[c];
            fp[k] = fh[c];
            const float f0p = f0s[c];
            float lens = 1.0f;
            if (p.lensShading > 0.0f) {
                const float cx = (float)(in.w / 2), cy = (float)(in.h / 2);
                const float dx = (float)x - cx, dy = (float)y - cy;
                const float d = sqrtf(dx * dx + dy * dy) / sqrtf(cx * cx + cy * cy);
                lens = 1.0f + p.lensShading * d * d;
            }
            // 1 / sigma per channel (hardware reciprocal square root)
            const float is0 = __builtin_amdgcn_rsqf(nlf->a[0] + nlf->b[0] * lens * v.x);
            const float is1 = __builtin_amdgcn_rsqf(nlf->a[1] + nlf->b[1] * lens * v.x);
            const float is2 = __builtin_amdgcn_rsqf(nlf->a[2] + nlf->b[2] * lens * v.x);
            const float dm0 = is0 * rthr0, dm1 = is1 * rthr1, dm2 = is2 * rthr2;
            float mag = 0.0f;
            if (grad.ptr) {
                const float2 gg = grad.at(x, y);
                mag = sqrtf(gg.x * gg.x + gg.y * gg.y);
            }
            const float edge = smoothstepf(2.0f, 16.0f, p.gradientThreshold * mag * is0);
            const float lm = dm0 * rscale * __builtin_amdgcn_rcpf(1.0f + p.gradientBoost * edge);
            kl[k] = -2.0f * kLog2e * lm * lm;  // per unit of scaled squared distance
            k2[k] = -2.0f * kl[k];
            cp[k] = fmaf(k2[k], v.w, lkroot);  // kl |f_p|^2 (v.w = -|f_p|^2 / 2) + log2 sqrt(2 log2 e)
            r1[k] = sqrtf(-kl[k]);             // kl t0^2 = -(r1 t0)^2
            pr[k] = -f0p * r1[k];
            sb[k] = kroot * (dm1 * rcb);
            sr[k] = kroot * (dm2 * rcb);
            sbc[k] = sb[k] * v.y;
            src[k] = sr[k] * v.z;
            fY[k] = fCb[k] = fCr[k] = nY[k] = nC[k] = 0.0f;
        }

        // One window row dy: every position q of the row against the outputs
        // k = KLO .. KHI of this thread (those whose window holds the row).
        auto window_row = [&](auto klo, auto khi, int dy, auto check) {
            constexpr int KLO = decltype(klo)::value, KHI = decltype(khi)::value;
            const int rowq = (ty * PV + R + dy) * T::FW + tx + R;
            // Operands of position dx are loaded while position dx - 1 is
            // computed; the scheduling barrier keeps the compiler from hoisting
            // the loads of the whole row (registers, and so occupancy).
            h8 fq = fh[rowq - R];
            float4 qy = ycn[rowq - R];
            float f0q = f0s[rowq - R];
#pragma unroll
            for (int dx = -R; dx <= R; dx++) {
                const h8 fq_c = fq;
                const float4 qy_c = qy;
                const float f0q_c = f0q;
                if (dx < R) {
                    fq = fh[rowq + dx + 1];
                    qy = ycn[rowq + dx + 1];
                    f0q = f0s[rowq + dx + 1];
                }
#pragma unroll
                for (int k = KLO; k <= KHI; k++) {
                    const int ddy = dy - k;
                    if (decltype(check)::value && (ddy < -R || ddy > R)) continue;  // wave-uniform
                    const int o2 = dx * dx + ddy * ddy;
                    const float wd = dir.w[o2], lwd = dir.lw[o2];
                    // acc = f_p . f_q - |f_q|^2 / 2 over components 1..7, so that
                    // k2 acc + kl |f_p|^2 = kl |f_q - f_p|^2 (Gram form)
                    float acc = P::dot2(h2{fq_c[0], fq_c[1]}, h2{fp[k][0], fp[k][1]}, qy_c.w);
                    acc = P::dot2(h2{fq_c[2], fq_c[3]}, h2{fp[k][2], fp[k][3]}, acc);
                    acc = P::dot2(h2{fq_c[4], fq_c[5]}, h2{fp[k][4], fp[k][5]}, acc);
                    acc = P::dot2(h2{fq_c[6], fq_c[7]}, h2{fp[k][6], fp[k][7]}, acc);
                    const float t0 = fmaf(f0q_c, r1[k], pr[k]);
                    // wl' = sqrt(2 log2 e) wl; wd holds w / sqrt(2 log2 e)
                    const float wl = __builtin_amdgcn_exp2f(fmaf(k2[k], acc, fmaf(-t0, t0, cp[k])));
                    const float wY = wd * wl;
                    const float tb = fmaf(sb[k], qy_c.y, -sbc[k]), tr = fmaf(sr[k], qy_c.z, -src[k]);
                    float e = fmaf(-tb, tb, lwd);
                    e = fmaf(-tr, tr, e);
                    e = fmaf(-wl, wl, e);
                    const float wC = __builtin_amdgcn_exp2f(e);
                    fY[k] = fmaf(wY, qy_c.x, fY[k]);
                    nY[k] += wY;
                    fCb[k] = fmaf(wC, qy_c.y, fCb[k]);
                    fCr[k] = fmaf(wC, qy_c.z, fCr[k]);
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
            constexpr int kMidLo = -R + PV - 1;                  // first row of every output
            constexpr int kBotLo = R + 1 > kMidLo ? R + 1 : kMidLo;  // first row after the band
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
            const float rY = __builtin_amdgcn_rcpf(nY[k]), rC = __builtin_amdgcn_rcpf(nC[k]);
            if (yb + k < out.h) out.at(x, yb + k) = make_float4(fY[k] * rY, fCb[k] * rC, fCr[k] * rC, nY[k]);
        }
    }
}

// One workgroup per TW x (TH * PV) output tile, tiles in XCD-contiguous order.
template <int R, int PV, int TW, int TH, class H>
__global__ __launch_bounds__(TW* TH) void k_block_match_lp(View<const float4> in, View<const float2> grad,
                                                           const float* __restrict__ basis,
                                                           const NlfDev* __restrict__ nlf, BmParams p,
                                                           DirTable dir, View<float4> out, XcdTiles order) {
    int tx, ty;
    if (!order.tile(tx, ty)) return;
    bm_lp_tile<R, PV, TW, TH, H>(in, grad, basis, nlf, p, dir, out, tx * TW, ty * TH * PV);
}

// Burst form over the tiles of every frame of a level.  kLoop false: one
// workgroup per tile (workgroup b takes tile b).  kLoop true, a persistent
// grid: schedule 0: workgroup b takes tiles b, b + G, b + 2G, ... (G = grid
// size); schedule 1: workgroups run on XCD b % 8, so XCD x takes the
// contiguous x-th eighth of the tiles (neighbouring tiles, whose aprons
// overlap, share that XCD's L2) and its workgroups stride through it.  Every
// workgroup leaves the loop when its tiles run out.  Across a tile loop the
// compiler would keep what it can compute once (addresses from the thread
// coordinates, the direction-table values of the unrolled window rows) in
// registers for the whole loop, which costs wavefronts per SIMD: an empty
// asm re-defines the thread coordinates and the table pointer per tile, and
// the one-tile form is an instance of its own.
template <int R, int PV, int TW, int TH, class H, bool kLoop>
__global__ __launch_bounds__(TW* TH) void k_burst_block_match_lp(const BurstTable<BurstBmFrame>* __restrict__ tab,
                                                                 const DirTable* __restrict__ dirp, int schedule) {
    if constexpr (!kLoop) {
        const int item = blockIdx.x;
        if (item >= tab->total) return;
        const BurstBmFrame& f = tab->frame[burst_tile_frame(tab, item)];
        const int lt = item - f.tile0;
        bm_lp_tile<R, PV, TW, TH, H>(f.in, f.grad, f.basis, f.nlf, f.p, *dirp, f.out, (lt % f.tiles_x) * TW,
                                     (lt / f.tiles_x) * (TH * PV));
        return;
    }
    const int total = tab->total, G = gridDim.x;
    int first = blockIdx.x, step = G, end = total;
    if (schedule == 1) {
        constexpr int kXcd = 8;
        const int x = blockIdx.x % kXcd, chunk = (total + kXcd - 1) / kXcd;
        first = x * chunk + blockIdx.x / kXcd;
        step = (G - x + kXcd - 1) / kXcd;
        end = min(total, (x + 1) * chunk);
    }
    for (int item = first; item < end; item += step) {
        int tx = threadIdx.x, ty = threadIdx.y;
        asm volatile("" : "+s"(dirp), "+v"(tx), "+v"(ty));
        const BurstBmFrame& f = tab->frame[burst_tile_frame(tab, item)];
        const int lt = item - f.tile0;
        bm_lp_tile<R, PV, TW, TH, H>(f.in, f.grad, f.basis, f.nlf, f.p, *dirp, f.out, (lt % f.tiles_x) * TW,
                                     (lt / f.tiles_x) * (TH * PV), tx, ty);
        __syncthreads();
    }
}

template <int R, int PV, int TW, int TH, class H>
gcam_status launch_lp_shape(View<const float4> in, View<const float2> grad, const float* basis, const NlfDev* nlf,
                            const BmParams& p, const DirTable& dir, View<float4> out, hipStream_t s) {
    const XcdTiles order = xcd_tiles(out.w, out.h, TW, TH * PV);
    k_block_match_lp<R, PV, TW, TH, H><<<order.grid(), dim3(TW, TH), 0, s>>>(in, grad, basis, nlf, p, dir, out, order);
    GCAM_LAUNCH_CHECK("blockMatchingDenoiseImage (low precision)");
    return GCAM_OK;
}

// Whether two workgroups of the one-row 32 x 16-thread tile (LpTile<R, 1, 32,
// 16>: the luma region, twice with BF16's remainders, and 36 B per feature
// pixel) fit a CU's 160 KB of LDS.
constexpr bool lp_tall_one_row_fits(int radius, bool split) {
    return (36 + 2 * radius) * (20 + 2 * radius) * (split ? 8 : 4) + (32 + 2 * radius) * (16 + 2 * radius) * 36 <=
           80 * 1024;
}
static_assert(LpTile<10, 1, 32, 16>::LN == (36 + 20) * (20 + 20) && LpTile<10, 1, 32, 16>::FN == 52 * 36,
              "lp_tall_one_row_fits follows LpTile");

// One output row per thread: 32 x 16 threads (32 x 16-output tiles) where the
// level has at least 1024 such tiles and two of these workgroups fit a CU's
// LDS, else 32 x 8 (smaller tiles, more workgroups for the small levels).  The
// LDS test keeps BF16 at radius 10 (85 KB a tile: one workgroup per CU) on
// 32 x 8, two workgroups per CU: 12 MP level 0 4.21 -> 3.79 ms.
template <int R, class H>
gcam_status launch_lp_one_row(View<const float4> in, View<const float2> grad, const float* basis, const NlfDev* nlf,
                              const BmParams& p, const DirTable& dir, View<float4> out, hipStream_t s) {
    constexpr bool kFits = lp_tall_one_row_fits(R, !std::is_same<H, _Float16>::value);
    if (kFits && (long)((out.w + 31) / 32) * ((out.h + 15) / 16) >= 1024)
        return launch_lp_shape<R, 1, 32, 16, H>(in, grad, basis, nlf, p, dir, out, s);
    return launch_lp_shape<R, 1, 32, 8, H>(in, grad, basis, nlf, p, dir, out, s);
}

template <int R, class H>
gcam_status launch_lp_r(View<const float4> in, View<const float2> grad, const float* basis, const NlfDev* nlf,
                        const BmParams& p, const DirTable& dir, int rows, View<float4> out, hipStream_t s) {
    if (rows == 2) return launch_lp_shape<R, 2, 32, 8, H>(in, grad, basis, nlf, p, dir, out, s);
    if constexpr (R <= 6) {
        if (rows == 4) return launch_lp_shape<R, 4, 32, 4, H>(in, grad, basis, nlf, p, dir, out, s);
    }
    // rows 1 (and 4 above radius 6, where four rows' registers do not fit)
    return launch_lp_one_row<R, H>(in, grad, basis, nlf, p, dir, out, s);
}

template <class H>
gcam_status dispatch(View<const float4> in, View<const float2> grad, const float* basis, const NlfDev* nlf,
                     const BmParams& p, int radius, int rows, View<float4> out, hipStream_t s) {
    static const DirTable dir = make_bm_dir_table();
    switch (radius) {
        case 1: return launch_lp_r<1, H>(in, grad, basis, nlf, p, dir, rows, out, s);
        case 2: return launch_lp_r<2, H>(in, grad, basis, nlf, p, dir, rows, out, s);
        case 3: return launch_lp_r<3, H>(in, grad, basis, nlf, p, dir, rows, out, s);
        case 4: return launch_lp_r<4, H>(in, grad, basis, nlf, p, dir, rows, out, s);
        case 5: return launch_lp_r<5, H>(in, grad, basis, nlf, p, dir, rows, out, s);
        case 6: return launch_lp_r<6, H>(in, grad, basis, nlf, p, dir, rows, out, s);
        case 7: return launch_lp_r<7, H>(in, grad, basis, nlf, p, dir, rows, out, s);
        case 8: return launch_lp_r<8, H>(in, grad, basis, nlf, p, dir, rows, out, s);
        case 9: return launch_lp_r<9, H>(in, grad, basis, nlf, p, dir, rows, out, s);
        case 10: return launch_lp_r<10, H>(in, grad, basis, nlf, p, dir, rows, out, s);
        default: return set_error(GCAM_ERR_INVALID_ARGUMENT, "search radius must be 1..10");
    }
}

template <int R, int PV, int TW, int TH, class H>
gcam_status launch_burst_shape(const BurstTable<BurstBmFrame>* dev, int total, const DirTable* dir,
                               const BurstBmConfig& cfg, hipStream_t s) {
    int grid = cfg.grid > 0 ? std::min(cfg.grid, total) : total;
    if (cfg.schedule == 1 && grid < 8) grid = std::min(8, total);  // every XCD chunk needs a workgroup
    if (grid < 1) return GCAM_OK;
    if (grid == total && cfg.schedule == 0)
        k_burst_block_match_lp<R, PV, TW, TH, H, false><<<grid, dim3(TW, TH), 0, s>>>(dev, dir, cfg.schedule);
    else
        k_burst_block_match_lp<R, PV, TW, TH, H, true><<<grid, dim3(TW, TH), 0, s>>>(dev, dir, cfg.schedule);
    GCAM_LAUNCH_CHECK("burst blockMatchingDenoiseImage");
    return GCAM_OK;
}

template <int R, class H>
gcam_status launch_burst_r(const BurstTable<BurstBmFrame>* dev, int total, const DirTable* dir,
                           const BurstBmConfig& cfg, hipStream_t s) {
    if constexpr (R <= 6) {
        if (cfg.rows == 4) return launch_burst_shape<R, 4, 32, 4, H>(dev, total, dir, cfg, s);
        if (cfg.rows == 2 && cfg.tile_w == 64) return launch_burst_shape<R, 2, 64, 8, H>(dev, total, dir, cfg, s);
        if (cfg.rows == 2) return launch_burst_shape<R, 2, 32, 8, H>(dev, total, dir, cfg, s);
        return launch_burst_shape<R, 1, 32, 8, H>(dev, total, dir, cfg, s);
    } else {
        if (cfg.rows == 2) return launch_burst_shape<R, 2, 32, 8, H>(dev, total, dir, cfg, s);
        if constexpr (lp_tall_one_row_fits(R, !std::is_same<H, _Float16>::value))
            return launch_burst_shape<R, 1, 32, 16, H>(dev, total, dir, cfg, s);
        return launch_burst_shape<R, 1, 32, 8, H>(dev, total, dir, cfg, s);
    }
}

template <class H>
gcam_status burst_dispatch(const BurstTable<BurstBmFrame>* dev, const DirTable* dir, int total, int radius,
                           const BurstBmConfig& cfg, hipStream_t s) {
    switch (radius) {
        case 1: return launch_burst_r<1, H>(dev, total, dir, cfg, s);
        case 2: return launch_burst_r<2, H>(dev, total, dir, cfg, s);
        case 3: return launch_burst_r<3, H>(dev, total, dir, cfg, s);
        case 4: return launch_burst_r<4, H>(dev, total, dir, cfg, s);
        case 5: return launch_burst_r<5, H>(dev, total, dir, cfg, s);
        case 6: return launch_burst_r<6, H>(dev, total, dir, cfg, s);
        case 7: return launch_burst_r<7, H>(dev, total, dir, cfg, s);
        case 8: return launch_burst_r<8, H>(dev, total, dir, cfg, s);
        case 9: return launch_burst_r<9, H>(dev, total, dir, cfg, s);
        case 10: return launch_burst_r<10, H>(dev, total, dir, cfg, s);
        default: return set_error(GCAM_ERR_INVALID_ARGUMENT, "search radius must be 1..10");
    }
}

}  // namespace

void burst_bm_tile(int radius, int rows, int tile_w, int precision, int* tw, int* th) {
    // the shapes of launch_burst_r: radius up to 6: 32 x 4 threads x 4 rows,
    // 32 (or 64) x 8 x 2 rows, 32 x 8 x 1 row; above: 32 x 8 x 2 rows, else
    // 32 x 16 x 1 row where two such workgroups fit a CU's LDS, 32 x 8 x 1 row
    // where not
    *tw = tile_w == 64 && rows == 2 && radius <= 6 ? 64 : 32;
    if (radius > 6) {
        *th = rows == 2 || lp_tall_one_row_fits(radius, precision == GCAM_PRECISION_BF16) ? 16 : 8;
    } else {
        *th = rows == 4 ? 16 : 8 * rows;
    }
}

// w: the direction weight divided by sqrt(2 log2 e) (the luma weight enters
// pre-scaled by that factor); lw: log2 of the direction weight itself.
BmDirTable make_bm_dir_table() {
    BmDirTable t{};
    const float kroot = std::sqrt(2.0f * kLog2e);
    for (int d2 = 0; d2 < (int)(sizeof(t.w) / sizeof(float)); d2++) {
        const float a = 0.1f * std::sqrt((float)d2);
        const float w = std::exp(-2.0f * a * a);
        t.w[d2] = w / kroot;
        t.lw[d2] = std::log2(w);
    }
    return t;
}

gcam_status launch_burst_block_match_lp(const BurstTable<BurstBmFrame>* dev, const BmDirTable* dir, int total_tiles,
                                        int radius, int precision, const BurstBmConfig& cfg, hipStream_t s) {
    if (precision == GCAM_PRECISION_FP16) return burst_dispatch<_Float16>(dev, dir, total_tiles, radius, cfg, s);
    if (precision == GCAM_PRECISION_BF16) return burst_dispatch<__bf16>(dev, dir, total_tiles, radius, cfg, s);
    return set_error(GCAM_ERR_INVALID_ARGUMENT, "the burst block matching needs FP16 or BF16");
}

gcam_status launch_block_match_lp(View<const float4> in, View<const float2> grad, const float* basis,
                                  const NlfDev* nlf_level, const BmParams& p, int radius, int rows_per_thread,
                                  int precision, View<float4> out, hipStream_t s) {
    int rows = rows_per_thread == 0 ? default_bm_rows(radius, out.w, out.h) : rows_per_thread;
    // BF16 also stages the luma's rounding remainders (one more LDS word per
    // luma position), which leaves three two-row workgroups per CU (three
    // wavefronts per SIMD) at radius 4; one row per thread on 32 x 16 threads
    // keeps six (12 MP level 0: 0.727 -> 0.692 ms).
    if (rows_per_thread == 0 && precision == GCAM_PRECISION_BF16) rows = 1;
    if (precision == GCAM_PRECISION_FP16) return dispatch<_Float16>(in, grad, basis, nlf_level, p, radius, rows, out, s);
    if (precision == GCAM_PRECISION_BF16) return dispatch<__bf16>(in, grad, basis, nlf_level, p, radius, rows, out, s);
    return set_error(GCAM_ERR_INVALID_ARGUMENT, "low-precision block matching needs FP16 or BF16");
}

}  // namespace gcam
