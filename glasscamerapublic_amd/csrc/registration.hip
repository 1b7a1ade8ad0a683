// Burst registration and fusion for gfx950.
//
//   registerAndFuse / registerImage / registerBayerImage
//                                 ImageRegistration/SURF.metal:489-559,
//                                 MetalPipeline/demosaic_kernels.hpp:690-749
//   FindHomography                ImageRegistration/Homography.cpp:72-116,
//                                 RTL/RANSAC.hpp:53-118
//   FindLeastSquaresHomography    ImageRegistration/LeastSquaresHomography.cpp:356-421
//
// The warps are streaming per-pixel kernels (HBM-bound: registerAndFuse
// moves 48 bytes per pixel).  Warped positions are FP32 evaluated in source
// order (no contraction); the linear sampler interpolates at the exact
// sub-texel position with clamp-to-edge addressing.
//
// RANSAC runs every hypothesis at once: lane k of a workgroup fits the
// minimal sample of hypothesis k in closed form and each of the workgroup's
// waves counts the outliers of those 64 hypotheses over its share of the
// matches.  One workgroup then selects the first hypothesis with the fewest
// outliers, compacts its inliers in index order and refits them: normalised
// DLT normal matrix reduced across the workgroup, smallest eigenvector by
// parallel-ordered cyclic Jacobi in LDS (FP64).
#pragma clang fp contract(off)

#include "gcam_internal.hpp"
#include "gcam/registration.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>

namespace gcam {

namespace {

struct Homography {
    float m[9];
};

Homography homography_of(const float h[9]) {
    Homography H;
    for (int i = 0; i < 9; i++) H.m[i] = h[i];
    return H;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <class P>
__device__ __forceinline__ P fetch(const View<const P>& v, int x, int y) {
    return v.at(clampi(x, 0, v.w - 1), clampi(y, 0, v.h - 1));
}

// p = (x, y, 1); (dot(m0, p), dot(m1, p), dot(m2, p)) as Metal's dot, left to right.
__device__ __forceinline__ float2 warp(const Homography& H, float x, float y) {
    const float u = H.m[0] * x + H.m[1] * y + H.m[2] * 1.0f;
    const float v = H.m[3] * x + H.m[4] * y + H.m[5] * 1.0f;
    const float w = H.m[6] * x + H.m[7] * y + H.m[8] * 1.0f;
    return make_float2(u / w, v / w);
}

__device__ __forceinline__ float lerp2(float a, float b, float c, float d, float fx, float fy) {
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    return (a * gx + b * fx) * gy + (c * gx + d * fx) * fy;
}

// The linear sampler at texel-space (tx, ty).  Positions beyond one texel
// outside the image read the edge whatever their distance, so they are
// clamped first (which also sends NaN to the edge).
struct Texel {
    int x0, y0;
    float fx, fy;
};
__device__ __forceinline__ Texel texel(float tx, float ty, int w, int h) {
    tx = fminf(fmaxf(tx, -1.0f), (float)w);
    ty = fminf(fmaxf(ty, -1.0f), (float)h);
    const float x0 = floorf(tx), y0 = floorf(ty);
    return Texel{(int)x0, (int)y0, tx - x0, ty - y0};
}

__device__ __forceinline__ float sample(const View<const float>& im, const Texel& t) {
    return lerp2(fetch(im, t.x0, t.y0), fetch(im, t.x0 + 1, t.y0), fetch(im, t.x0, t.y0 + 1),
                 fetch(im, t.x0 + 1, t.y0 + 1), t.fx, t.fy);
}
__device__ __forceinline__ float4 sample(const View<const float4>& im, const Texel& t) {
    const float4 a = fetch(im, t.x0, t.y0), b = fetch(im, t.x0 + 1, t.y0), c = fetch(im, t.x0, t.y0 + 1),
                 d = fetch(im, t.x0 + 1, t.y0 + 1);
    return make_float4(lerp2(a.x, b.x, c.x, d.x, t.fx, t.fy), lerp2(a.y, b.y, c.y, d.y, t.fx, t.fy),
                       lerp2(a.z, b.z, c.z, d.z, t.fx, t.fy), lerp2(a.w, b.w, c.w, d.w, t.fx, t.fy));
}

// Texel-space position of the sampler's read at (xx, yy): the reference reads
// at normalised ((xx, yy) + 0.5) / output size, which the sampler scales by
// the input size.
__device__ __forceinline__ Texel warped_texel(const Homography& H, int x, int y, float nx, float ny, int iw, int ih) {
    const float2 p = warp(H, (float)x, (float)y);
    return texel((p.x + 0.5f) * nx * (float)iw - 0.5f, (p.y + 0.5f) * ny * (float)ih - 0.5f, iw, ih);
}

// out may alias fused: each pixel reads and writes only its own entry.
__global__ __launch_bounds__(256) void k_register_and_fuse(View<const float4> fused, View<const float4> in,
                                                           View<float4> out, Homography H, float nx, float ny,
                                                           int count) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    const float4 a = fused.at(x, y);
    const float4 b = sample(in, warped_texel(H, x, y, nx, ny, in.w, in.h));
    const float c1 = (float)(count - 1), c = (float)count;
    out.at(x, y) = make_float4((c1 * a.x + b.x) / c, (c1 * a.y + b.y) / c, (c1 * a.z + b.z) / c, (c1 * a.w + b.w) / c);
}

template <class P>
__global__ __launch_bounds__(256) void k_register_image(View<const P> in, View<P> out, Homography H, float nx,
                                                        float ny) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    out.at(x, y) = sample(in, warped_texel(H, x, y, nx, ny, in.w, in.h));
}

// (int) of a float, saturated to +-2^29 so that 2 v + 1 stays an int; NaN -> 0.
__device__ __forceinline__ int to_int_sat(float v) {
    if (!(v == v)) return 0;
    return (int)fminf(fmaxf(v, -536870912.0f), 536870912.0f);
}

__global__ __launch_bounds__(256) void k_register_bayer(View<const uint16_t> in, View<uint16_t> out, Homography H) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    const float2 p = warp(H, (float)(x / 2), (float)(y / 2));
    const int ix = 2 * to_int_sat(roundf(p.x)) + (x & 1), iy = 2 * to_int_sat(roundf(p.y)) + (y & 1);
    out.at(x, y) = fetch(in, ix, iy);
}

// ---- homography fit ---------------------------------------------------------

constexpr int kRansacWaves = 8;     // waves per workgroup sharing 64 hypotheses
constexpr int kFitThreads = 256;    // the select / compact / refit workgroup
constexpr int kJacobiSweeps = 30;   // more than a 9 x 9 Jacobi ever needs
constexpr int kMaxMatches = 1 << 22;
constexpr int kMaxIterations = 1 << 20;

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Floyd's algorithm: 4 distinct indices of [0, n), uniform over subsets,
// sorted ascending (the order of the reference's std::set).
__host__ __device__ inline void ransac_sample(uint64_t seed, int k, int n, int s[4]) {
    for (int d = 0; d < 4; d++) {
        const int j = n - 4 + d;
        const uint64_t r = splitmix64(seed + ((uint64_t)(uint32_t)k << 32) + (uint64_t)d);
        const int t = (int)(((r >> 32) * (uint64_t)(uint32_t)(j + 1)) >> 32);
        bool seen = false;
        for (int q = 0; q < d; q++) seen = seen || s[q] == t;
        s[d] = seen ? j : t;
    }
    for (int a = 1; a < 4; a++)
        for (int b = a; b > 0 && s[b - 1] > s[b]; b--) {
            const int t = s[b];
            s[b] = s[b - 1];
            s[b - 1] = t;
        }
}

// Squared transfer error of match q = (x1, y1, x2, y2) under h (FP32).
__device__ __forceinline__ float transfer_error(const float h[9], float4 q) {
    const float u = h[0] * q.x + h[1] * q.y + h[2];
    const float v = h[3] * q.x + h[4] * q.y + h[5];
    const float w = h[6] * q.x + h[7] * q.y + h[8];
    const float dx = u / w - q.z, dy = v / w - q.w;
    return dx * dx + dy * dy;
}

// Point normalisation of LeastSquaresHomography.cpp:361-387: centroids cM, cm
// and mean absolute deviations sM, sm of the first and second points.
struct Norm {
    double cM[2], cm[2], sM[2], sm[2];
    __host__ __device__ bool degenerate() const {
        return fabs(sm[0]) < DBL_EPSILON || fabs(sm[1]) < DBL_EPSILON || fabs(sM[0]) < DBL_EPSILON ||
               fabs(sM[1]) < DBL_EPSILON;
    }
};

// H = invHnorm H0 Hnorm2 (LeastSquaresHomography.cpp:414-418), then the FP32
// matrix divided by its FP32 corner (:421).
__device__ void denormalise(const Norm& z, const double H0[9], float out[9]) {
    const double in[9] = {z.sm[0], 0, z.cm[0], 0, z.sm[1], z.cm[1], 0, 0, 1};
    const double n2[9] = {1 / z.sM[0], 0, -z.cM[0] / z.sM[0], 0, 1 / z.sM[1], -z.cM[1] / z.sM[1], 0, 0, 1};
    double t[9], H[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t[3 * i + j] = in[3 * i] * H0[j] + in[3 * i + 1] * H0[3 + j] + in[3 * i + 2] * H0[6 + j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H[3 * i + j] = t[3 * i] * n2[j] + t[3 * i + 1] * n2[3 + j] + t[3 * i + 2] * n2[6 + j];
    const float c = (float)H[8];
    for (int i = 0; i < 9; i++) out[i] = (float)H[i] / c;
}

__device__ __forceinline__ void adjugate3(const double m[9], double a[9]) {
    a[0] = m[4] * m[8] - m[5] * m[7];
    a[1] = m[2] * m[7] - m[1] * m[8];
    a[2] = m[1] * m[5] - m[2] * m[4];
    a[3] = m[5] * m[6] - m[3] * m[8];
    a[4] = m[0] * m[8] - m[2] * m[6];
    a[5] = m[2] * m[3] - m[0] * m[5];
    a[6] = m[3] * m[7] - m[4] * m[6];
    a[7] = m[1] * m[6] - m[0] * m[7];
    a[8] = m[0] * m[4] - m[1] * m[3];
}

// Model of a minimal sample (the reference's ComputeModel, Homography.cpp:42-61,
// which fits the 4 matches by FindLeastSquaresHomography).  For 4 points in
// general position the DLT null vector is the homography through them, here
// in closed form on the normalised points: with A = [a1 a2 a3] the columns of
// the first points, la = adj(A) a4 and lb = adj(B) b4 likewise,
// H0 = B diag(lb / la) adj(A) maps each a_i to b_i.  Returns 0 with the model,
// 1 for the identity (no spread in x or y: the reference's range_error) and 2
// for a sample with three collinear points (no unique model; discarded).
__device__ int minimal_model(const float* m, const int s[4], float h[9]) {
    float4 q[4];
    for (int i = 0; i < 4; i++) q[i] = *reinterpret_cast<const float4*>(m + 4 * (size_t)s[i]);
    Norm z;
    double c[4] = {0, 0, 0, 0}, d[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        c[0] += q[i].x, c[1] += q[i].y, c[2] += q[i].z, c[3] += q[i].w;
    }
    for (int k = 0; k < 4; k++) c[k] /= 4.0;
    for (int i = 0; i < 4; i++) {
        d[0] += fabs(q[i].x - c[0]), d[1] += fabs(q[i].y - c[1]), d[2] += fabs(q[i].z - c[2]),
            d[3] += fabs(q[i].w - c[3]);
    }
    for (int k = 0; k < 4; k++) d[k] /= 4.0;
    z.cM[0] = c[0], z.cM[1] = c[1], z.cm[0] = c[2], z.cm[1] = c[3];
    z.sM[0] = d[0], z.sM[1] = d[1], z.sm[0] = d[2], z.sm[1] = d[3];
    if (z.degenerate()) {
        for (int i = 0; i < 9; i++) h[i] = (i % 4 == 0) ? 1.0f : 0.0f;
        return 1;
    }
    double A[9], B[9], a4[3], b4[3];
    for (int i = 0; i < 4; i++) {
        const double ax = (q[i].x - z.cM[0]) / z.sM[0], ay = (q[i].y - z.cM[1]) / z.sM[1];
        const double bx = (q[i].z - z.cm[0]) / z.sm[0], by = (q[i].w - z.cm[1]) / z.sm[1];
        if (i < 3) {
            A[i] = ax, A[3 + i] = ay, A[6 + i] = 1;
            B[i] = bx, B[3 + i] = by, B[6 + i] = 1;
        } else {
            a4[0] = ax, a4[1] = ay, a4[2] = 1;
            b4[0] = bx, b4[1] = by, b4[2] = 1;
        }
    }
    double adjA[9], adjB[9];
    adjugate3(A, adjA);
    adjugate3(B, adjB);
    double r[3];
    for (int i = 0; i < 3; i++) {
        const double la = adjA[3 * i] * a4[0] + adjA[3 * i + 1] * a4[1] + adjA[3 * i + 2] * a4[2];
        const double lb = adjB[3 * i] * b4[0] + adjB[3 * i + 1] * b4[1] + adjB[3 * i + 2] * b4[2];
        if (!(fabs(la) > 1e-12) || !(fabs(lb) > 1e-12)) return 2;
        r[i] = lb / la;
    }
    double H0[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            H0[3 * i + j] = B[3 * i] * r[0] * adjA[j] + B[3 * i + 1] * r[1] * adjA[3 + j] + B[3 * i + 2] * r[2] * adjA[6 + j];
    denormalise(z, H0, h);
    return 0;
}

// Lane k of the workgroup owns hypothesis blockIdx.x * 64 + k; wave v counts
// its outliers over the v-th share of the matches.  loss = outliers, or
// count + 1 for a discarded sample.
__global__ __launch_bounds__(64 * kRansacWaves) void k_ransac_hypotheses(const float* m, int n, uint64_t seed,
                                                                        int iters, float thr, float* models,
                                                                        int* loss) {
    __shared__ int part[kRansacWaves][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane;
    float h[9];
    int kind = 2;
    if (k < iters) {
        int s[4];
        ransac_sample(seed, k, n, s);
        kind = minimal_model(m, s, h);
    }
    int outliers = 0;
    if (kind != 2) {
        const int per = (n + kRansacWaves - 1) / kRansacWaves;
        const int i0 = wv * per, i1 = min(n, i0 + per);
        for (int i = i0; i < i1; i++) {
            const float4 q = *reinterpret_cast<const float4*>(m + 4 * (size_t)i);
            outliers += fabsf(transfer_error(h, q)) > thr;
        }
    }
    part[wv][lane] = outliers;
    __syncthreads();
    if (wv == 0 && k < iters) {
        int total = 0;
        for (int v = 0; v < kRansacWaves; v++) total += part[v][lane];
        loss[k] = kind == 2 ? n + 1 : total;
        for (int i = 0; i < 9; i++) models[9 * (size_t)k + i] = kind == 2 ? 0.0f : h[i];
    }
}

// ---- workgroup-wide pieces of the fit (kFitThreads threads) ----

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums of K per-thread values over the workgroup, into tot[K] (LDS).
template <int K>
__device__ void block_sum(double (&v)[K], double (*red)[kFitThreads / 64], double* tot) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double s = wave_sum(v[k]);
        if (lane == 0) red[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0;
        for (int w = 0; w < kFitThreads / 64; w++) s += red[threadIdx.x][w];
        tot[threadIdx.x] = s;
    }
    __syncthreads();
}

// Round r of the round-robin schedule of 10 indices (9 and a dummy 9): the
// partner of i, or -1 when i is paired with the dummy.
__device__ __forceinline__ int partner(int r, int i) {
    if (i == 9) return -1;
    int p;
    if (i == r) p = 9;
    else p = (2 * r - i + 18) % 9;  // pairs (r - d, r + d) mod 9
    return p == 9 ? -1 : p;
}

struct FitShared {
    double red[45][kFitThreads / 64];
    double tot[45];
    double A[9][9], V[9][9];
    double cs[9][2];  // per index: (c, s) of its rotation this round
    double off;
    Norm z;
    int count;
};

// LtL of LeastSquaresHomography.cpp:389-402 over the points idx(0 .. c-1),
// eigenvector of its smallest eigenvalue, denormalised into h.  Returns false
// (h untouched) where the reference throws.
template <class Idx>
__device__ bool least_squares(const float* m, int c, Idx idx, FitShared& S, float h[9]) {
    const int t = threadIdx.x;
    {
        double v[4] = {0, 0, 0, 0};
        for (int j = t; j < c; j += kFitThreads) {
            const float4 q = *reinterpret_cast<const float4*>(m + 4 * (size_t)idx(j));
            v[0] += q.x, v[1] += q.y, v[2] += q.z, v[3] += q.w;
        }
        block_sum<4>(v, S.red, S.tot);
    }
    const double cM0 = S.tot[0] / c, cM1 = S.tot[1] / c, cm0 = S.tot[2] / c, cm1 = S.tot[3] / c;
    __syncthreads();
    {
        double v[4] = {0, 0, 0, 0};
        for (int j = t; j < c; j += kFitThreads) {
            const float4 q = *reinterpret_cast<const float4*>(m + 4 * (size_t)idx(j));
            v[0] += fabs(q.x - cM0), v[1] += fabs(q.y - cM1), v[2] += fabs(q.z - cm0), v[3] += fabs(q.w - cm1);
        }
        block_sum<4>(v, S.red, S.tot);
    }
    Norm z;
    z.cM[0] = cM0, z.cM[1] = cM1, z.cm[0] = cm0, z.cm[1] = cm1;
    z.sM[0] = S.tot[0] / c, z.sM[1] = S.tot[1] / c, z.sm[0] = S.tot[2] / c, z.sm[1] = S.tot[3] / c;
    __syncthreads();
    if (z.degenerate()) return false;
    {
        double v[45];
#pragma unroll
        for (int k = 0; k < 45; k++) v[k] = 0;
        for (int j = t; j < c; j += kFitThreads) {
            const float4 q = *reinterpret_cast<const float4*>(m + 4 * (size_t)idx(j));
            const double p1x = (q.x - z.cM[0]) / z.sM[0], p1y = (q.y - z.cM[1]) / z.sM[1];
            const double p2x = (q.z - z.cm[0]) / z.sm[0], p2y = (q.w - z.cm[1]) / z.sm[1];
            const double Lx[9] = {p1x, p1y, 1, 0, 0, 0, -p2x * p1x, -p2x * p1y, -p2x};
            const double Ly[9] = {0, 0, 0, p1x, p1y, 1, -p2y * p1x, -p2y * p1y, -p2y};
            int k = 0;
#pragma unroll
            for (int a = 0; a < 9; a++)
#pragma unroll
                for (int b = a; b < 9; b++) v[k++] += Lx[a] * Lx[b] + Ly[a] * Ly[b];
        }
        block_sum<45>(v, S.red, S.tot);
    }
    // Symmetric A from the upper triangle, V = I.
    if (t < 81) {
        const int i = t / 9, j = t % 9, a = min(i, j), b = max(i, j);
        S.A[i][j] = S.tot[a * 9 - a * (a - 1) / 2 + (b - a)];
        S.V[i][j] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    // Cyclic Jacobi, 9 rounds of 4 disjoint rotations per sweep.
    for (int sweep = 0; sweep < kJacobiSweeps; sweep++) {
        if (t == 0) {
            double off = 0, all = 0;
            for (int i = 0; i < 9; i++)
                for (int j = 0; j < 9; j++) {
                    const double e = S.A[i][j] * S.A[i][j];
                    all += e;
                    if (i != j) off += e;
                }
            S.off = off <= 1e-30 * all ? 0.0 : off;
        }
        __syncthreads();
        if (S.off == 0.0) break;
        for (int r = 0; r < 9; r++) {
            if (t < 9) {
                const int p = partner(r, t);
                double c = 1, s = 0;
                if (p >= 0) {
                    const int lo = min(t, p), hi = max(t, p);
                    const double apq = S.A[lo][hi];
                    if (apq != 0.0) {
                        const double tau = (S.A[hi][hi] - S.A[lo][lo]) / (2 * apq);
                        const double tt = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
                        c = 1 / sqrt(1 + tt * tt);
                        s = tt * c;
                    }
                }
                S.cs[t][0] = c, S.cs[t][1] = s;
            }
            __syncthreads();
            // A' = J^T A J and V' = V J with J[lo][lo] = J[hi][hi] = c,
            // J[lo][hi] = s, J[hi][lo] = -s for each pair (lo, hi).
            double na = 0, nv = 0;
            const int i = t / 9, j = t % 9;
            if (t < 81) {
                const int pi = partner(r, i), pj = partner(r, j);
                // J[a][b] for a != b in a pair (lo, hi): J[lo][hi] = s, J[hi][lo] = -s.
                auto Jo = [&](int a, int b) { return a < b ? S.cs[b][1] : -S.cs[b][1]; };
                const double ci = S.cs[i][0], cj = S.cs[j][0];
                na = ci * cj * S.A[i][j];
                if (pi >= 0) na += Jo(pi, i) * cj * S.A[pi][j];
                if (pj >= 0) na += ci * Jo(pj, j) * S.A[i][pj];
                if (pi >= 0 && pj >= 0) na += Jo(pi, i) * Jo(pj, j) * S.A[pi][pj];
                nv = S.V[i][j] * cj;
                if (pj >= 0) nv += S.V[i][pj] * Jo(pj, j);
            }
            __syncthreads();
            if (t < 81) {
                S.A[i][j] = na;
                S.V[i][j] = nv;
            }
            __syncthreads();
        }
    }
    if (t == 0) {
        int r = 0;
        for (int i = 1; i < 9; i++)
            if (S.A[i][i] < S.A[r][r]) r = i;
        double H0[9];
        for (int i = 0; i < 9; i++) H0[i] = S.V[i][r];
        denormalise(z, H0, h);
    }
    __syncthreads();
    return true;
}

struct AllIdx {
    __device__ int operator()(int j) const { return j; }
};
struct ListIdx {
    const int* list;
    __device__ int operator()(int j) const { return list[j]; }
};

// out: 9 floats of model, then (as int) the status: 0 fitted, 1 degenerate.
__global__ __launch_bounds__(kFitThreads) void k_least_squares(const float* m, int n, float* out) {
    __shared__ FitShared S;
    float h[9];
    const bool ok = least_squares(m, n, AllIdx{}, S, h);
    if (threadIdx.x == 0) {
        for (int i = 0; i < 9; i++) out[i] = ok ? h[i] : 0.0f;
        reinterpret_cast<int*>(out)[9] = ok ? 0 : 1;
    }
}

// Select, compact, refit.  out: 9 floats of model, then the inlier count (int).
__global__ __launch_bounds__(kFitThreads) void k_ransac_fit(const float* m, int n, int iters, float thr,
                                                            const float* models, const int* loss, int* inliers,
                                                            float* out) {
    __shared__ FitShared S;
    __shared__ int bl[kFitThreads / 64], bk[kFitThreads / 64], wcount[kFitThreads / 64];
    __shared__ float best[9];
    __shared__ int base;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    // First hypothesis with the fewest outliers (RTL/RANSAC.hpp:63-66: strict <).
    int l = INT_MAX, k = INT_MAX;
    for (int j = t; j < iters; j += kFitThreads) {
        const int lj = loss[j];
        if (lj < l) l = lj, k = j;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int l2 = __shfl_xor(l, o, 64), k2 = __shfl_xor(k, o, 64);
        if (l2 < l || (l2 == l && k2 < k)) l = l2, k = k2;
    }
    if (lane == 0) bl[wv] = l, bk[wv] = k;
    __syncthreads();
    if (t == 0) {
        int L = bl[0], K = bk[0];
        for (int w = 1; w < kFitThreads / 64; w++)
            if (bl[w] < L || (bl[w] == L && bk[w] < K)) L = bl[w], K = bk[w];
        for (int i = 0; i < 9; i++) best[i] = models[9 * (size_t)K + i];
        base = 0;
    }
    __syncthreads();
    float h[9];
    for (int i = 0; i < 9; i++) h[i] = best[i];
    // Inliers in ascending index order (FindInliers, RTL/RANSAC.hpp:74-81).
    for (int c0 = 0; c0 < n; c0 += kFitThreads) {
        const int i = c0 + t;
        bool in = false;
        if (i < n) in = fabsf(transfer_error(h, *reinterpret_cast<const float4*>(m + 4 * (size_t)i))) < thr;
        const uint64_t mask = __ballot(in);
        const int before = __popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) wcount[wv] = __popcll(mask);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wv; w++) off += wcount[w];
        if (in) inliers[off + before] = i;
        __syncthreads();
        if (t == 0) {
            int s = 0;
            for (int w = 0; w < kFitThreads / 64; w++) s += wcount[w];
            base += s;
        }
        __syncthreads();
    }
    const int c = base;
    float r[9];
    bool refit = false;
    if (c >= 4) refit = least_squares(m, c, ListIdx{inliers}, S, r);
    if (t == 0) {
        for (int i = 0; i < 9; i++) out[i] = refit ? r[i] : h[i];
        reinterpret_cast<int*>(out)[9] = c;
    }
}

}  // namespace

}  // namespace gcam

using namespace gcam;

namespace {

const dim3 kBlock(64, 4);

dim3 grid_for(int w, int h) { return dim3((w + kBlock.x - 1) / kBlock.x, (h + kBlock.y - 1) / kBlock.y); }

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

bool finite9(const float h[9]) {
    for (int i = 0; i < 9; i++)
        if (!std::isfinite(h[i])) return false;
    return true;
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// Copies count bytes from the device to host, synchronizing s.
gcam_status to_host(void* host, const void* dev, size_t bytes, hipStream_t s, const char* what) {
    hipError_t e = hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_error(GCAM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return GCAM_OK;
}

}  // namespace

extern "C" {

gcam_status gcam_register_and_fuse_image(const gcam_image* fused, const gcam_image* input, gcam_image* out,
                                         const float homography[9], int32_t count, void* stream) {
    if (!valid_image(fused, GCAM_PIXEL_RGBA32F) || !valid_image(input, GCAM_PIXEL_RGBA32F) ||
        !valid_image(out, GCAM_PIXEL_RGBA32F) || !same_size(fused, out))
        return bad("gcam_register_and_fuse_image: RGBA32F images, fused and out of one size");
    if (!homography || !finite9(homography)) return bad("gcam_register_and_fuse_image: homography must be finite");
    if (count < 1) return bad("gcam_register_and_fuse_image: count must be >= 1");
    if (out->data != fused->data && out->data == input->data)
        return bad("gcam_register_and_fuse_image: out may alias fused, not input");
    if (out->data == fused->data && out->stride != fused->stride)
        return bad("gcam_register_and_fuse_image: in-place fuse needs one stride");
    k_register_and_fuse<<<grid_for(out->width, out->height), kBlock, 0, (hipStream_t)stream>>>(
        view<const float4>(*fused), view<const float4>(*input), view<float4>(*out), homography_of(homography),
        1.0f / (float)out->width, 1.0f / (float)out->height, count);
    return launch_status("registerAndFuse");
}

gcam_status gcam_register_image(const gcam_image* input, gcam_image* out, const float homography[9], void* stream) {
    if (!input || !out || input->format != out->format ||
        (input->format != GCAM_PIXEL_R32F && input->format != GCAM_PIXEL_RGBA32F) ||
        !valid_image(input, input->format) || !valid_image(out, out->format) || !same_size(input, out))
        return bad("gcam_register_image: R32F or RGBA32F input and out of one size and format");
    if (input->data == out->data) return bad("gcam_register_image: input and out must differ");
    if (!homography || !finite9(homography)) return bad("gcam_register_image: homography must be finite");
    const Homography H = homography_of(homography);
    const float nx = 1.0f / (float)out->width, ny = 1.0f / (float)out->height;
    const hipStream_t s = (hipStream_t)stream;
    if (input->format == GCAM_PIXEL_R32F)
        k_register_image<float><<<grid_for(out->width, out->height), kBlock, 0, s>>>(
            view<const float>(*input), view<float>(*out), H, nx, ny);
    else
        k_register_image<float4><<<grid_for(out->width, out->height), kBlock, 0, s>>>(
            view<const float4>(*input), view<float4>(*out), H, nx, ny);
    return launch_status("registerImage");
}

gcam_status gcam_register_bayer_image(const gcam_image* input, gcam_image* out, const float homography[9],
                                      void* stream) {
    if (!valid_image(input, GCAM_PIXEL_R16U) || !valid_image(out, GCAM_PIXEL_R16U) || !same_size(input, out))
        return bad("gcam_register_bayer_image: R16U input and out of one size");
    if (input->data == out->data) return bad("gcam_register_bayer_image: input and out must differ");
    if (!homography || !finite9(homography)) return bad("gcam_register_bayer_image: homography must be finite");
    k_register_bayer<<<grid_for(out->width, out->height), kBlock, 0, (hipStream_t)stream>>>(
        view<const uint16_t>(*input), view<uint16_t>(*out), homography_of(homography));
    return launch_status("registerBayerImage");
}

gcam_status gcam_find_least_squares_homography(const float* matches, int32_t count, float homography[9],
                                               void* stream) {
    if (!matches || !homography || !aligned16(matches))
        return bad("gcam_find_least_squares_homography: 16-byte aligned device matches and an output");
    if (count < 4 || count > kMaxMatches)
        return bad("gcam_find_least_squares_homography: count must be in [4, 2^22]");
    const hipStream_t s = (hipStream_t)stream;
    float* out = nullptr;
    if (hipMalloc(&out, 16 * sizeof(float)) != hipSuccess)
        return set_error(GCAM_ERR_OUT_OF_MEMORY, "gcam_find_least_squares_homography: device allocation failed");
    k_least_squares<<<1, kFitThreads, 0, s>>>(matches, count, out);
    gcam_status st = launch_status("FindLeastSquaresHomography");
    float host[10];
    if (st == GCAM_OK) st = to_host(host, out, sizeof(host), s, "gcam_find_least_squares_homography");
    hipFree(out);
    if (st != GCAM_OK) return st;
    int status;
    std::memcpy(&status, &host[9], sizeof(int));
    if (status != 0) return bad("gcam_find_least_squares_homography: can't generate homography from input points");
    std::copy(host, host + 9, homography);
    return GCAM_OK;
}

gcam_status gcam_find_homography(const float* matches, int32_t count, float threshold, int32_t max_iterations,
                                 uint64_t seed, float homography[9], int32_t* inlier_indices, int32_t* inlier_count,
                                 void* stream) {
    if (!matches || !homography || !inlier_count || !aligned16(matches))
        return bad("gcam_find_homography: 16-byte aligned device matches, homography and inlier_count");
    if (count < 4 || count > kMaxMatches) return bad("gcam_find_homography: count must be in [4, 2^22]");
    if (max_iterations < 1 || max_iterations > kMaxIterations)
        return bad("gcam_find_homography: max_iterations must be in [1, 2^20]");
    if (!std::isfinite(threshold)) return bad("gcam_find_homography: threshold must be finite");
    const hipStream_t s = (hipStream_t)stream;
    const size_t models = (size_t)max_iterations * 9 * sizeof(float), losses = (size_t)max_iterations * sizeof(int);
    const size_t list = (size_t)count * sizeof(int);
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    char* buf = nullptr;
    if (hipMalloc(&buf, up(models) + up(losses) + up(list) + 64) != hipSuccess)
        return set_error(GCAM_ERR_OUT_OF_MEMORY, "gcam_find_homography: device allocation failed");
    float* d_models = reinterpret_cast<float*>(buf);
    int* d_loss = reinterpret_cast<int*>(buf + up(models));
    int* d_list = reinterpret_cast<int*>(buf + up(models) + up(losses));
    float* d_out = reinterpret_cast<float*>(buf + up(models) + up(losses) + up(list));
    k_ransac_hypotheses<<<(max_iterations + 63) / 64, 64 * kRansacWaves, 0, s>>>(matches, count, seed,
                                                                                max_iterations, threshold,
                                                                                d_models, d_loss);
    k_ransac_fit<<<1, kFitThreads, 0, s>>>(matches, count, max_iterations, threshold, d_models, d_loss, d_list,
                                           d_out);
    gcam_status st = launch_status("FindHomography");
    float host[10];
    if (st == GCAM_OK) st = to_host(host, d_out, sizeof(host), s, "gcam_find_homography");
    int c = 0;
    if (st == GCAM_OK) {
        std::memcpy(&c, &host[9], sizeof(int));
        if (inlier_indices && c > 0) st = to_host(inlier_indices, d_list, (size_t)c * sizeof(int), s, "gcam_find_homography");
    }
    hipFree(buf);
    if (st != GCAM_OK) return st;
    std::copy(host, host + 9, homography);
    *inlier_count = c;
    return GCAM_OK;
}

gcam_status gcam_ransac_samples(uint64_t seed, int32_t count, int32_t first, int32_t n, int32_t* out) {
    if (count < 4 || n < 0 || first < 0 || !out) return bad("gcam_ransac_samples: count >= 4, n >= 0, out");
    for (int32_t k = 0; k < n; k++) ransac_sample(seed, first + k, count, out + 4 * (size_t)k);
    return GCAM_OK;
}

void gcam_scale_homography(const float h[9], float scale, float out[9]) {
    const float s[3] = {scale, scale, 1.0f}, inv[3] = {1.0f / scale, 1.0f / scale, 1.0f};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[3 * i + j] = (s[i] * h[3 * i + j]) * inv[j];
}

}  // extern "C"
