// Seeded synthetic raw frames for tests and benchmarks (no counterpart in the
// reference, which reads DNG files).
//
// scene     multi-octave 2-D simplex gradient noise (Perlin's simplex lattice,
//           as described in S. Gustavson, "Simplex noise demystified", 2005),
//           persistence 1/2, lacunarity 2, one independent seed per R/G/B,
//           mapped affinely onto [lo, hi] in linear units
// mosaic    RGGB: R at (0,0), B at (1,1) (MetalPipeline/demosaic.hpp:131)
// noise     Poisson-Gaussian y = a * Poisson(x / a) + N(0, b^2): Poisson by
//           multiplication of uniforms below lambda 10, otherwise W. Hoermann's
//           transformed rejection with squeeze (PTRS, "The transformed
//           rejection method for generating Poisson random variables", 1993);
//           normal by Box-Muller
// quantize  DN = clamp(round(4095 y), 0, 4095), 12 bit
// demosaic  bilinear, reflected at the borders (keeps the CFA parity)
// colour    sRGB -> YCbCr with srgb_ycbcr (MetalPipeline/demosaic_utils.cpp:730-734)
// RNG       Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as
//           1, 2, 3", SC'11), key = noise seed, counter = (pixel, draw)
#include "gcam_internal.hpp"

namespace gcam {

namespace {

struct U4 {
    uint32_t v[4];
};

__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                            uint32_t k1) {
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{{c0, c1, c2, c3}};
}

// Uniform in (0, 1) with 32 random bits.
__device__ __forceinline__ double u01(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }

__device__ __forceinline__ uint32_t hash3(int i, int j, uint32_t seed) {
    uint32_t h = (uint32_t)i * 0x8da6b343u ^ (uint32_t)j * 0xd8163841u ^ seed * 0xcb1ab31fu;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

__device__ float simplex2(float x, float y, uint32_t seed) {
    const float F2 = 0.36602540378f, G2 = 0.21132486540f;
    const float s = (x + y) * F2;
    const int i = (int)floorf(x + s), j = (int)floorf(y + s);
    const float t = (float)(i + j) * G2;
    const float x0 = x - ((float)i - t), y0 = y - ((float)j - t);
    const int i1 = x0 > y0 ? 1 : 0, j1 = 1 - i1;
    const float xs[3] = {x0, x0 - (float)i1 + G2, x0 - 1.0f + 2.0f * G2};
    const float ys[3] = {y0, y0 - (float)j1 + G2, y0 - 1.0f + 2.0f * G2};
    const int is[3] = {i, i + i1, i + 1}, js[3] = {j, j + j1, j + 1};
    float n = 0.0f;
    for (int k = 0; k < 3; k++) {
        float tt = 0.5f - xs[k] * xs[k] - ys[k] * ys[k];
        if (tt > 0.0f) {
            const float ang = (float)hash3(is[k], js[k], seed) * (6.28318530718f / 4294967296.0f);
            float sn, cs;
            sincosf(ang, &sn, &cs);
            tt *= tt;
            n += tt * tt * (cs * xs[k] + sn * ys[k]);
        }
    }
    return 70.0f * n;
}

__device__ float scene_value(int x, int y, int c, const gcam_synth_params& p) {
    float sum = 0.0f, amp = 1.0f, norm = 0.0f, freq = 1.0f / p.period;
    const uint32_t seed = (uint32_t)(p.scene_seed + (uint64_t)c) * 0x9E3779B1u + (uint32_t)((p.scene_seed + c) >> 32);
    for (int o = 0; o < p.octaves; o++) {
        sum += amp * simplex2((float)x * freq, (float)y * freq, seed + (uint32_t)o * 0x68E31DA4u);
        norm += amp;
        amp *= 0.5f;
        freq *= 2.0f;
    }
    float v = sum / norm;
    v = fminf(fmaxf(v, -1.0f), 1.0f);
    return p.lo + (p.hi - p.lo) * (0.5f + 0.5f * v);
}

struct Rng {
    uint32_t k0, k1, c0, c1, draw;
    U4 buf;
    int used;
    __device__ double next() {
        if (used == 4) {
            buf = philox4x32_10(c0, c1, draw++, 0x5ca1ab1eu, k0, k1);
            used = 0;
        }
        return u01(buf.v[used++]);
    }
};

__device__ double poisson(double lam, Rng& rng) {
    if (lam <= 0.0) return 0.0;
    if (lam < 10.0) {
        const double L = exp(-lam);
        double prod = rng.next();
        int k = 0;
        while (prod > L && k < 1000) {
            prod *= rng.next();
            k++;
        }
        return (double)k;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < 1000; it++) {
        const double U = rng.next() - 0.5, V = rng.next();
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return k;
    }
    return floor(lam);
}

__global__ __launch_bounds__(256) void k_synth_raw(int w, int h, gcam_synth_params p, uint16_t* raw,
                                                   View<float4> clean) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= w || y >= h) return;
    const int cfa = (x & 1) + (y & 1);  // 0 = R, 1 = G, 2 = B for RGGB
    const float v = scene_value(x, y, cfa, p);
    const uint64_t idx = (uint64_t)y * w + x;
    Rng rng{(uint32_t)p.noise_seed, (uint32_t)(p.noise_seed >> 32), (uint32_t)idx, (uint32_t)(idx >> 32), 0, {}, 4};
    const double a = p.shot_gain;
    double yv = a > 0.0 ? a * poisson((double)v / a, rng) : (double)v;
    const double u1 = rng.next(), u2 = rng.next();
    yv += (double)p.read_sigma * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    double dn = rint(4095.0 * yv);
    dn = fmin(fmax(dn, 0.0), 4095.0);
    raw[idx] = (uint16_t)dn;
    if (clean.ptr) {
        const float r = scene_value(x, y, 0, p), g = scene_value(x, y, 1, p), b = scene_value(x, y, 2, p);
        clean.at(x, y) = make_float4(0.2126f * r + 0.7152f * g + 0.0722f * b, -0.1146f * r - 0.3854f * g + 0.5f * b,
                                     0.5f * r - 0.4542f * g - 0.0458f * b, 0.0f);
    }
}

__device__ __forceinline__ int reflect(int v, int n) { return v < 0 ? -v : (v >= n ? 2 * n - 2 - v : v); }

__global__ __launch_bounds__(256) void k_demosaic_bilinear(int w, int h, const uint16_t* raw, View<float4> ycc) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= w || y >= h) return;
    auto R = [&](int dx, int dy) {
        return (float)raw[(size_t)reflect(y + dy, h) * w + reflect(x + dx, w)] * (1.0f / 4095.0f);
    };
    const float c = R(0, 0);
    const float cross = 0.25f * (R(-1, 0) + R(1, 0) + R(0, -1) + R(0, 1));
    const float diag = 0.25f * (R(-1, -1) + R(1, -1) + R(-1, 1) + R(1, 1));
    const float horz = 0.5f * (R(-1, 0) + R(1, 0)), vert = 0.5f * (R(0, -1) + R(0, 1));
    float r, g, b;
    const int px = x & 1, py = y & 1;
    if (px == 0 && py == 0) {
        r = c, g = cross, b = diag;
    } else if (px == 1 && py == 1) {
        r = diag, g = cross, b = c;
    } else if (py == 0) {  // green on a red row
        r = horz, g = c, b = vert;
    } else {  // green on a blue row
        r = vert, g = c, b = horz;
    }
    ycc.at(x, y) = make_float4(0.2126f * r + 0.7152f * g + 0.0722f * b, -0.1146f * r - 0.3854f * g + 0.5f * b,
                               0.5f * r - 0.4542f * g - 0.0458f * b, 0.0f);
}

}  // namespace

gcam_status launch_synth(int w, int h, const gcam_synth_params& p, uint16_t* raw, View<float4> ycc,
                         View<float4> clean, hipStream_t s) {
    const dim3 grid((w + 63) / 64, (h + 3) / 4), block(64, 4);
    k_synth_raw<<<grid, block, 0, s>>>(w, h, p, raw, clean);
    GCAM_LAUNCH_CHECK("synthetic raw");
    if (ycc.ptr) {
        k_demosaic_bilinear<<<grid, block, 0, s>>>(w, h, raw, ycc);
        GCAM_LAUNCH_CHECK("bilinear demosaic");
    }
    return GCAM_OK;
}

}  // namespace gcam
