// CPU oracle for the raw front end that produces the denoise's inputs.
//
// TEST INFRASTRUCTURE ONLY.  A plain FP32 restatement of the steps of
// RawConverter::demosaic (MetalPipeline/raw_converter.cpp:148-213) that turn a
// Bayer raw frame into the YCbCr image and the gradient image given to
// PyramidProcessor::denoise, of the despeckle that RawConverter::denoise runs
// first (raw_converter.cpp:70-78), and of the raw noise-model fit
// (MeasureRawNLF, raw_converter.cpp:339-490, FP64 as in the reference).  Only
// the test suite loads it.
//
// Conventions fixed here (the reference leaves them to Metal):
//   * every out-of-bounds texture.read() clamps to the edge;
//   * the reference computes scaleRawData, rawImageSobel and the despeckle in
//     half precision and reads the 16-bit raw as a half; this oracle works in
//     FP32 throughout (raw / 65535);
//   * the linear sampler of sampledConvolution is evaluated exactly: the
//     sample at texel position x + o (o a kernel offset) interpolates texels
//     x + floor(o) and x + floor(o) + 1 with weight frac(o), the same for every
//     x (Metal's sampler quantises the sub-texel position);
//   * Metal's smoothstep / mix / clamp / length / dot are evaluated in the
//     order written below (left to right), without contraction; min, max and
//     clamp are fmin / fmax (a NaN operand yields the other operand), which
//     is how the Metal specification defines clamp.  A NaN arises where a
//     noise model makes a variance negative (the reference leaves
//     smoothstep of a NaN undefined): the despeckle then clamps the centre to
//     the second extremes.
#include "front_end.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

// MetalPipeline/demosaic.metal:24-38 (BayerPattern, bayerOffsets): the
// (x, y) offset of the red, green, blue and second green sample of a 2x2 quad.
const int kBayer[4][4][2] = {
    {{1, 0}, {0, 0}, {0, 1}, {1, 1}},  // grbg
    {{0, 1}, {0, 0}, {1, 0}, {1, 1}},  // gbrg
    {{0, 0}, {0, 1}, {1, 1}, {1, 0}},  // rggb
    {{1, 1}, {0, 1}, {0, 0}, {1, 0}},  // bggr
};
enum { kRed = 0, kGreen = 1, kBlue = 2, kGreen2 = 3 };

int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

float fmn(float a, float b) { return std::fmin(a, b); }
float fmx(float a, float b) { return std::fmax(a, b); }

// Metal: smoothstep(e0, e1, x) = t^2 (3 - 2t), t = clamp((x - e0) / (e1 - e0), 0, 1).
float smoothstepf(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = fmn(fmx(t, 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
// Metal: mix(a, b, t) = a + (b - a) t.
float mixf(float a, float b, float t) { return a + (b - a) * t; }
// Metal: clamp(x, lo, hi) = fmin(fmax(x, lo), hi).
float clampf(float x, float lo, float hi) { return fmn(fmx(x, lo), hi); }

const float kPi = 3.14159265358979323846f;

// A single-channel or multi-channel FP32 image with clamp-to-edge reads.
struct Img {
    const float* p;
    int w, h, ch;
    float at(int x, int y, int c = 0) const {
        return p[((size_t)clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1)) * ch + c];
    }
};

// Bilinear sample of channel c of a 4-channel image at texel position
// (x + ox, y + oy), texel i centred at integer i.
float bilinear(const Img& im, int x, int y, float ox, float oy, int c) {
    const float fx0 = std::floor(ox), fy0 = std::floor(oy);
    const float fx = ox - fx0, fy = oy - fy0;
    const int x0 = x + (int)fx0, y0 = y + (int)fy0;
    const float s0 = im.at(x0, y0, c) * (1.0f - fx) + im.at(x0 + 1, y0, c) * fx;
    const float s1 = im.at(x0, y0 + 1, c) * (1.0f - fx) + im.at(x0 + 1, y0 + 1, c) * fx;
    return s0 * (1.0f - fy) + s1 * fy;
}

// MetalPipeline/demosaic.metal:170-186 (sampledConvolution): the weighted
// mean of the bilinear samples at the kernel's offsets.
void sampled_convolution(const Img& im, int x, int y, const float* wts, int n, float out[4]) {
    float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f}, norm = 0.0f;
    for (int i = 0; i < n; i++) {
        const float wt = wts[3 * i], ox = wts[3 * i + 1], oy = wts[3 * i + 2];
        for (int c = 0; c < 4; c++) sum[c] = sum[c] + wt * bilinear(im, x, y, ox, oy, c);
        norm = norm + wt;
    }
    for (int c = 0; c < 4; c++) out[c] = sum[c] / norm;
}

}  // namespace

extern "C" {

// MetalPipeline/demosaic.metal:95-117 (scaleRawData), grid of one thread per
// 2x2 quad (demosaic_kernels.hpp:44).  The lens-shading distance is taken at
// the quad's first pixel, like the reference.
void oracle_scale_raw(const uint16_t* raw, int w, int h, int bayer, const float* scale_mul, float black_level,
                      float lens_correction, float* out) {
    const float cx = (float)(w / 2), cy = (float)(h / 2);
    const float rc = std::sqrt(cx * cx + cy * cy);
    for (int qy = 0; qy < h / 2; qy++) {
        for (int qx = 0; qx < w / 2; qx++) {
            const int x = 2 * qx, y = 2 * qy;
            float lens = 1.0f;
            if (lens_correction > 0.0f) {
                const float dx = (float)x - cx, dy = (float)y - cy;
                const float d = std::sqrt(dx * dx + dy * dy) / rc;
                lens = 1.0f + lens_correction * d * d;
            }
            for (int c = 0; c < 4; c++) {
                const int px = x + kBayer[bayer][c][0], py = y + kBayer[bayer][c][1];
                const float v = (float)raw[(size_t)py * w + px] / 65535.0f;
                const float s = lens * scale_mul[c] * (v - black_level) * 0.9f + 0.1f;
                out[(size_t)py * w + px] = fmx(s, 0.0f);
            }
        }
    }
}

// MetalPipeline/demosaic.metal:40-44 (sobelKernel2D), :120-139 (sobel,
// rawImageSobel): (gx, gy, |gx|, |gy|), gx = left - right and gy = top -
// bottom columns/rows weighted 1-2-1, divided by sqrt(4.5).
void oracle_raw_sobel(const float* in, int w, int h, float* out4) {
    static const float kx[3][3] = {{1, 0, -1}, {2, 0, -2}, {1, 0, -1}};
    static const float ky[3][3] = {{1, 2, 1}, {0, 0, 0}, {-1, -2, -1}};
    const Img im{in, w, h, 1};
    const float norm = std::sqrt(4.5f);
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            float gx = 0.0f, gy = 0.0f;
            for (int j = -1; j <= 1; j++) {
                for (int i = -1; i <= 1; i++) {
                    const float s = im.at(x + i, y + j);
                    gx = gx + kx[j + 1][i + 1] * s;
                    gy = gy + ky[j + 1][i + 1] * s;
                }
            }
            gx = gx / norm;
            gy = gy / norm;
            float* o = out4 + ((size_t)y * w + x) * 4;
            o[0] = gx;
            o[1] = gy;
            o[2] = std::fabs(gx);
            o[3] = std::fabs(gy);
        }
    }
}

// MetalPipeline/demosaic_utils.cpp:987-1006 (gaussianKernelBilinearWeights)
// and :939-985 (KernelOptimizeBilinear2d, after J. Chapman's bilinear
// convolution note): the k x k Gaussian (k = max(ceil(2r), 2), made odd) is
// folded into ((k/2)+1)^2 bilinear taps {weight, x offset, y offset}: each
// 2x2 block of weights becomes one tap at the weighted position of the block;
// the last column and row fold pairs; the corner keeps its single weight at
// offset (k/2 + 1/2, k/2 + 1/2) as the reference writes it (width / 2.0f).
// Returns the number of taps (written when cap allows).
int oracle_gaussian_bilinear_weights(float radius, float* out, int cap) {
    int k = std::max((int)std::ceil(2 * radius), 2);
    if (k % 2 == 0) k++;
    const int half = k / 2, ow = half + 1;
    std::vector<float> g((size_t)k * k);
    for (int y = -half; y <= half; y++)
        for (int x = -half; x <= half; x++)
            g[(size_t)(y + half) * k + (x + half)] = std::exp(-((float)(x * x + y * y) / (2 * radius * radius)));
    auto G = [&](int r, int c) { return g[(size_t)r * k + c]; };
    std::vector<float> taps((size_t)ow * ow * 3);
    auto put = [&](int r, int c, float wt, float ox, float oy) {
        float* t = &taps[((size_t)(r / 2) * ow + (c / 2)) * 3];
        t[0] = wt;
        t[1] = ox;
        t[2] = oy;
    };
    int r = 0, c = 0;
    for (r = 0; r < k - 1; r += 2) {
        for (c = 0; c < k - 1; c += 2) {
            const float w1 = G(r, c), w2 = G(r, c + 1), w3 = G(r + 1, c), w4 = G(r + 1, c + 1);
            const float x1 = (float)(c - half), x2 = x1 + 1.0f, y1 = (float)(r - half), y2 = y1 + 1.0f;
            put(r, c, w1 + w2 + w3 + w4, (x1 * w1 + x2 * w2) / (w1 + w2), (y1 * w1 + y2 * w3) / (w1 + w3));
        }
        const float w1 = G(r, c), w2 = G(r + 1, c);
        const float y1 = (float)(r - half), y2 = y1 + 1.0f;
        put(r, c, w1 + w2, (float)(c - half), (y1 * w1 + y2 * w2) / (w1 + w2));
    }
    for (c = 0; c < k - 1; c += 2) {
        const float w1 = G(r, c), w2 = G(r, c + 1);
        const float x1 = (float)(c - half), x2 = x1 + 1.0f;
        put(r, c, w1 + w2, (x1 * w1 + x2 * w2) / (w1 + w2), (float)(r - half));
    }
    put(r, c, G(r, c), k / 2.0f, k / 2.0f);
    const int n = ow * ow;
    if (out && cap >= n) std::copy(taps.begin(), taps.end(), out);
    return n;
}

// MetalPipeline/demosaic.metal:188-210 (sampledConvolutionSobel): blur the
// Sobel image with the narrow kernel; where the blurred gradient is below 4
// sigma of the raw noise (var = a + b * raw), use the wide kernel instead.
// Output: the blurred magnitudes (z, w) with the signs of the blurred signed
// gradient (x, y).
void oracle_blur_sobel(const float* raw, const float* sobel4, int w, int h, const float* w1, int n1,
                       const float* w2, int n2, float var_a, float var_b, float* out2) {
    const Img rs{raw, w, h, 1}, sb{sobel4, w, h, 4};
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            float res[4];
            sampled_convolution(sb, x, y, w1, n1, res);
            const float sigma = std::sqrt(var_a + var_b * rs.at(x, y));
            if (std::sqrt(res[0] * res[0] + res[1] * res[1]) < 4.0f * sigma)
                sampled_convolution(sb, x, y, w2, n2, res);
            float* o = out2 + ((size_t)y * w + x) * 2;
            o[0] = std::copysign(res[2], res[0]);
            o[1] = std::copysign(res[3], res[1]);
        }
    }
}

// MetalPipeline/demosaic.metal:214-300 (interpolateGreen): modified
// Hamilton-Adams green at red and blue sites; green sites copy the raw value.
void oracle_interpolate_green(const float* raw, const float* grad2, int w, int h, int bayer, float gvar_a,
                              float gvar_b, float* green) {
    const Img R{raw, w, h, 1}, G{grad2, w, h, 2};
    const int* r = kBayer[bayer][kRed];
    const int* b = kBayer[bayer][kBlue];
    const float low_noise = 1.0f - smoothstepf(3.5e-4f, 2e-3f, gvar_b);
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            const bool red = (r[0] & 1) == (x & 1) && (r[1] & 1) == (y & 1);
            const bool blue = (b[0] & 1) == (x & 1) && (b[1] & 1) == (y & 1);
            float& out = green[(size_t)y * w + x];
            if (!(red || blue)) {
                out = R.at(x, y);
                continue;
            }
            const float gl = R.at(x - 1, y), gr = R.at(x + 1, y), gu = R.at(x, y - 1), gd = R.at(x, y + 1);
            const float c = R.at(x, y);
            const float cl = R.at(x - 2, y), cr = R.at(x + 2, y), cu = R.at(x, y - 2), cd = R.at(x, y + 2);
            const float d_ave = (R.at(x - 1, y - 1) + R.at(x + 1, y - 1) + R.at(x - 1, y + 1) + R.at(x + 1, y + 1)) / 4;
            const float g_ave = (gl + gr + gu + gd) / 4;
            const float ax = std::fabs(G.at(x, y, 0)), ay = std::fabs(G.at(x, y, 1));
            const float lf_h = (gl + gr) / 2, lf_v = (gu + gd) / 2;
            const float hf_h = ((cl + cr) - 2 * c) / 4, hf_v = ((cu + cd) - 2 * c) / 4;
            const float sd = std::sqrt(gvar_a + gvar_b * g_ave);
            const float glen = std::sqrt(ax * ax + ay * ay);
            const float thr = smoothstepf(sd, 4 * sd, glen);
            const float low_thr = 1.0f - smoothstepf(2 * sd, 8 * sd, glen);
            const float sharpening = (0.5f + 0.5f * low_noise) * thr * (1 + low_noise * low_thr);
            const float highlights = 1.0f - smoothstepf(0.25f, 1.0f, fmx(c, fmx(fmx(cl, cr), fmx(cu, cd))));
            float dir = 2 * std::atan2(ay, ax) / kPi;
            if (gvar_b < 1e-3f) {
                dir = dir < 0.5f ? mixf(dir, 0.0f, 1.0f - smoothstepf(0.3f, 0.45f, dir))
                                 : mixf(dir, 1.0f, smoothstepf(1.0f - 0.45f, 1.0f - 0.3f, dir));
            }
            const float cmin = fmn(c, fmn(g_ave, d_ave)), cmax = fmx(c, fmx(g_ave, d_ave));
            const float white = cmin / cmax;
            const float k = highlights * white * sharpening;
            const float est_h = lf_h - k * hf_h, est_v = lf_v - k * hf_v;
            float g = mixf(est_v, est_h, dir);
            const float gmax = fmx(fmx(gl, gr), fmx(gu, gd));
            const float gmin = fmn(fmn(gl, gr), fmn(gu, gd));
            g = clampf(g, mixf(1.0f, 0.5f, white) * gmin, mixf(1.0f, 1.5f, white) * gmax);
            out = g;
        }
    }
}

namespace {

// MetalPipeline/demosaic.metal:309-380 (interpolateRedBluePixel): the missing
// colour at a red (red_site) or blue site from its diagonal neighbours'
// colour-difference, steered by the gradient direction.
void red_blue_site(const Img& R, const Img& Gn, const Img& Gr, int x, int y, bool red_site, const float* rvar,
                   const float* bvar, float* rgb4) {
    const float g = Gn.at(x, y), c1 = R.at(x, y);
    const float g_tl = Gn.at(x - 1, y - 1), g_tr = Gn.at(x + 1, y - 1), g_bl = Gn.at(x - 1, y + 1), g_br = Gn.at(x + 1, y + 1);
    const float c_tl = R.at(x - 1, y - 1), c_tr = R.at(x + 1, y - 1), c_bl = R.at(x - 1, y + 1), c_br = R.at(x + 1, y + 1);
    const float c2_ave = (c_tl + c_tr + c_bl + c_br) / 4;
    const float d_tl = g_tl - c_tl, d_tr = g_tr - c_tr, d_bl = g_bl - c_bl, d_br = g_br - c_br;
    const float g2_tl = Gn.at(x - 2, y - 2), g2_tr = Gn.at(x + 2, y - 2), g2_bl = Gn.at(x - 2, y + 2), g2_br = Gn.at(x + 2, y + 2);
    const float d2_tl = g2_tl - R.at(x - 2, y - 2), d2_tr = g2_tr - R.at(x + 2, y - 2);
    const float d2_bl = g2_bl - R.at(x - 2, y + 2), d2_br = g2_br - R.at(x + 2, y + 2);
    const float* var = red_site ? rvar : bvar;
    const float sd = std::sqrt(var[0] + var[1] * c2_ave);
    const float ax = std::fabs(Gr.at(x, y, 0)), ay = std::fabs(Gr.at(x, y, 1));
    const float dir = 1.0f - 2 * std::atan2(ay, ax) / kPi;
    const float thr = smoothstepf(sd, 4 * sd, std::sqrt(ax * ax + ay * ay));
    const float alpha = mixf(0.5f, dir, thr);
    const float highlights = 1.0f - smoothstepf(0.25f, 1.0f, fmx(g, fmx(fmx(g2_tr, g2_bl), fmx(g2_tl, g2_br))));
    const float centre = 2 * (g - c1);
    const float anti = (d_tr + d_bl) / 2 + highlights * (d2_tr + d2_bl - centre) / 8;
    const float main = (d_tl + d_br) / 2 + highlights * (d2_tl + d2_br - centre) / 8;
    float c2 = g - mixf(anti, main, alpha);
    const float c2max = fmx(fmx(c_tl, c_tr), fmx(c_bl, c_br));
    const float c2min = fmn(fmn(c_tl, c_tr), fmn(c_bl, c_br));
    c2 = clampf(c2, c2min, c2max);
    rgb4[0] = red_site ? c1 : c2;
    rgb4[1] = g;
    rgb4[2] = red_site ? c2 : c1;
    rgb4[3] = 0.0f;
}

// MetalPipeline/demosaic.metal:420-502 (interpolateRedBlueAtGreenPixel).
void red_blue_at_green_site(const Img& C, const Img& Gr, int x, int y, const float* rvar, const float* bvar,
                            float* rgb4) {
    auto px = [&](int i, int j, int c) { return C.at(x + i, y + j, c); };
    const float gc = px(0, 0, 1);
    const float r_l = px(-1, 0, 0), r_r = px(1, 0, 0), r_u = px(0, -1, 0), r_d = px(0, 1, 0);
    const float b_l = px(-1, 0, 2), b_r = px(1, 0, 2), b_u = px(0, -1, 2), b_d = px(0, 1, 2);
    const float g_l = px(-1, 0, 1), g_r = px(1, 0, 1), g_u = px(0, -1, 1), g_d = px(0, 1, 1);
    const float red_ave = (r_l + r_r + r_u + r_d) / 4, blue_ave = (b_l + b_r + b_u + b_d) / 4;
    // colour differences g - r and g - b at distance 1 and 3
    const float dr_l = g_l - r_l, dr_r = g_r - r_r, dr_u = g_u - r_u, dr_d = g_d - r_d;
    const float db_l = g_l - b_l, db_r = g_r - b_r, db_u = g_u - b_u, db_d = g_d - b_d;
    const float dr_l3 = px(-3, 0, 1) - px(-3, 0, 0), dr_r3 = px(3, 0, 1) - px(3, 0, 0);
    const float dr_u3 = px(0, -3, 1) - px(0, -3, 0), dr_d3 = px(0, 3, 1) - px(0, 3, 0);
    const float db_l3 = px(-3, 0, 1) - px(-3, 0, 2), db_r3 = px(3, 0, 1) - px(3, 0, 2);
    const float db_u3 = px(0, -3, 1) - px(0, -3, 2), db_d3 = px(0, 3, 1) - px(0, 3, 2);
    const float ax = std::fabs(Gr.at(x, y, 0)), ay = std::fabs(Gr.at(x, y, 1));
    const float dir = 2 * std::atan2(ay, ax) / kPi;
    const float glen = std::sqrt(ax * ax + ay * ay);
    const float rsd = std::sqrt(rvar[0] + rvar[1] * red_ave), bsd = std::sqrt(bvar[0] + bvar[1] * blue_ave);
    const float ralpha = mixf(0.5f, dir, smoothstepf(rsd, 4 * rsd, glen));
    const float balpha = mixf(0.5f, dir, smoothstepf(bsd, 4 * bsd, glen));
    const float highlights = 1.0f - smoothstepf(0.25f, 1.0f, fmx(gc, fmx(fmx(g_r, g_l), fmx(g_d, g_u))));
    const float red = gc - mixf((dr_u + dr_d) / 2 - highlights * ((dr_d3 - dr_u) - (dr_d - dr_u3)) / 8,
                                (dr_l + dr_r) / 2 - highlights * ((dr_r3 - dr_l) - (dr_r - dr_l3)) / 8, ralpha);
    const float blue = gc - mixf((db_u + db_d) / 2 - highlights * ((db_d3 - db_u) - (db_d - db_u3)) / 8,
                                 (db_l + db_r) / 2 - highlights * ((db_r3 - db_l) - (db_r - db_l3)) / 8, balpha);
    rgb4[0] = clampf(red, fmn(fmn(r_l, r_r), fmn(r_u, r_d)), fmx(fmx(r_l, r_r), fmx(r_u, r_d)));
    rgb4[1] = gc;
    rgb4[2] = clampf(blue, fmn(fmn(b_l, b_r), fmn(b_u, b_d)), fmx(fmx(b_l, b_r), fmx(b_u, b_d)));
    rgb4[3] = 0.0f;
}

}  // namespace

// MetalPipeline/demosaic.metal:389-416 (interpolateRedBlue), one quad per
// thread: red and blue sites get all three colours, green sites (0, G, 0).
void oracle_interpolate_red_blue(const float* raw, const float* green, const float* grad2, int w, int h, int bayer,
                                 const float* rvar, const float* bvar, float* rgb4) {
    const Img R{raw, w, h, 1}, Gn{green, w, h, 1}, Gr{grad2, w, h, 2};
    const int(*o)[2] = kBayer[bayer];
    for (int qy = 0; qy < h / 2; qy++) {
        for (int qx = 0; qx < w / 2; qx++) {
            const int x = 2 * qx, y = 2 * qy;
            auto at = [&](int c) { return rgb4 + ((size_t)(y + o[c][1]) * w + (x + o[c][0])) * 4; };
            red_blue_site(R, Gn, Gr, x + o[kRed][0], y + o[kRed][1], true, rvar, bvar, at(kRed));
            red_blue_site(R, Gn, Gr, x + o[kBlue][0], y + o[kBlue][1], false, rvar, bvar, at(kBlue));
            for (int c : {kGreen, kGreen2}) {
                float* p = at(c);
                p[0] = 0.0f;
                p[1] = Gn.at(x + o[c][0], y + o[c][1]);
                p[2] = 0.0f;
                p[3] = 0.0f;
            }
        }
    }
}

// MetalPipeline/demosaic.metal:504-521 (interpolateRedBlueAtGreen), one quad
// per thread: green sites are interpolated, red and blue sites copied.
void oracle_interpolate_red_blue_at_green(const float* rgb4_in, const float* grad2, int w, int h, int bayer,
                                          const float* rvar, const float* bvar, float* rgb4_out) {
    const Img C{rgb4_in, w, h, 4}, Gr{grad2, w, h, 2};
    const int(*o)[2] = kBayer[bayer];
    for (int qy = 0; qy < h / 2; qy++) {
        for (int qx = 0; qx < w / 2; qx++) {
            const int x = 2 * qx, y = 2 * qy;
            for (int c : {kGreen, kGreen2})
                red_blue_at_green_site(C, Gr, x + o[c][0], y + o[c][1], rvar, bvar,
                                       rgb4_out + ((size_t)(y + o[c][1]) * w + (x + o[c][0])) * 4);
            for (int c : {kRed, kBlue}) {
                const size_t e = ((size_t)(y + o[c][1]) * w + (x + o[c][0])) * 4;
                for (int k = 0; k < 4; k++) rgb4_out[e + k] = rgb4_in[e + k];
            }
        }
    }
}

// MetalPipeline/demosaic.metal:523-564 (trans, itrans, blendHighlightsImage):
// a pixel with a channel above `clip` keeps its sum of channels and takes the
// chroma length of its clipped copy, in the opponent space (trans).
void oracle_blend_highlights(const float* in4, int w, int h, float clip, float* out4) {
    const float s3 = 1.7320508f;
    const float T[3][3] = {{1, 1, 1}, {s3, -s3, 0}, {-1, -1, 2}};
    const float Ti[3][3] = {{1, s3 / 2, -0.5f}, {1, -s3 / 2, -0.5f}, {1, 0, 1}};
    auto dot3 = [](const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    for (size_t e = 0; e < (size_t)w * h; e++) {
        const float* p = in4 + e * 4;
        float px[3] = {p[0], p[1], p[2]};
        if (px[0] > clip || px[1] > clip || px[2] > clip) {
            const float cl[3] = {fmn(px[0], clip), fmn(px[1], clip), fmn(px[2], clip)};
            float lab0[3], lab1[3];
            for (int i = 0; i < 3; i++) {
                lab0[i] = dot3(T[i], px);
                lab1[i] = dot3(T[i], cl);
            }
            const float sum0 = lab0[1] * lab0[1] + lab0[2] * lab0[2], sum1 = lab1[1] * lab1[1] + lab1[2] * lab1[2];
            const float ratio = sum0 > 0 ? std::sqrt(sum1 / sum0) : 1.0f;
            lab0[1] *= ratio;
            lab0[2] *= ratio;
            for (int i = 0; i < 3; i++) px[i] = dot3(Ti[i], lab0) / 3;
        }
        float* o = out4 + e * 4;
        o[0] = px[0];
        o[1] = px[1];
        o[2] = px[2];
        o[3] = 0.0f;
    }
}

// MetalPipeline/demosaic.metal:649-662 (transformImage): out = M in (rows of
// M dotted with the pixel), fourth channel 0.
void oracle_transform(const float* in4, int w, int h, const float* m, float* out4) {
    for (size_t e = 0; e < (size_t)w * h; e++) {
        const float* p = in4 + e * 4;
        float* o = out4 + e * 4;
        for (int i = 0; i < 3; i++) o[i] = m[3 * i] * p[0] + m[3 * i + 1] * p[1] + m[3 * i + 2] * p[2];
        o[3] = 0.0f;
    }
}

// MetalPipeline/demosaic.metal:1160-1192 (fast_median3x3) and :1242-1283
// (despeckleLumaMedianChromaImage): chroma = per-channel median of the 3x3
// neighbourhood; luma = the centre clamped to [lo, hi], lo between the second
// smallest and the smallest luma of the neighbourhood, hi between the second
// largest and the largest, each moved to the extreme as the gap between them
// grows from 1 to 4 sigma (sigma^2 = var_a.Y + var_b.Y * centre).  The running
// extremes start at 0 (maxima) and 65535 (minima), as in the reference.
void oracle_despeckle(const float* in4, int w, int h, const float* var_a, const float* var_b, float* out4) {
    const Img I{in4, w, h, 4};
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            std::vector<float> lum, cb, cr;
            for (int j = -1; j <= 1; j++)
                for (int i = -1; i <= 1; i++) {
                    lum.push_back(I.at(x + i, y + j, 0));
                    cb.push_back(I.at(x + i, y + j, 1));
                    cr.push_back(I.at(x + i, y + j, 2));
                }
            float max1 = 0.0f, max2 = 0.0f, min1 = 65535.0f, min2 = 65535.0f;
            for (float v : lum) {
                if (v > max1) {
                    max2 = max1;
                    max1 = v;
                } else if (v > max2) {
                    max2 = v;
                }
                if (v < min1) {
                    min2 = min1;
                    min1 = v;
                } else if (v < min2) {
                    min2 = v;
                }
            }
            std::nth_element(cb.begin(), cb.begin() + 4, cb.end());
            std::nth_element(cr.begin(), cr.begin() + 4, cr.end());
            const float centre = I.at(x, y, 0);
            const float sigma = std::sqrt(var_a[0] + var_b[0] * centre);
            const float lo = mixf(min2, min1, smoothstepf(sigma, 4 * sigma, min2 - min1));
            const float hi = mixf(max2, max1, smoothstepf(sigma, 4 * sigma, max1 - max2));
            float* o = out4 + ((size_t)y * w + x) * 4;
            o[0] = clampf(centre, lo, hi);
            o[1] = cb[4];
            o[2] = cr[4];
            o[3] = 0.0f;
        }
    }
}

// MetalPipeline/demosaic_utils.cpp:703-720 (cam_ycbcr) with xyz_rgb = xyz_sRGB
// (demosaic.hpp:138-142, raw_converter.hpp:154): K = row 1 of xyz_sRGB times
// rgb_cam (a row vector times the matrix), then the YCbCr rows of K.
void oracle_cam_ycbcr(const float* rgb_cam, float* out) {
    const float y_row[3] = {0.2126f, 0.7152f, 0.0722f};
    float K[3];
    for (int j = 0; j < 3; j++) K[j] = y_row[0] * rgb_cam[j] + y_row[1] * rgb_cam[3 + j] + y_row[2] * rgb_cam[6 + j];
    const float kr = K[0], kg = K[1], kb = K[2];
    const float m[9] = {kr, kg, kb, -0.5f * kr / (1.0f - kb), -0.5f * kg / (1 - kb), 0.5f,
                        0.5f, -0.5f * kg / (1 - kr), -0.5f * kb / (1 - kr)};
    std::copy(m, m + 9, out);
}

// MetalPipeline/demosaic.metal:1096-1136 (readRAWQuad, basicRawNoiseStatistics):
// per quad (w/2 x h/2), the mean and population variance of the four quad
// samples read at 2 (qx, qy) + (x, y) + offset[c] for x, y in [-4, 4] (the
// window steps by one pixel, so odd steps read the quad shifted by a pixel,
// as the reference does), in FP32, summed row by row.
void oracle_raw_noise_stats(const float* raw, int w, int h, int bayer, float* mean4, float* var4) {
    const Img R{raw, w, h, 1};
    const int qw = w / 2, qh = h / 2;
    const float count = 81.0f;
    for (int qy = 0; qy < qh; qy++) {
        for (int qx = 0; qx < qw; qx++) {
            float sum[4] = {0, 0, 0, 0}, sq[4] = {0, 0, 0, 0};
            for (int y = -4; y <= 4; y++) {
                for (int x = -4; x <= 4; x++) {
                    for (int c = 0; c < 4; c++) {
                        const float v = R.at(2 * qx + x + kBayer[bayer][c][0], 2 * qy + y + kBayer[bayer][c][1]);
                        sum[c] = sum[c] + v;
                        sq[c] = sq[c] + v * v;
                    }
                }
            }
            for (int c = 0; c < 4; c++) {
                mean4[((size_t)qy * qw + qx) * 4 + c] = sum[c] / count;
                var4[((size_t)qy * qw + qx) * 4 + c] = (sq[c] - (sum[c] * sum[c]) / count) / count;
            }
        }
    }
}

// MetalPipeline/raw_converter.cpp:339-490 (MeasureRawNLF), FP64 on the
// statistics images (qw x qh quads; npx = the raw frame's pixel count):
// 1. log10 histogram of the largest channel variance (quads with four positive
//    variances) picks varianceMax;
// 2. per-channel least squares var = A + B mean over quads whose four means
//    are in [0.001, 0.9] and four variances <= varianceMax; B, A >= 1e-8;
// 3. the fit's mean squared error per channel; varianceMax = B;
// 4. refit on the quads whose squared residuals are all <= err2 / 2, kept if
//    they are more than 0.1 % of npx, their error is finite and smaller in
//    every channel (the reference writes newErr2 < err2 on a vector type);
// 5. A and B times exposure_multiplier^2.
// diag (nullable, 8): N1, N2, accepted, varianceMax of pass 1, then the four
// err2 values of the kept fit.
void oracle_fit_raw_nlf(const float* mean4, const float* var4, int qw, int qh, double exposure_multiplier,
                        double npx, double* a4, double* b4, double* diag) {
    const size_t n = (size_t)qw * qh;
    auto nan4 = [](const float* p) { return std::isnan(p[0]) || std::isnan(p[1]) || std::isnan(p[2]) || std::isnan(p[3]); };
    double hist[6] = {0, 0, 0, 0, 0, 0};
    for (size_t k = 0; k < n; k++) {
        const float* v = var4 + 4 * k;
        // log10 needs positive variances: quads with a NaN, zero or negative
        // (rounding) variance are not counted, as in the YCbCr fit.
        if (nan4(v) || !(v[0] > 0 && v[1] > 0 && v[2] > 0 && v[3] > 0)) continue;
        double mx = -INFINITY;
        for (int c = 0; c < 4; c++) mx = std::fmax(mx, std::log10((double)v[c]));
        const int entry = 6 + (int)std::min(std::max(mx, -6.0), -1.0);
        hist[entry] += 1;
    }
    double vmax[4];
    const double vm = hist[0] > 1e3 ? 1.0e-5 : hist[1] > 1e3 ? 1.0e-4 : hist[2] > 1e3 ? 1.0e-3
                    : hist[3] > 1e3 ? 1.0e-2 : hist[4] > 1e3 ? 1.0e-1 : 1.0;
    for (int c = 0; c < 4; c++) vmax[c] = vm;
    auto selected = [&](size_t k) {
        const float* m = mean4 + 4 * k;
        const float* v = var4 + 4 * k;
        if (nan4(m) || nan4(v)) return false;
        for (int c = 0; c < 4; c++)
            if (!((double)m[c] >= 0.001 && (double)m[c] <= 0.9 && (double)v[c] <= vmax[c])) return false;
        return true;
    };
    auto floor_max = [](double x, double f) { return std::isnan(x) ? f : std::max(x, f); };
    double sx[4] = {0}, sy[4] = {0}, sxx[4] = {0}, sxy[4] = {0}, N1 = 0;
    for (size_t k = 0; k < n; k++) {
        if (!selected(k)) continue;
        for (int c = 0; c < 4; c++) {
            const double m = mean4[4 * k + c], v = var4[4 * k + c];
            sx[c] += m;
            sy[c] += v;
            sxx[c] += m * m;
            sxy[c] += m * v;
        }
        N1 += 1;
    }
    double A[4], B[4];
    for (int c = 0; c < 4; c++) {
        B[c] = floor_max((N1 * sxy[c] - sx[c] * sy[c]) / (N1 * sxx[c] - sx[c] * sx[c]), 1e-8);
        A[c] = floor_max((sy[c] - B[c] * sx[c]) / N1, 1e-8);
    }
    double err2[4] = {0, 0, 0, 0};
    for (size_t k = 0; k < n; k++) {
        if (!selected(k)) continue;
        for (int c = 0; c < 4; c++) {
            const double d = A[c] + B[c] * mean4[4 * k + c] - var4[4 * k + c];
            err2[c] += d * d;
        }
    }
    for (int c = 0; c < 4; c++) err2[c] /= N1;
    const double vmax1 = vm;
    for (int c = 0; c < 4; c++) vmax[c] = B[c];
    double tx[4] = {0}, ty[4] = {0}, txx[4] = {0}, txy[4] = {0}, e2n[4] = {0}, N2 = 0;
    for (size_t k = 0; k < n; k++) {
        if (!selected(k)) continue;
        double dsq[4];
        bool fits = true;
        for (int c = 0; c < 4; c++) {
            const double d = std::fabs(A[c] + B[c] * mean4[4 * k + c] - var4[4 * k + c]);
            dsq[c] = d * d;
            fits = fits && dsq[c] <= 0.5 * err2[c];
        }
        if (!fits) continue;
        for (int c = 0; c < 4; c++) {
            const double m = mean4[4 * k + c], v = var4[4 * k + c];
            tx[c] += m;
            ty[c] += v;
            txx[c] += m * m;
            txy[c] += m * v;
            e2n[c] += dsq[c];
        }
        N2 += 1;
    }
    bool better = N2 > 0.001 * npx;
    for (int c = 0; c < 4; c++) {
        e2n[c] /= N2;
        better = better && !std::isnan(e2n[c]) && e2n[c] < err2[c];
    }
    if (better) {
        for (int c = 0; c < 4; c++) {
            err2[c] = e2n[c];
            B[c] = floor_max((N2 * txy[c] - tx[c] * ty[c]) / (N2 * txx[c] - tx[c] * tx[c]), 1e-8);
            A[c] = floor_max((ty[c] - B[c] * tx[c]) / N2, 1e-8);
        }
    }
    const double adj = exposure_multiplier * exposure_multiplier;
    for (int c = 0; c < 4; c++) {
        a4[c] = A[c] * adj;
        b4[c] = B[c] * adj;
    }
    if (diag) {
        diag[0] = N1;
        diag[1] = N2;
        diag[2] = better ? 1.0 : 0.0;
        diag[3] = vmax1;
        for (int c = 0; c < 4; c++) diag[4 + c] = err2[c];
    }
}

// MetalPipeline/demosaic.metal:1357-1384 (histogram_data, histogramImage):
// the 256-bin histogram of index = clamp((int)(255 sqrt(Y / 2)), 0, 255) over
// the image's luma (a NaN index, from negative luma, counts in bin 0).
// Accumulates into hist (the caller zeroes it, like histogramImageKernel::reset).
void oracle_histogram(const float* in4, int w, int h, uint32_t* hist) {
    for (size_t e = 0; e < (size_t)w * h; e++) {
        const float v = std::sqrt(in4[4 * e] / 2.0f) * 255.0f;
        int idx = v >= 0.0f ? (int)v : 0;
        hist[idx < 0 ? 0 : (idx > 255 ? 255 : idx)] += 1;
    }
}

// MetalPipeline/demosaic.metal:1386-1453 (histogramStatistics), one thread:
// bands (8 x 32 bins, added to bands), mean, the bins where the cumulative
// count first reaches half, 0.1 % and 99.9 % of the pixels, and from them
// black / white level, median, mean, highlights, shadows (FP32, in the
// reference's order).  stats: black_level, white_level, shadows, highlights,
// mean, median.
void oracle_histogram_statistics(const uint32_t* hist, uint32_t* bands, int w, int h, float* stats) {
    const uint32_t image_size = (uint32_t)w * (uint32_t)h;
    float mean = 0.0f;
    bool found_median = false, found_black = false, found_white = false;
    int median_index = 0, black_index = 0, white_index = 0;
    uint32_t sum = 0;
    for (int band = 0; band < 8; band++) {
        for (int i = 0; i < 32; i++) {
            const int index = band * 32 + i;
            const uint32_t entry = hist[index];
            bands[band] += entry;
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
    const float black = 2.0f * vb * vb, white = 2.0f * vw * vw;
    stats[0] = black;
    stats[1] = white;
    stats[5] = (float)median_index / 255.0f - black;
    stats[4] = mean - black;
    stats[3] = 1.0f + 2.0f * smoothstepf(0.01f, 0.1f, (float)(bands[5] + bands[6] + bands[7]) / (float)image_size);
    stats[2] = 0.8f - 0.5f * smoothstepf(0.25f, 0.5f, (float)(bands[0] + bands[1]) / (float)image_size);
}

// MetalPipeline/demosaic.metal:1295-1323 (GuidedFilterABImage): per guide
// pixel, the 5x5 box mean and population variance of channel 0 (the linear
// sampler at texel centres reads the texels themselves, clamped), then
// a = var / (var + eps), b = mean (1 - a).
void oracle_guided_filter_ab(const float* guide4, int w, int h, float eps, float* ab2) {
    const Img G{guide4, w, h, 4};
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            float sum = 0.0f, sq = 0.0f;
            for (int j = -2; j <= 2; j++)
                for (int i = -2; i <= 2; i++) {
                    const float v = G.at(x + i, y + j, 0);
                    sum = sum + v;
                    sq = sq + v * v;
                }
            const float mean = sum / 25, var = (sq - sum * sum / 25) / 25;
            const float a = var / (var + eps);
            ab2[((size_t)y * w + x) * 2] = a;
            ab2[((size_t)y * w + x) * 2 + 1] = mean * (1 - a);
        }
    }
}

namespace {
// The 5x5 box as 9 bilinear taps {weight, offset} (demosaic.metal:1329-1339).
const float kBox[9][3] = {{0.16f, -1.5f, -1.5f}, {0.16f, 0.5f, -1.5f}, {0.08f, 2.0f, -1.5f},
                          {0.16f, -1.5f, 0.5f},  {0.16f, 0.5f, 0.5f},  {0.08f, 2.0f, 0.5f},
                          {0.08f, -1.5f, 2.0f},  {0.08f, 0.5f, 2.0f},  {0.04f, 2.0f, 2.0f}};

// Bilinear sample of a 2-channel image (iw x ih) at the texel-space position
// x0 + fx, y0 + fy (texel i centred at i), clamped.
void sample2(const float* img, int iw, int ih, int x0, float fx, int y0, float fy, float out[2]) {
    const Img I{img, iw, ih, 2};
    for (int c = 0; c < 2; c++) {
        const float s0 = I.at(x0, y0, c) * (1.0f - fx) + I.at(x0 + 1, y0, c) * fx;
        const float s1 = I.at(x0, y0 + 1, c) * (1.0f - fx) + I.at(x0 + 1, y0 + 1, c) * fx;
        out[c] = s0 * (1.0f - fy) + s1 * fy;
    }
}

// Texel-space position in an image of size `in` of the centre of pixel `o`
// of an image of size `out`: ((2 o + 1) in - out) / (2 out) = i0 + frac, in
// integers (exact).
void centre_map(int o, int in, int out, int& i0, float& frac) {
    const long long num = (long long)(2 * o + 1) * in - out, den = 2LL * out;
    long long q = num / den;
    if (num % den != 0 && num < 0) q -= 1;
    i0 = (int)q;
    frac = (float)(num - q * den) / (float)den;
}

// computeLtmMultiplier (demosaic.metal:1456-1472).
float ltm_multiplier(float y, float ill, float shadows, float highlights, float detail) {
    const float mid = 0.22f;
    const float reflectance = y / ill;
    const float adj = ill <= mid ? mid * std::pow(ill / mid, shadows)
                                 : (1 - mid) * std::pow((ill - mid) / (1 - mid), highlights) + mid;
    return adj * std::pow(reflectance, detail) / y;
}
}  // namespace

// MetalPipeline/demosaic.metal:1325-1355 (BoxFilterGFImage): the 9 taps in
// table order, sum of weight x sample.
void oracle_box_filter_gf(const float* in2, int w, int h, float* out2) {
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            float acc[2] = {0.0f, 0.0f};
            for (int t = 0; t < 9; t++) {
                const float ox = kBox[t][1], oy = kBox[t][2];
                const float fx0 = std::floor(ox), fy0 = std::floor(oy);
                float s[2];
                sample2(in2, w, h, x + (int)fx0, ox - fx0, y + (int)fy0, oy - fy0, s);
                acc[0] = acc[0] + kBox[t][0] * s[0];
                acc[1] = acc[1] + kBox[t][0] * s[1];
            }
            out2[((size_t)y * w + x) * 2] = acc[0];
            out2[((size_t)y * w + x) * 2 + 1] = acc[1];
        }
    }
}

// MetalPipeline/demosaic.metal:1474-1535 (LTMParameters,
// localToneMappingMaskImage): per pixel, the illuminance a Y + b from each
// band's filtered (a, b), sampled bilinearly at the pixel centre; the
// multipliers of the low band (shadows and highlights from the histogram),
// the mid band and the high band (highlights 1), with the detail gains
// faded by the ringing and noise guards; their product.  params: eps,
// shadows, highlights, detail[3] (only detail enters here).
void oracle_ltm_mask(const float* in4, const float* grad2, int w, int h, const float* lf2, int lw, int lh,
                     const float* mf2, int mw, int mh, const float* hf2, int hw, int hh, const float* params,
                     float nlf_a, float nlf_b, float hist_shadows, float hist_highlights, float* mask) {
    const float* bands[3] = {lf2, mf2, hf2};
    const int bw[3] = {lw, mw, hw}, bh[3] = {lh, mh, hh};
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            const float Y = in4[((size_t)y * w + x) * 4];
            float m = 1.0f;
            for (int k = 0; k < 3; k++) {
                int x0, y0;
                float fx, fy, ab[2];
                centre_map(x, bw[k], w, x0, fx);
                centre_map(y, bh[k], h, y0, fy);
                sample2(bands[k], bw[k], bh[k], x0, fx, y0, fy, ab);
                const float ill = ab[0] * Y + ab[1];
                float detail = params[3 + k];
                if (detail > 1.0f) {
                    float d = (1.0f - smoothstepf(0.75f, 0.95f, ill)) * smoothstepf(0.05f, 0.1f, ill);
                    if (k == 2) {
                        const float gx = grad2[((size_t)y * w + x) * 2], gy = grad2[((size_t)y * w + x) * 2 + 1];
                        const float g = std::sqrt(gx * gx + gy * gy);
                        const float sigma = std::sqrt(nlf_a + nlf_b * ill);
                        d = smoothstepf(2 * sigma, 8 * sigma, g) * (1 - smoothstepf(32 * sigma, 128 * sigma, g)) *
                            (1.0f - smoothstepf(0.75f, 0.95f, ill)) * smoothstepf(0.05f, 0.1f, ill);
                    }
                    detail = 1 + (detail - 1) * d;
                }
                const float mk = ltm_multiplier(Y, ill, hist_shadows, k == 0 ? hist_highlights : 1.0f, detail);
                m = k == 0 ? mk : m * mk;
            }
            mask[(size_t)y * w + x] = m;
        }
    }
}

}  // extern "C"
