// CPU oracle for burst registration and fusion (test infrastructure only; the
// product library never links, loads or calls it).
//
// The warps restate ImageRegistration/SURF.metal:489-559 in FP32 with the
// same conventions as the HIP path: source-order arithmetic (no contraction),
// the linear sampler at exact sub-texel positions, clamp-to-edge reads.
// The homography fit restates LeastSquaresHomography.cpp:189-421 (normalised
// DLT, symmetric Jacobi with the reference's pivoting and stopping rule) and
// FindHomography's RANSAC (Homography.cpp:72-116, RTL/RANSAC.hpp:53-118) over
// the build's own minimal-sample generator (SplitMix64 + Floyd), since the
// reference's std::mt19937 draws depend on the C++ library.
#include "registration.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

namespace {

int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

struct Warped {
    float x, y;
};

// Metal's dot(m_r, (x, y, 1)), left to right; then the perspective divide.
Warped warp(const float* H, float x, float y) {
    const float u = H[0] * x + H[1] * y + H[2] * 1.0f;
    const float v = H[3] * x + H[4] * y + H[5] * 1.0f;
    const float w = H[6] * x + H[7] * y + H[8] * 1.0f;
    return {u / w, v / w};
}

// Linear sampler (clamp to edge) of channel c of an image with ch channels at
// texel-space (tx, ty); positions beyond one texel outside read the edge.
float sample(const float* im, int ch, int c, int w, int h, float tx, float ty) {
    tx = std::fmin(std::fmax(tx, -1.0f), (float)w);
    ty = std::fmin(std::fmax(ty, -1.0f), (float)h);
    const float fx0 = std::floor(tx), fy0 = std::floor(ty);
    const int x0 = (int)fx0, y0 = (int)fy0;
    const float fx = tx - fx0, fy = ty - fy0;
    auto at = [&](int x, int y) { return im[((size_t)clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1)) * ch + c]; };
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    return (at(x0, y0) * gx + at(x0 + 1, y0) * fx) * gy + (at(x0, y0 + 1) * gx + at(x0 + 1, y0 + 1) * fx) * fy;
}

// The sampler's read of SURF.metal:512-513: normalised ((xx, yy) + 0.5) / output size.
void warped_texel(const float* H, int x, int y, int ow, int oh, int iw, int ih, float& tx, float& ty) {
    const Warped p = warp(H, (float)x, (float)y);
    const float nx = 1.0f / (float)ow, ny = 1.0f / (float)oh;
    tx = (p.x + 0.5f) * nx * (float)iw - 0.5f;
    ty = (p.y + 0.5f) * ny * (float)ih - 0.5f;
}

int to_int_sat(float v) {
    if (!(v == v)) return 0;
    return (int)std::fmin(std::fmax(v, -536870912.0f), 536870912.0f);
}

// ---- FindLeastSquaresHomography ----

// Largest |a[k][i]|, i > k (row k's candidate pivot column).
int row_pivot(const double a[9][9], int k) {
    int m = k + 1;
    double mv = std::fabs(a[k][m]);
    for (int i = k + 2; i < 9; i++)
        if (mv < std::fabs(a[k][i])) mv = std::fabs(a[k][i]), m = i;
    return m;
}
// Largest |a[i][k]|, i < k (column k's candidate pivot row).
int col_pivot(const double a[9][9], int k) {
    int m = 0;
    double mv = std::fabs(a[0][k]);
    for (int i = 1; i < k; i++)
        if (mv < std::fabs(a[i][k])) mv = std::fabs(a[i][k]), m = i;
    return m;
}

void givens(double& x, double& y, double c, double s) {
    const double a = x, b = y;
    x = a * c - b * s;
    y = a * s + b * c;
}

// LeastSquaresHomography.cpp:216-349: classic Jacobi on the upper triangle,
// pivot = the largest off-diagonal entry (tracked per row and column), stop
// when it is at most DBL_EPSILON or after 30 N^2 rotations; eigenvalues sorted
// descending with their eigenvectors (rows of v).
void jacobi9(const double A[9][9], double w[9], double v[9][9]) {
    const int N = 9;
    double a[9][9];
    std::copy(&A[0][0], &A[0][0] + 81, &a[0][0]);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) v[i][j] = i == j ? 1.0 : 0.0;
    int rp[9] = {0}, cp[9] = {0};
    for (int k = 0; k < N; k++) {
        w[k] = a[k][k];
        if (k < N - 1) rp[k] = row_pivot(a, k);
        if (k > 0) cp[k] = col_pivot(a, k);
    }
    for (int it = 0; it < N * N * 30; it++) {
        int k = 0, l = rp[0];
        double mv = std::fabs(a[0][rp[0]]);
        for (int i = 1; i < N - 1; i++)
            if (mv < std::fabs(a[i][rp[i]])) mv = std::fabs(a[i][rp[i]]), k = i, l = rp[i];
        for (int i = 1; i < N; i++)
            if (mv < std::fabs(a[cp[i]][i])) mv = std::fabs(a[cp[i]][i]), k = cp[i], l = i;
        const double p = a[k][l];
        if (std::fabs(p) <= DBL_EPSILON) break;
        const double y = (w[l] - w[k]) * 0.5;
        double t = std::fabs(y) + std::hypot(p, y);
        double s = std::hypot(p, t);
        const double c = t / s;
        s = p / s;
        t = (p / t) * p;
        if (y < 0) s = -s, t = -t;
        a[k][l] = 0;
        w[k] -= t;
        w[l] += t;
        for (int i = 0; i < k; i++) givens(a[i][k], a[i][l], c, s);
        for (int i = k + 1; i < l; i++) givens(a[k][i], a[i][l], c, s);
        for (int i = l + 1; i < N; i++) givens(a[k][i], a[l][i], c, s);
        for (int i = 0; i < N; i++) givens(v[k][i], v[l][i], c, s);
        for (int idx : {k, l}) {
            if (idx < N - 1) rp[idx] = row_pivot(a, idx);
            if (idx > 0) cp[idx] = col_pivot(a, idx);
        }
    }
    for (int k = 0; k < N - 1; k++) {
        int m = k;
        for (int i = k + 1; i < N; i++)
            if (w[m] < w[i]) m = i;
        if (m != k) {
            std::swap(w[m], w[k]);
            for (int i = 0; i < N; i++) std::swap(v[m][i], v[k][i]);
        }
    }
}

void mul3(const double a[9], const double b[9], double out[9]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += a[3 * i + k] * b[3 * k + j];
            out[3 * i + j] = s;
        }
}

// Points (x1, y1) -> (x2, y2) of the matches idx[0 .. n-1].
int least_squares(const float* m, const int* idx, int n, float* H) {
    auto P = [&](int i) { return m + 4 * (size_t)(idx ? idx[i] : i); };
    double cM[2] = {0, 0}, cm[2] = {0, 0}, sM[2] = {0, 0}, sm[2] = {0, 0};
    for (int i = 0; i < n; i++) {
        cm[0] += P(i)[2], cm[1] += P(i)[3];
        cM[0] += P(i)[0], cM[1] += P(i)[1];
    }
    for (int k = 0; k < 2; k++) cm[k] /= (double)n, cM[k] /= (double)n;
    for (int i = 0; i < n; i++) {
        sm[0] += std::fabs(P(i)[2] - cm[0]), sm[1] += std::fabs(P(i)[3] - cm[1]);
        sM[0] += std::fabs(P(i)[0] - cM[0]), sM[1] += std::fabs(P(i)[1] - cM[1]);
    }
    for (int k = 0; k < 2; k++) sm[k] /= (double)n, sM[k] /= (double)n;
    if (std::fabs(sm[0]) < DBL_EPSILON || std::fabs(sm[1]) < DBL_EPSILON || std::fabs(sM[0]) < DBL_EPSILON ||
        std::fabs(sM[1]) < DBL_EPSILON)
        return 1;
    double L[9][9] = {};
    for (int i = 0; i < n; i++) {
        const double p1x = (P(i)[0] - cM[0]) / sM[0], p1y = (P(i)[1] - cM[1]) / sM[1];
        const double p2x = (P(i)[2] - cm[0]) / sm[0], p2y = (P(i)[3] - cm[1]) / sm[1];
        const double Lx[9] = {p1x, p1y, 1, 0, 0, 0, -p2x * p1x, -p2x * p1y, -p2x};
        const double Ly[9] = {0, 0, 0, p1x, p1y, 1, -p2y * p1x, -p2y * p1y, -p2y};
        for (int j = 0; j < 9; j++)
            for (int k = j; k < 9; k++) L[j][k] += Lx[j] * Lx[k] + Ly[j] * Ly[k];
    }
    for (int i = 0; i < 9; i++)
        for (int j = 0; j < i; j++) L[i][j] = L[j][i];
    double w[9], v[9][9];
    jacobi9(L, w, v);
    const double inv[9] = {sm[0], 0, cm[0], 0, sm[1], cm[1], 0, 0, 1};
    const double n2[9] = {1 / sM[0], 0, -cM[0] / sM[0], 0, 1 / sM[1], -cM[1] / sM[1], 0, 0, 1};
    double t[9], Hd[9];
    mul3(inv, v[8], t);
    mul3(t, n2, Hd);
    const float c = (float)Hd[8];
    for (int i = 0; i < 9; i++) H[i] = (float)Hd[i] / c;
    return 0;
}

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Floyd's subset sampling: draw d of hypothesis k picks t in [0, j] with
// j = n - 4 + d, taking j itself when t was already drawn.
void sample4(uint64_t seed, int k, int n, int s[4]) {
    for (int d = 0; d < 4; d++) {
        const int j = n - 4 + d;
        const uint64_t r = splitmix64(seed + ((uint64_t)(uint32_t)k << 32) + (uint64_t)d);
        const int t = (int)(((r >> 32) * (uint64_t)(uint32_t)(j + 1)) >> 32);
        s[d] = std::find(s, s + d, t) != s + d ? j : t;
    }
    std::sort(s, s + 4);
}

float transfer_error(const float* h, const float* q) {
    const float u = h[0] * q[0] + h[1] * q[1] + h[2];
    const float v = h[3] * q[0] + h[4] * q[1] + h[5];
    const float w = h[6] * q[0] + h[7] * q[1] + h[8];
    const float dx = u / w - q[2], dy = v / w - q[3];
    return dx * dx + dy * dy;
}

}  // namespace

extern "C" {

void oracle_register_and_fuse(const float* fused4, const float* in4, int ow, int oh, int iw, int ih, const float* H,
                              int count, float* out4) {
    const float c1 = (float)(count - 1), c = (float)count;
    for (int y = 0; y < oh; y++)
        for (int x = 0; x < ow; x++) {
            float tx, ty;
            warped_texel(H, x, y, ow, oh, iw, ih, tx, ty);
            for (int ch = 0; ch < 4; ch++) {
                const size_t o = ((size_t)y * ow + x) * 4 + ch;
                out4[o] = (c1 * fused4[o] + sample(in4, 4, ch, iw, ih, tx, ty)) / c;
            }
        }
}

void oracle_register_image(const float* in, int channels, int ow, int oh, int iw, int ih, const float* H, float* out) {
    for (int y = 0; y < oh; y++)
        for (int x = 0; x < ow; x++) {
            float tx, ty;
            warped_texel(H, x, y, ow, oh, iw, ih, tx, ty);
            for (int ch = 0; ch < channels; ch++)
                out[((size_t)y * ow + x) * channels + ch] = sample(in, channels, ch, iw, ih, tx, ty);
        }
}

void oracle_register_bayer(const uint16_t* in, int w, int h, const float* H, uint16_t* out) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const Warped p = warp(H, (float)(x / 2), (float)(y / 2));
            const int ix = 2 * to_int_sat(std::round(p.x)) + x % 2, iy = 2 * to_int_sat(std::round(p.y)) + y % 2;
            out[(size_t)y * w + x] = in[(size_t)clampi(iy, 0, h - 1) * w + clampi(ix, 0, w - 1)];
        }
}

void oracle_jacobi9(const double* a, double* w, double* v) {
    double A[9][9], V[9][9];
    std::copy(a, a + 81, &A[0][0]);
    jacobi9(A, w, V);
    std::copy(&V[0][0], &V[0][0] + 81, v);
}

int oracle_least_squares_homography(const float* matches, int n, float* H) {
    return least_squares(matches, nullptr, n, H);
}

void oracle_ransac_samples(uint64_t seed, int count, int first, int n, int* out) {
    for (int k = 0; k < n; k++) sample4(seed, first + k, count, out + 4 * (size_t)k);
}

void oracle_find_homography(const float* m, int n, float threshold, int iters, uint64_t seed, float* H,
                            int* inliers, int* inlier_count, int* best_out) {
    float best[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float best_loss = HUGE_VALF;
    int best_k = -1;
    for (int k = 0; k < iters; k++) {
        int s[4];
        sample4(seed, k, n, s);
        float model[9];
        if (least_squares(m, s, 4, model) != 0) {  // ComputeModel's range_error: identity
            const float id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            std::copy(id, id + 9, model);
        }
        float loss = 0;
        for (int i = 0; i < n; i++) loss += (std::fabs(transfer_error(model, m + 4 * (size_t)i)) > threshold);
        if (loss < best_loss) {
            best_loss = loss;
            best_k = k;
            std::copy(model, model + 9, best);
        }
    }
    int c = 0;
    for (int i = 0; i < n; i++)
        if (std::fabs(transfer_error(best, m + 4 * (size_t)i)) < threshold) inliers[c++] = i;
    std::copy(best, best + 9, H);
    if (c >= 4) {
        float r[9];
        if (least_squares(m, inliers, c, r) == 0) std::copy(r, r + 9, H);
    }
    *inlier_count = c;
    best_out[0] = best_k;
    best_out[1] = (int)best_loss;
}

}  // extern "C"
