// CPU oracle for the pyramid PCA block-matching denoise -- TEST INFRASTRUCTURE.
// Restates MetalPipeline/demosaic.metal, MetalPipeline/PCA.cpp and
// MetalPipeline/pyramid_processor.cpp on the CPU; see gcam_oracle.hpp for the
// semantics this file fixes where the reference leaves them undefined.
#include "gcam_oracle.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace {

int g_threads = 1;

// Row-sliced parallel loop (the reference's CPU paths slice rows over a
// ThreadPool, MetalPipeline/PCA.cpp:151-176).
void parallel_rows(int h, const std::function<void(int, int)>& fn) {
    int nt = std::max(1, std::min(g_threads, h));
    if (nt == 1) {
        fn(0, h);
        return;
    }
    std::vector<std::thread> pool;
    std::atomic<int> next{0};
    const int chunk = std::max(1, h / (nt * 8));
    for (int t = 0; t < nt; t++) {
        pool.emplace_back([&] {
            for (;;) {
                int y0 = next.fetch_add(chunk);
                if (y0 >= h) break;
                fn(y0, std::min(h, y0 + chunk));
            }
        });
    }
    for (auto& th : pool) th.join();
}

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <class T>
inline T smoothstep(T e0, T e1, T x) {
    // Metal smoothstep: t = clamp((x - e0) / (e1 - e0), 0, 1); t * t * (3 - 2 t)
    T t = (x - e0) / (e1 - e0);
    t = t < T(0) ? T(0) : (t > T(1) ? T(1) : t);
    return t * t * (T(3) - T(2) * t);
}

// demosaic.metal:87-89
template <class T>
inline T gaussian(T x) {
    return std::exp(-T(2) * x * x);
}

// One axis of a linear-filtered, normalized-coordinate, clamp-to-edge sample
// whose position in texel units is num/den (i.e. the normalized coordinate
// times the texture size).  The texel-centre offset of 0.5 is applied exactly:
// u = num/den - 1/2 = (2 num - den) / (2 den).
template <class T>
inline void lin_taps(long num, long den, int size, int* i0, int* i1, T* frac) {
    long n2 = 2 * num - den, d2 = 2 * den;
    long f = n2 >= 0 ? n2 / d2 : -((-n2 + d2 - 1) / d2);  // floor division
    long rem = n2 - f * d2;
    *frac = T(rem) / T(d2);
    *i0 = clampi((int)f, 0, size - 1);
    *i1 = clampi((int)f + 1, 0, size - 1);
}

template <class T>
inline void bilinear(const T* img, int w, int /*h*/, int ch, int x0, int x1, T a, int y0, int y1, T b, T* out, int nout) {
    for (int c = 0; c < nout; c++) {
        T t00 = img[((size_t)y0 * w + x0) * ch + c], t01 = img[((size_t)y0 * w + x1) * ch + c];
        T t10 = img[((size_t)y1 * w + x0) * ch + c], t11 = img[((size_t)y1 * w + x1) * ch + c];
        T top = (T(1) - a) * t00 + a * t01;
        T bot = (T(1) - a) * t10 + a * t11;
        out[c] = (T(1) - b) * top + b * bot;
    }
}

// demosaic.metal:886-920.  Output o samples the input at normalized positions
// (o + 1/2 -+ 1/2) / dim_out, i.e. at texel positions o*win/wout and
// (o+1)*win/wout, with the linear sampler (clamp_to_edge), and averages the 4
// samples.
template <class T>
void downsample(const T* in, int win, int hin, int ch, T* out, int wout, int hout) {
    const int nc = ch == 4 ? 3 : 2;
    parallel_rows(hout, [&](int ya, int yb) {
        for (int oy = ya; oy < yb; oy++) {
            int yi[2][2];
            T yf[2];
            for (int k = 0; k < 2; k++) lin_taps<T>((long)(oy + k) * hin, hout, hin, &yi[k][0], &yi[k][1], &yf[k]);
            for (int ox = 0; ox < wout; ox++) {
                int xi[2][2];
                T xf[2];
                for (int k = 0; k < 2; k++)
                    lin_taps<T>((long)(ox + k) * win, wout, win, &xi[k][0], &xi[k][1], &xf[k]);
                T acc[3] = {0, 0, 0};
                // sample order of the reference: (-s,-s), (+s,-s), (-s,+s), (+s,+s)
                for (int ky = 0; ky < 2; ky++) {
                    for (int kx = 0; kx < 2; kx++) {
                        T s[3];
                        bilinear(in, win, hin, ch, xi[kx][0], xi[kx][1], xf[kx], yi[ky][0], yi[ky][1], yf[ky], s, nc);
                        for (int c = 0; c < nc; c++) acc[c] += s[c];
                    }
                }
                T* o = out + ((size_t)oy * wout + ox) * ch;
                for (int c = 0; c < nc; c++) o[c] = T(0.25) * acc[c];
                if (ch == 4) o[3] = 0;
            }
        }
    });
}

// demosaic.metal:922-959.  in: level i (w x h), in1 / den1: the noisy and the
// denoised level i+1 (w1 x h1), sampled bilinearly at (o + 1/2) / dim_i.
template <class T>
void subtract_noise(const T* in, int w, int h, const T* in1, const T* den1, int w1, int h1, const T* grad,
                    T luma_weight, T sharpening, T nlf_x, T nlf_y, T* out) {
    parallel_rows(h, [&](int ya, int yb) {
        for (int oy = ya; oy < yb; oy++) {
            int y0, y1;
            T fy;
            lin_taps<T>((long)(2 * oy + 1) * h1, 2L * h, h1, &y0, &y1, &fy);
            for (int ox = 0; ox < w; ox++) {
                int x0, x1;
                T fx;
                lin_taps<T>((long)(2 * ox + 1) * w1, 2L * w, w1, &x0, &x1, &fx);
                const T* ip = in + ((size_t)oy * w + ox) * 4;
                T p1[3], d1[3];
                bilinear(in1, w1, h1, 4, x0, x1, fx, y0, y1, fy, p1, 3);
                bilinear(den1, w1, h1, 4, x0, x1, fx, y0, y1, fy, d1, 3);
                const T lw[3] = {luma_weight, T(1), T(1)};
                T dp[3];
                for (int c = 0; c < 3; c++) dp[c] = ip[c] - lw[c] * (p1[c] - d1[c]);
                T alpha = sharpening;
                if (alpha > T(1)) {
                    T gx = 0, gy = 0;
                    if (grad) {
                        gx = grad[((size_t)oy * w + ox) * 2 + 0];
                        gy = grad[((size_t)oy * w + ox) * 2 + 1];
                    }
                    T gradient = std::sqrt(gx * gx + gy * gy);
                    T sigma = std::sqrt(nlf_x + nlf_y * d1[0]);
                    T detail = smoothstep(sigma, T(4) * sigma, gradient) *
                               (T(1) - smoothstep(T(0.75), T(0.95), d1[0])) * smoothstep(T(0.05), T(0.1), d1[0]);
                    alpha = T(1) + (alpha - T(1)) * detail;
                }
                T* o = out + ((size_t)oy * w + ox) * 4;
                for (int c = 0; c < 3; c++) o[c] = d1[c] + (dp[c] - d1[c]) * alpha;  // Metal mix()
                o[0] = std::max(o[0], T(0));
                o[3] = ip[3];
            }
        }
    });
}

// demosaic.metal:755-772, grid (w/8, h/8) (demosaic_kernels.hpp:388), n from
// pyramid_processor.cpp:142.  Slot = y*w/8 + x with C precedence ((y*w)/8).
template <class T>
int collect_patches(const T* in, int w, int h, T* patches) {
    const int n = (int)(((long)w * h) / 64);
    std::memset(patches, 0, sizeof(T) * 25 * (size_t)n);
    for (int y = 0; y < h / 8; y++) {
        for (int x = 0; x < w / 8; x++) {
            const long slot = (long)y * w / 8 + x;
            for (int j = -2; j <= 2; j++) {
                for (int i = -2; i <= 2; i++) {
                    int sx = clampi(8 * x + i, 0, w - 1), sy = clampi(8 * y + j, 0, h - 1);
                    if (slot < n) patches[slot * 25 + (j + 2) * 5 + (i + 2)] = in[((size_t)sy * w + sx) * 4];
                }
            }
        }
    }
    return n;
}

// PCA.cpp:31-32: centred = X - mean(X); C = centred^T centred / (n - 1).
template <class T>
void patch_covariance(const T* patches, int n, double* cov) {
    double mean[25] = {0};
    for (int k = 0; k < n; k++)
        for (int r = 0; r < 25; r++) mean[r] += (double)patches[(size_t)k * 25 + r];
    for (int r = 0; r < 25; r++) mean[r] /= n;
    std::vector<double> acc(625, 0.0);
    for (int k = 0; k < n; k++) {
        double c[25];
        for (int r = 0; r < 25; r++) c[r] = (double)patches[(size_t)k * 25 + r] - mean[r];
        for (int r = 0; r < 25; r++)
            for (int s = r; s < 25; s++) acc[r * 25 + s] += c[r] * c[s];
    }
    for (int r = 0; r < 25; r++)
        for (int s = r; s < 25; s++) cov[r * 25 + s] = cov[s * 25 + r] = acc[r * 25 + s] / (n - 1);
}

}  // namespace

// Fixed sign-reference vector: every basis vector is oriented to have a
// non-negative dot product with it.  The device code uses the same formula.
static inline double sign_ref(int r) { return 1.0 + (double)((r * r * 7 + r * 3 + 5) % 23) / 16.0; }

extern "C" {

void oracle_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int oracle_get_threads(void) { return g_threads; }

void oracle_downsample_f32(const float* in, int win, int hin, int ch, float* out, int wout, int hout) {
    downsample(in, win, hin, ch, out, wout, hout);
}
void oracle_downsample_f64(const double* in, int win, int hin, int ch, double* out, int wout, int hout) {
    downsample(in, win, hin, ch, out, wout, hout);
}

void oracle_subtract_noise_f32(const float* in, int w, int h, const float* in1, const float* den1, int w1, int h1,
                               const float* grad, float lw, float sh, float nx, float ny, float* out) {
    subtract_noise(in, w, h, in1, den1, w1, h1, grad, lw, sh, nx, ny, out);
}
void oracle_subtract_noise_f64(const double* in, int w, int h, const double* in1, const double* den1, int w1, int h1,
                               const double* grad, double lw, double sh, double nx, double ny, double* out) {
    subtract_noise(in, w, h, in1, den1, w1, h1, grad, lw, sh, nx, ny, out);
}

int oracle_collect_patches_f32(const float* in, int w, int h, float* patches) {
    return collect_patches(in, w, h, patches);
}
int oracle_collect_patches_f64(const double* in, int w, int h, double* patches) {
    return collect_patches(in, w, h, patches);
}

void oracle_patch_covariance_f32(const float* patches, int n, double* cov) { patch_covariance(patches, n, cov); }
void oracle_patch_covariance_f64(const double* patches, int n, double* cov) { patch_covariance(patches, n, cov); }

// Cyclic Jacobi eigen-decomposition of a symmetric matrix (Golub & Van Loan,
// "Matrix Computations", 4th ed., Alg. 8.5.3, with the stable rotation of
// 8.5.2).  Stands in for Eigen's SelfAdjointEigenSolver (PCA.cpp:35).  The
// eigenpairs are returned sorted by decreasing eigenvalue (ties keep index
// order); eigenvector j is column j of eigvecs (row-major n x n).
void oracle_eigh_jacobi(const double* cov, int n, double* eigvals, double* eigvecs) {
    std::vector<double> a(cov, cov + (size_t)n * n), v((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) v[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = 0, tot = 0;
        for (int p = 0; p < n; p++)
            for (int q = 0; q < n; q++) {
                double e = a[(size_t)p * n + q] * a[(size_t)p * n + q];
                tot += e;
                if (p != q) off += e;
            }
        // stop once the off-diagonal energy is below 1e-26 of the total
        // (entries ~1e-13 of the norm, above FP64 rounding noise)
        if (off <= 1e-26 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; p++) {
            for (int q = p + 1; q < n; q++) {
                double apq = a[(size_t)p * n + q];
                if (apq == 0.0) continue;
                double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {  // A <- A J (columns p, q)
                    double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
                    a[(size_t)k * n + p] = c * akp - s * akq;
                    a[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {  // A <- J^T A (rows p, q)
                    double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
                    a[(size_t)p * n + k] = c * apk - s * aqk;
                    a[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {  // V <- V J
                    double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
                    v[(size_t)k * n + p] = c * vkp - s * vkq;
                    v[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
        }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int i, int j) { return a[(size_t)i * n + i] > a[(size_t)j * n + j]; });
    for (int j = 0; j < n; j++) {
        int src = idx[j];
        if (eigvals) eigvals[j] = a[(size_t)src * n + src];
        for (int k = 0; k < n; k++) eigvecs[(size_t)k * n + j] = v[(size_t)k * n + src];
    }
}

static void basis_from_cov(const double* cov, double* basis, double* eigvals) {
    double vals[25], vecs[625];
    oracle_eigh_jacobi(cov, 25, vals, vecs);
    // PCA.cpp:49-55: the 8 largest eigenvectors, largest first, as columns.
    for (int c = 0; c < 8; c++) {
        double dot = 0;
        for (int r = 0; r < 25; r++) dot += vecs[r * 25 + c] * sign_ref(r);
        double sgn = dot < 0 ? -1.0 : 1.0;
        for (int r = 0; r < 25; r++) basis[r * 8 + c] = sgn * vecs[r * 25 + c];
    }
    if (eigvals)
        for (int i = 0; i < 25; i++) eigvals[i] = vals[i];
}

void oracle_build_pca_space_f32(const float* patches, int n, double* basis, double* eigvals) {
    double cov[625];
    patch_covariance(patches, n, cov);
    basis_from_cov(cov, basis, eigvals);
}
void oracle_build_pca_space_f64(const double* patches, int n, double* basis, double* eigvals) {
    double cov[625];
    patch_covariance(patches, n, cov);
    basis_from_cov(cov, basis, eigvals);
}

}  // extern "C"

namespace {

// demosaic.metal:774-796: f[c] = sum_row basis[row][c] * Y(p + off(row)),
// rows j-outer / i-inner, clamp-to-edge reads.
template <class T>
void pca_project(const T* in, int w, int h, const T* basis, T* feat) {
    parallel_rows(h, [&](int ya, int yb) {
        for (int y = ya; y < yb; y++) {
            for (int x = 0; x < w; x++) {
                T f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                int row = 0;
                for (int j = -2; j <= 2; j++) {
                    for (int i = -2; i <= 2; i++) {
                        int sx = clampi(x + i, 0, w - 1), sy = clampi(y + j, 0, h - 1);
                        T val = in[((size_t)sy * w + sx) * 4];
                        for (int c = 0; c < 8; c++) f[c] += basis[row * 8 + c] * val;
                        row++;
                    }
                }
                std::memcpy(feat + ((size_t)y * w + x) * 8, f, sizeof(f));
            }
        }
    });
}

// demosaic.metal:806-884 with gaussian() (:87-89) and lensShading() (:91-93).
template <class T>
void block_match(const T* in, const T* grad, const T* feat, int w, int h, const T var_a[3], const T var_b[3],
                 const T thr[3], T chromaBoost, T gradientBoost, T gradientThreshold, T lensShadingCorrection,
                 int radius, T* out) {
    parallel_rows(h, [&](int ya, int yb) {
        for (int py = ya; py < yb; py++) {
            for (int px = 0; px < w; px++) {
                const T* ip = in + ((size_t)py * w + px) * 4;
                const T* fp = feat + ((size_t)py * w + px) * 8;
                T lens_shading = 1;
                if (lensShadingCorrection > 0) {
                    // get_image_dim() is int2, so the centre is (w/2, h/2) in integers.
                    T cx = T(w / 2), cy = T(h / 2);
                    T dx = T(px) - cx, dy = T(py) - cy;
                    T d = std::sqrt(dx * dx + dy * dy) / std::sqrt(cx * cx + cy * cy);
                    lens_shading = T(1) + lensShadingCorrection * d * d;
                }
                T sigma[3], dm[3];
                for (int c = 0; c < 3; c++) {
                    sigma[c] = std::sqrt(var_a[c] + var_b[c] * lens_shading * ip[0]);
                    dm[c] = T(1) / (thr[c] * sigma[c]);
                }
                T gx = 0, gy = 0;
                if (grad) {
                    gx = grad[((size_t)py * w + px) * 2];
                    gy = grad[((size_t)py * w + px) * 2 + 1];
                }
                T magnitude = std::sqrt(gx * gx + gy * gy);
                T edge = smoothstep(T(2), T(16), gradientThreshold * magnitude / sigma[0]);
                T fil[3] = {0, 0, 0}, norm[3] = {0, 0, 0};
                for (int y = -radius; y <= radius; y++) {
                    for (int x = -radius; x <= radius; x++) {
                        int sx = clampi(px + x, 0, w - 1), sy = clampi(py + y, 0, h - 1);
                        const T* sp = in + ((size_t)sy * w + sx) * 4;
                        const T* fq = feat + ((size_t)sy * w + sx) * 8;
                        T sum = 0;
                        for (int c = 0; c < 8; c++) {
                            T d = fq[c] - fp[c];
                            sum += d * d;
                        }
                        T pcaDiff = std::sqrt(sum) * dm[0];
                        T cdy = (sp[1] - ip[1]) * dm[1], cdz = (sp[2] - ip[2]) * dm[2];
                        T lumaWeight = gaussian(pcaDiff / (T(1) + gradientBoost * edge));
                        T cy = cdy / chromaBoost, cz = cdz / chromaBoost;
                        T chromaWeight = gaussian(std::sqrt(lumaWeight * lumaWeight + cy * cy + cz * cz));
                        T directionWeight = gaussian(T(0.1) * std::sqrt(T(x * x + y * y)));
                        T sw[3] = {directionWeight * lumaWeight, directionWeight * chromaWeight,
                                   directionWeight * chromaWeight};
                        for (int c = 0; c < 3; c++) {
                            fil[c] += sw[c] * sp[c];
                            norm[c] += sw[c];
                        }
                    }
                }
                T* o = out + ((size_t)py * w + px) * 4;
                for (int c = 0; c < 3; c++) o[c] = fil[c] / norm[c];
                o[3] = norm[0];
            }
        }
    });
}

// demosaic.metal:664-719: tunnel() and denoiseImage.  Metal's step(e, x) is
// x < e ? 0 : 1; length() sums x, y, z in order.
template <class T>
void denoise_image(const T* in, const T* grad, int w, int h, const T var_a[3], const T var_b[3], const T thr[3],
                   T chromaBoost, T gradientBoost, T gradientThreshold, T* out) {
    parallel_rows(h, [&](int ya, int yb) {
        for (int py = ya; py < yb; py++) {
            for (int px = 0; px < w; px++) {
                const T* ip = in + ((size_t)py * w + px) * 4;
                T sigma[3], dm[3];
                for (int c = 0; c < 3; c++) {
                    sigma[c] = std::sqrt(var_a[c] + var_b[c] * ip[0]);
                    dm[c] = T(1) / (thr[c] * sigma[c]);
                }
                T gx = 0, gy = 0;
                if (grad) {
                    gx = grad[((size_t)py * w + px) * 2];
                    gy = grad[((size_t)py * w + px) * 2 + 1];
                }
                const T angle = std::atan2(gy, gx);
                const T magnitude = std::sqrt(gx * gx + gy * gy);
                const T edge = smoothstep(T(4), T(16), gradientThreshold * magnitude / sigma[0]);
                const T ca = std::cos(angle), sa = std::sin(angle);
                const int size = gradientBoost > 0 ? 4 : 2;
                T fil[3] = {0, 0, 0}, norm[3] = {0, 0, 0};
                for (int y = -size; y <= size; y++) {
                    for (int x = -size; x <= size; x++) {
                        const int sx = clampi(px + x, 0, w - 1), sy = clampi(py + y, 0, h - 1);
                        const T* sp = in + ((size_t)sy * w + sx) * 4;
                        T sgx = 0, sgy = 0;
                        if (grad) {
                            sgx = grad[((size_t)sy * w + sx) * 2];
                            sgy = grad[((size_t)sy * w + sx) * 2 + 1];
                        }
                        T d[3];
                        for (int c = 0; c < 3; c++) d[c] = (sp[c] - ip[c]) * dm[c];
                        const T gdx = (sgx - gx) / sigma[0], gdy = (sgy - gy) / sigma[0];
                        const T a = T(x) * ca + T(y) * sa;
                        const T tunnel = std::exp(-(a * a) / T(0.25));
                        const T directionWeight = T(1) + (tunnel - T(1)) * edge;  // mix(1, tunnel, edge)
                        const T gradientWeight = T(1) - smoothstep(T(2), T(8), std::sqrt(gdx * gdx + gdy * gdy));
                        const T lumaWeight = std::fabs(d[0]) < T(1) + gradientBoost * edge ? T(1) : T(0);
                        const T chromaWeight =
                            std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) < chromaBoost ? T(1) : T(0);
                        const T sw[3] = {directionWeight * gradientWeight * lumaWeight, chromaWeight, chromaWeight};
                        for (int c = 0; c < 3; c++) {
                            fil[c] += sw[c] * sp[c];
                            norm[c] += sw[c];
                        }
                    }
                }
                T* o = out + ((size_t)py * w + px) * 4;
                for (int c = 0; c < 3; c++) o[c] = fil[c] / norm[c];
                o[3] = magnitude;
            }
        }
    });
}

// demosaic.metal:1138-1159, clamp-to-edge reads.
template <class T>
void noise_stats(const T* in, int w, int h, T* stats) {
    parallel_rows(h, [&](int ya, int yb) {
        for (int py = ya; py < yb; py++) {
            for (int px = 0; px < w; px++) {
                T sum[3] = {0, 0, 0}, sq[3] = {0, 0, 0};
                for (int y = -2; y <= 2; y++) {
                    for (int x = -2; x <= 2; x++) {
                        int sx = clampi(px + x, 0, w - 1), sy = clampi(py + y, 0, h - 1);
                        const T* s = in + ((size_t)sy * w + sx) * 4;
                        for (int c = 0; c < 3; c++) {
                            sum[c] += s[c];
                            sq[c] += s[c] * s[c];
                        }
                    }
                }
                T* o = stats + ((size_t)py * w + px) * 4;
                T var[3];
                for (int c = 0; c < 3; c++) var[c] = (sq[c] - (sum[c] * sum[c]) / T(25)) / T(25);
                o[0] = sum[0] / T(25);
                o[1] = var[0];
                o[2] = var[1];
                o[3] = var[2];
            }
        }
    });
}

inline double floor_max(double v, double floor) { return std::isnan(v) ? floor : std::max(v, floor); }

}  // namespace

extern "C" {

void oracle_pca_project_f32(const float* in, int w, int h, const float* basis, float* feat) {
    pca_project(in, w, h, basis, feat);
}
void oracle_pca_project_f64(const double* in, int w, int h, const double* basis, double* feat) {
    pca_project(in, w, h, basis, feat);
}

void oracle_block_match_f32(const float* in, const float* grad, const float* feat, int w, int h, const float var_a[3],
                            const float var_b[3], const float thr[3], float cB, float gB, float gT, float ls,
                            int radius, float* out) {
    block_match(in, grad, feat, w, h, var_a, var_b, thr, cB, gB, gT, ls, radius, out);
}
void oracle_block_match_f64(const double* in, const double* grad, const double* feat, int w, int h,
                            const double var_a[3], const double var_b[3], const double thr[3], double cB, double gB,
                            double gT, double ls, int radius, double* out) {
    block_match(in, grad, feat, w, h, var_a, var_b, thr, cB, gB, gT, ls, radius, out);
}

void oracle_denoise_image_f32(const float* in, const float* grad, int w, int h, const float var_a[3],
                              const float var_b[3], const float thr[3], float cB, float gB, float gT, float* out) {
    denoise_image(in, grad, w, h, var_a, var_b, thr, cB, gB, gT, out);
}
void oracle_denoise_image_f64(const double* in, const double* grad, int w, int h, const double var_a[3],
                              const double var_b[3], const double thr[3], double cB, double gB, double gT,
                              double* out) {
    denoise_image(in, grad, w, h, var_a, var_b, thr, cB, gB, gT, out);
}

void oracle_noise_stats_f32(const float* in, int w, int h, float* stats) { noise_stats(in, w, h, stats); }
void oracle_noise_stats_f64(const double* in, int w, int h, double* stats) { noise_stats(in, w, h, stats); }

// pyramid_processor.cpp:186-323.
int oracle_fit_ycbcr_nlf(const float* stats, int w, int h, double exposure_multiplier, double A[3], double B[3],
                         double* diag) {
    const size_t npx = (size_t)w * h;
    double hist[6] = {0, 0, 0, 0, 0, 0};
    for (size_t k = 0; k < npx; k++) {  // :190-201
        const float* ns = stats + k * 4;
        double v[3] = {ns[1], ns[2], ns[3]};
        bool valid = true;
        for (int c = 0; c < 3; c++) valid = valid && !std::isnan(v[c]) && v[c] > 0.0;
        if (!valid) continue;
        double mx = -INFINITY;
        for (int c = 0; c < 3; c++) mx = std::max(mx, std::log10(v[c]));
        int entry = 6 + (int)std::clamp(mx, -6.0, -1.0);
        hist[entry]++;
    }
    double vmax1 = hist[0] > 1e3 ? 1.0e-5 : hist[1] > 1e3 ? 1.0e-4 : hist[2] > 1e3 ? 1.0e-3
                 : hist[3] > 1e3 ? 1.0e-2 : hist[4] > 1e3 ? 1.0e-1 : 1;  // :204-208
    double varianceMax[3] = {vmax1, vmax1, vmax1};
    const double maxValue = 0.9, minValue = 0.001;

    auto selected = [&](const float* ns, double& m, double v[3]) {
        m = ns[0];
        v[0] = ns[1];
        v[1] = ns[2];
        v[2] = ns[3];
        bool valid = !(std::isnan(m) || std::isnan(v[0]) || std::isnan(v[1]) || std::isnan(v[2]));
        return valid && m >= minValue && m <= maxValue && v[0] <= varianceMax[0] && v[1] <= varianceMax[1] &&
               v[2] <= varianceMax[2];
    };

    double s_x = 0, s_xx = 0, s_y[3] = {0, 0, 0}, s_xy[3] = {0, 0, 0}, N = 0;
    for (size_t k = 0; k < npx; k++) {  // :223-236
        double m, v[3];
        if (!selected(stats + k * 4, m, v)) continue;
        s_x += m;
        s_xx += m * m;
        for (int c = 0; c < 3; c++) {
            s_y[c] += v[c];
            s_xy[c] += m * v[c];
        }
        N++;
    }
    double nlfA[3], nlfB[3];
    for (int c = 0; c < 3; c++) {  // :239-240
        nlfB[c] = floor_max((N * s_xy[c] - s_x * s_y[c]) / (N * s_xx - s_x * s_x), 1e-8);
        nlfA[c] = floor_max((s_y[c] - nlfB[c] * s_x) / N, 1e-8);
    }
    const double N1 = N;
    double err2[3] = {0, 0, 0};
    for (size_t k = 0; k < npx; k++) {  // :243-256
        double m, v[3];
        if (!selected(stats + k * 4, m, v)) continue;
        for (int c = 0; c < 3; c++) {
            double diff = nlfA[c] + nlfB[c] * m - v[c];
            err2[c] += diff * diff;
        }
    }
    for (int c = 0; c < 3; c++) err2[c] /= N;

    for (int c = 0; c < 3; c++) varianceMax[c] = nlfB[c];  // :259

    s_x = s_xx = N = 0;  // :262-292
    double newErr2[3] = {0, 0, 0};
    int discarded = 0;
    for (int c = 0; c < 3; c++) s_y[c] = s_xy[c] = 0;
    for (size_t k = 0; k < npx; k++) {
        double m, v[3];
        if (!selected(stats + k * 4, m, v)) continue;
        double dsq[3];
        bool fits = true;
        for (int c = 0; c < 3; c++) {
            double diff = std::fabs(nlfA[c] + nlfB[c] * m - v[c]);
            dsq[c] = diff * diff;
            fits = fits && dsq[c] <= err2[c];
        }
        if (fits) {
            s_x += m;
            s_xx += m * m;
            for (int c = 0; c < 3; c++) {
                s_y[c] += v[c];
                s_xy[c] += m * v[c];
                newErr2[c] += dsq[c];
            }
            N++;
        } else {
            discarded++;
        }
    }
    for (int c = 0; c < 3; c++) newErr2[c] /= N;

    int accepted = 0;
    if (newErr2[0] <= err2[0] && newErr2[1] <= err2[1] && newErr2[2] <= err2[2] && N / (double)npx > 0.01) {
        for (int c = 0; c < 3; c++) {  // :296-297
            nlfB[c] = floor_max((N * s_xy[c] - s_x * s_y[c]) / (N * s_xx - s_x * s_x), 1e-8);
            nlfA[c] = floor_max((s_y[c] - nlfB[c] * s_x) / N, 1e-8);
        }
        accepted = 1;
    }
    const double adj = exposure_multiplier * exposure_multiplier;  // :316-319
    for (int c = 0; c < 3; c++) {
        A[c] = nlfA[c] * adj;
        B[c] = nlfB[c] * adj;
    }
    if (diag) {
        for (int i = 0; i < 6; i++) diag[i] = hist[i];
        diag[6] = N1;
        diag[7] = N;
        for (int c = 0; c < 3; c++) {
            diag[8 + c] = err2[c];
            diag[11 + c] = newErr2[c];
        }
        diag[14] = discarded;
        diag[15] = vmax1;
    }
    return accepted;
}

}  // extern "C"

namespace {

template <class T>
int pyramid_denoise(const T* image, const T* gradient, int w, int h, int levels, int radius, int patch_similarity,
                    const oracle_denoise_params* params, double* nlf, T exposure, T lensShading, int calibrate,
                    T* denoised, T* subtracted_out, double* bases, double* eigvals) {
    std::vector<int> W(levels), H(levels);
    std::vector<size_t> off(levels + 1, 0);
    for (int i = 0; i < levels; i++) {  // pyramid_processor.cpp:38-47: sizes from the base, width / 2^i
        W[i] = w >> i;
        H[i] = h >> i;
        off[i + 1] = off[i] + (size_t)W[i] * H[i];
        if (((long)W[i] * H[i]) / 64 < 2) return -1;
    }
    // layer[i]: level-i noisy image (level 0 = input); grad[i]: gradient pyramid
    std::vector<std::vector<T>> layer(levels), gradl(levels), sub(levels);
    for (int i = 1; i < levels; i++) layer[i].resize((size_t)W[i] * H[i] * 4);
    const bool haveGrad = gradient != nullptr;
    if (haveGrad)
        for (int i = 1; i < levels; i++) gradl[i].resize((size_t)W[i] * H[i] * 2);
    auto L = [&](int i) -> const T* { return i == 0 ? image : layer[i].data(); };
    auto G = [&](int i) -> const T* { return !haveGrad ? nullptr : (i == 0 ? gradient : gradl[i].data()); };

    T thr[16][3];
    for (int i = 0; i < levels; i++) {  // loop A, :103-120
        if (i < levels - 1) {
            downsample(L(i), W[i], H[i], 4, layer[i + 1].data(), W[i + 1], H[i + 1]);
            if (haveGrad) downsample(G(i), W[i], H[i], 2, gradl[i + 1].data(), W[i + 1], H[i + 1]);
        }
        if (calibrate) {
            std::vector<T> st((size_t)W[i] * H[i] * 4);
            noise_stats(L(i), W[i], H[i], st.data());
            std::vector<float> stf(st.begin(), st.end());
            oracle_fit_ycbcr_nlf(stf.data(), W[i], H[i], (double)exposure, nlf + i * 6, nlf + i * 6 + 3, nullptr);
        }
        thr[i][0] = params[i].luma;  // nflMultiplier, :52-56
        thr[i][1] = params[i].chroma;
        thr[i][2] = params[i].chroma;
    }
    for (int i = levels - 1; i >= 0; i--) {  // loop B, :123-169
        const T* layerImg;
        if (i < levels - 1) {
            sub[i].resize((size_t)W[i] * H[i] * 4);
            T npx = (T)nlf[i * 6 + 0] * thr[i][0], npy = (T)nlf[i * 6 + 3] * thr[i][0];
            subtract_noise(L(i), W[i], H[i], L(i + 1), denoised + off[i + 1] * 4, W[i + 1], H[i + 1], G(i), T(1),
                           (T)params[i].sharpening, npx, npy, sub[i].data());
            layerImg = sub[i].data();
        } else {
            layerImg = L(i);
        }
        if (subtracted_out) std::memcpy(subtracted_out + off[i] * 4, layerImg, sizeof(T) * W[i] * H[i] * 4);
        if (!patch_similarity) {  // :162-168
            T va[3], vb[3];
            for (int c = 0; c < 3; c++) {
                va[c] = (T)nlf[i * 6 + c];
                vb[c] = (T)nlf[i * 6 + 3 + c];
            }
            denoise_image(layerImg, G(i), W[i], H[i], va, vb, thr[i], (T)params[i].chromaBoost,
                          (T)params[i].gradientBoost, (T)params[i].gradientThreshold, denoised + off[i] * 4);
            continue;
        }
        std::vector<T> patches((size_t)W[i] * H[i] / 64 * 25 + 25);
        int n = collect_patches(layerImg, W[i], H[i], patches.data());
        double basis_d[200], ev[25], cov[625];
        patch_covariance(patches.data(), n, cov);
        basis_from_cov(cov, basis_d, ev);
        if (bases) std::memcpy(bases + i * 200, basis_d, sizeof(basis_d));
        if (eigvals) std::memcpy(eigvals + i * 25, ev, sizeof(ev));
        T basis[200];
        for (int k = 0; k < 200; k++) basis[k] = (T)basis_d[k];
        std::vector<T> feat((size_t)W[i] * H[i] * 8);
        pca_project(layerImg, W[i], H[i], basis, feat.data());
        T va[3], vb[3];
        for (int c = 0; c < 3; c++) {
            va[c] = (T)nlf[i * 6 + c];
            vb[c] = (T)nlf[i * 6 + 3 + c];
        }
        block_match(layerImg, G(i), feat.data(), W[i], H[i], va, vb, thr[i], (T)params[i].chromaBoost,
                    (T)params[i].gradientBoost, (T)params[i].gradientThreshold, lensShading, radius,
                    denoised + off[i] * 4);
    }
    return 0;
}

}  // namespace

extern "C" {

int oracle_pyramid_denoise_f32(const float* image, const float* gradient, int w, int h, int levels, int radius,
                               int patch_similarity,
                               const oracle_denoise_params* params, double* nlf, float exposure, float ls,
                               int calibrate, float* denoised, float* subtracted, double* bases, double* eigvals) {
    return pyramid_denoise(image, gradient, w, h, levels, radius, patch_similarity, params, nlf, exposure, ls, calibrate, denoised,
                           subtracted, bases, eigvals);
}
int oracle_pyramid_denoise_f64(const double* image, const double* gradient, int w, int h, int levels, int radius,
                               int patch_similarity,
                               const oracle_denoise_params* params, double* nlf, double exposure, double ls,
                               int calibrate, double* denoised, double* subtracted, double* bases, double* eigvals) {
    return pyramid_denoise(image, gradient, w, h, levels, radius, patch_similarity, params, nlf, exposure, ls, calibrate, denoised,
                           subtracted, bases, eigvals);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Differentiation form (FP64): forward + reverse-mode gradient w.r.t. the
// per-level, per-channel thresholds.  Forward as pyramid_processor.cpp:96-172
// with gradientBoost = 0, sharpening = 1, lens shading 0 and no gradient
// image; every stage is differentiated by scatter (the device adjoint gathers,
// so the two formulations differ).
// ---------------------------------------------------------------------------
namespace {

struct LevelState {
    int w = 0, h = 0, n = 0;
    std::vector<double> X, S, F, D;  // noisy, layer (subtracted), features, denoised
    double V[625], lam[25], B[200], sgn[8], mu[25];
};

// Full eigen-decomposition of the patch covariance and the oriented basis.
void pca_full(const double* patches, int n, double* V, double* lam, double* B, double* sgn, double* mu) {
    double cov[625];
    patch_covariance(patches, n, cov);
    for (int r = 0; r < 25; r++) {
        double m = 0;
        for (int k = 0; k < n; k++) m += patches[(size_t)k * 25 + r];
        mu[r] = m / n;
    }
    oracle_eigh_jacobi(cov, 25, lam, V);
    for (int c = 0; c < 8; c++) {
        double dot = 0;
        for (int r = 0; r < 25; r++) dot += V[r * 25 + c] * sign_ref(r);
        sgn[c] = dot < 0 ? -1.0 : 1.0;
        for (int r = 0; r < 25; r++) B[r * 8 + c] = sgn[c] * V[r * 25 + c];
    }
}

// Reverse of block_match (demosaic.metal:806-884) with gB = 0, no lens shading.
void bm_backward(const LevelState& L, const double thr[3], double cB, const double va[3], const double vb[3],
                 int R, const double* gD, double* gS, double* gF, double gthr[3]) {
    const int w = L.w, h = L.h;
    const double* S = L.S.data();
    const double* F = L.F.data();
    const double* D = L.D.data();
    for (int py = 0; py < h; py++) {
        for (int px = 0; px < w; px++) {
            const size_t p = (size_t)py * w + px;
            const double* ip = S + p * 4;
            double sigma[3], dm[3];
            for (int c = 0; c < 3; c++) {
                sigma[c] = std::sqrt(va[c] + vb[c] * ip[0]);
                dm[c] = 1.0 / (thr[c] * sigma[c]);
            }
            // denominators of this pixel (recomputed)
            double DY = 0, DC = 0;
            for (int y = -R; y <= R; y++)
                for (int x = -R; x <= R; x++) {
                    const size_t q = (size_t)clampi(py + y, 0, h - 1) * w + clampi(px + x, 0, w - 1);
                    double d2 = 0;
                    for (int c = 0; c < 8; c++) d2 += (F[q * 8 + c] - F[p * 8 + c]) * (F[q * 8 + c] - F[p * 8 + c]);
                    const double wl = std::exp(-2.0 * d2 * dm[0] * dm[0]);
                    const double db = (S[q * 4 + 1] - ip[1]) * dm[1] / cB, dr = (S[q * 4 + 2] - ip[2]) * dm[2] / cB;
                    const double wc = std::exp(-2.0 * (wl * wl + db * db + dr * dr));
                    const double wd = std::exp(-2.0 * 0.01 * (x * x + y * y));
                    DY += wd * wl;
                    DC += wd * wc;
                }
            const double* o = D + p * 4;
            const double g0 = gD[p * 4], g1 = gD[p * 4 + 1], g2 = gD[p * 4 + 2];
            double gdm[3] = {0, 0, 0};
            for (int y = -R; y <= R; y++)
                for (int x = -R; x <= R; x++) {
                    const size_t q = (size_t)clampi(py + y, 0, h - 1) * w + clampi(px + x, 0, w - 1);
                    const double* sq = S + q * 4;
                    double d2 = 0;
                    for (int c = 0; c < 8; c++) d2 += (F[q * 8 + c] - F[p * 8 + c]) * (F[q * 8 + c] - F[p * 8 + c]);
                    const double wl = std::exp(-2.0 * d2 * dm[0] * dm[0]);
                    const double Db = sq[1] - ip[1], Dr = sq[2] - ip[2];
                    const double kb = dm[1] * dm[1] / (cB * cB), kr = dm[2] * dm[2] / (cB * cB);
                    const double v = Db * Db * kb + Dr * Dr * kr;
                    const double wd = std::exp(-2.0 * 0.01 * (x * x + y * y));
                    const double WY = wd * wl, WC = wd * std::exp(-2.0 * (wl * wl + v));
                    // output = sum W value / sum W
                    const double gWY = g0 * (sq[0] - o[0]) / DY;
                    const double gWC = (g1 * (sq[1] - o[1]) + g2 * (sq[2] - o[2])) / DC;
                    gS[q * 4 + 0] += g0 * WY / DY;
                    gS[q * 4 + 1] += g1 * WC / DC;
                    gS[q * 4 + 2] += g2 * WC / DC;
                    const double gwl = gWY * wd + gWC * WC * (-4.0 * wl);
                    const double gv = gWC * WC * (-2.0);
                    const double gu = gwl * (-2.0 * wl);  // wl = exp(-2 u), u = d2 dm0^2
                    const double gd2 = gu * dm[0] * dm[0];
                    gdm[0] += gu * 2.0 * d2 * dm[0];
                    const double gDb = gv * 2.0 * Db * kb, gDr = gv * 2.0 * Dr * kr;
                    gdm[1] += gv * 2.0 * Db * Db * dm[1] / (cB * cB);
                    gdm[2] += gv * 2.0 * Dr * Dr * dm[2] / (cB * cB);
                    gS[q * 4 + 1] += gDb;
                    gS[p * 4 + 1] -= gDb;
                    gS[q * 4 + 2] += gDr;
                    gS[p * 4 + 2] -= gDr;
                    for (int c = 0; c < 8; c++) {
                        const double t = 2.0 * gd2 * (F[q * 8 + c] - F[p * 8 + c]);
                        gF[q * 8 + c] += t;
                        gF[p * 8 + c] -= t;
                    }
                }
            for (int c = 0; c < 3; c++) {
                gthr[c] += gdm[c] * (-dm[c] / thr[c]);
                const double gsig = gdm[c] * (-dm[c] / sigma[c]);
                gS[p * 4] += gsig * vb[c] / (2.0 * sigma[c]);
            }
        }
    }
}

}  // namespace

extern "C" double oracle_pyramid_loss_grad_f64(const double* image, const double* clean, int w, int h, int levels,
                                               int radius, const double* thr, const double* chroma_boost,
                                               const double* nlf, double* grad_thr, double* denoised0) {
    std::vector<LevelState> Ls(levels);
    for (int i = 0; i < levels; i++) {
        Ls[i].w = w >> i;
        Ls[i].h = h >> i;
        Ls[i].n = (int)(((long)Ls[i].w * Ls[i].h) / 64);
        if (Ls[i].n < 2) return NAN;
    }
    Ls[0].X.assign(image, image + (size_t)w * h * 4);
    for (int i = 1; i < levels; i++) {
        Ls[i].X.resize((size_t)Ls[i].w * Ls[i].h * 4);
        downsample(Ls[i - 1].X.data(), Ls[i - 1].w, Ls[i - 1].h, 4, Ls[i].X.data(), Ls[i].w, Ls[i].h);
    }
    for (int i = levels - 1; i >= 0; i--) {
        LevelState& L = Ls[i];
        const size_t np = (size_t)L.w * L.h;
        if (i == levels - 1) {
            L.S = L.X;
        } else {
            L.S.resize(np * 4);
            subtract_noise(L.X.data(), L.w, L.h, Ls[i + 1].X.data(), Ls[i + 1].D.data(), Ls[i + 1].w, Ls[i + 1].h,
                           (const double*)nullptr, 1.0, 1.0, 0.0, 0.0, L.S.data());
        }
        std::vector<double> patches(np / 64 * 25 + 25);
        collect_patches(L.S.data(), L.w, L.h, patches.data());
        pca_full(patches.data(), L.n, L.V, L.lam, L.B, L.sgn, L.mu);
        L.F.resize(np * 8);
        pca_project(L.S.data(), L.w, L.h, L.B, L.F.data());
        L.D.resize(np * 4);
        block_match(L.S.data(), (const double*)nullptr, L.F.data(), L.w, L.h, nlf + i * 6, nlf + i * 6 + 3,
                    thr + i * 3, chroma_boost[i], 0.0, 1.0, 0.0, radius, L.D.data());
    }
    double loss = 0;
    const size_t np0 = (size_t)w * h;
    std::vector<double> gD(np0 * 4, 0.0);
    for (size_t p = 0; p < np0; p++)
        for (int c = 0; c < 3; c++) {
            const double d = Ls[0].D[p * 4 + c] - clean[p * 4 + c];
            loss += 0.5 * d * d;
            gD[p * 4 + c] = d;
        }
    if (denoised0) std::memcpy(denoised0, Ls[0].D.data(), sizeof(double) * np0 * 4);
    if (!grad_thr) return loss;

    for (int i = 0; i < levels; i++) {
        LevelState& L = Ls[i];
        const size_t np = (size_t)L.w * L.h;
        std::vector<double> gS(np * 4, 0.0), gF(np * 8, 0.0);
        double gthr[3] = {0, 0, 0};
        bm_backward(L, thr + i * 3, chroma_boost[i], nlf + i * 6, nlf + i * 6 + 3, radius, gD.data(), gS.data(),
                    gF.data(), gthr);
        for (int c = 0; c < 3; c++) grad_thr[i * 3 + c] = gthr[c];
        if (i == levels - 1) break;
        // projection F = P B (demosaic.metal:774-796): gB and the patch scatter
        double gB[200] = {0};
        for (int py = 0; py < L.h; py++)
            for (int px = 0; px < L.w; px++) {
                const size_t p = (size_t)py * L.w + px;
                int row = 0;
                for (int j = -2; j <= 2; j++)
                    for (int ii = -2; ii <= 2; ii++, row++) {
                        const size_t t = (size_t)clampi(py + j, 0, L.h - 1) * L.w + clampi(px + ii, 0, L.w - 1);
                        double acc = 0;
                        for (int c = 0; c < 8; c++) {
                            gB[row * 8 + c] += L.S[t * 4] * gF[p * 8 + c];
                            acc += gF[p * 8 + c] * L.B[row * 8 + c];
                        }
                        gS[t * 4] += acc;
                    }
            }
        // eigenvector derivative: gC = sym(V M V^T), M_jk = v_j^T g_k / (lambda_k - lambda_j)
        double M[625] = {0};
        for (int k = 0; k < 8; k++)
            for (int j = 0; j < 25; j++) {
                if (j == k || L.lam[k] == L.lam[j]) continue;
                double vg = 0;
                for (int r = 0; r < 25; r++) vg += L.V[r * 25 + j] * L.sgn[k] * gB[r * 8 + k];
                M[j * 25 + k] = vg / (L.lam[k] - L.lam[j]);
            }
        double G[625] = {0}, gC[625];
        for (int a = 0; a < 25; a++)
            for (int b = 0; b < 25; b++) {
                double s = 0;
                for (int j = 0; j < 25; j++)
                    for (int k = 0; k < 8; k++) s += L.V[a * 25 + j] * M[j * 25 + k] * L.V[b * 25 + k];
                G[a * 25 + b] = s;
            }
        for (int a = 0; a < 25; a++)
            for (int b = 0; b < 25; b++) gC[a * 25 + b] = 0.5 * (G[a * 25 + b] + G[b * 25 + a]);
        // covariance -> sampled patches (PCA.cpp:31-32): dL/dx_k = 2 gC (x_k - mu) / (n - 1)
        for (int y = 0; y < L.h / 8; y++)
            for (int x = 0; x < L.w / 8; x++) {
                const long slot = (long)y * L.w / 8 + x;
                if (slot >= L.n) continue;
                double xk[25];
                size_t idx[25];
                int e = 0;
                for (int j = -2; j <= 2; j++)
                    for (int ii = -2; ii <= 2; ii++, e++) {
                        idx[e] = (size_t)clampi(8 * y + j, 0, L.h - 1) * L.w + clampi(8 * x + ii, 0, L.w - 1);
                        xk[e] = L.S[idx[e] * 4] - L.mu[e];
                    }
                for (int a = 0; a < 25; a++) {
                    double s = 0;
                    for (int b = 0; b < 25; b++) s += gC[a * 25 + b] * xk[b];
                    gS[idx[a] * 4] += 2.0 * s / (L.n - 1);
                }
            }
        // subtract (demosaic.metal:922-959) with luma_weight 1, sharpening 1:
        // S = max-clamped(X - up(X1) + up(D1)) -> gD1 = up^T(gS masked)
        LevelState& C = Ls[i + 1];
        std::vector<double> gD1((size_t)C.w * C.h * 4, 0.0);
        for (int oy = 0; oy < L.h; oy++) {
            int y0, y1;
            double fy;
            lin_taps<double>((long)(2 * oy + 1) * C.h, 2L * L.h, C.h, &y0, &y1, &fy);
            for (int ox = 0; ox < L.w; ox++) {
                int x0, x1;
                double fx;
                lin_taps<double>((long)(2 * ox + 1) * C.w, 2L * L.w, C.w, &x0, &x1, &fx);
                const size_t o = (size_t)oy * L.w + ox;
                for (int c = 0; c < 3; c++) {
                    double g = gS[o * 4 + c];
                    if (c == 0 && !(L.S[o * 4] > 0.0)) g = 0.0;
                    gD1[((size_t)y0 * C.w + x0) * 4 + c] += (1 - fx) * (1 - fy) * g;
                    gD1[((size_t)y0 * C.w + x1) * 4 + c] += fx * (1 - fy) * g;
                    gD1[((size_t)y1 * C.w + x0) * 4 + c] += (1 - fx) * fy * g;
                    gD1[((size_t)y1 * C.w + x1) * 4 + c] += fx * fy * g;
                }
            }
        }
        gD.swap(gD1);
    }
    return loss;
}
