// Reverse pass of the multi-level denoise (differentiation form) for gfx950.
//
// The forward is PyramidProcessor::denoise (MetalPipeline/pyramid_processor.cpp:96-172)
// with independent per-channel thresholds, gradientBoost 0, sharpening 1, no
// lens shading and no gradient image.  The reverse pass walks the levels fine
// to coarse; per level it differentiates
//   block matching  demosaic.metal:806-884  (weights recomputed, never stored)
//   projection      demosaic.metal:774-796  (transpose of the 5x5 im2col, and
//                                            dL/dbasis as an FP64 reduction)
//   PCA basis       PCA.cpp:23-56           (eigenvector derivative, FP64)
//   covariance      PCA.cpp:31-32           (back to the sampled patches)
//   subtract        demosaic.metal:922-959  (transpose of the bilinear upsample,
//                                            luma clamp mask)
// Every kernel gathers into the pixel it owns (no atomics): clamp-to-edge
// reads make the forward's index map many-to-one at the borders, and the
// per-axis set of offsets that clamp onto a pixel is enumerated exactly.
#include "gcam_internal.hpp"

#include <cmath>

namespace gcam {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Offsets o in [-R, R] with clamp(p + o, 0, n - 1) == t, for a source p
// within R of t: one offset inside the image, a range at either border.
__device__ __forceinline__ void clamp_preimage(int t, int p, int n, int R, int& lo, int& hi) {
    if (t == 0 && t == n - 1) {
        lo = -R;
        hi = R;
    } else if (t == 0) {
        lo = -R;
        hi = -p;
    } else if (t == n - 1) {
        lo = n - 1 - p;
        hi = R;
    } else {
        lo = hi = t - p;
    }
    lo = max(lo, -R);
    hi = min(hi, R);
}

struct DirTable {
    float w[2 * GCAM_MAX_SEARCH_RADIUS * GCAM_MAX_SEARCH_RADIUS + 1];
};

DirTable make_dir_table() {
    DirTable t{};
    for (int d2 = 0; d2 < (int)(sizeof(t.w) / sizeof(float)); d2++) {
        const float a = 0.1f * std::sqrt((float)d2);
        t.w[d2] = std::exp(-2.0f * a * a);
    }
    return t;
}

__device__ __forceinline__ float dist2(const float4* F, size_t p, size_t q) {
    const float4 a0 = F[2 * q], a1 = F[2 * q + 1], b0 = F[2 * p], b1 = F[2 * p + 1];
    float t, d;
    t = a0.x - b0.x; d = t * t;
    t = a0.y - b0.y; d = fmaf(t, t, d);
    t = a0.z - b0.z; d = fmaf(t, t, d);
    t = a0.w - b0.w; d = fmaf(t, t, d);
    t = a1.x - b1.x; d = fmaf(t, t, d);
    t = a1.y - b1.y; d = fmaf(t, t, d);
    t = a1.z - b1.z; d = fmaf(t, t, d);
    t = a1.w - b1.w; d = fmaf(t, t, d);
    return d;
}

// Per-pixel coefficients of the block-matching backward (the pixel as centre):
//   c0 = (gY/DY, outY, (gCb/DC), outCb)
//   c1 = ((gCr/DC), outCr, dm0^2, (dm1/cB)^2)
//   c2 = ((dm2/cB)^2, dm0, dm1, dm2)
struct BwdParams {
    float thr[3];
    float chromaBoost;
    int radius;
};

__global__ __launch_bounds__(256) void k_bm_bwd_prep(View<const float4> S, const float4* __restrict__ F,
                                                     View<const float4> D, View<const float4> gD,
                                                     const NlfDev* __restrict__ nlf, BwdParams bp, DirTable dir,
                                                     float4* __restrict__ coef) {
    const int px = blockIdx.x * 64 + threadIdx.x, py = blockIdx.y * 4 + threadIdx.y;
    if (px >= S.w || py >= S.h) return;
    const size_t p = (size_t)py * S.w + px;
    const float4 ip = S.at(px, py);
    float dm[3];
    for (int c = 0; c < 3; c++) dm[c] = 1.0f / (bp.thr[c] * sqrtf(nlf->a[c] + nlf->b[c] * ip.x));
    const float k0 = dm[0] * dm[0], kb = (dm[1] / bp.chromaBoost) * (dm[1] / bp.chromaBoost),
                kr = (dm[2] / bp.chromaBoost) * (dm[2] / bp.chromaBoost);
    const int R = bp.radius;
    float DY = 0.0f, DC = 0.0f;
    for (int dy = -R; dy <= R; dy++) {
        const int qy = clampi(py + dy, 0, S.h - 1);
        for (int dx = -R; dx <= R; dx++) {
            const int qx = clampi(px + dx, 0, S.w - 1);
            const size_t q = (size_t)qy * S.w + qx;
            const float4 sq = S.at(qx, qy);
            const float wl = __expf(-2.0f * dist2(F, p, q) * k0);
            const float db = sq.y - ip.y, dr = sq.z - ip.z;
            const float wc = __expf(-2.0f * (wl * wl + db * db * kb + dr * dr * kr));
            const float wd = dir.w[dx * dx + dy * dy];
            DY += wd * wl;
            DC += wd * wc;
        }
    }
    const float4 o = D.at(px, py), g = gD.at(px, py);
    coef[3 * p + 0] = make_float4(g.x / DY, o.x, g.y / DC, o.y);
    coef[3 * p + 1] = make_float4(g.z / DC, o.z, k0, kb);
    coef[3 * p + 2] = make_float4(kr, dm[0], dm[1], dm[2]);
}

// Shared per-pair reverse step.  Given the pair (centre p, neighbour q) with
// direction weight wd, returns gd2 (dL/d d2), gDb, gDr (dL/d(Cq - Cp)) and
// the direct value weights wY = gY/DY * WY, wCb, wCr; accumulates the centre's
// dL/d(coefficients) when asked.
struct PairGrad {
    float gd2, gDb, gDr, vY, vCb, vCr;
    float gk0, gkb, gkr;  // dL/d dm0^2, dL/d (dm1/cB)^2, dL/d (dm2/cB)^2
};

__device__ __forceinline__ PairGrad pair_grad(const float4 c0, const float4 c1, const float4 c2, float d2,
                                              const float4 sq, const float4 sp, float wd) {
    PairGrad r;
    const float k0 = c1.z, kb = c1.w, kr = c2.x;
    const float wl = __expf(-2.0f * d2 * k0);
    const float Db = sq.y - sp.y, Dr = sq.z - sp.z;
    const float v = Db * Db * kb + Dr * Dr * kr;
    const float WY = wd * wl, WC = wd * __expf(-2.0f * (wl * wl + v));
    const float gWY = c0.x * (sq.x - c0.y);
    const float gWC = c0.z * (sq.y - c0.w) + c1.x * (sq.z - c1.y);
    r.vY = c0.x * WY;
    r.vCb = c0.z * WC;
    r.vCr = c1.x * WC;
    const float gwl = gWY * wd - 4.0f * gWC * WC * wl;
    const float gv = -2.0f * gWC * WC;
    const float gu = -2.0f * gwl * wl;  // wl = exp(-2 u), u = d2 k0
    r.gd2 = gu * k0;
    r.gk0 = gu * d2;
    r.gDb = 2.0f * gv * Db * kb;
    r.gDr = 2.0f * gv * Dr * kr;
    r.gkb = gv * Db * Db;
    r.gkr = gv * Dr * Dr;
    return r;
}

// Block matching backward, one thread per pixel t, both roles gathered:
// as centre (its own window) and as neighbour of every pixel whose window
// reads it (clamped preimage).  Writes dL/dS (Y, Cb, Cr), dL/dF and per-block
// partial dL/dthr.
__global__ __launch_bounds__(256) void k_bm_bwd(View<const float4> S, const float4* __restrict__ F,
                                                const float4* __restrict__ coef, const NlfDev* __restrict__ nlf,
                                                BwdParams bp, DirTable dir, float4* __restrict__ gS,
                                                float4* __restrict__ gF, float* __restrict__ gthr_part) {
    const int tx = blockIdx.x * 64 + threadIdx.x, ty = blockIdx.y * 4 + threadIdx.y;
    const bool live = tx < S.w && ty < S.h;
    const int px = min(tx, S.w - 1), py = min(ty, S.h - 1);
    const size_t t = (size_t)py * S.w + px;
    const int R = bp.radius;
    const float4 st = S.at(px, py);
    float gY = 0.0f, gCb = 0.0f, gCr = 0.0f;
    float gf[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    float gthr[3] = {0.0f, 0.0f, 0.0f};
    const float4 f0 = F[2 * t], f1 = F[2 * t + 1];
    if (live) {
        // centre role
        const float4 c0 = coef[3 * t], c1 = coef[3 * t + 1], c2 = coef[3 * t + 2];
        float gk0 = 0.0f, gkb = 0.0f, gkr = 0.0f;
        for (int dy = -R; dy <= R; dy++) {
            const int qy = clampi(py + dy, 0, S.h - 1);
            for (int dx = -R; dx <= R; dx++) {
                const int qx = clampi(px + dx, 0, S.w - 1);
                const size_t q = (size_t)qy * S.w + qx;
                const float4 sq = S.at(qx, qy);
                const PairGrad g = pair_grad(c0, c1, c2, dist2(F, t, q), sq, st, dir.w[dx * dx + dy * dy]);
                gk0 += g.gk0;
                gkb += g.gkb;
                gkr += g.gkr;
                gCb -= g.gDb;
                gCr -= g.gDr;
                const float4 q0 = F[2 * q], q1 = F[2 * q + 1];
                const float s2 = 2.0f * g.gd2;
                gf[0] -= s2 * (q0.x - f0.x); gf[1] -= s2 * (q0.y - f0.y);
                gf[2] -= s2 * (q0.z - f0.z); gf[3] -= s2 * (q0.w - f0.w);
                gf[4] -= s2 * (q1.x - f1.x); gf[5] -= s2 * (q1.y - f1.y);
                gf[6] -= s2 * (q1.z - f1.z); gf[7] -= s2 * (q1.w - f1.w);
            }
        }
        // k0 = dm0^2, kb = (dm1/cB)^2, kr = (dm2/cB)^2, dm_c = 1 / (thr_c sigma_c), sigma_c = sqrt(a_c + b_c Y)
        const float dm[3] = {c2.y, c2.z, c2.w};
        const float cb2 = bp.chromaBoost * bp.chromaBoost;
        const float gdm[3] = {gk0 * 2.0f * dm[0], gkb * 2.0f * dm[1] / cb2, gkr * 2.0f * dm[2] / cb2};
        for (int c = 0; c < 3; c++) {
            gthr[c] = gdm[c] * (-dm[c] / bp.thr[c]);
            const float sigma = 1.0f / (bp.thr[c] * dm[c]);
            gY += gdm[c] * (-dm[c] / sigma) * nlf->b[c] / (2.0f * sigma);
        }
        // neighbour role: every (p, offset) whose clamped read lands on t
        const int pyl = max(py - R, 0), pyh = min(py + R, S.h - 1);
        const int pxl = max(px - R, 0), pxh = min(px + R, S.w - 1);
        for (int sy = pyl; sy <= pyh; sy++) {
            int oyl, oyh;
            clamp_preimage(py, sy, S.h, R, oyl, oyh);
            for (int sx = pxl; sx <= pxh; sx++) {
                int oxl, oxh;
                clamp_preimage(px, sx, S.w, R, oxl, oxh);
                if (oyl > oyh || oxl > oxh) continue;
                const size_t p = (size_t)sy * S.w + sx;
                const float4 pc0 = coef[3 * p], pc1 = coef[3 * p + 1], pc2 = coef[3 * p + 2];
                const float4 sp = S.at(sx, sy);
                const float d2 = dist2(F, p, t);
                const float4 p0 = F[2 * p], p1 = F[2 * p + 1];
                for (int oy = oyl; oy <= oyh; oy++)
                    for (int ox = oxl; ox <= oxh; ox++) {
                        const PairGrad g = pair_grad(pc0, pc1, pc2, d2, st, sp, dir.w[ox * ox + oy * oy]);
                        gY += g.vY;
                        gCb += g.vCb + g.gDb;
                        gCr += g.vCr + g.gDr;
                        const float s2 = 2.0f * g.gd2;
                        gf[0] += s2 * (f0.x - p0.x); gf[1] += s2 * (f0.y - p0.y);
                        gf[2] += s2 * (f0.z - p0.z); gf[3] += s2 * (f0.w - p0.w);
                        gf[4] += s2 * (f1.x - p1.x); gf[5] += s2 * (f1.y - p1.y);
                        gf[6] += s2 * (f1.z - p1.z); gf[7] += s2 * (f1.w - p1.w);
                    }
            }
        }
        gS[t] = make_float4(gY, gCb, gCr, 0.0f);
        gF[2 * t] = make_float4(gf[0], gf[1], gf[2], gf[3]);
        gF[2 * t + 1] = make_float4(gf[4], gf[5], gf[6], gf[7]);
    }
    // per-block partial dL/dthr (fixed tree: deterministic)
    __shared__ float red[3][256];
    const int tid = threadIdx.y * 64 + threadIdx.x;
    for (int c = 0; c < 3; c++) red[c][tid] = gthr[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s)
            for (int c = 0; c < 3; c++) red[c][tid] += red[c][tid + s];
        __syncthreads();
    }
    if (tid == 0) {
        const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        for (int c = 0; c < 3; c++) gthr_part[b * 3 + c] = red[c][0];
    }
}

// Projection backward, luma part: dL/dS_Y(t) += sum over (p, row) with
// clamp(p + off(row)) == t of sum_c gF[p][c] basis[row][c].
__global__ __launch_bounds__(256) void k_proj_bwd(View<const float4> S, const float4* __restrict__ gF,
                                                  const float* __restrict__ basis, float4* __restrict__ gS) {
    const int px = blockIdx.x * 64 + threadIdx.x, py = blockIdx.y * 4 + threadIdx.y;
    if (px >= S.w || py >= S.h) return;
    float acc = 0.0f;
    for (int sy = max(py - 2, 0); sy <= min(py + 2, S.h - 1); sy++) {
        int jl, jh;
        clamp_preimage(py, sy, S.h, 2, jl, jh);
        for (int sx = max(px - 2, 0); sx <= min(px + 2, S.w - 1); sx++) {
            int il, ih;
            clamp_preimage(px, sx, S.w, 2, il, ih);
            const size_t p = (size_t)sy * S.w + sx;
            const float4 g0 = gF[2 * p], g1 = gF[2 * p + 1];
            for (int j = jl; j <= jh; j++)
                for (int i = il; i <= ih; i++) {
                    const float* b = basis + ((j + 2) * 5 + (i + 2)) * kComponents;
                    acc += g0.x * b[0] + g0.y * b[1] + g0.z * b[2] + g0.w * b[3] + g1.x * b[4] + g1.y * b[5] +
                           g1.z * b[6] + g1.w * b[7];
                }
        }
    }
    gS[(size_t)py * S.w + px].x += acc;
}

// dL/dbasis[row][c] = sum_p Y(clamp(p + off(row))) gF[p][c]: per-workgroup FP64
// partials over a fixed grid (200 values per workgroup).
__global__ __launch_bounds__(256) void k_basis_grad(View<const float4> S, const float4* __restrict__ gF,
                                                    double* __restrict__ partials) {
    __shared__ float pt[256][kPatch];
    __shared__ float gf[256][kComponents];
    const int tid = threadIdx.x;
    const int e = tid < kPatch * kComponents ? tid : 0;
    const int row = e / kComponents, comp = e % kComponents;
    double acc = 0.0;
    const int64_t npx = (int64_t)S.w * S.h;
    for (int64_t base = (int64_t)blockIdx.x * 256; base < npx; base += (int64_t)gridDim.x * 256) {
        const int64_t k = base + tid;
        if (k < npx) {
            const int x = (int)(k % S.w), y = (int)(k / S.w);
            for (int j = -2; j <= 2; j++)
                for (int i = -2; i <= 2; i++)
                    pt[tid][(j + 2) * 5 + i + 2] = S.at(clampi(x + i, 0, S.w - 1), clampi(y + j, 0, S.h - 1)).x;
            const float4 g0 = gF[2 * k], g1 = gF[2 * k + 1];
            gf[tid][0] = g0.x; gf[tid][1] = g0.y; gf[tid][2] = g0.z; gf[tid][3] = g0.w;
            gf[tid][4] = g1.x; gf[tid][5] = g1.y; gf[tid][6] = g1.z; gf[tid][7] = g1.w;
        } else {
            for (int r = 0; r < kPatch; r++) pt[tid][r] = 0.0f;
            for (int c = 0; c < kComponents; c++) gf[tid][c] = 0.0f;
        }
        __syncthreads();
        const int cnt = (int)min<int64_t>(256, npx - base);
        for (int i = 0; i < cnt; i++) acc += (double)pt[i][row] * (double)gf[i][comp];
        __syncthreads();
    }
    if (tid < kPatch * kComponents) partials[(size_t)blockIdx.x * kPatch * kComponents + tid] = acc;
}

// Eigenvector derivative (one workgroup, FP64): gB -> gV (sign-oriented
// columns) -> M_jk = v_j^T g_k / (lambda_k - lambda_j) -> gC = sym(V M V^T).
// eig: V (25 x 25, column j = eigenvector j, descending), lambda (25),
// sign (8), mean patch (25).
__global__ __launch_bounds__(256) void k_eig_bwd(const double* __restrict__ partials, int nblocks,
                                                 const double* __restrict__ eig, double* __restrict__ gC) {
    __shared__ double gB[kPatch * kComponents];
    __shared__ double M[kPatch * kComponents];  // M[j][k], k < 8
    __shared__ double VM[kPatch * kComponents];
    const double* V = eig;
    const double* lam = eig + 625;
    const double* sgn = eig + 650;
    const int tid = threadIdx.x;
    if (tid < kPatch * kComponents) {
        double a = 0.0;
        for (int b = 0; b < nblocks; b++) a += partials[(size_t)b * kPatch * kComponents + tid];
        gB[tid] = a;
    }
    __syncthreads();
    if (tid < kPatch * kComponents) {
        const int j = tid / kComponents, k = tid % kComponents;
        double m = 0.0;
        if (j != k && lam[k] != lam[j]) {
            double vg = 0.0;
            for (int r = 0; r < kPatch; r++) vg += V[r * 25 + j] * sgn[k] * gB[r * kComponents + k];
            m = vg / (lam[k] - lam[j]);
        }
        M[tid] = m;
    }
    __syncthreads();
    if (tid < kPatch * kComponents) {  // (V M)[a][k]
        const int a = tid / kComponents, k = tid % kComponents;
        double s = 0.0;
        for (int j = 0; j < kPatch; j++) s += V[a * 25 + j] * M[j * kComponents + k];
        VM[tid] = s;
    }
    __syncthreads();
    for (int e = tid; e < 625; e += 256) {  // G = (V M) V[:, :8]^T, gC = sym(G)
        const int a = e / 25, b = e % 25;
        double gab = 0.0, gba = 0.0;
        for (int k = 0; k < kComponents; k++) {
            gab += VM[a * kComponents + k] * V[b * 25 + k];
            gba += VM[b * kComponents + k] * V[a * 25 + k];
        }
        gC[e] = 0.5 * (gab + gba);
    }
}

// Covariance -> sampled patches: dL/dx_k = 2 gC (x_k - mu) / (n - 1), added
// to dL/dS_Y at the patch's clamped positions.  Patches of the 8-pixel grid
// do not overlap, so each pixel has one writer.
__global__ __launch_bounds__(256) void k_patch_bwd(View<const float4> S, const double* __restrict__ gC,
                                                   const double* __restrict__ eig, double n, float4* __restrict__ gS) {
    const int gw = S.w / 8, gh = S.h / 8;
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= gw * gh) return;
    const int x = k % gw, y = k / gw;
    if ((int64_t)y * S.w / 8 + x >= (int64_t)n) return;
    const double* mu = eig + 658;
    double xc[kPatch];
    int idx[kPatch];
    for (int j = -2; j <= 2; j++)
        for (int i = -2; i <= 2; i++) {
            const int e = (j + 2) * 5 + i + 2;
            const int sx = clampi(8 * x + i, 0, S.w - 1), sy = clampi(8 * y + j, 0, S.h - 1);
            idx[e] = sy * S.w + sx;
            xc[e] = (double)S.at(sx, sy).x - mu[e];
        }
    for (int a = 0; a < kPatch; a++) {
        double s = 0.0;
        for (int b = 0; b < kPatch; b++) s += gC[a * 25 + b] * xc[b];
        gS[idx[a]].x += (float)(2.0 * s / (n - 1.0));
    }
}

// Subtract backward: S_i = X_i - up(X_{i+1}) + up(D_{i+1}), luma clamped at 0.
// dL/dD_{i+1}(c) gathers every fine pixel o whose bilinear taps include c.
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

__device__ __forceinline__ float tap_weight(const Tap& t, int c) {
    return (t.i0 == c ? 1.0f - t.f : 0.0f) + (t.i1 == c ? t.f : 0.0f);
}

__global__ __launch_bounds__(256) void k_sub_bwd(View<const float4> S, const float4* __restrict__ gS, int w1, int h1,
                                                 float4* __restrict__ gD1) {
    const int cx = blockIdx.x * 64 + threadIdx.x, cy = blockIdx.y * 4 + threadIdx.y;
    if (cx >= w1 || cy >= h1) return;
    // fine pixels whose sample position (2o + 1) w1 / (2 w) lies within one texel of c
    const int oxl = max(0, (int)floorf(((float)cx - 1.0f) * S.w / w1) - 1);
    const int oxh = min(S.w - 1, (int)ceilf(((float)cx + 2.0f) * S.w / w1) + 1);
    const int oyl = max(0, (int)floorf(((float)cy - 1.0f) * S.h / h1) - 1);
    const int oyh = min(S.h - 1, (int)ceilf(((float)cy + 2.0f) * S.h / h1) + 1);
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int oy = oyl; oy <= oyh; oy++) {
        const float wy = tap_weight(lin_tap((2 * oy + 1) * h1, 2 * S.h, h1), cy);
        if (wy == 0.0f) continue;
        for (int ox = oxl; ox <= oxh; ox++) {
            const float wx = tap_weight(lin_tap((2 * ox + 1) * w1, 2 * S.w, w1), cx);
            if (wx == 0.0f) continue;
            const float4 g = gS[(size_t)oy * S.w + ox];
            const float wgt = wx * wy;
            acc.x += S.at(ox, oy).x > 0.0f ? wgt * g.x : 0.0f;
            acc.y += wgt * g.y;
            acc.z += wgt * g.z;
        }
    }
    gD1[(size_t)cy * w1 + cx] = acc;
}

// Sum the per-block dL/dthr partials of a level in FP64 (fixed order).
__global__ __launch_bounds__(64) void k_thr_reduce(const float* __restrict__ part, int nblocks, float* __restrict__ out) {
    if (threadIdx.x < 3) {
        double a = 0.0;
        for (int b = 0; b < nblocks; b++) a += (double)part[(size_t)b * 3 + threadIdx.x];
        out[threadIdx.x] = (float)a;
    }
}

inline dim3 grid64x4(int w, int h) { return dim3((w + 63) / 64, (h + 3) / 4); }

}  // namespace

int adjoint_thr_blocks(int w, int h) { return ((w + 63) / 64) * ((h + 3) / 4); }

gcam_status launch_level_backward(const AdjointLevel& L, hipStream_t s) {
    static const DirTable dir = make_dir_table();
    BwdParams bp;
    for (int c = 0; c < 3; c++) bp.thr[c] = L.thr[c];
    bp.chromaBoost = L.chromaBoost;
    bp.radius = L.radius;
    const dim3 g = grid64x4(L.S.w, L.S.h), b(64, 4);
    const float4* F = reinterpret_cast<const float4*>(L.feat);
    k_bm_bwd_prep<<<g, b, 0, s>>>(L.S, F, L.D, L.gD, L.nlf, bp, dir, L.coef);
    k_bm_bwd<<<g, b, 0, s>>>(L.S, F, L.coef, L.nlf, bp, dir, L.gS, L.gF, L.gthr_part);
    k_thr_reduce<<<1, 64, 0, s>>>(L.gthr_part, (int)(g.x * g.y), L.gthr);
    GCAM_LAUNCH_CHECK("block-matching backward");
    if (!L.gD1) return GCAM_OK;  // coarsest level: nothing upstream depends on thresholds
    k_proj_bwd<<<g, b, 0, s>>>(L.S, L.gF, L.basis, L.gS);
    k_basis_grad<<<kBasisGradBlocks, 256, 0, s>>>(L.S, L.gF, L.gB_part);
    k_eig_bwd<<<1, 256, 0, s>>>(L.gB_part, kBasisGradBlocks, L.eig, L.gC);
    const int gw = L.S.w / 8, gh = L.S.h / 8;
    if (gw * gh > 0) k_patch_bwd<<<(gw * gh + 255) / 256, 256, 0, s>>>(L.S, L.gC, L.eig, L.n, L.gS);
    k_sub_bwd<<<grid64x4(L.w1, L.h1), b, 0, s>>>(L.S, L.gS, L.w1, L.h1, L.gD1);
    GCAM_LAUNCH_CHECK("pyramid level backward");
    return GCAM_OK;
}

}  // namespace gcam
