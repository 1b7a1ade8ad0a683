// PCA patch basis on the device: patch sampling, FP64 moments, FP64 Jacobi
// eigensolve and the 25 -> 8 projection, for gfx950.
//
//   collect_patches  MetalPipeline/demosaic.metal:755-772 (grid demosaic_kernels.hpp:388)
//   moments + solve  MetalPipeline/PCA.cpp:23-56 (build_pca_space<25,8>)
//   pca_project      MetalPipeline/demosaic.metal:774-796
//
// The reference maps the patch buffer to the CPU and runs Eigen there, with a
// host synchronization per level (pyramid_processor.cpp:146-151).  Here the
// basis never leaves the device: the patches are reduced straight from the
// image into FP64 first and second moments (sum x, sum x x^T) by a fixed grid
// of workgroups, one workgroup turns them into the covariance and runs a
// parallel-ordered cyclic Jacobi, and the projection reads the basis from
// device memory, so a whole denoise call stays stream-ordered.
#include "gcam_internal.hpp"

namespace gcam {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Sign reference of the basis vectors: each vector is oriented so that its dot
// product with this fixed vector is non-negative (any eigenvector solver
// leaves the sign free; fixing it makes the features reproducible).
__device__ __forceinline__ double sign_ref(int r) { return 1.0 + (double)((r * r * 7 + r * 3 + 5) % 23) / 16.0; }

// Row/column of moment entry e: entries 0..24 are sum x_r (r = e, s = -1),
// entries 25.. enumerate the upper triangle r <= s row by row.
__device__ __forceinline__ void moment_rs(int e, int& r, int& s) {
    if (e < kPatch) {
        r = e;
        s = -1;
        return;
    }
    int t = e - kPatch;
    r = 0;
    while (t >= kPatch - r) {
        t -= kPatch - r;
        r++;
    }
    s = r + t;
}

// Moment accumulation over rounds of 256 patches staged in LDS.  Each thread
// owns moment entries tid and tid + 256 (350 in all).  `fetch(k, lds_row)`
// writes the 25 values of patch k (zeros when k is out of range).
template <class Fetch>
__device__ void moments_body(int64_t npatch, double* partials, Fetch fetch) {
    __shared__ float pt[256 * kPatch];
    int r0, s0, r1 = 0, s1 = -1;
    moment_rs(threadIdx.x, r0, s0);
    const bool has1 = threadIdx.x + 256 < kMoments;
    if (has1) moment_rs(threadIdx.x + 256, r1, s1);
    double acc0 = 0.0, acc1 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t base = (int64_t)blockIdx.x * 256; base < npatch; base += stride) {
        const int64_t k = base + threadIdx.x;
        fetch(k, &pt[threadIdx.x * kPatch]);
        __syncthreads();
        const int cnt = (int)min<int64_t>(256, npatch - base);
        for (int i = 0; i < cnt; i++) {
            const float* row = &pt[i * kPatch];
            const double a0 = row[r0];
            acc0 += s0 < 0 ? a0 : a0 * (double)row[s0];
            if (has1) {
                const double a1 = row[r1];
                acc1 += s1 < 0 ? a1 : a1 * (double)row[s1];
            }
        }
        __syncthreads();
    }
    partials[(size_t)blockIdx.x * kMoments + threadIdx.x] = acc0;
    if (has1) partials[(size_t)blockIdx.x * kMoments + threadIdx.x + 256] = acc1;
}

// Patches sampled from the image at 8*(x, y), x < w/8, y < h/8, luma only,
// clamp-to-edge reads (collectPatches).  Slots the grid does not reach are
// zero rows: they add nothing to the moments but count in n.
__global__ __launch_bounds__(256) void k_image_moments(View<const float4> in, double* partials) {
    const int gw = in.w / 8, gh = in.h / 8;
    const int64_t npatch = (int64_t)gw * gh;
    moments_body(npatch, partials, [&](int64_t k, float* row) {
        if (k < npatch) {
            const int x = (int)(k % gw), y = (int)(k / gw);
            for (int j = -2; j <= 2; j++) {
                const int sy = clampi(8 * y + j, 0, in.h - 1);
                for (int i = -2; i <= 2; i++) row[(j + 2) * 5 + (i + 2)] = in.at(clampi(8 * x + i, 0, in.w - 1), sy).x;
            }
        } else {
            for (int e = 0; e < kPatch; e++) row[e] = 0.0f;
        }
    });
}

__global__ __launch_bounds__(256) void k_buffer_moments(const float* patches, int64_t n, double* partials) {
    moments_body(n, partials, [&](int64_t k, float* row) {
        for (int e = 0; e < kPatch; e++) row[e] = k < n ? patches[k * kPatch + e] : 0.0f;
    });
}

// collectPatches materialized (per-operation entry point): slot = y*w/8 + x.
__global__ __launch_bounds__(256) void k_collect_patches(View<const float4> in, float* patches, int64_t n) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= in.w / 8 || y >= in.h / 8) return;
    const int64_t slot = (int64_t)y * in.w / 8 + x;
    if (slot >= n) return;
    for (int j = -2; j <= 2; j++) {
        const int sy = clampi(8 * y + j, 0, in.h - 1);
        for (int i = -2; i <= 2; i++)
            patches[slot * kPatch + (j + 2) * 5 + (i + 2)] = in.at(clampi(8 * x + i, 0, in.w - 1), sy).x;
    }
}

__global__ __launch_bounds__(256) void k_fill_zero(float* p, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) p[i] = 0.0f;
}

// Synchronization point of one wavefront's lanes over LDS.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FP64 reciprocal square root and reciprocal from the hardware estimates
// (v_rsq_f64, v_rcp_f64) refined by two Newton steps each.
__device__ __forceinline__ double rsqrt_f64(double x) {
    double r = __builtin_amdgcn_rsq(x);
    r = r * (1.5 - 0.5 * x * r * r);
    return r * (1.5 - 0.5 * x * r * r);
}
__device__ __forceinline__ double rcp_f64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = r * (2.0 - x * r);
    return r * (2.0 - x * r);
}

// Covariance + cyclic Jacobi (parallel circle ordering) + basis selection.
//
// C = (sum x x^T - (sum x)(sum x)^T / n) / (n - 1)   -- PCA.cpp:31-32
// Jacobi: each sweep runs 25 rounds; round r rotates the 12 disjoint pairs
// ((r + k) mod 25, (r - k) mod 25), k = 1..12, so every off-diagonal pair is
// annihilated once per sweep (index r idles).  Rotation as in Golub & Van Loan
// Alg. 8.5.2 (c^2 + s^2 = 1 to FP64 precision, so V stays orthogonal).  Sweeps
// stop when the off-diagonal energy is below kJacobiTol of the total.  The 8
// largest eigenpairs, largest first (PCA.cpp:49-55), are written as
// basis[r * 8 + c]; ties in eigenvalue keep index order.
__global__ __launch_bounds__(256) void k_pca_solve(const double* partials, int nblocks, double n, float* basis,
                                                   double* eigvals, double* eig_full) {
    constexpr int N = kPatch;
    constexpr int LD = N + 1;
    __shared__ double mom[kMoments];
    __shared__ double A[N * LD];
    __shared__ double V[N * LD];
    __shared__ double cs[12][2];
    __shared__ int pq[12][2];
    const int tid = threadIdx.x;

    for (int e = tid; e < kMoments; e += 256) {
        double a = 0.0;
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // 8 loads in flight; fixed summation order
        int b = 0;
        for (; b + 8 <= nblocks; b += 8)
#pragma unroll
            for (int u = 0; u < 8; u++) acc[u] += partials[(size_t)(b + u) * kMoments + e];
        for (; b < nblocks; b++) acc[0] += partials[(size_t)b * kMoments + e];
        for (int u = 0; u < 8; u++) a += acc[u];
        mom[e] = a;
    }
    __syncthreads();
    for (int e = tid; e < kMoments; e += 256) {
        if (e < N) continue;
        int r, s;
        moment_rs(e, r, s);
        const double c = (mom[e] - mom[r] * mom[s] / n) / (n - 1.0);
        A[r * LD + s] = c;
        A[s * LD + r] = c;
    }
    for (int e = tid; e < N * N; e += 256) {
        const int r = e / N, s = e % N;
        V[r * LD + s] = r == s ? 1.0 : 0.0;
    }
    __syncthreads();

    // The rotations run on wavefront 0 alone: its lanes share LDS without
    // workgroup barriers (a wavefront's LDS accesses complete in order; the
    // fences keep the compiler from reordering them across the sync points).
    if (tid < 64) {
        const int lane = tid;
        for (int sweep = 0; sweep < 40; sweep++) {
            double off = 0.0, tot = 0.0;
            for (int e = lane; e < N * N; e += 64) {
                const int r = e / N, s = e % N;
                const double v = A[r * LD + s] * A[r * LD + s];
                tot += v;
                if (r != s) off += v;
            }
            for (int o = 32; o > 0; o >>= 1) {
                off += __shfl_xor(off, o);
                tot += __shfl_xor(tot, o);
            }
            if (off == 0.0 || off <= kJacobiTol * tot) break;  // wave-uniform

            for (int round = 0; round < N; round++) {
                if (lane < 12) {
                    int p = (round + lane + 1) % N, q = (round - lane - 1 + N) % N;
                    if (p > q) {
                        const int t = p;
                        p = q;
                        q = t;
                    }
                    // t = sign(theta) / (|theta| + sqrt(theta^2 + 1)), theta = d / e with
                    // d = a_qq - a_pp, e = 2 a_pq:  t = sign(d e) |e| / (|d| + hypot(d, e)).
                    const double e = 2.0 * A[p * LD + q], d = A[q * LD + q] - A[p * LD + p];
                    double c = 1.0, s = 0.0;
                    if (e != 0.0) {
                        const double h2 = d * d + e * e;
                        const double hyp = h2 * rsqrt_f64(h2);
                        double t = fabs(e) * rcp_f64(fabs(d) + hyp);
                        if (d * e < 0.0) t = -t;
                        c = rsqrt_f64(1.0 + t * t);
                        s = t * c;
                    }
                    cs[lane][0] = c;
                    cs[lane][1] = s;
                    pq[lane][0] = p;
                    pq[lane][1] = q;
                }
                wave_sync();
                // A <- J^T A: rows p, q of every pair; all loads first, then the stores.
                {
                    constexpr int ITEMS = (12 * N + 63) / 64;
                    double ap[ITEMS], aq[ITEMS], cc[ITEMS], ss[ITEMS];
                    int ip[ITEMS], iq[ITEMS];
#pragma unroll
                    for (int m = 0; m < ITEMS; m++) {
                        const int t = min(lane + 64 * m, 12 * N - 1);
                        const int k = t / N, j = t % N;
                        ip[m] = pq[k][0] * LD + j;
                        iq[m] = pq[k][1] * LD + j;
                        cc[m] = cs[k][0];
                        ss[m] = cs[k][1];
                        ap[m] = A[ip[m]];
                        aq[m] = A[iq[m]];
                    }
                    wave_sync();
#pragma unroll
                    for (int m = 0; m < ITEMS; m++) {
                        if (lane + 64 * m < 12 * N) {
                            A[ip[m]] = cc[m] * ap[m] - ss[m] * aq[m];
                            A[iq[m]] = ss[m] * ap[m] + cc[m] * aq[m];
                        }
                    }
                }
                wave_sync();
                // A <- A J and V <- V J: columns p, q of every pair.
                {
                    constexpr int ITEMS = (24 * N + 63) / 64;
                    double mp[ITEMS], mq[ITEMS], cc[ITEMS], ss[ITEMS];
                    int ip[ITEMS], iq[ITEMS];
                    bool isA[ITEMS];
#pragma unroll
                    for (int m = 0; m < ITEMS; m++) {
                        const int t = min(lane + 64 * m, 24 * N - 1);
                        const int u = t / N, j = t % N, k = u % 12;
                        isA[m] = u < 12;
                        ip[m] = j * LD + pq[k][0];
                        iq[m] = j * LD + pq[k][1];
                        cc[m] = cs[k][0];
                        ss[m] = cs[k][1];
                        const double* M = isA[m] ? A : V;
                        mp[m] = M[ip[m]];
                        mq[m] = M[iq[m]];
                    }
                    wave_sync();
#pragma unroll
                    for (int m = 0; m < ITEMS; m++) {
                        if (lane + 64 * m < 24 * N) {
                            double* M = isA[m] ? A : V;
                            M[ip[m]] = cc[m] * mp[m] - ss[m] * mq[m];
                            M[iq[m]] = ss[m] * mp[m] + cc[m] * mq[m];
                        }
                    }
                }
                wave_sync();
            }
        }
    }
    __syncthreads();

    if (tid < N) {
        const double di = A[tid * LD + tid];
        int rank = 0;
        for (int j = 0; j < N; j++) {
            const double dj = A[j * LD + j];
            rank += (dj > di || (dj == di && j < tid)) ? 1 : 0;
        }
        if (eigvals) eigvals[rank] = di;
        double sg = 1.0;
        if (rank < kComponents) {
            double dot = 0.0;
            for (int r = 0; r < N; r++) dot += V[r * LD + tid] * sign_ref(r);
            sg = dot < 0.0 ? -1.0 : 1.0;
            for (int r = 0; r < N; r++) basis[r * kComponents + rank] = (float)(sg * V[r * LD + tid]);
        }
        // Full decomposition for the reverse pass: V (column = eigenvector,
        // descending), eigenvalues, the 8 orientations, the mean patch.
        if (eig_full) {
            for (int r = 0; r < N; r++) eig_full[r * N + rank] = V[r * LD + tid];
            eig_full[625 + rank] = di;
            if (rank < kComponents) eig_full[650 + rank] = sg;
            eig_full[658 + tid] = mom[tid] / n;
        }
    }
}

// pcaProjection: f[c] = sum_row basis[row][c] * Y(p + off(row)), rows j-outer /
// i-inner, clamp-to-edge.  A 64 x 4 tile plus its 2-pixel apron of luma is
// staged in LDS; the basis is wave-uniform and comes in through scalar loads.
__global__ __launch_bounds__(256) void k_pca_project(View<const float4> in, const float* __restrict__ basis,
                                                     float* __restrict__ feat, int feat_stride) {
    constexpr int TW = 64, TH = 4, LW = TW + 4, LH = TH + 4;
    __shared__ float lum[LH * LW];
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    const int tid = threadIdx.y * TW + threadIdx.x;
    for (int e = tid; e < LH * LW; e += 256) {
        const int ly = e / LW, lx = e % LW;
        lum[e] = in.at(clampi(x0 + lx - 2, 0, in.w - 1), clampi(y0 + ly - 2, 0, in.h - 1)).x;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= in.w || y >= in.h) return;
    float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 5; j++) {
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const float v = lum[(threadIdx.y + j) * LW + threadIdx.x + i];
            const int row = j * 5 + i;
#pragma unroll
            for (int c = 0; c < 8; c++) f[c] += basis[row * 8 + c] * v;
        }
    }
    float4* o = reinterpret_cast<float4*>(feat + ((size_t)y * feat_stride + x) * 8);
    o[0] = make_float4(f[0], f[1], f[2], f[3]);
    o[1] = make_float4(f[4], f[5], f[6], f[7]);
}

}  // namespace

gcam_status launch_collect_patches(View<const float4> in, float* patches, int64_t n, hipStream_t s) {
    const int64_t count = n * kPatch;
    int blocks = (int)std::min<int64_t>((count + 255) / 256, 4096);
    if (blocks < 1) blocks = 1;
    k_fill_zero<<<blocks, 256, 0, s>>>(patches, count);
    const int gw = in.w / 8, gh = in.h / 8;
    if (gw > 0 && gh > 0) k_collect_patches<<<dim3((gw + 63) / 64, (gh + 3) / 4), dim3(64, 4), 0, s>>>(in, patches, n);
    GCAM_LAUNCH_CHECK("collectPatches");
    return GCAM_OK;
}

// Workgroups of the moment reduction: one per 256 patches, at most kMomentBlocks.
static int moment_blocks(int64_t npatch) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(kMomentBlocks, (npatch + 255) / 256));
}

gcam_status launch_image_moments(View<const float4> in, double* partials, int* nblocks, hipStream_t s) {
    *nblocks = moment_blocks((int64_t)(in.w / 8) * (in.h / 8));
    k_image_moments<<<*nblocks, 256, 0, s>>>(in, partials);
    GCAM_LAUNCH_CHECK("patch moments");
    return GCAM_OK;
}

gcam_status launch_buffer_moments(const float* patches, int64_t n, double* partials, int* nblocks, hipStream_t s) {
    *nblocks = moment_blocks(n);
    k_buffer_moments<<<*nblocks, 256, 0, s>>>(patches, n, partials);
    GCAM_LAUNCH_CHECK("patch moments");
    return GCAM_OK;
}

gcam_status launch_pca_solve(const double* partials, int nblocks, int64_t n, float* basis, double* eigvals,
                             double* eig_full, hipStream_t s) {
    k_pca_solve<<<1, 256, 0, s>>>(partials, nblocks, (double)n, basis, eigvals, eig_full);
    GCAM_LAUNCH_CHECK("PCA eigensolve");
    return GCAM_OK;
}

gcam_status launch_pca_project(View<const float4> in, const float* basis, View<float> feat8, hipStream_t s) {
    k_pca_project<<<dim3((in.w + 63) / 64, (in.h + 3) / 4), dim3(64, 4), 0, s>>>(in, basis, feat8.ptr, feat8.stride);
    GCAM_LAUNCH_CHECK("pcaProjection");
    return GCAM_OK;
}

}  // namespace gcam
