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
#include "burst_tables.hpp"

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

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Moments of workgroup `blk` of `nblk` as a Gram product on the FP64 matrix
// cores: with X the npatch x 32 patch matrix (25 luma values, 7 zero
// columns), X^T X is accumulated 4 patches at a time by
// v_mfma_f64_16x16x4f64 in three 16 x 16 tiles (00, 01, 11; 10 is the
// transpose of 01).  Operand layout: lane l supplies X[4 g + l/16][l%16 (+16)]
// as both the A (16 x 4) and the B (4 x 16) operand; the result tile holds
// C[l/16 + 4 r][l%16] in register r of lane l (measured on gfx950; unlike
// the FP32 16 x 16 forms, rows interleave by 4).  sum x accumulates on the
// vector ALU.  The four wavefronts combine in LDS in a fixed order, so the
// partial of a workgroup is deterministic.  `fetch(k, f)` returns feature f
// of patch k (0 for k >= npatch or f >= 25).
template <class Fetch>
__device__ void moments_body(int blk, int nblk, int64_t npatch, double* partial, Fetch fetch) {
    __shared__ double gram[4][32][33];
    __shared__ double sums[4][64][2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int f = lane & 15, kq = lane >> 4;
    f64x4 c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
    double slo = 0.0, shi = 0.0;
    const int64_t ngroups = (npatch + 3) / 4;
    const int64_t stride = (int64_t)nblk * 4;
    constexpr int U = 8;  // groups in flight per wavefront
    for (int64_t g0 = (int64_t)blk * 4 + wave; g0 < ngroups; g0 += U * stride) {
        float lo[U], hi[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t k = (g0 + u * stride) * 4 + kq;
            lo[u] = fetch(k, f);
            hi[u] = fetch(k, f + 16);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const double dl = lo[u], dh = hi[u];
            slo += dl;
            shi += dh;
            c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(dl, dl, c00, 0, 0, 0);
            c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(dl, dh, c01, 0, 0, 0);
            c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(dh, dh, c11, 0, 0, 0);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = kq + 4 * r;
        gram[wave][i][f] = c00[r];
        gram[wave][i][16 + f] = c01[r];
        gram[wave][16 + i][16 + f] = c11[r];
    }
    sums[wave][lane][0] = slo;
    sums[wave][lane][1] = shi;
    __syncthreads();
    for (int e = tid; e < kMoments; e += 256) {
        int r, s;
        moment_rs(e, r, s);
        double acc = 0.0;
        if (s < 0) {
            for (int w = 0; w < 4; w++)
                for (int q = 0; q < 4; q++) acc += sums[w][16 * q + (r & 15)][r >> 4];
        } else {
            for (int w = 0; w < 4; w++) acc += gram[w][r][s];
        }
        partial[e] = acc;
    }
}

// Patches sampled from the image at 8*(x, y), x < w/8, y < h/8, luma only,
// clamp-to-edge reads (collectPatches).  Slots the grid does not reach are
// zero rows: they add nothing to the moments but count in n.
__device__ __forceinline__ float image_patch_value(const View<const float4>& in, int gw, int64_t npatch, int64_t k,
                                                   int f) {
    if (k >= npatch || f >= kPatch) return 0.0f;
    const int kk = (int)k;  // npatch <= 16384^2 / 64
    const int x = kk % gw, y = kk / gw;
    const int j = f / 5 - 2, i = f % 5 - 2;
    return in.at(clampi(8 * x + i, 0, in.w - 1), clampi(8 * y + j, 0, in.h - 1)).x;
}

__global__ __launch_bounds__(256) void k_image_moments(View<const float4> in, double* partials) {
    const int gw = in.w / 8, gh = in.h / 8;
    const int64_t npatch = (int64_t)gw * gh;
    moments_body(blockIdx.x, gridDim.x, npatch, partials + (size_t)blockIdx.x * kMoments,
                 [&](int64_t k, int f) { return image_patch_value(in, gw, npatch, k, f); });
}

__global__ __launch_bounds__(256) void k_buffer_moments(const float* patches, int64_t n, double* partials) {
    moments_body(blockIdx.x, gridDim.x, n, partials + (size_t)blockIdx.x * kMoments, [&](int64_t k, int f) {
        return (k < n && f < kPatch) ? patches[k * kPatch + f] : 0.0f;
    });
}

// Sum of the per-workgroup partials, entry by entry: workgroup x owns 7
// entries; thread t adds the partials b = t/7, t/7 + 36, ... of entry
// 7 x + t%7, and the 36 sums are added in index order (deterministic).
constexpr int kRedEntries = 7, kRedLanes = 36;
static_assert(kMoments % kRedEntries == 0, "moment reduction tiling");
__device__ void moments_reduce_body(int blk, const double* partials, int nblocks, double* mom) {
    __shared__ double part[kRedLanes][kRedEntries];
    const int t = threadIdx.x;
    const int e = blk * kRedEntries + t % kRedEntries, sub = t / kRedEntries;
    if (sub < kRedLanes) {
        double acc = 0.0;
        for (int b = sub; b < nblocks; b += kRedLanes) acc += partials[(size_t)b * kMoments + e];
        part[sub][t % kRedEntries] = acc;
    }
    __syncthreads();
    if (t < kRedEntries) {
        double acc = 0.0;
        for (int q = 0; q < kRedLanes; q++) acc += part[q][t];
        mom[blk * kRedEntries + t] = acc;
    }
}

__global__ __launch_bounds__(256) void k_moments_reduce(const double* partials, int nblocks, double* mom) {
    moments_reduce_body(blockIdx.x, partials, nblocks, mom);
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

// FP64 reciprocal square root: the hardware estimate (v_rsq_f64) refined by
// two Newton steps y (1.5 - 0.5 x y^2), each written as y * fma(x y, -0.5 y, 1.5).
__device__ __forceinline__ double rsq_refined(double x) {
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(x * y, -0.5 * y, 1.5);
    return y * fma(x * y, -0.5 * y, 1.5);
}

// Cyclic Jacobi sweeps on the symmetric 25 x 25 matrix in Ac (scalar type T),
// accumulating the rotations into V, until the off-diagonal energy is at most
// tol of the total or max_sweeps have run.  Every thread of the 256-thread
// workgroup calls it; on return Ac holds the rotated matrix.
//
// A round applies its 12 rotations at once, A <- J^T A J and V <- V J.
// Round r pairs the indices into super-indices u = 0..11, (p_u, q_u) = the
// sorted ((r + u + 1) mod 25, (r - u - 1) mod 25), and u = 12 = {r}
// (identity).  J acts on the 2 x 2 blocks A[{p_u, q_u}][{p_v, q_v}]:
// Y = J_u^T X J_v with J = [[c, s], [-s, c]] (rows/columns p, q), and the pair
// entry of a diagonal block is set to zero.  Every entry belongs to exactly one
// block: threads 0..90 (wavefronts 0 and 1) own one block each (both
// triangles), reading A from one buffer and writing the other, so a round needs
// one workgroup barrier.  Threads 128..244 (wavefronts 2 and 3) own the rows
// {j0, j0 + 9, j0 + 18} of V at one super-index v each and update them in
// place.  Every thread computes the rotations it needs from the round's A
// (bit-identical everywhere).  Rotation: the Golub & Van Loan angle
// (|phi| <= pi/4, tan(2 phi) = 2 a_pq / (a_qq - a_pp)) from its half-angle
// form, c^2 = (1 + |d| / h) / 2, s = sign(d e) |e| / (2 h c), h = hypot(d, e),
// d = a_qq - a_pp, e = 2 a_pq: one reciprocal square root for 1/h, one for 1/c.
// Row pitch of the 25 x 25 matrices in LDS: odd, so the rows of a column fall
// in different banks, and p * 33 = (p << 5) + p.
constexpr int kLD = 33;

template <class T>
__device__ void jacobi_sweeps(T*& Ac, T*& An, T* V, double tol, int max_sweeps, double2* energy) {
    constexpr int N = kPatch, LD = kLD;
    constexpr int NU = N * (N + 1) / 2;  // upper triangle, a <= b
    constexpr int IU = (NU + 255) / 256;
    constexpr int NS = 13, NB = NS * (NS + 1) / 2;  // super-indices, blocks
    constexpr int V0T = 128, NVT = NS * 9;           // V threads: wavefronts 2 and 3
    const T one = T(1), half = T(0.5), guard = sizeof(T) == 8 ? T(1e-300) : T(1e-30f);
    const int tid = threadIdx.x;
    int ua[IU], ub[IU];
#pragma unroll
    for (int m = 0; m < IU; m++) {
        int t = min(tid + 256 * m, NU - 1), a = 0;
        while (t >= N - a) {
            t -= N - a;
            a++;
        }
        ua[m] = a;
        ub[m] = a + t;
    }
    const bool is_block = tid < NB, is_v = tid >= V0T && tid < V0T + NVT;
    int bu = NS - 1, bv = NS - 1;  // block of a block thread; bu = the super-index of a V thread
    if (is_block) {
        int t = tid;
        bu = 0;
        while (t >= NS - bu) {
            t -= NS - bu;
            bu++;
        }
        bv = bu + t;
    } else if (is_v) {
        bu = (tid - V0T) % NS;
    }
    const int vj0 = is_v ? (tid - V0T) / NS : 0;  // rows vj0, vj0 + 9, vj0 + 18
    // The pair of super-index u in round r is the sorted ((r + u + 1) mod 25,
    // (r - u - 1) mod 25) (u = 12: (r, r)); both indices advance by one per
    // round, so they are kept and stepped instead of recomputed.
    auto first = [&](int u, int& i, int& j) {
        i = u == NS - 1 ? 0 : u + 1;
        j = u == NS - 1 ? 0 : N - 1 - u;
    };
    auto step = [&](int& i) { i = i == N - 1 ? 0 : i + 1; };

    // Straight-line code (no select on the expensive operands, which the
    // compiler would turn into branches).  The identity (p == q, or a_pq == 0)
    // is blended in arithmetically: c + m (1 - c) = 1 and s (1 - m) = 0 exactly
    // for m = 1 (1 - c is exact for c in [1/2, 1]).  rotation2 advances two
    // independent rotations step by step, so their dependent chains overlap.
    auto rotation2 = [&](bool id0, T app0, T aqq0, T apq0, bool id1, T app1, T aqq1, T apq1, T& c0, T& s0, T& c1,
                         T& s1) {
        const T e0 = T(2) * apq0, d0 = aqq0 - app0, e1 = T(2) * apq1, d1 = aqq1 - app1;
        const T m0 = (id0 || e0 == T(0)) ? one : T(0), m1 = (id1 || e1 == T(0)) ? one : T(0);
        const T y0 = rsq_refined(fma(d0, d0, e0 * e0) + guard), y1 = rsq_refined(fma(d1, d1, e1 * e1) + guard);
        const T c20 = half * fma(fabs(d0), y0, one), c21 = half * fma(fabs(d1), y1, one);  // cos^2 in [1/2, 1]
        const T z0 = rsq_refined(c20), z1 = rsq_refined(c21);                              // 1 / c
        const T cc0 = c20 * z0, cc1 = c21 * z1;
        c0 = fma(m0, one - cc0, cc0);
        c1 = fma(m1, one - cc1, cc1);
        s0 = copysign((half * fabs(e0) * y0) * z0, d0 * e0) * (one - m0);
        s1 = copysign((half * fabs(e1) * y1) * z1, d1 * e1) * (one - m1);
    };

    for (int sweep = 0; sweep < max_sweeps; sweep++) {
        // Off-diagonal and total energy (FP64 sums): per-thread entries,
        // wavefront sums, then the 4 wavefront sums in a fixed order (uniform).
        double off = 0.0, tot = 0.0;
#pragma unroll
        for (int m = 0; m < IU; m++) {
            if (tid + 256 * m < NU) {
                const double x = (double)Ac[ua[m] * LD + ub[m]];
                if (ua[m] == ub[m]) {
                    tot += x * x;
                } else {
                    off += 2.0 * x * x;
                }
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            off += __shfl_xor(off, o);
            tot += __shfl_xor(tot, o);
        }
        if ((tid & 63) == 0) energy[tid >> 6] = make_double2(off, tot);
        __syncthreads();
        off = energy[0].x + energy[1].x + energy[2].x + energy[3].x;
        tot = energy[0].y + energy[1].y + energy[2].y + energy[3].y + off;
        __syncthreads();  // energy[] is rewritten by the next sweep
        if (off == 0.0 || off <= tol * tot) break;

        int iu, ju, iv, jv;
        first(bu, iu, ju);
        first(bv, iv, jv);
        for (int round = 0; round < N; round++, step(iu), step(ju), step(iv), step(jv)) {
            if (is_block) {
                const int pu = min(iu, ju), qu = max(iu, ju), pv = min(iv, jv), qv = max(iv, jv);
                const T uu0 = Ac[pu * LD + pu], uu1 = Ac[qu * LD + qu], uu2 = Ac[pu * LD + qu];
                const T vv0 = Ac[pv * LD + pv], vv1 = Ac[qv * LD + qv], vv2 = Ac[pv * LD + qv];
                const T x00 = Ac[pu * LD + pv], x01 = Ac[pu * LD + qv];
                const T x10 = Ac[qu * LD + pv], x11 = Ac[qu * LD + qv];
                T cu, su, cv, sv;  // (cv, sv) = (cu, su) on a diagonal block
                rotation2(pu == qu, uu0, uu1, uu2, pv == qv, vv0, vv1, vv2, cu, su, cv, sv);
                // Z = X J_v, Y = J_u^T Z
                const T z00 = x00 * cv - x01 * sv, z01 = x00 * sv + x01 * cv;
                const T z10 = x10 * cv - x11 * sv, z11 = x10 * sv + x11 * cv;
                const T y00 = cu * z00 - su * z10, y11 = su * z01 + cu * z11;
                T y01 = cu * z01 - su * z11, y10 = su * z00 + cu * z10;
                if (bu == bv && bu != NS - 1) y01 = y10 = T(0);  // the annihilated pair
                An[pu * LD + pv] = y00;
                An[pu * LD + qv] = y01;
                An[qu * LD + pv] = y10;
                An[qu * LD + qv] = y11;
                An[pv * LD + pu] = y00;
                An[qv * LD + pu] = y01;
                An[pv * LD + qu] = y10;
                An[qv * LD + qu] = y11;
            } else if (is_v) {
                const int p = min(iu, ju), q = max(iu, ju);
                const T a0 = Ac[p * LD + p], a1 = Ac[q * LD + q], a2 = Ac[p * LD + q];
                T v0[3], v1[3];
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    const int j = min(vj0 + 9 * m, N - 1);
                    v0[m] = V[j * LD + p];
                    v1[m] = V[j * LD + q];
                }
                T c, s, c_unused, s_unused;
                rotation2(p == q, a0, a1, a2, p == q, a0, a1, a2, c, s, c_unused, s_unused);
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    const int j = vj0 + 9 * m;
                    if (j < N) {
                        V[j * LD + p] = c * v0[m] - s * v1[m];
                        V[j * LD + q] = s * v0[m] + c * v1[m];
                    }
                }
            }
            __syncthreads();
            T* sw = Ac;
            Ac = An;
            An = sw;
        }
    }
}

// Covariance + Jacobi eigensolve + basis selection.
//
// C = (sum x x^T - (sum x)(sum x)^T / n) / (n - 1)   -- PCA.cpp:31-32
// FP64 Jacobi sweeps until the off-diagonal energy is kJacobiTol of the total.
// The 8 largest eigenpairs, largest first (PCA.cpp:49-55), are written as
// basis[r * 8 + c]; ties in eigenvalue keep index order; each vector is
// oriented to a non-negative dot product with sign_ref.
__device__ void pca_solve_body(const double* moments, double n, float* basis, double* eigvals, double* eig_full) {
    constexpr int N = kPatch;
    constexpr int LD = kLD;
    __shared__ double mom[kMoments];
    __shared__ double A0[N * LD], A1[N * LD];  // current and next A
    __shared__ double V[N * LD];
    __shared__ double2 energy[4];
    const int tid = threadIdx.x;

    for (int e = tid; e < kMoments; e += 256) mom[e] = moments[e];
    __syncthreads();
    for (int e = tid; e < kMoments; e += 256) {
        if (e < N) continue;
        int r, s;
        moment_rs(e, r, s);
        const double c = (mom[e] - mom[r] * mom[s] / n) / (n - 1.0);
        A0[r * LD + s] = c;
        A0[s * LD + r] = c;
    }
    for (int e = tid; e < N * N; e += 256) {
        const int r = e / N, s = e % N;
        V[r * LD + s] = r == s ? 1.0 : 0.0;
    }
    __syncthreads();
    double *Ac = A0, *An = A1;
    jacobi_sweeps<double>(Ac, An, V, kJacobiTol, 40, energy);
    __syncthreads();

    if (tid < N) {
        const double di = Ac[tid * LD + tid];
        int rank = 0;
        for (int j = 0; j < N; j++) {
            const double dj = Ac[j * LD + j];
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

__global__ __launch_bounds__(256) void k_pca_solve(const double* moments, double n, float* basis, double* eigvals,
                                                   double* eig_full) {
    pca_solve_body(moments, n, basis, eigvals, eig_full);
}

// Burst forms: moments of every frame of a level in one launch (frame f owns
// workgroups block0 .. block0 + nblocks - 1), the reductions on a
// (25 x frames) grid, one eigensolve workgroup per frame.
__global__ __launch_bounds__(256) void k_burst_moments(const BurstTable<BurstMomFrame>* __restrict__ tab) {
    const BurstMomFrame& f = tab->frame[burst_frame_of(tab, blockIdx.x)];
    const int lb = blockIdx.x - f.block0;
    const View<const float4> in = f.in;
    const int gw = in.w / 8, gh = in.h / 8;
    const int64_t npatch = (int64_t)gw * gh;
    moments_body(lb, f.nblocks, npatch, f.partials + (size_t)lb * kMoments,
                 [&](int64_t k, int c) { return image_patch_value(in, gw, npatch, k, c); });
}

__global__ __launch_bounds__(256) void k_burst_moments_reduce(const BurstTable<BurstMomFrame>* __restrict__ tab) {
    const BurstMomFrame& f = tab->frame[blockIdx.y];
    moments_reduce_body(blockIdx.x, f.partials, f.nblocks, f.partials + (size_t)kMomentBlocks * kMoments);
}

__global__ __launch_bounds__(256) void k_burst_pca_solve(const BurstTable<BurstMomFrame>* __restrict__ tab) {
    const BurstMomFrame& f = tab->frame[blockIdx.x];
    pca_solve_body(f.partials + (size_t)kMomentBlocks * kMoments, f.n, f.basis, f.eigvals, nullptr);
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

// Workgroups of the moment reduction: one per 64 patches (each of its four
// wavefronts then takes four groups of four patches: a short latency chain),
// at most kMomentBlocks.
static int moment_blocks(int64_t npatch) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(kMomentBlocks, (npatch + 63) / 64));
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

gcam_status launch_pca_solve(double* partials, int nblocks, int64_t n, float* basis, double* eigvals,
                             double* eig_full, hipStream_t s) {
    double* mom = partials + (size_t)kMomentBlocks * kMoments;
    k_moments_reduce<<<kMoments / kRedEntries, 256, 0, s>>>(partials, nblocks, mom);
    k_pca_solve<<<1, 256, 0, s>>>(mom, (double)n, basis, eigvals, eig_full);
    GCAM_LAUNCH_CHECK("PCA eigensolve");
    return GCAM_OK;
}

gcam_status launch_burst_pca(const BurstTable<BurstMomFrame>* dev, int total_blocks, int nframes, hipStream_t s) {
    if (nframes < 1) return GCAM_OK;
    k_burst_moments<<<total_blocks, 256, 0, s>>>(dev);
    k_burst_moments_reduce<<<dim3(kMoments / kRedEntries, nframes), 256, 0, s>>>(dev);
    k_burst_pca_solve<<<nframes, 256, 0, s>>>(dev);
    GCAM_LAUNCH_CHECK("burst PCA basis");
    return GCAM_OK;
}

int moment_block_count(int64_t npatch) { return moment_blocks(npatch); }

gcam_status launch_pca_project(View<const float4> in, const float* basis, View<float> feat8, hipStream_t s) {
    k_pca_project<<<dim3((in.w + 63) / 64, (in.h + 3) / 4), dim3(64, 4), 0, s>>>(in, basis, feat8.ptr, feat8.stride);
    GCAM_LAUNCH_CHECK("pcaProjection");
    return GCAM_OK;
}

}  // namespace gcam
