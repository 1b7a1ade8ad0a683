// PCA patch basis on the device: patch sampling, FP64 moments, FP64
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
// of workgroups, one workgroup turns them into the covariance and finds its
// leading eigenpairs (Householder tridiagonalisation, multisection and
// inverse iteration, with parallel-ordered cyclic Jacobi as the checked
// fallback), and the projection reads the basis from device memory, so a
// whole denoise call stays stream-ordered.
#include "gcam_internal.hpp"
#include "burst_tables.hpp"
#include "moments.hpp"

namespace gcam {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Sign reference of the basis vectors: each vector is oriented so that its dot
// product with this fixed vector is non-negative (any eigenvector solver
// leaves the sign free; fixing it makes the features reproducible).
__device__ __forceinline__ double sign_ref(int r) { return 1.0 + (double)((r * r * 7 + r * 3 + 5) % 23) / 16.0; }


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
        for (int u = 0; u < U; u++) moments_group(lo[u], hi[u], c00, c01, c11, slo, shi);
    }
    moments_store<4>(tid, 256, c00, c01, c11, slo, shi, gram, sums, partial);
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
        // Unrolled so that the loads of 8 partials are in flight together; the
        // additions stay in index order.
#pragma unroll 8
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

// FP64 reciprocal square root for the rotations: the hardware estimate
// (v_rsq_f64, about 23 correct bits) refined by one Newton step
// y (1.5 - 0.5 x y^2), written as y * fma(x y, -0.5 y, 1.5): about 46 bits.
// A rotation is then orthogonal to ~1e-14, three orders below the sweeps'
// stopping point (off-diagonal entries ~1e-11 of the norm).
__device__ __forceinline__ double rsq_refined(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    return y * fma(x * y, -0.5 * y, 1.5);
}

// Cyclic Jacobi sweeps on the symmetric 25 x 25 matrix C, in the parallel
// (circle) ordering: round r of a sweep pairs the indices into super-indices
// u = 0..11, {(r + u + 1) mod 25, (r - u - 1) mod 25}, and u = 12 = {r}, and
// applies the 12 rotations at once, A <- J^T A J and V <- V J.
//
// Slot frame: in round r, slot s stands for index (s + r) mod 25, so the pairs
// are the same slots every round: super-index u < 12 is slots f = u + 1 and
// g = 24 - u, and u = 12 is slot 0.  A is stored as the 91 2 x 2 blocks
// A[{f_u, g_u}][{f_v, g_v}], u <= v (lo = row f, hi = row g, 16 bytes each; a
// diagonal block keeps both copies of its off-diagonal entry, absent rows and
// columns of u = 12 stay zero), and V as the row pairs (V[j][f_u], V[j][g_u]).
// Moving to the next round shifts every slot down by one, so each rotated
// entry goes to a destination that is the same in every round: the addresses
// are computed once.  Threads 0..90 (wavefronts 0 and 1) own one block each
// and compute the two rotations they need from the diagonal blocks; threads
// 128..244 (wavefronts 2 and 3) own rows {j0, j0 + 9, j0 + 18} of V at one
// super-index.  Every read is of the current buffer and every write of the
// other one, so a round needs one barrier.  After a sweep (25 rounds) the frame
// is the identity again.  Rotation: the Golub & Van Loan angle
// (|phi| <= pi/4, tan(2 phi) = 2 a_fg / (a_gg - a_ff)) from its half-angle
// form, c^2 = (1 + |d| / h) / 2, s = sign(d e) |e| / (2 h c), h = hypot(d, e),
// d = a_gg - a_ff, e = 2 a_fg: one reciprocal square root for 1/h, one for
// 1/c.  The rotation does not depend on which of a pair is called f.
//
// On entry C is the matrix (row pitch kLD); on return its diagonal holds the
// eigenvalues and V (row pitch kLD) the eigenvectors as columns.  Stops when
// the off-diagonal energy is at most tol of the total, or after max_sweeps.
// Row pitch of the 25 x 25 matrices in LDS: odd, so the rows of a column fall
// in different banks.
constexpr int kLD = 33;

namespace slot {
constexpr int NS = 13, NB = NS * (NS + 1) / 2;  // super-indices, blocks
__device__ __forceinline__ int of(int u, int i) { return u == NS - 1 ? (i == 0 ? 0 : -1) : (i == 0 ? u + 1 : 24 - u); }
__device__ __forceinline__ void loc(int s, int& u, int& i) {
    u = s == 0 ? NS - 1 : (s <= 12 ? s - 1 : 24 - s);
    i = s <= 12 ? 0 : 1;
}
__device__ __forceinline__ int block(int u, int v) { return u * NS - u * (u - 1) / 2 + (v - u); }  // u <= v
// Element (i, j) of block b in a buffer of 4 NB doubles: lo rows first.
__device__ __forceinline__ int elem(int b, int i, int j) { return (i == 0 ? 0 : 2 * NB) + 2 * b + j; }
}  // namespace slot

__device__ void jacobi_sweeps(double* C, double* V, double2 (*Ab)[2 * slot::NB], double2 (*Vb)[kPatch * slot::NS],
                              double tol, int max_sweeps, double2* energy) {
    using namespace slot;
    constexpr int N = kPatch, LD = kLD;
    constexpr int V0T = 128, NVT = NS * 9;  // V threads: wavefronts 2 and 3
    const double one = 1.0, half = 0.5, guard = 1e-300;
    const int tid = threadIdx.x;
    const bool is_block = tid < NB, is_v = tid >= V0T && tid < V0T + NVT;
    int bu = NS - 1, bv = NS - 1;
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
    const int du = block(bu, bu), dv = block(bv, bv);

    // Destinations in the next round of this thread's entries (-1: none) and,
    // for an entry landing off the diagonal of a diagonal block, of its mirror.
    int dst[4], mir[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int i = k >> 1, j = k & 1;
        const int s = of(bu, i), t = of(bv, j);
        dst[k] = mir[k] = -1;
        if (!is_block || s < 0 || t < 0 || (bu == bv && i == 1 && j == 0)) continue;
        int u1, i1, v1, j1;
        loc((s + 24) % 25, u1, i1);
        loc((t + 24) % 25, v1, j1);
        if (u1 > v1) {
            const int tu = u1, ti = i1;
            u1 = v1;
            i1 = j1;
            v1 = tu;
            j1 = ti;
        }
        const int b2 = block(u1, v1);
        dst[k] = elem(b2, i1, j1);
        if (u1 == v1 && i1 != j1) mir[k] = elem(b2, j1, i1);
    }
    int vd[3][2];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const int j = vj0 + 9 * m;
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int s = of(bu, i);
            vd[m][i] = -1;
            if (!is_v || j >= N || s < 0) continue;
            int u1, i1;
            loc((s + 24) % 25, u1, i1);
            vd[m][i] = (j * NS + u1) * 2 + i1;
        }
    }

    // Load C and V = I into slot blocks (frame of round 0: slot = index).
    for (int e = tid; e < 2 * 2 * NB; e += 256) reinterpret_cast<double2*>(Ab)[e] = make_double2(0.0, 0.0);
    for (int e = tid; e < 2 * N * NS; e += 256) reinterpret_cast<double2*>(Vb)[e] = make_double2(0.0, 0.0);
    __syncthreads();
    if (is_block) {
        double* A = reinterpret_cast<double*>(Ab[0]);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int s = of(bu, k >> 1), t = of(bv, k & 1);
            if (s >= 0 && t >= 0) A[elem(tid, k >> 1, k & 1)] = C[s * LD + t];
        }
    }
    for (int e = tid; e < N * NS; e += 256) {
        const int j = e / NS, u = e % NS;
        Vb[0][e] = make_double2(of(u, 0) == j ? 1.0 : 0.0, of(u, 1) == j ? 1.0 : 0.0);
    }
    __syncthreads();

    // Straight-line code (no select on the expensive operands, which the
    // compiler would turn into branches).  The identity (u = 12, or a_fg == 0)
    // is blended in arithmetically: c + m (1 - c) = 1 and s (1 - m) = 0 exactly
    // for m = 1 (1 - c is exact for c in [1/2, 1]).  rotation2 advances two
    // independent rotations step by step, so their dependent chains overlap.
    auto rotation2 = [&](bool id0, double aff0, double agg0, double afg0, bool id1, double aff1, double agg1,
                         double afg1, double& c0, double& s0, double& c1, double& s1) {
        const double e0 = 2.0 * afg0, d0 = agg0 - aff0, e1 = 2.0 * afg1, d1 = agg1 - aff1;
        const double m0 = (id0 || e0 == 0.0) ? one : 0.0, m1 = (id1 || e1 == 0.0) ? one : 0.0;
        const double y0 = rsq_refined(fma(d0, d0, e0 * e0) + guard), y1 = rsq_refined(fma(d1, d1, e1 * e1) + guard);
        const double c20 = half * fma(fabs(d0), y0, one), c21 = half * fma(fabs(d1), y1, one);  // cos^2
        const double z0 = rsq_refined(c20), z1 = rsq_refined(c21);                              // 1 / c
        const double cc0 = c20 * z0, cc1 = c21 * z1;
        c0 = fma(m0, one - cc0, cc0);
        c1 = fma(m1, one - cc1, cc1);
        s0 = copysign((half * fabs(e0) * y0) * z0, d0 * e0) * (one - m0);
        s1 = copysign((half * fabs(e1) * y1) * z1, d1 * e1) * (one - m1);
    };

    int cur = 0;
    for (int sweep = 0; sweep < max_sweeps; sweep++) {
        // Off-diagonal and total energy: per-block entries, wavefront sums,
        // then the 4 wavefront sums in a fixed order (uniform).
        double off = 0.0, tot = 0.0;
        if (is_block) {
            const double2 lo = Ab[cur][tid], hi = Ab[cur][NB + tid];
            const double q00 = lo.x * lo.x, q01 = lo.y * lo.y, q10 = hi.x * hi.x, q11 = hi.y * hi.y;
            if (bu != bv) {
                off = 2.0 * (q00 + q01 + q10 + q11);
            } else {
                tot = q00 + q11;
                off = q01 + q10;
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

        for (int round = 0; round < N; round++) {
            double* An = reinterpret_cast<double*>(Ab[cur ^ 1]);
            double* Vn = reinterpret_cast<double*>(Vb[cur ^ 1]);
            if (is_block) {
                const double2 xl = Ab[cur][tid], xh = Ab[cur][NB + tid];
                const double2 ul = Ab[cur][du], uh = Ab[cur][NB + du], vl = Ab[cur][dv], vh = Ab[cur][NB + dv];
                double cu, su, cv, sv;
                rotation2(bu == NS - 1, ul.x, uh.y, ul.y, bv == NS - 1, vl.x, vh.y, vl.y, cu, su, cv, sv);
                // Z = X J_v, Y = J_u^T Z
                const double z00 = xl.x * cv - xl.y * sv, z01 = xl.x * sv + xl.y * cv;
                const double z10 = xh.x * cv - xh.y * sv, z11 = xh.x * sv + xh.y * cv;
                double y[4] = {cu * z00 - su * z10, cu * z01 - su * z11, su * z00 + cu * z10, su * z01 + cu * z11};
                if (bu == bv && bu != NS - 1) y[1] = y[2] = 0.0;  // the annihilated pair
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    if (dst[k] >= 0) An[dst[k]] = y[k];
                    if (mir[k] >= 0) An[mir[k]] = y[k];
                }
            } else if (is_v) {
                const double2 ul = Ab[cur][du], uh = Ab[cur][NB + du];
                double2 p[3];
#pragma unroll
                for (int m = 0; m < 3; m++) p[m] = Vb[cur][min(vj0 + 9 * m, N - 1) * NS + bu];
                double c, s, c_unused, s_unused;
                rotation2(bu == NS - 1, ul.x, uh.y, ul.y, bu == NS - 1, ul.x, uh.y, ul.y, c, s, c_unused, s_unused);
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    if (vd[m][0] >= 0) Vn[vd[m][0]] = c * p[m].x - s * p[m].y;
                    if (vd[m][1] >= 0) Vn[vd[m][1]] = s * p[m].x + c * p[m].y;
                }
            }
            __syncthreads();
            cur ^= 1;
        }
    }

    // Back to the plain layout (the frame is the identity at a sweep boundary).
    const double* A = reinterpret_cast<const double*>(Ab[cur]);
    const double* Vs = reinterpret_cast<const double*>(Vb[cur]);
    if (tid < N) {
        int u, i;
        loc(tid, u, i);
        C[tid * LD + tid] = A[elem(block(u, u), i, i)];
    }
    for (int e = tid; e < N * N; e += 256) {
        const int j = e / N, a = e % N;
        int u, i;
        loc(a, u, i);
        V[j * LD + a] = Vs[(j * NS + u) * 2 + i];
    }
}

// ---- leading eigenpairs without Jacobi sweeps ----
//
// The forward pass needs the 8 leading eigenvectors and the 25 eigenvalues.
// Wave 0 reduces the covariance (scaled by its largest entry) to tridiagonal
// form T = Q^T C Q with 23 Householder reflections H_k = I - u_k u_k^T,
// accumulating Q = H_0 H_1 .. H_22 in registers as it goes; every eigenvalue
// of T is bracketed by 17 rounds of 9-way multisection (8 lanes per
// eigenvalue; Sturm counts from the division-free recurrence of the leading
// minors, p_i = (d_i - x) p_{i-1} - e_{i-1}^2 p_{i-2}, rescaled by powers of
// two); 8 lanes run one solve of inverse iteration on T - lambda I (Gaussian
// elimination with partial pivoting, from a fixed start vector, all in
// registers) and the vectors are mapped back as v = Q y.  The result is
// accepted when every residual |C v - lambda v| and every
// |v_i . v_j - delta_ij| is at most 1e-10 of the matrix scale; otherwise
// (eigenvalues too close for inverse iteration to separate, a degenerate
// matrix) the caller runs the Jacobi sweeps.
struct Top8 {
    double W[kPatch][kPatch + 1];  // working matrix on entry; Q after the reduction
    double bu[64], bp[64];         // u and p = W u of a reduction step, one entry per lane
    double d[kPatch], e[kPatch];
    double2 de2[kPatch];           // (d_i, e_{i-1}^2), e_{-1} = 0
    double lam[32];                // eigenvalues of T, descending
    double Y[kComponents][kPatch + 1];  // eigenvectors of T (the eliminated b before)
    double ur[kComponents][kPatch], u1[kComponents][kPatch], u2[kComponents][kPatch];  // U of T - lambda I
    double X[kComponents][kPatch];      // eigenvectors of C
    double red[4];
    double lo, hi;
    double scale;  // C = scale * W on entry (W's largest entry is 1)
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The value of v held by lane l (l wave-uniform), in every lane.
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// v plus the value of the same lane in the other 32-lane half of the
// wavefront (v_permlane32_swap on both words); bit-identical in both halves.
__device__ __forceinline__ double sum_halves(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
    const double a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    const double c = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
    return a + c;
}

// Eigenvalues of T below x: sign changes along p_0 = 1, p_1, .., p_25
// (p_{-1} = 0); d[i] = T_ii, e2[i] = T_{i,i-1}^2 (e2[0] = 0), held in
// registers; the pair is rescaled by a power of two every five terms.
__device__ __forceinline__ int sturm_below(const double (&d)[kPatch], const double (&e2)[kPatch], double x) {
    double pm2 = 0.0, pm1 = 1.0;
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < kPatch; i++) {
        const double p = fma(d[i] - x, pm1, -e2[i] * pm2);
        cnt += (p < 0.0) != (pm1 < 0.0) ? 1 : 0;
        pm2 = pm1;
        pm1 = p;
        if (i % 5 == 4) {
            const double a = fmax(fabs(pm1), fabs(pm2));
            const double f = a > 0x1p+400 ? 0x1p-400 : (a < 0x1p-400 ? 0x1p+400 : 1.0);
            pm1 *= f;
            pm2 *= f;
        }
    }
    return cnt;
}

__device__ __forceinline__ double rcp_refined(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = fma(y, fma(-x, y, 1.0), y);
    return fma(y, fma(-x, y, 1.0), y);
}

// 1 / sqrt(x) for x > 0 to full double precision (two Newton steps).
__device__ __forceinline__ double rsq_full(double x) {
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(x * y, -0.5 * y, 1.5);
    return y * fma(x * y, -0.5 * y, 1.5);
}

// The Gershgorin interval of the tridiagonal T (diagonal d, off-diagonal e),
// widened by a relative 1e-12, and the Sturm recurrence terms
// de2[i] = (d_i, e_{i-1}^2) (e_{-1} = 0).
__device__ __forceinline__ void tridiag_bracket(const double* d, const double* e, double2* de2, double& lo,
                                                double& hi) {
    constexpr int N = kPatch;
    lo = 1e300;
    hi = -1e300;
    for (int i = 0; i < N; i++) {
        const double rad = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < N - 1 ? fabs(e[i]) : 0.0);
        lo = fmin(lo, d[i] - rad);
        hi = fmax(hi, d[i] + rad);
        de2[i] = make_double2(d[i], i > 0 ? e[i - 1] * e[i - 1] : 0.0);
    }
    const double pad = 1e-12 * fmax(fabs(lo), fabs(hi)) + 1e-300;
    lo -= pad;
    hi += pad;
}

// Eigenvalue j (the j-th largest) of T by `rounds` rounds of (NL + 1)-way
// multisection of [lo, hi] on a group of NL lanes (lane m of the group tests
// lo + (m + 1) w / (NL + 1)); every lane of the group returns the midpoint of
// the final bracket.  17 rounds at NL = 8 and 11 at NL = 32 both shrink the
// bracket by more than 2^53.
template <int NL>
__device__ __forceinline__ double multisect(const double (&dd)[kPatch], const double (&ee)[kPatch], double lo,
                                            double hi, int j, int m, int rounds) {
    const int target = kPatch - 1 - j;  // x is above the j-th largest iff more than target lie below x
    const int lane = threadIdx.x & 63;
    constexpr uint64_t kMask = NL >= 64 ? ~0ull : ((1ull << NL) - 1);
    const double at = (double)(m + 1) / (double)(NL + 1);
    for (int it = 0; it < rounds; it++) {
        const double w = hi - lo;
        const double x = fma(w, at, lo);
        const bool above = sturm_below(dd, ee, x) > target;
        const uint64_t grp = (__ballot(above) >> (lane & ~(NL - 1))) & kMask;
        const int ms = grp ? __builtin_ctzll(grp) : NL;
        const double nlo = ms == 0 ? lo : fma(w, (double)ms / (double)(NL + 1), lo);
        const double nhi = ms == NL ? hi : fma(w, (double)(ms + 1) / (double)(NL + 1), lo);
        lo = nlo;
        hi = nhi;
    }
    return 0.5 * (lo + hi);
}

// Writes S.lam (scaled back: eigenvalues of C, descending) and S.X (the 8
// leading eigenvectors); returns whether they pass the checks.  kAll: all 25
// eigenvalues (8 lanes each, 17 rounds); otherwise the 8 leading only (32
// lanes each, 11 rounds) and S.lam[8..24] = NaN.  A0: the
// covariance (leading dimension kLD).  Every thread of the workgroup calls it.
template <bool kAll>
__device__ bool top8_eigen(const double* A0, Top8& S) {
    constexpr int N = kPatch, LD = kLD;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    double mx = 0.0;
    for (int e = tid; e < N * N; e += 256) mx = fmax(mx, fabs(A0[(e / N) * LD + e % N]));
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if (lane == 0) S.red[wave] = mx;
    __syncthreads();
    const double scale = fmax(fmax(S.red[0], S.red[1]), fmax(S.red[2], S.red[3]));
    if (tid == 0) S.scale = scale;
    if (!(scale > 0.0) || !(scale < 1e300)) return false;
    const double rscale = 1.0 / scale;
    for (int e = tid; e < N * N; e += 256) S.W[e / N][e % N] = A0[(e / N) * LD + e % N] * rscale;
    __syncthreads();

    // Householder tridiagonalisation on wave 0, every step unrolled.  Lane l
    // holds row r = l % 32 (r < N) of W and of Q, columns c = 2 j + h (h = l / 32,
    // the lane's half, j = 0 .. 12) in registers: the halves split every row's
    // columns by parity, so the active columns of a step (c > k) fall evenly on
    // both and a lane does half of its row's work; sums over a row's columns
    // are completed by one v_permlane32_swap exchange between the halves (the
    // same value in both).  Step k: the lanes of row k form
    // alpha = -sign(x_0) |x|, v_0 = x_0 - alpha and f = sqrt(2 / |v|^2)
    // (x = W[k][k+1 ..] = column k below the diagonal), broadcast with
    // readlane; the half holding column k forms u_r = v_r f from its W[r][k].
    // Two exchanges through LDS: u, then p = W u; with K = u . p,
    // W <- W - u p^T - p u^T + K u u^T (row r: w_c -= u_r p_c + (p_r - K u_r)
    // u_c) and Q <- Q - (Q u) u^T.  Entries of columns c <= k are stale and
    // are masked where they would enter a sum (u_c = p_c = 0 there, so the
    // updates leave them unchanged).
    if (wave == 0) {
        constexpr int J = (N + 1) / 2;
        const int h = lane >> 5, r = lane & 31;
        const bool row_ok = r < N;
        double w[J], q[J];
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int c = 2 * j + h;
            w[j] = row_ok && c < N ? S.W[r][c] : 0.0;
            q[j] = row_ok && c == r ? 1.0 : 0.0;
        }
#pragma unroll
        for (int k = 0; k < N - 2; k++) {
            const int j0 = k / 2;  // first register slot with an active column in either half
            // |x|^2 of row k over c > k (every row forms its own; row k's is used)
            double na = 0.0, nb = 0.0;
#pragma unroll
            for (int j = j0; j < J; j++) {
                const double x = (j > j0 || 2 * j + h > k) ? w[j] : 0.0;
                if ((j - j0) & 1) nb = fma(x, x, nb);
                else na = fma(x, x, na);
            }
            const double n2 = sum_halves(na + nb);
            const int kx = k + 1, xl = k + 32 * (kx & 1);  // lane holding x_0 = W[k][k + 1]
            const double x0 = readlane_f64(w[kx >> 1], xl);
            const double nrm = n2 > 0.0 ? n2 * rsq_full(n2) : 0.0;
            const double al = x0 < 0.0 ? nrm : -nrm;
            const double v0l = x0 - al;
            const double nv = fma(v0l, v0l, fma(-x0, x0, n2));
            const double fl = nv > 0.0 ? 1.4142135623730951 * rsq_full(nv) : 0.0;
            const double alpha = readlane_f64(al, k), v0 = readlane_f64(v0l, k), f = readlane_f64(fl, k);
            // column k lives in half k & 1, slot k / 2
            const bool below = row_ok && r > k;
            if (h == (k & 1) && row_ok) {
                S.bu[r] = below ? (r == k + 1 ? v0 : w[k >> 1]) * f : 0.0;
                if (r == k) S.d[k] = w[k >> 1];
            }
            if (lane == 0) S.e[k] = alpha;
            wave_sync();
            double uc[J];
            double pa = 0.0, pb = 0.0, qa = 0.0, qb = 0.0;
#pragma unroll
            for (int j = j0; j < J; j++) {
                uc[j] = 2 * j + h < N ? S.bu[2 * j + h] : 0.0;
                if ((j - j0) & 1) {
                    pb = fma(w[j], uc[j], pb);
                    qb = fma(q[j], uc[j], qb);
                } else {
                    pa = fma(w[j], uc[j], pa);
                    qa = fma(q[j], uc[j], qa);
                }
            }
            const double u = row_ok ? S.bu[r] : 0.0;
            const double psum = sum_halves(pa + pb);
            const double qu = sum_halves(qa + qb);
            const double p = below ? psum : 0.0;
            if (h == 0 && row_ok) S.bp[r] = p;
            wave_sync();
            double pv[J];
            double ka = 0.0, kb = 0.0;
#pragma unroll
            for (int j = j0; j < J; j++) {
                pv[j] = 2 * j + h < N ? S.bp[2 * j + h] : 0.0;
                if ((j - j0) & 1) kb = fma(uc[j], pv[j], kb);
                else ka = fma(uc[j], pv[j], ka);
            }
            const double K = sum_halves(ka + kb);
            const double pp = fma(-K, u, p);
#pragma unroll
            for (int j = j0; j < J; j++) {
                w[j] = fma(-u, pv[j], fma(-pp, uc[j], w[j]));
                q[j] = fma(-qu, uc[j], q[j]);
            }
            wave_sync();
        }
        // d[N-2] = W[N-2][N-2] (half 1, slot 11), d[N-1] = W[N-1][N-1] (half 0,
        // slot 12), e[N-2] = W[N-1][N-2] (half 1, slot 11)
        if (lane == (N - 2) + 32) S.d[N - 2] = w[(N - 2) >> 1];
        if (lane == N - 1) S.d[N - 1] = w[(N - 1) >> 1];
        if (lane == (N - 1) + 32) S.e[N - 2] = w[(N - 2) >> 1];
        // Q replaces W in LDS (W is no longer needed).
        if (row_ok) {
#pragma unroll
            for (int j = 0; j < J; j++)
                if (2 * j + h < N) S.W[r][2 * j + h] = q[j];
        }
        wave_sync();
        if (lane == 0) tridiag_bracket(S.d, S.e, S.de2, S.lo, S.hi);
    }
    __syncthreads();

    // Multisection, all eigenvalues: eigenvalue j on lanes 8j' .. 8j' + 7 of
    // wave j / 8; the 8 leading only: eigenvalue j on lanes 32 (j % 2) ..
    // 32 (j % 2) + 31 of wave j / 2.
    {
        double dd[N], ee[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const double2 t = S.de2[i];
            dd[i] = t.x;
            ee[i] = t.y;
        }
        if constexpr (kAll) {
            const int j = wave * 8 + (lane >> 3), m = lane & 7;
            const double lam = multisect<8>(dd, ee, S.lo, S.hi, j, m, 17);
            if (m == 0 && j < N) S.lam[j] = lam;
        } else {
            const int j = wave * 2 + (lane >> 5), m = lane & 31;
            const double lam = multisect<32>(dd, ee, S.lo, S.hi, j, m, 11);
            if (m == 0) S.lam[j] = lam;
            if (tid >= kComponents && tid < N) S.lam[tid] = __longlong_as_double(0x7ff8000000000000ll);
        }
    }
    __syncthreads();

    // Inverse iteration, eigenvector j on lane j of wave 0: one solve of
    // (T - lambda I) y = b from a fixed start vector b (lambda is accurate to
    // rounding, so one solve suffices), by Gaussian elimination with partial
    // pivoting applied to b as it goes (U has two superdiagonals; pivots
    // below 1e-15 |T| are raised to that size), then back substitution and
    // normalisation, unrolled (U and the eliminated b in LDS, the rest in
    // registers).
    if (wave == 0 && lane < kComponents) {
        const int j = lane;
        const double lam = S.lam[j];
        const double tiny = 1e-15 * fmax(fabs(S.lo), fabs(S.hi)) + 1e-300;
        double* ur = S.ur[j];
        double* u1 = S.u1[j];
        double* u2 = S.u2[j];
        double* bb = S.Y[j];
        // Start vector b_i = 1 + m_i / 44, m_i = (7 i + 3 j) mod 11: fixed, positive and
        // different per eigenvector (one solve from any such vector suffices), stepped
        // incrementally so that no division or modulo enters the elimination.
        int m = 3 * j >= 11 ? 3 * j - 11 : 3 * j;
        auto next_start = [&]() {
            const double b = fma((double)m, 1.0 / 44.0, 1.0);
            m = m >= 4 ? m - 4 : m + 7;
            return b;
        };
        double c0 = S.d[0] - lam, c1 = S.e[0], cb = next_start();
#pragma unroll
        for (int i = 0; i < N - 1; i++) {
            const double nl = S.e[i], n0 = S.d[i + 1] - lam, n1 = i + 1 < N - 1 ? S.e[i + 1] : 0.0;
            const double nb = next_start();
            const bool sw = fabs(c0) < fabs(nl);
            const double pv = sw ? nl : (fabs(c0) < tiny ? copysign(tiny, c0) : c0);
            const double rp = rcp_refined(pv);
            const double mu = (sw ? c0 : nl) * rp;
            ur[i] = rp;
            u1[i] = sw ? n0 : c1;
            u2[i] = sw ? n1 : 0.0;
            bb[i] = sw ? nb : cb;
            const double nc0 = sw ? fma(-mu, n0, c1) : fma(-mu, c1, n0);
            const double ncb = sw ? fma(-mu, nb, cb) : fma(-mu, cb, nb);
            c1 = sw ? -mu * n1 : n1;
            c0 = nc0;
            cb = ncb;
        }
        double x1 = cb * rcp_refined(fabs(c0) < tiny ? copysign(tiny, c0) : c0);
        double x0 = fma(-u1[N - 2], x1, bb[N - 2]) * ur[N - 2];
        bb[N - 1] = x1;
        bb[N - 2] = x0;
        double n2 = fma(x1, x1, x0 * x0);
#pragma unroll
        for (int i = N - 3; i >= 0; i--) {
            const double xi = fma(-u2[i], x1, fma(-u1[i], x0, bb[i])) * ur[i];
            bb[i] = xi;
            n2 = fma(xi, xi, n2);
            x1 = x0;
            x0 = xi;
        }
        const double rn = rsq_full(n2);
#pragma unroll
        for (int i = 0; i < N; i++) bb[i] *= rn;
    }
    __syncthreads();

    // Back to C's basis: X[j][r] = sum_c Q[r][c] Y[j][c] (threads 0 .. 199).
    if (tid < kComponents * N) {
        const int j = tid / N, r = tid % N;
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int c = 0; c < N; c += 2) {
            a = fma(S.W[r][c], S.Y[j][c], a);
            if (c + 1 < N) b = fma(S.W[r][c + 1], S.Y[j][c + 1], b);
        }
        S.X[j][r] = a + b;
    }
    if (tid < N) S.lam[tid] *= scale;
    __syncthreads();

    // Checks: residuals (threads 0 .. 199) and the Gram matrix (200 .. 235).
    double err = 0.0;
    if (tid < kComponents * N) {
        const int j = tid / N, r = tid % N;
        double acc = -S.lam[j] * S.X[j][r];
        for (int c = 0; c < N; c++) acc = fma(A0[r * LD + c], S.X[j][c], acc);
        err = fabs(acc) * rscale;
    } else if (tid < kComponents * N + 36) {
        int pidx = tid - kComponents * N, a = 0;
        while (pidx >= kComponents - a) {
            pidx -= kComponents - a;
            a++;
        }
        const int b2 = a + pidx;
        double g = a == b2 ? -1.0 : 0.0;
        for (int i = 0; i < N; i++) g = fma(S.X[a][i], S.X[b2][i], g);
        err = fabs(g);
    }
    if (!(err <= 1e-10)) err = 1.0;  // NaN too
    for (int o = 32; o > 0; o >>= 1) err = fmax(err, __shfl_xor(err, o));
    if (lane == 0) S.red[wave] = err;
    __syncthreads();
    const bool ok = fmax(fmax(S.red[0], S.red[1]), fmax(S.red[2], S.red[3])) <= 1e-10;
    __syncthreads();
    return ok;
}

// Covariance + eigensolve + basis selection.
//
// C = (sum x x^T - (sum x)(sum x)^T / n) / (n - 1)   -- PCA.cpp:31-32
// Leading eigenpairs by top8_eigen; where its checks fail, FP64 Jacobi sweeps
// until the off-diagonal energy is kJacobiTol of the total.
// The 8 largest eigenpairs, largest first (PCA.cpp:49-55), are written as
// basis[r * 8 + c]; ties in eigenvalue keep index order; each vector is
// oriented to a non-negative dot product with sign_ref.
// eigvals (nullable): the 25 eigenvalues.  tri (nullable): with eigvals set,
// the solve brackets only the 8 leading eigenvalues (the basis needs no more),
// writes NaN for eigenvalues 9..25 and keeps the tridiagonal form in tri
// (diagonal 0..24, off-diagonal 25..48, scale 49) for k_tridiag_eigvals to
// complete them on demand; without tri the solve brackets all 25.
__device__ __forceinline__ void pca_solve_body(const double* moments, double n, float* basis, double* eigvals,
                                               double* eig_full, double* tri) {
    constexpr int N = kPatch;
    constexpr int LD = kLD;
    __shared__ double mom[kMoments];
    __shared__ double A0[N * LD];  // the covariance; its diagonal the eigenvalues after the sweeps
    __shared__ double V[N * LD];
    __shared__ double2 Ab[2][2 * slot::NB];  // slot blocks, current and next
    __shared__ double2 Vb[2][N * slot::NS];  // V row pairs, current and next
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
    __syncthreads();
    // The basis and eigenvalues come from the leading-eigenpair solver when
    // its checks pass; the Jacobi sweeps run when they fail, and for the full
    // decomposition of the reverse pass (which then only fills eig_full).
    __shared__ Top8 t8;
    const bool all = eigvals && !tri;
    const bool fast = all ? top8_eigen<true>(A0, t8) : top8_eigen<false>(A0, t8);
    if (fast) {
        if (tid < N && eigvals) eigvals[tid] = t8.lam[tid];
        if (tri && eigvals) {
            if (tid < N) tri[tid] = t8.d[tid];
            if (tid < N - 1) tri[N + tid] = t8.e[tid];
            if (tid == 0) tri[2 * N - 1] = t8.scale;
        }
        if (tid < kComponents) {
            double dot = 0.0;
            for (int r = 0; r < N; r++) dot += t8.X[tid][r] * sign_ref(r);
            const double sg = dot < 0.0 ? -1.0 : 1.0;
            for (int r = 0; r < N; r++) basis[r * kComponents + tid] = (float)(sg * t8.X[tid][r]);
        }
        if (!eig_full) return;
    }
    jacobi_sweeps(A0, V, Ab, Vb, kJacobiTol, 40, energy);
    __syncthreads();
    const double* Ac = A0;

    if (tid < N) {
        const double di = Ac[tid * LD + tid];
        int rank = 0;
        for (int j = 0; j < N; j++) {
            const double dj = Ac[j * LD + j];
            rank += (dj > di || (dj == di && j < tid)) ? 1 : 0;
        }
        if (eigvals && !fast) eigvals[rank] = di;
        double sg = 1.0;
        if (rank < kComponents) {
            double dot = 0.0;
            for (int r = 0; r < N; r++) dot += V[r * LD + tid] * sign_ref(r);
            sg = dot < 0.0 ? -1.0 : 1.0;
            if (!fast)
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
                                                   double* eig_full, double* tri) {
    pca_solve_body(moments, n, basis, eigvals, eig_full, tri);
}

// Eigenvalues 9..25 of a level whose solve kept the tridiagonal form and left
// them NaN: the multisection of the all-eigenvalue solve on the same T.
// Nothing to do when eigvals[8] is a number (the all-eigenvalue solve or the
// Jacobi sweeps wrote all 25).
__global__ __launch_bounds__(256) void k_tridiag_eigvals(const double* __restrict__ tri, double* __restrict__ eigvals) {
    constexpr int N = kPatch;
    if (!isnan(eigvals[kComponents])) return;
    __shared__ double d[N], e[N];
    __shared__ double2 de2[N];
    __shared__ double lohi[2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < N) d[tid] = tri[tid];
    if (tid < N - 1) e[tid] = tri[N + tid];
    __syncthreads();
    if (tid == 0) tridiag_bracket(d, e, de2, lohi[0], lohi[1]);
    __syncthreads();
    double dd[N], ee[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        dd[i] = de2[i].x;
        ee[i] = de2[i].y;
    }
    const int j = wave * 8 + (lane >> 3), m = lane & 7;
    const double lam = multisect<8>(dd, ee, lohi[0], lohi[1], j, m, 17);
    if (m == 0 && j >= kComponents && j < N) eigvals[j] = lam * tri[2 * N - 1];
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
    pca_solve_body(f.partials + (size_t)kMomentBlocks * kMoments, f.n, f.basis, f.eigvals, nullptr, f.tri);
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
                             double* eig_full, double* tri, hipStream_t s) {
    double* mom = partials + (size_t)kMomentBlocks * kMoments;
    k_moments_reduce<<<kMoments / kRedEntries, 256, 0, s>>>(partials, nblocks, mom);
    k_pca_solve<<<1, 256, 0, s>>>(mom, (double)n, basis, eigvals, eig_full, tri);
    GCAM_LAUNCH_CHECK("PCA eigensolve");
    return GCAM_OK;
}

gcam_status launch_tridiag_eigvals(const double* tri, double* eigvals, hipStream_t s) {
    k_tridiag_eigvals<<<1, 256, 0, s>>>(tri, eigvals);
    GCAM_LAUNCH_CHECK("PCA eigenvalues from the tridiagonal form");
    return GCAM_OK;
}

gcam_status launch_burst_pca_solve(const BurstTable<BurstMomFrame>* dev, int nframes, hipStream_t s) {
    if (nframes < 1) return GCAM_OK;
    k_burst_moments_reduce<<<dim3(kMoments / kRedEntries, nframes), 256, 0, s>>>(dev);
    k_burst_pca_solve<<<nframes, 256, 0, s>>>(dev);
    GCAM_LAUNCH_CHECK("burst PCA basis");
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
