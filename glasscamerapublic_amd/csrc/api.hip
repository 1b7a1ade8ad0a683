// C ABI of the gcam pyramid denoise: argument checking, the per-level buffers
// of a PyramidProcessor handle and the stream-ordered orchestration of
// PyramidProcessor<levels>::denoise (MetalPipeline/pyramid_processor.cpp:96-172).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "gcam_internal.hpp"
#include "handle.hpp"

namespace gcam {

namespace {
thread_local std::string g_last_error;
}

gcam_status set_error(gcam_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

}  // namespace gcam

using namespace gcam;

namespace {

constexpr int kMaxDim = 16384;

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

gcam_status check_image(const gcam_image* im, int fmt, const char* name, int w = -1, int h = -1) {
    if (!im) return set_error(GCAM_ERR_INVALID_ARGUMENT, std::string(name) + " is NULL");
    if (!im->data) return set_error(GCAM_ERR_INVALID_ARGUMENT, std::string(name) + ".data is NULL");
    if (im->format != fmt)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, std::string(name) + " has pixel format " +
                                                        std::to_string(im->format) + ", expected " + std::to_string(fmt));
    if (im->width < 1 || im->height < 1 || im->width > kMaxDim || im->height > kMaxDim)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, std::string(name) + " size out of range 1.." + std::to_string(kMaxDim));
    if (im->stride < im->width) return set_error(GCAM_ERR_INVALID_ARGUMENT, std::string(name) + ".stride < width");
    if ((w >= 0 && im->width != w) || (h >= 0 && im->height != h))
        return set_error(GCAM_ERR_INVALID_ARGUMENT, std::string(name) + " is " + std::to_string(im->width) + "x" +
                                                        std::to_string(im->height) + ", expected " + std::to_string(w) +
                                                        "x" + std::to_string(h));
    return GCAM_OK;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

gcam_image make_image(void* data, int w, int h, int fmt) { return gcam_image{data, w, h, w, fmt}; }

}  // namespace

extern "C" {

int32_t gcam_abi_version(void) { return GCAM_ABI_VERSION; }

const char* gcam_status_string(gcam_status status) {
    switch (status) {
        case GCAM_OK: return "ok";
        case GCAM_ERR_INVALID_ARGUMENT: return "invalid argument";
        case GCAM_ERR_HIP: return "HIP runtime error";
        case GCAM_ERR_OUT_OF_MEMORY: return "out of device memory";
        case GCAM_ERR_UNSUPPORTED: return "unsupported";
        case GCAM_ERR_NO_DEVICE: return "no HIP device";
    }
    return "unknown status";
}

const char* gcam_last_error(void) { return g_last_error.c_str(); }

gcam_status gcam_pyramid_create(int32_t width, int32_t height, int32_t levels, const gcam_config* config,
                                gcam_pyramid** out) {
    if (!out) return set_error(GCAM_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (levels < 1 || levels > GCAM_MAX_LEVELS)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "levels must be 1.." + std::to_string(GCAM_MAX_LEVELS));
    if (width < 1 || height < 1 || width > kMaxDim || height > kMaxDim)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "image size out of range 1.." + std::to_string(kMaxDim));
    gcam_config cfg{10, GCAM_PRECISION_FP32, 0, 0, {0, 0, 0, 0}};
    if (config) cfg = *config;
    if (cfg.bilateral != 0 && cfg.bilateral != 1) return set_error(GCAM_ERR_INVALID_ARGUMENT, "bilateral must be 0 or 1");
    if (cfg.search_radius < 1 || cfg.search_radius > GCAM_MAX_SEARCH_RADIUS)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "search_radius must be 1..10");
    if (cfg.precision != GCAM_PRECISION_FP32 && cfg.precision != GCAM_PRECISION_FP16 &&
        cfg.precision != GCAM_PRECISION_BF16)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "unknown precision " + std::to_string(cfg.precision));
    if (cfg.bm_rows_per_thread != 0 && cfg.bm_rows_per_thread != 1 && cfg.bm_rows_per_thread != 2 &&
        cfg.bm_rows_per_thread != 4)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "bm_rows_per_thread must be 0, 1, 2 or 4");
    for (int i = 0; i < 4; i++)
        if (cfg.reserved[i] != 0) return set_error(GCAM_ERR_INVALID_ARGUMENT, "gcam_config.reserved must be zero");
    for (int i = 0; i < levels; i++) {
        const long w = width >> i, h = height >> i;
        if (w * h / 64 < 2)
            return set_error(GCAM_ERR_INVALID_ARGUMENT,
                             "pyramid level " + std::to_string(i) + " (" + std::to_string(w) + "x" + std::to_string(h) +
                                 ") is too small to sample two 5x5 PCA patches");
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return set_error(GCAM_ERR_NO_DEVICE, "no HIP device");

    auto* p = new gcam_pyramid();
    p->width = width;
    p->height = height;
    p->levels = levels;
    p->radius = cfg.search_radius;
    p->precision = cfg.precision;
    p->bm_rows = cfg.bm_rows_per_thread;
    p->bilateral = cfg.bilateral != 0;
    p->device = dev;
    for (int i = 0; i < levels; i++) {
        p->W[i] = width >> i;
        p->H[i] = height >> i;
    }
    // Carve every fixed-size buffer out of one arena.
    std::vector<std::pair<void**, size_t>> parts;
    for (int i = 1; i < levels; i++) parts.push_back({(void**)&p->layer[i], sizeof(float4) * p->W[i] * p->H[i]});
    for (int i = 0; i < levels - 1; i++) parts.push_back({(void**)&p->sub[i], sizeof(float4) * p->W[i] * p->H[i]});
    for (int i = 0; i < levels; i++) parts.push_back({(void**)&p->den[i], sizeof(float4) * p->W[i] * p->H[i]});
    // Materialized features: only the FP32 PCA form reads them (the precision
    // form projects inside the block-matching tiles).
    if (cfg.precision == GCAM_PRECISION_FP32 && !cfg.bilateral)
        parts.push_back({(void**)&p->feat, sizeof(float) * 8 * (size_t)width * height});
    parts.push_back({(void**)&p->moment_partials, sizeof(double) * kMomentSlots * kMoments});
    parts.push_back({(void**)&p->nlf_partials, sizeof(double) * kNlfBlocks * 13});
    parts.push_back({(void**)&p->nlf_state, sizeof(NlfFitState)});
    parts.push_back({(void**)&p->nlf_dev, sizeof(NlfDev) * GCAM_MAX_LEVELS});
    parts.push_back({(void**)&p->basis, sizeof(float) * 200 * levels});
    parts.push_back({(void**)&p->eigvals, sizeof(double) * 25 * levels});
    parts.push_back({(void**)&p->tri, sizeof(double) * 50 * levels});
    size_t total = 0;
    for (auto& pr : parts) total += align256(pr.second);
    hipError_t e = hipMalloc((void**)&p->arena, total);
    if (e != hipSuccess) {
        delete p;
        return set_error(e == hipErrorOutOfMemory ? GCAM_ERR_OUT_OF_MEMORY : GCAM_ERR_HIP,
                         std::string("hipMalloc(") + std::to_string(total) + "): " + hipGetErrorString(e));
    }
    size_t off = 0;
    for (auto& pr : parts) {
        *pr.first = p->arena + off;
        off += align256(pr.second);
    }
    *out = p;
    return GCAM_OK;
}

void gcam_pyramid_destroy(gcam_pyramid* p) {
    if (!p) return;
    DeviceGuard g(p->device);
    for (auto& lv : p->ev)
        for (auto& st : lv)
            for (auto& e : st)
                if (e) (void)hipEventDestroy(e);
    if (p->arena) (void)hipFree(p->arena);
    if (p->grad_arena) (void)hipFree(p->grad_arena);
    if (p->adj_arena) (void)hipFree(p->adj_arena);
    delete p;
}

// The orchestration behind gcam_pyramid_denoise and gcam_pyramid_denoise_fwd.
// thr (nullable): per-level (Y, Cb, Cr) thresholds replacing (luma, chroma,
// chroma); keep_eig: keep each level's full eigen-decomposition.
static gcam_status denoise_impl(gcam_pyramid* p, const gcam_denoise_params* params, const float (*thr)[3],
                                const gcam_image* image, const gcam_image* gradient, gcam_ycbcr_nlf* nlf,
                                float exposure_multiplier, float lens_shading_correction, int32_t calibrate_from_image,
                                void* stream, gcam_image* out, bool keep_eig) {
    if (!p) return set_error(GCAM_ERR_INVALID_ARGUMENT, "pyramid is NULL");
    if (!params) return set_error(GCAM_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!nlf) return set_error(GCAM_ERR_INVALID_ARGUMENT, "nlf is NULL");
    gcam_status st = check_image(image, GCAM_PIXEL_RGBA32F, "image", p->width, p->height);
    if (st != GCAM_OK) return st;
    if (gradient) {
        st = check_image(gradient, GCAM_PIXEL_RG32F, "gradient", p->width, p->height);
        if (st != GCAM_OK) return st;
    }
    for (int i = 0; i < p->levels; i++) {
        const gcam_denoise_params& q = params[i];
        if (!(q.luma > 0.0f) || !(q.chroma > 0.0f) || !(q.chromaBoost > 0.0f))
            return set_error(GCAM_ERR_INVALID_ARGUMENT,
                             "level " + std::to_string(i) + ": luma, chroma and chromaBoost must be > 0");
        if (thr && !(thr[i][0] > 0.0f && thr[i][1] > 0.0f && thr[i][2] > 0.0f))
            return set_error(GCAM_ERR_INVALID_ARGUMENT, "level " + std::to_string(i) + ": thresholds must be > 0");
    }
    p->fwd_ready = false;
    DeviceGuard guard(p->device);
    hipStream_t s = as_stream(stream);
    const int L = p->levels;
    std::memset(p->ran, 0, sizeof(p->ran));
    auto mark = [&](int level, int stage, int end) -> gcam_status {
        if (!p->profiling) return GCAM_OK;
        GCAM_HIP_TRY(hipEventRecord(p->ev[level][stage][end], s));
        if (end) p->ran[level][stage] = true;
        return GCAM_OK;
    };
#define GCAM_MARK(level, stage, end)                   \
    do {                                               \
        gcam_status _m = mark(level, stage, end);      \
        if (_m != GCAM_OK) return _m;                  \
    } while (0)

    if (gradient && !p->grad_arena && L > 1) {
        size_t total = 0;
        for (int i = 1; i < L; i++) total += align256(sizeof(float2) * p->W[i] * p->H[i]);
        GCAM_HIP_TRY(hipMalloc((void**)&p->grad_arena, total));
        size_t off = 0;
        for (int i = 1; i < L; i++) {
            p->gradl[i] = reinterpret_cast<float2*>(reinterpret_cast<char*>(p->grad_arena) + off);
            off += align256(sizeof(float2) * p->W[i] * p->H[i]);
        }
    }
    if (!calibrate_from_image) {
        st = launch_set_nlf(nlf, L, p->nlf_dev, s);
        if (st != GCAM_OK) return st;
    }

    const View<const float4> in0 = view<const float4>(*image);
    const View<const float2> g0 = gradient ? view<const float2>(*gradient) : View<const float2>{nullptr, 0, 0, 0};
    auto noisy = [&](int i) { return i == 0 ? in0 : view<const float4>(p->layer[i], p->W[i], p->H[i]); };
    auto gradv = [&](int i) {
        if (!gradient) return View<const float2>{nullptr, 0, 0, 0};
        return i == 0 ? g0 : view<const float2>(p->gradl[i], p->W[i], p->H[i]);
    };

    // Loop A (pyramid_processor.cpp:103-120): Gaussian pyramid and noise model.
    for (int i = 0; i < L; i++) {
        if (i < L - 1) {
            GCAM_MARK(i, GCAM_STAGE_DOWNSAMPLE, 0);
            st = launch_downsample4(noisy(i), view<float4>(p->layer[i + 1], p->W[i + 1], p->H[i + 1]), s);
            if (st != GCAM_OK) return st;
            if (gradient) {
                st = launch_downsample2(gradv(i), view<float2>(p->gradl[i + 1], p->W[i + 1], p->H[i + 1]), s);
                if (st != GCAM_OK) return st;
            }
            GCAM_MARK(i, GCAM_STAGE_DOWNSAMPLE, 1);
        }
        if (calibrate_from_image) {  // statistics go to denoisedImagePyramid[i] as scratch (:116)
            GCAM_MARK(i, GCAM_STAGE_NLF, 0);
            const View<float4> stats = view<float4>(p->den[i], p->W[i], p->H[i]);
            st = launch_noise_stats(noisy(i), stats, s);
            if (st != GCAM_OK) return st;
            st = launch_nlf_fit(View<const float4>{stats.ptr, stats.w, stats.h, stats.stride}, exposure_multiplier,
                                p->nlf_partials, p->nlf_state, p->nlf_dev + i, s);
            if (st != GCAM_OK) return st;
            GCAM_MARK(i, GCAM_STAGE_NLF, 1);
        }
    }

    // Loop B (pyramid_processor.cpp:123-169): coarse to fine.
    for (int i = L - 1; i >= 0; i--) {
        View<const float4> layer = noisy(i);
        int sub_blocks = 0;
        if (i < L - 1) {
            GCAM_MARK(i, GCAM_STAGE_SUBTRACT, 0);
            const View<float4> sv = view<float4>(p->sub[i], p->W[i], p->H[i]);
            const View<const float4> den1 = view<const float4>(p->den[i + 1], p->W[i + 1], p->H[i + 1]);
            // With patch similarity the subtract also forms the level's PCA
            // moment partials (the sample is never written).
            if (p->bilateral)
                st = launch_subtract_noise(noisy(i), noisy(i + 1), den1, gradv(i), /*luma_weight=*/1.0f,
                                           params[i].sharpening, 0.0f, 0.0f, p->nlf_dev + i, params[i].luma, sv, s);
            else
                st = launch_subtract_moments(noisy(i), noisy(i + 1), den1, gradv(i), /*luma_weight=*/1.0f,
                                             params[i].sharpening, 0.0f, 0.0f, p->nlf_dev + i, params[i].luma, sv,
                                             p->moment_partials, &sub_blocks, s);
            if (st != GCAM_OK) return st;
            layer = View<const float4>{sv.ptr, sv.w, sv.h, sv.stride};
            GCAM_MARK(i, GCAM_STAGE_SUBTRACT, 1);
        }
        BmParams bp;
        bp.thr[0] = thr ? thr[i][0] : params[i].luma;  // nflMultiplier (:52-56)
        bp.thr[1] = thr ? thr[i][1] : params[i].chroma;
        bp.thr[2] = thr ? thr[i][2] : params[i].chroma;
        bp.chromaBoost = params[i].chromaBoost;
        bp.gradientBoost = params[i].gradientBoost;
        bp.gradientThreshold = params[i].gradientThreshold;
        bp.lensShading = lens_shading_correction;
        if (p->bilateral) {  // usePatchSimiliarity == false (:162-168)
            GCAM_MARK(i, GCAM_STAGE_BLOCK_MATCH, 0);
            st = launch_denoise_image(layer, gradv(i), p->nlf_dev + i, bp, view<float4>(p->den[i], p->W[i], p->H[i]), s);
            if (st != GCAM_OK) return st;
            GCAM_MARK(i, GCAM_STAGE_BLOCK_MATCH, 1);
            continue;
        }
        GCAM_MARK(i, GCAM_STAGE_PCA, 0);
        const int64_t n = ((int64_t)p->W[i] * p->H[i]) / 64;  // :142
        // Levels below the top: their subtract wrote the moment partials.
        int nblocks = sub_blocks;
        GCAM_MARK(i, GCAM_STAGE_MOMENTS, 0);
        if (i == L - 1) st = launch_image_moments(layer, p->moment_partials, &nblocks, s);
        if (st != GCAM_OK) return st;
        GCAM_MARK(i, GCAM_STAGE_MOMENTS, 1);
        st = launch_pca_solve(p->moment_partials, nblocks, n, p->basis + 200 * i, p->eigvals + 25 * i,
                              keep_eig ? p->eig_full + kEigStride * i : nullptr, p->tri + 50 * i, s);
        if (st != GCAM_OK) return st;
        const View<float> fv = view<float>(p->feat, p->W[i], p->H[i]);
        if (p->precision == GCAM_PRECISION_FP32) {
            st = launch_pca_project(layer, p->basis + 200 * i, fv, s);
            if (st != GCAM_OK) return st;
        }
        GCAM_MARK(i, GCAM_STAGE_PCA, 1);
        GCAM_MARK(i, GCAM_STAGE_BLOCK_MATCH, 0);
        if (p->precision == GCAM_PRECISION_FP32)
            st = launch_block_match(layer, gradv(i), View<const float>{fv.ptr, fv.w, fv.h, fv.stride}, p->nlf_dev + i,
                                    bp, p->radius, p->bm_rows, view<float4>(p->den[i], p->W[i], p->H[i]), s);
        else
            st = launch_block_match_lp(layer, gradv(i), p->basis + 200 * i, p->nlf_dev + i, bp, p->radius, p->bm_rows,
                                       p->precision, view<float4>(p->den[i], p->W[i], p->H[i]), s);
        if (st != GCAM_OK) return st;
        GCAM_MARK(i, GCAM_STAGE_BLOCK_MATCH, 1);
    }
#undef GCAM_MARK
    p->feature_level = (p->precision == GCAM_PRECISION_FP32 && !p->bilateral) ? 0 : -1;

    if (calibrate_from_image) {
        NlfDev host[GCAM_MAX_LEVELS];
        GCAM_HIP_TRY(hipMemcpyAsync(host, p->nlf_dev, sizeof(NlfDev) * L, hipMemcpyDeviceToHost, s));
        GCAM_HIP_TRY(hipStreamSynchronize(s));
        for (int i = 0; i < L; i++)
            for (int c = 0; c < 3; c++) {
                nlf[i].a[c] = host[i].a[c];
                nlf[i].b[c] = host[i].b[c];
            }
    }
    if (out) *out = make_image(p->den[0], p->W[0], p->H[0], GCAM_PIXEL_RGBA32F);
    return GCAM_OK;
}

gcam_status gcam_pyramid_denoise(gcam_pyramid* p, const gcam_denoise_params* params, const gcam_image* image,
                                 const gcam_image* gradient, gcam_ycbcr_nlf* nlf, float exposure_multiplier,
                                 float lens_shading_correction, int32_t calibrate_from_image, void* stream,
                                 gcam_image* out) {
    return denoise_impl(p, params, nullptr, image, gradient, nlf, exposure_multiplier, lens_shading_correction,
                        calibrate_from_image, stream, out, false);
}

static gcam_status alloc_adjoint(gcam_pyramid* p) {
    if (p->adj_arena) return GCAM_OK;
    const size_t n0 = (size_t)p->W[0] * p->H[0];
    const size_t n1 = p->levels > 1 ? (size_t)p->W[1] * p->H[1] : 1;
    std::vector<std::pair<void**, size_t>> parts = {
        {(void**)&p->eig_full, sizeof(double) * kEigStride * p->levels},
        {(void**)&p->adj_coef, sizeof(float4) * 3 * n0},
        {(void**)&p->adj_gS, sizeof(float4) * n0},
        {(void**)&p->adj_gF, sizeof(float4) * 2 * n0},
        {(void**)&p->adj_gD[0], sizeof(float4) * n1},
        {(void**)&p->adj_gD[1], sizeof(float4) * n1},
        {(void**)&p->adj_gthr_part, sizeof(float) * 3 * adjoint_thr_blocks(p->W[0], p->H[0])},
        {(void**)&p->adj_gthr, sizeof(float) * 3 * p->levels},
        {(void**)&p->adj_gB_part, sizeof(double) * kBasisGradBlocks * kPatch * kComponents},
        {(void**)&p->adj_gC, sizeof(double) * 625},
    };
    size_t total = 0;
    for (auto& pr : parts) total += align256(pr.second);
    GCAM_HIP_TRY(hipMalloc((void**)&p->adj_arena, total));
    size_t off = 0;
    for (auto& pr : parts) {
        *pr.first = p->adj_arena + off;
        off += align256(pr.second);
    }
    return GCAM_OK;
}

gcam_status gcam_pyramid_denoise_fwd(gcam_pyramid* p, const gcam_denoise_params* params, const float* thresholds,
                                     const gcam_image* image, const gcam_ycbcr_nlf* nlf, void* stream,
                                     gcam_image* out) {
    if (!p || !params || !thresholds || !nlf) return set_error(GCAM_ERR_INVALID_ARGUMENT, "NULL argument");
    if (p->precision != GCAM_PRECISION_FP32 || p->bilateral)
        return set_error(GCAM_ERR_UNSUPPORTED, "the differentiable form runs the FP32 PCA path");
    for (int i = 0; i < p->levels; i++)
        if (params[i].gradientBoost != 0.0f || params[i].sharpening > 1.0f)
            return set_error(GCAM_ERR_UNSUPPORTED,
                             "the differentiable form needs gradientBoost == 0 and sharpening <= 1 on every level");
    DeviceGuard guard(p->device);
    gcam_status st = alloc_adjoint(p);
    if (st != GCAM_OK) return st;
    gcam_ycbcr_nlf nl[GCAM_MAX_LEVELS];
    std::memcpy(nl, nlf, sizeof(gcam_ycbcr_nlf) * p->levels);
    const float(*thr)[3] = reinterpret_cast<const float(*)[3]>(thresholds);
    st = denoise_impl(p, params, thr, image, nullptr, nl, 1.0f, 0.0f, 0, stream, out, true);
    if (st != GCAM_OK) return st;
    for (int i = 0; i < p->levels; i++) {
        for (int c = 0; c < 3; c++) p->fwd_thr[i][c] = thr[i][c];
        p->fwd_cb[i] = params[i].chromaBoost;
    }
    p->fwd_image = *image;
    p->fwd_ready = true;
    return GCAM_OK;
}

gcam_status gcam_pyramid_denoise_bwd(gcam_pyramid* p, const gcam_image* grad_output, float* grad_thresholds,
                                     void* stream) {
    if (!p || !grad_thresholds) return set_error(GCAM_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!p->fwd_ready) return set_error(GCAM_ERR_INVALID_ARGUMENT, "no gcam_pyramid_denoise_fwd to differentiate");
    gcam_status st = check_image(grad_output, GCAM_PIXEL_RGBA32F, "grad_output", p->width, p->height);
    if (st != GCAM_OK) return st;
    DeviceGuard guard(p->device);
    hipStream_t s = as_stream(stream);
    const int L = p->levels;
    for (int i = 0; i < L; i++) {
        const int w = p->W[i], h = p->H[i];
        View<const float4> S;
        if (i < L - 1)
            S = view<const float4>(p->sub[i], w, h);
        else if (i > 0)
            S = view<const float4>(p->layer[i], w, h);
        else
            S = view<const float4>(p->fwd_image);
        st = launch_pca_project(S, p->basis + 200 * i, view<float>(p->feat, w, h), s);  // features recomputed
        if (st != GCAM_OK) return st;
        AdjointLevel a;
        a.S = S;
        a.D = view<const float4>(p->den[i], w, h);
        a.gD = i == 0 ? view<const float4>(*grad_output) : view<const float4>(p->adj_gD[(i - 1) % 2], w, h);
        a.feat = p->feat;
        a.basis = p->basis + 200 * i;
        a.eig = p->eig_full + kEigStride * i;
        a.nlf = p->nlf_dev + i;
        for (int c = 0; c < 3; c++) a.thr[c] = p->fwd_thr[i][c];
        a.chromaBoost = p->fwd_cb[i];
        a.radius = p->radius;
        a.n = (double)(((int64_t)w * h) / 64);
        a.coef = p->adj_coef;
        a.gS = p->adj_gS;
        a.gF = p->adj_gF;
        a.gthr_part = p->adj_gthr_part;
        a.gthr = p->adj_gthr + 3 * i;
        a.gB_part = p->adj_gB_part;
        a.gC = p->adj_gC;
        a.gD1 = i < L - 1 ? p->adj_gD[i % 2] : nullptr;
        a.w1 = i < L - 1 ? p->W[i + 1] : 0;
        a.h1 = i < L - 1 ? p->H[i + 1] : 0;
        st = launch_level_backward(a, s);
        if (st != GCAM_OK) return st;
    }
    GCAM_HIP_TRY(hipMemcpyAsync(grad_thresholds, p->adj_gthr, sizeof(float) * 3 * L, hipMemcpyDeviceToHost, s));
    GCAM_HIP_TRY(hipStreamSynchronize(s));
    return GCAM_OK;
}

gcam_status gcam_pyramid_set_profiling(gcam_pyramid* p, int32_t enable) {
    if (!p) return set_error(GCAM_ERR_INVALID_ARGUMENT, "pyramid is NULL");
    DeviceGuard g(p->device);
    if (enable && !p->ev[0][0][0]) {
        for (auto& lv : p->ev)
            for (auto& st : lv)
                for (auto& e : st) GCAM_HIP_TRY(hipEventCreate(&e));
    }
    p->profiling = enable != 0;
    return GCAM_OK;
}

gcam_status gcam_pyramid_stage_times(const gcam_pyramid* p, float* ms) {
    if (!p || !ms) return set_error(GCAM_ERR_INVALID_ARGUMENT, "NULL argument");
    DeviceGuard g(p->device);
    for (int i = 0; i < p->levels; i++)
        for (int k = 0; k < GCAM_STAGE_COUNT; k++) {
            float v = 0.0f;
            if (p->profiling && p->ran[i][k]) {
                GCAM_HIP_TRY(hipEventSynchronize(p->ev[i][k][1]));
                GCAM_HIP_TRY(hipEventElapsedTime(&v, p->ev[i][k][0], p->ev[i][k][1]));
            }
            ms[i * GCAM_STAGE_COUNT + k] = v;
        }
    return GCAM_OK;
}

gcam_status gcam_pyramid_level(const gcam_pyramid* p, int32_t level, int32_t which, gcam_image* out) {
    if (!p || !out) return set_error(GCAM_ERR_INVALID_ARGUMENT, "NULL argument");
    if (level < 0 || level >= p->levels) return set_error(GCAM_ERR_INVALID_ARGUMENT, "level out of range");
    const int w = p->W[level], h = p->H[level];
    switch (which) {
        case GCAM_LEVEL_DENOISED: *out = make_image(p->den[level], w, h, GCAM_PIXEL_RGBA32F); return GCAM_OK;
        case GCAM_LEVEL_SUBTRACTED:
            if (level == p->levels - 1)
                return set_error(GCAM_ERR_INVALID_ARGUMENT, "the coarsest level has no subtracted image");
            *out = make_image(p->sub[level], w, h, GCAM_PIXEL_RGBA32F);
            return GCAM_OK;
        case GCAM_LEVEL_NOISY:
            if (level == 0) return set_error(GCAM_ERR_INVALID_ARGUMENT, "level 0 is the caller's input image");
            *out = make_image(p->layer[level], w, h, GCAM_PIXEL_RGBA32F);
            return GCAM_OK;
        case GCAM_LEVEL_FEATURES:
            if (level != p->feature_level)
                return set_error(GCAM_ERR_INVALID_ARGUMENT, "features are kept for the last denoised level only");
            *out = make_image(p->feat, w, h, GCAM_PIXEL_PCA8_F32);
            return GCAM_OK;
    }
    return set_error(GCAM_ERR_INVALID_ARGUMENT, "unknown level image");
}

gcam_status gcam_pyramid_basis(const gcam_pyramid* p, int32_t level, float* basis, double* eigenvalues) {
    if (!p) return set_error(GCAM_ERR_INVALID_ARGUMENT, "pyramid is NULL");
    if (level < 0 || level >= p->levels) return set_error(GCAM_ERR_INVALID_ARGUMENT, "level out of range");
    DeviceGuard g(p->device);
    GCAM_HIP_TRY(hipDeviceSynchronize());
    if (basis) GCAM_HIP_TRY(hipMemcpy(basis, p->basis + 200 * level, sizeof(float) * 200, hipMemcpyDeviceToHost));
    if (eigenvalues) {
        // The denoise's solve brackets only the 8 leading eigenvalues; the rest
        // come from the tridiagonal form it kept.
        const gcam_status st = launch_tridiag_eigvals(p->tri + 50 * level, p->eigvals + 25 * level, nullptr);
        if (st != GCAM_OK) return st;
        GCAM_HIP_TRY(hipDeviceSynchronize());
        GCAM_HIP_TRY(hipMemcpy(eigenvalues, p->eigvals + 25 * level, sizeof(double) * 25, hipMemcpyDeviceToHost));
    }
    return GCAM_OK;
}

gcam_status gcam_downsample_image(const gcam_image* input, gcam_image* output, void* stream) {
    if (!input || !output) return set_error(GCAM_ERR_INVALID_ARGUMENT, "NULL image");
    if (input->format == GCAM_PIXEL_RGBA32F) {
        gcam_status st = check_image(input, GCAM_PIXEL_RGBA32F, "input");
        if (st == GCAM_OK) st = check_image(output, GCAM_PIXEL_RGBA32F, "output");
        if (st != GCAM_OK) return st;
        return launch_downsample4(view<const float4>(*input), view<float4>(*output), as_stream(stream));
    }
    if (input->format == GCAM_PIXEL_RG32F) {
        gcam_status st = check_image(input, GCAM_PIXEL_RG32F, "input");
        if (st == GCAM_OK) st = check_image(output, GCAM_PIXEL_RG32F, "output");
        if (st != GCAM_OK) return st;
        return launch_downsample2(view<const float2>(*input), view<float2>(*output), as_stream(stream));
    }
    return set_error(GCAM_ERR_INVALID_ARGUMENT, "downsample takes RGBA32F or RG32F images");
}

gcam_status gcam_subtract_noise_image(const gcam_image* input, const gcam_image* input1,
                                      const gcam_image* input_denoised1, const gcam_image* gradient,
                                      float luma_weight, float sharpening, float nlf_x, float nlf_y,
                                      gcam_image* output, void* stream) {
    gcam_status st = check_image(input, GCAM_PIXEL_RGBA32F, "input");
    if (st == GCAM_OK) st = check_image(input1, GCAM_PIXEL_RGBA32F, "input1");
    if (st == GCAM_OK) st = check_image(input_denoised1, GCAM_PIXEL_RGBA32F, "input_denoised1", input1->width, input1->height);
    if (st == GCAM_OK) st = check_image(output, GCAM_PIXEL_RGBA32F, "output", input->width, input->height);
    if (st == GCAM_OK && gradient) st = check_image(gradient, GCAM_PIXEL_RG32F, "gradient", input->width, input->height);
    if (st != GCAM_OK) return st;
    return launch_subtract_noise(view<const float4>(*input), view<const float4>(*input1),
                                 view<const float4>(*input_denoised1),
                                 gradient ? view<const float2>(*gradient) : View<const float2>{nullptr, 0, 0, 0},
                                 luma_weight, sharpening, nlf_x, nlf_y, nullptr, 0.0f, view<float4>(*output),
                                 as_stream(stream));
}

gcam_status gcam_collect_patches(const gcam_image* input, float* patches, int64_t capacity, int64_t* count,
                                 void* stream) {
    gcam_status st = check_image(input, GCAM_PIXEL_RGBA32F, "input");
    if (st != GCAM_OK) return st;
    if (!patches) return set_error(GCAM_ERR_INVALID_ARGUMENT, "patches is NULL");
    const int64_t n = ((int64_t)input->width * input->height) / 64;
    if (capacity < n)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "patch buffer holds " + std::to_string(capacity) +
                                                        " patches, the image needs " + std::to_string(n));
    if (count) *count = n;
    if (n == 0) return GCAM_OK;
    return launch_collect_patches(view<const float4>(*input), patches, n, as_stream(stream));
}

gcam_status gcam_build_pca_space(const float* patches, int64_t n, float* basis, double* eigenvalues, void* stream) {
    if (!patches || !basis) return set_error(GCAM_ERR_INVALID_ARGUMENT, "NULL pointer");
    if (n < 2) return set_error(GCAM_ERR_INVALID_ARGUMENT, "build_pca_space needs at least 2 patches");
    hipStream_t s = as_stream(stream);
    double* partials = nullptr;
    GCAM_HIP_TRY(hipMallocAsync((void**)&partials, sizeof(double) * kMomentSlots * kMoments, s));
    int nblocks = 0;
    gcam_status st = launch_buffer_moments(patches, n, partials, &nblocks, s);
    if (st == GCAM_OK) st = launch_pca_solve(partials, nblocks, n, basis, eigenvalues, nullptr, nullptr, s);
    GCAM_HIP_TRY(hipFreeAsync(partials, s));
    return st;
}

gcam_status gcam_pca_projection(const gcam_image* input, const float* basis, gcam_image* features, void* stream) {
    gcam_status st = check_image(input, GCAM_PIXEL_RGBA32F, "input");
    if (st == GCAM_OK) st = check_image(features, GCAM_PIXEL_PCA8_F32, "features", input->width, input->height);
    if (st != GCAM_OK) return st;
    if (!basis) return set_error(GCAM_ERR_INVALID_ARGUMENT, "basis is NULL");
    return launch_pca_project(view<const float4>(*input), basis,
                              View<float>{reinterpret_cast<float*>(features->data), features->width, features->height,
                                          features->stride},
                              as_stream(stream));
}

gcam_status gcam_block_matching_denoise_image(const gcam_image* input, const gcam_image* gradient,
                                              const gcam_image* features, const float var_a[3],
                                              const float var_b[3], const float threshold_multipliers[3],
                                              float chroma_boost, float gradient_boost, float gradient_threshold,
                                              float lens_shading_correction, int32_t search_radius,
                                              gcam_image* output, void* stream) {
    gcam_status st = check_image(input, GCAM_PIXEL_RGBA32F, "input");
    if (st == GCAM_OK) st = check_image(features, GCAM_PIXEL_PCA8_F32, "features", input->width, input->height);
    if (st == GCAM_OK) st = check_image(output, GCAM_PIXEL_RGBA32F, "output", input->width, input->height);
    if (st == GCAM_OK && gradient) st = check_image(gradient, GCAM_PIXEL_RG32F, "gradient", input->width, input->height);
    if (st != GCAM_OK) return st;
    if (!var_a || !var_b || !threshold_multipliers) return set_error(GCAM_ERR_INVALID_ARGUMENT, "NULL parameter");
    if (search_radius < 1 || search_radius > GCAM_MAX_SEARCH_RADIUS)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "search radius must be 1..10");
    hipStream_t s = as_stream(stream);
    NlfDev* nd = nullptr;
    GCAM_HIP_TRY(hipMallocAsync((void**)&nd, sizeof(NlfDev), s));
    gcam_ycbcr_nlf hn;
    for (int c = 0; c < 3; c++) {
        hn.a[c] = var_a[c];
        hn.b[c] = var_b[c];
    }
    st = launch_set_nlf(&hn, 1, nd, s);
    BmParams bp;
    for (int c = 0; c < 3; c++) bp.thr[c] = threshold_multipliers[c];
    bp.chromaBoost = chroma_boost;
    bp.gradientBoost = gradient_boost;
    bp.gradientThreshold = gradient_threshold;
    bp.lensShading = lens_shading_correction;
    if (st == GCAM_OK)
        st = launch_block_match(view<const float4>(*input),
                                gradient ? view<const float2>(*gradient) : View<const float2>{nullptr, 0, 0, 0},
                                View<const float>{reinterpret_cast<const float*>(features->data), features->width,
                                                  features->height, features->stride},
                                nd, bp, search_radius, 0, view<float4>(*output), s);
    GCAM_HIP_TRY(hipFreeAsync(nd, s));
    return st;
}

gcam_status gcam_block_matching_denoise_image_lp(const gcam_image* input, const gcam_image* gradient,
                                                 const float* basis, const float var_a[3], const float var_b[3],
                                                 const float threshold_multipliers[3], float chroma_boost,
                                                 float gradient_boost, float gradient_threshold,
                                                 float lens_shading_correction, int32_t search_radius,
                                                 int32_t precision, gcam_image* output, void* stream) {
    gcam_status st = check_image(input, GCAM_PIXEL_RGBA32F, "input");
    if (st == GCAM_OK) st = check_image(output, GCAM_PIXEL_RGBA32F, "output", input->width, input->height);
    if (st == GCAM_OK && gradient) st = check_image(gradient, GCAM_PIXEL_RG32F, "gradient", input->width, input->height);
    if (st != GCAM_OK) return st;
    if (!basis || !var_a || !var_b || !threshold_multipliers)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "NULL parameter");
    if (search_radius < 1 || search_radius > GCAM_MAX_SEARCH_RADIUS)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "search radius must be 1..10");
    if (precision != GCAM_PRECISION_FP16 && precision != GCAM_PRECISION_BF16)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "precision must be GCAM_PRECISION_FP16 or GCAM_PRECISION_BF16");
    hipStream_t s = as_stream(stream);
    NlfDev* nd = nullptr;
    GCAM_HIP_TRY(hipMallocAsync((void**)&nd, sizeof(NlfDev), s));
    gcam_ycbcr_nlf hn;
    for (int c = 0; c < 3; c++) {
        hn.a[c] = var_a[c];
        hn.b[c] = var_b[c];
    }
    st = launch_set_nlf(&hn, 1, nd, s);
    BmParams bp;
    for (int c = 0; c < 3; c++) bp.thr[c] = threshold_multipliers[c];
    bp.chromaBoost = chroma_boost;
    bp.gradientBoost = gradient_boost;
    bp.gradientThreshold = gradient_threshold;
    bp.lensShading = lens_shading_correction;
    if (st == GCAM_OK)
        st = launch_block_match_lp(view<const float4>(*input),
                                   gradient ? view<const float2>(*gradient) : View<const float2>{nullptr, 0, 0, 0},
                                   basis, nd, bp, search_radius, 0, precision, view<float4>(*output), s);
    GCAM_HIP_TRY(hipFreeAsync(nd, s));
    return st;
}

gcam_status gcam_denoise_image(const gcam_image* input, const gcam_image* gradient, const float var_a[3],
                               const float var_b[3], const float threshold_multipliers[3], float chroma_boost,
                               float gradient_boost, float gradient_threshold, gcam_image* output, void* stream) {
    gcam_status st = check_image(input, GCAM_PIXEL_RGBA32F, "input");
    if (st == GCAM_OK) st = check_image(output, GCAM_PIXEL_RGBA32F, "output", input->width, input->height);
    if (st == GCAM_OK && gradient) st = check_image(gradient, GCAM_PIXEL_RG32F, "gradient", input->width, input->height);
    if (st != GCAM_OK) return st;
    if (!var_a || !var_b || !threshold_multipliers) return set_error(GCAM_ERR_INVALID_ARGUMENT, "NULL parameter");
    hipStream_t s = as_stream(stream);
    NlfDev* nd = nullptr;
    GCAM_HIP_TRY(hipMallocAsync((void**)&nd, sizeof(NlfDev), s));
    gcam_ycbcr_nlf hn;
    for (int c = 0; c < 3; c++) {
        hn.a[c] = var_a[c];
        hn.b[c] = var_b[c];
    }
    st = launch_set_nlf(&hn, 1, nd, s);
    BmParams bp;
    for (int c = 0; c < 3; c++) bp.thr[c] = threshold_multipliers[c];
    bp.chromaBoost = chroma_boost;
    bp.gradientBoost = gradient_boost;
    bp.gradientThreshold = gradient_threshold;
    bp.lensShading = 0.0f;
    if (st == GCAM_OK)
        st = launch_denoise_image(view<const float4>(*input),
                                  gradient ? view<const float2>(*gradient) : View<const float2>{nullptr, 0, 0, 0}, nd,
                                  bp, view<float4>(*output), s);
    GCAM_HIP_TRY(hipFreeAsync(nd, s));
    return st;
}

gcam_status gcam_basic_noise_statistics(const gcam_image* input, gcam_image* statistics, void* stream) {
    gcam_status st = check_image(input, GCAM_PIXEL_RGBA32F, "input");
    if (st == GCAM_OK) st = check_image(statistics, GCAM_PIXEL_RGBA32F, "statistics", input->width, input->height);
    if (st != GCAM_OK) return st;
    return launch_noise_stats(view<const float4>(*input), view<float4>(*statistics), as_stream(stream));
}

gcam_status gcam_measure_ycbcr_nlf(const gcam_image* input, gcam_image* scratch, float exposure_multiplier,
                                   gcam_ycbcr_nlf* nlf, void* stream) {
    gcam_status st = check_image(input, GCAM_PIXEL_RGBA32F, "input");
    if (st == GCAM_OK) st = check_image(scratch, GCAM_PIXEL_RGBA32F, "scratch", input->width, input->height);
    if (st != GCAM_OK) return st;
    if (!nlf) return set_error(GCAM_ERR_INVALID_ARGUMENT, "nlf is NULL");
    hipStream_t s = as_stream(stream);
    char* buf = nullptr;
    const size_t part = align256(sizeof(double) * kNlfBlocks * 13), stsz = align256(sizeof(NlfFitState));
    GCAM_HIP_TRY(hipMallocAsync((void**)&buf, part + stsz + sizeof(NlfDev), s));
    st = launch_noise_stats(view<const float4>(*input), view<float4>(*scratch), s);
    if (st == GCAM_OK)
        st = launch_nlf_fit(view<const float4>(*scratch), exposure_multiplier, reinterpret_cast<double*>(buf),
                            reinterpret_cast<NlfFitState*>(buf + part), reinterpret_cast<NlfDev*>(buf + part + stsz), s);
    NlfDev host;
    if (st == GCAM_OK) {
        GCAM_HIP_TRY(hipMemcpyAsync(&host, buf + part + stsz, sizeof(NlfDev), hipMemcpyDeviceToHost, s));
    }
    GCAM_HIP_TRY(hipFreeAsync(buf, s));
    GCAM_HIP_TRY(hipStreamSynchronize(s));
    if (st != GCAM_OK) return st;
    for (int c = 0; c < 3; c++) {
        nlf->a[c] = host.a[c];
        nlf->b[c] = host.b[c];
    }
    return GCAM_OK;
}

gcam_status gcam_synth_frame(int32_t width, int32_t height, const gcam_synth_params* params, uint16_t* raw,
                             gcam_image* ycc, gcam_image* clean_ycc, void* stream) {
    if (!params) return set_error(GCAM_ERR_INVALID_ARGUMENT, "params is NULL");
    if (width < 2 || height < 2 || width > kMaxDim || height > kMaxDim)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "frame size out of range");
    if (!(params->period > 0.0f) || params->octaves < 1 || params->octaves > 16 || params->shot_gain < 0.0f ||
        params->read_sigma < 0.0f)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "bad synthesis parameters");
    gcam_status st;
    if (ycc && (st = check_image(ycc, GCAM_PIXEL_RGBA32F, "ycc", width, height)) != GCAM_OK) return st;
    if (clean_ycc && (st = check_image(clean_ycc, GCAM_PIXEL_RGBA32F, "clean_ycc", width, height)) != GCAM_OK)
        return st;
    hipStream_t s = as_stream(stream);
    uint16_t* tmp = nullptr;
    if (!raw) GCAM_HIP_TRY(hipMallocAsync((void**)&tmp, sizeof(uint16_t) * (size_t)width * height, s));
    st = launch_synth(width, height, *params, raw ? raw : tmp,
                      ycc ? view<float4>(*ycc) : View<float4>{nullptr, 0, 0, 0},
                      clean_ycc ? view<float4>(*clean_ycc) : View<float4>{nullptr, 0, 0, 0}, s);
    if (tmp) GCAM_HIP_TRY(hipFreeAsync(tmp, s));
    return st;
}

}  // extern "C"
