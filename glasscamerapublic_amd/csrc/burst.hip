// Batched (burst) pyramid denoise: many frames of uneven sizes, every stage of
// a level in one launch over all frames, the block matching as one persistent
// launch over the tiles of every frame, the whole call one hipGraph.
//
// The algorithm per frame is PyramidProcessor<levels>::denoise
// (MetalPipeline/pyramid_processor.cpp:96-172) with the noise model given
// (no calibration); the kernels are the single-frame ones behind frame
// tables (burst_tables.hpp).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "burst_tables.hpp"
#include "gcam_internal.hpp"
#include "handle.hpp"
#include "moments.hpp"

using namespace gcam;

namespace {

// Everything the captured kernels read, in one device buffer (and one
// pinned host staging copy).
struct BurstTables {
    NlfDev nlf[kMaxBurst][GCAM_MAX_LEVELS];
    BurstTable<BurstDownFrame> down[GCAM_MAX_LEVELS];
    BurstTable<BurstSubFrame> sub[GCAM_MAX_LEVELS];
    BurstTable<BurstMomFrame> mom[GCAM_MAX_LEVELS];
    BurstTable<BurstBmFrame> bm[GCAM_MAX_LEVELS];
    BmDirTable dir;  // constant; set once at creation
};

}  // namespace

struct gcam_burst {
    int n = 0, levels = 0, radius = 10, precision = GCAM_PRECISION_FP16, device = 0, cus = 256;
    std::vector<gcam_pyramid*> frames;
    BurstTables* dev = nullptr;   // device copy
    BurstTables* host = nullptr;  // pinned staging
    hipEvent_t staged = nullptr;  // completion of the last staging copy
    hipStream_t cap = nullptr;    // capture stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    gcam_burst_tuning tuning = {};
    int down_blocks[GCAM_MAX_LEVELS] = {}, sub_blocks[GCAM_MAX_LEVELS] = {}, mom_blocks[GCAM_MAX_LEVELS] = {};
    int bm_tiles[GCAM_MAX_LEVELS] = {};
};

namespace {

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

int blocks64x4(int w, int h) { return ((w + 63) / 64) * ((h + 3) / 4); }

void drop_graph(gcam_burst* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

gcam_status check_frame_image(const gcam_image* im, int fmt, const char* what, int f, int w, int h) {
    const std::string name = std::string(what) + "[" + std::to_string(f) + "]";
    if (!im->data) return set_error(GCAM_ERR_INVALID_ARGUMENT, name + ".data is NULL");
    if (im->format != fmt) return set_error(GCAM_ERR_INVALID_ARGUMENT, name + " has the wrong pixel format");
    if (im->width != w || im->height != h)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, name + " is " + std::to_string(im->width) + "x" +
                                                        std::to_string(im->height) + ", expected " +
                                                        std::to_string(w) + "x" + std::to_string(h));
    if (im->stride < im->width) return set_error(GCAM_ERR_INVALID_ARGUMENT, name + ".stride < width");
    return GCAM_OK;
}

gcam_status check_call(gcam_burst* b, const gcam_denoise_params* params, const gcam_image* frames,
                       const gcam_image* gradients, const gcam_ycbcr_nlf* nlf) {
    if (!b) return set_error(GCAM_ERR_INVALID_ARGUMENT, "burst is NULL");
    if (!params || !frames || !nlf) return set_error(GCAM_ERR_INVALID_ARGUMENT, "params, frames and nlf are required");
    for (int f = 0; f < b->n; f++) {
        const gcam_pyramid* p = b->frames[f];
        gcam_status st = check_frame_image(&frames[f], GCAM_PIXEL_RGBA32F, "frames", f, p->width, p->height);
        if (st != GCAM_OK) return st;
        if (gradients && gradients[f].data) {
            st = check_frame_image(&gradients[f], GCAM_PIXEL_RG32F, "gradients", f, p->width, p->height);
            if (st != GCAM_OK) return st;
        }
        for (int i = 0; i < b->levels; i++) {
            const gcam_denoise_params& q = params[f * b->levels + i];
            if (!(q.luma > 0.0f) || !(q.chroma > 0.0f) || !(q.chromaBoost > 0.0f))
                return set_error(GCAM_ERR_INVALID_ARGUMENT, "frame " + std::to_string(f) + " level " +
                                                                std::to_string(i) +
                                                                ": luma, chroma and chromaBoost must be > 0");
        }
    }
    return GCAM_OK;
}

// Gradient pyramid of frame f, allocated the first time it has a gradient.
gcam_status ensure_gradient(gcam_pyramid* p) {
    if (p->grad_arena || p->levels < 2) return GCAM_OK;
    size_t total = 0;
    auto a256 = [](size_t v) { return (v + 255) & ~(size_t)255; };
    for (int i = 1; i < p->levels; i++) total += a256(sizeof(float2) * p->W[i] * p->H[i]);
    GCAM_HIP_TRY(hipMalloc((void**)&p->grad_arena, total));
    size_t off = 0;
    for (int i = 1; i < p->levels; i++) {
        p->gradl[i] = reinterpret_cast<float2*>(reinterpret_cast<char*>(p->grad_arena) + off);
        off += a256(sizeof(float2) * p->W[i] * p->H[i]);
    }
    return GCAM_OK;
}

BurstBmConfig level_config(const gcam_burst* b, int i) {
    return BurstBmConfig{b->tuning.rows_per_thread[i], b->tuning.grid[i], b->tuning.schedule[i],
                         b->tuning.tile_width[i]};
}

// Block-matching table of level i for `rows` rows per thread (tile geometry).
void fill_bm_table(gcam_burst* b, BurstTables* t, int i, int rows, int tile_w, const gcam_denoise_params* params,
                   const gcam_image* frames, const gcam_image* gradients, float lens) {
    int tw = 0, th = 0;
    burst_bm_tile(b->radius, rows, tile_w, b->precision, &tw, &th);
    BurstTable<BurstBmFrame>& bt = t->bm[i];
    const BurstTables* d = b->dev;
    int tiles = 0;
    bt.nframes = b->n;
    for (int f = 0; f < b->n; f++) {
        gcam_pyramid* p = b->frames[f];
        const int w = p->W[i], h = p->H[i], L = b->levels;
        BurstBmFrame& e = bt.frame[f];
        const gcam_denoise_params& q = params[f * L + i];
        // level-i layer: the subtracted image below the coarsest level, the
        // noisy level at the coarsest (the input frame itself when L == 1)
        if (i < L - 1) {
            e.in = View<const float4>{p->sub[i], w, h, w};
        } else if (i == 0) {
            e.in = view<const float4>(frames[f]);
        } else {
            e.in = View<const float4>{p->layer[i], w, h, w};
        }
        e.grad = View<const float2>{nullptr, 0, 0, 0};
        if (gradients && gradients[f].data)
            e.grad = i == 0 ? view<const float2>(gradients[f]) : View<const float2>{p->gradl[i], w, h, w};
        e.basis = p->basis + 200 * i;
        e.nlf = &d->nlf[f][i];
        e.p.thr[0] = q.luma;  // nflMultiplier (pyramid_processor.cpp:52-56)
        e.p.thr[1] = q.chroma;
        e.p.thr[2] = q.chroma;
        e.p.chromaBoost = q.chromaBoost;
        e.p.gradientBoost = q.gradientBoost;
        e.p.gradientThreshold = q.gradientThreshold;
        e.p.lensShading = lens;
        e.out = View<float4>{p->den[i], w, h, w};
        e.tiles_x = (w + tw - 1) / tw;
        e.tile0 = tiles;
        tiles += e.tiles_x * ((h + th - 1) / th);
    }
    bt.total = tiles;
    b->bm_tiles[i] = tiles;
}

// Host tables of one call (frames, gradients, parameters, noise models).
void fill_tables(gcam_burst* b, BurstTables* t, const gcam_denoise_params* params, const gcam_image* frames,
                 const gcam_image* gradients, const gcam_ycbcr_nlf* nlf, float lens) {
    const int L = b->levels;
    for (int f = 0; f < b->n; f++)
        for (int i = 0; i < L; i++)
            for (int c = 0; c < 3; c++) {
                t->nlf[f][i].a[c] = nlf[f * L + i].a[c];
                t->nlf[f][i].b[c] = nlf[f * L + i].b[c];
            }
    const BurstTables* d = b->dev;
    for (int i = 0; i < L; i++) {
        int down = 0, sub = 0, mom = 0;
        t->down[i].nframes = t->sub[i].nframes = t->mom[i].nframes = b->n;
        for (int f = 0; f < b->n; f++) {
            gcam_pyramid* p = b->frames[f];
            const int w = p->W[i], h = p->H[i];
            const bool has_grad = gradients && gradients[f].data;
            const View<const float4> noisy =
                i == 0 ? view<const float4>(frames[f]) : View<const float4>{p->layer[i], w, h, w};
            const View<const float2> grad =
                !has_grad ? View<const float2>{nullptr, 0, 0, 0}
                          : (i == 0 ? view<const float2>(gradients[f]) : View<const float2>{p->gradl[i], w, h, w});
            if (i < L - 1) {
                const int w1 = p->W[i + 1], h1 = p->H[i + 1];
                BurstDownFrame& dn = t->down[i].frame[f];
                dn.in = noisy;
                dn.out = View<float4>{p->layer[i + 1], w1, h1, w1};
                dn.gin = grad;
                dn.gout = has_grad ? View<float2>{p->gradl[i + 1], w1, h1, w1} : View<float2>{nullptr, 0, 0, 0};
                dn.bx = (w1 + 63) / 64;
                dn.block0 = down;
                down += blocks64x4(w1, h1);

                BurstSubFrame& sb = t->sub[i].frame[f];
                sb.in = noisy;
                sb.in1 = View<const float4>{p->layer[i + 1], w1, h1, w1};
                sb.den1 = View<const float4>{p->den[i + 1], w1, h1, w1};
                sb.grad = grad;
                sb.nlf = &d->nlf[f][i];
                sb.sharpening = params[f * L + i].sharpening;
                sb.thr_y = params[f * L + i].luma;
                sb.out = View<float4>{p->sub[i], w, h, w};
                sb.partials = p->moment_partials;
                sb.nblk = submom_blocks(w, h);
                sub = (sub + 7) & ~7;  // the frame's workgroup b runs on XCD b % 8, as in the single-frame pass
                sb.block0 = sub;
                sub += sb.nblk;
            }
            // Levels below the top: the fused pass wrote the partials; the top
            // level is sampled by the moment kernel.
            BurstMomFrame& m = t->mom[i].frame[f];
            m.in = noisy;
            m.partials = p->moment_partials;
            m.nblocks = i < L - 1 ? submom_blocks(w, h) : moment_block_count((int64_t)(w / 8) * (h / 8));
            m.block0 = mom;
            m.n = (double)(((int64_t)w * h) / 64);  // pyramid_processor.cpp:142
            m.basis = p->basis + 200 * i;
            m.eigvals = p->eigvals + 25 * i;
            m.tri = p->tri + 50 * i;
            mom += m.nblocks;
        }
        t->down[i].total = down;
        t->sub[i].total = sub;
        t->mom[i].total = mom;
        b->down_blocks[i] = down;
        b->sub_blocks[i] = sub;
        b->mom_blocks[i] = mom;
        fill_bm_table(b, t, i, b->tuning.rows_per_thread[i], b->tuning.tile_width[i], params, frames, gradients, lens);
    }
}

// The launches of one call, in stream order (eager or under capture).
gcam_status enqueue(gcam_burst* b, hipStream_t s) {
    const int L = b->levels;
    BurstTables* d = b->dev;
    gcam_status st = GCAM_OK;
    for (int i = 0; i < L - 1 && st == GCAM_OK; i++) st = launch_burst_downsample(&d->down[i], b->down_blocks[i], s);
    for (int i = L - 1; i >= 0 && st == GCAM_OK; i--) {
        if (i < L - 1) {
            st = launch_burst_subtract_moments(&d->sub[i], b->sub_blocks[i], s);
            if (st == GCAM_OK) st = launch_burst_pca_solve(&d->mom[i], b->n, s);
        } else {
            st = launch_burst_pca(&d->mom[i], b->mom_blocks[i], b->n, s);
        }
        if (st == GCAM_OK)
            st = launch_burst_block_match_lp(&d->bm[i], &d->dir, b->bm_tiles[i], b->radius, b->precision,
                                             level_config(b, i), s);
    }
    return st;
}

gcam_status stage(gcam_burst* b, const gcam_denoise_params* params, const gcam_image* frames,
                  const gcam_image* gradients, const gcam_ycbcr_nlf* nlf, float lens, hipStream_t s) {
    GCAM_HIP_TRY(hipEventSynchronize(b->staged));  // the previous copy has left the staging buffer
    fill_tables(b, b->host, params, frames, gradients, nlf, lens);
    GCAM_HIP_TRY(hipMemcpyAsync(b->dev, b->host, sizeof(BurstTables), hipMemcpyHostToDevice, s));
    GCAM_HIP_TRY(hipEventRecord(b->staged, s));
    return GCAM_OK;
}

// Mean time of `reps` launches of level i's block matching under cfg.
gcam_status time_bm(gcam_burst* b, int i, const BurstBmConfig& cfg, int reps, hipStream_t s, float* ms) {
    hipEvent_t e0, e1;
    GCAM_HIP_TRY(hipEventCreate(&e0));
    GCAM_HIP_TRY(hipEventCreate(&e1));
    gcam_status st =
        launch_burst_block_match_lp(&b->dev->bm[i], &b->dev->dir, b->bm_tiles[i], b->radius, b->precision, cfg, s);
    if (st == GCAM_OK && hipEventRecord(e0, s) != hipSuccess) st = set_error(GCAM_ERR_HIP, "hipEventRecord");
    for (int r = 0; r < reps && st == GCAM_OK; r++)
        st = launch_burst_block_match_lp(&b->dev->bm[i], &b->dev->dir, b->bm_tiles[i], b->radius, b->precision, cfg,
                                         s);
    if (st == GCAM_OK && hipEventRecord(e1, s) != hipSuccess) st = set_error(GCAM_ERR_HIP, "hipEventRecord");
    if (st == GCAM_OK && hipEventSynchronize(e1) != hipSuccess) st = set_error(GCAM_ERR_HIP, "hipEventSynchronize");
    float t = 0.0f;
    if (st == GCAM_OK && hipEventElapsedTime(&t, e0, e1) != hipSuccess) st = set_error(GCAM_ERR_HIP, "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / reps;
    return st;
}

// Autotuner: run the burst once with the default launch, then per level time
// the candidate block-matching launches on the level's real inputs (the
// launch is idempotent) and keep the fastest: rows per thread first (one
// workgroup per tile, 32-wide tiles), then 64-wide tiles for two rows per
// thread, then the persistent grid size, then the tile order.
gcam_status autotune(gcam_burst* b, const gcam_denoise_params* params, const gcam_image* frames,
                     const gcam_image* gradients, const gcam_ycbcr_nlf* nlf, float lens, hipStream_t s) {
    const int L = b->levels;
    const int def_rows = b->radius <= 6 ? 2 : 1;
    for (int i = 0; i < L; i++) {
        b->tuning.rows_per_thread[i] = def_rows;
        b->tuning.grid[i] = 0;
        b->tuning.schedule[i] = 0;
        b->tuning.tile_width[i] = 32;
    }
    gcam_status st = stage(b, params, frames, gradients, nlf, lens, s);
    if (st == GCAM_OK) st = enqueue(b, s);
    if (st != GCAM_OK) return st;
    for (int i = 0; i < L; i++) {
        const int reps = i == 0 ? 2 : 4;
        BurstBmConfig best{def_rows, 0, 0, 32};
        float best_ms = 1e30f;
        auto consider = [&](const BurstBmConfig& c) -> gcam_status {
            // the tile table depends on the rows per thread
            GCAM_HIP_TRY(hipEventSynchronize(b->staged));
            fill_bm_table(b, b->host, i, c.rows, c.tile_w, params, frames, gradients, lens);
            GCAM_HIP_TRY(hipMemcpyAsync(&b->dev->bm[i], &b->host->bm[i], sizeof(b->host->bm[i]),
                                        hipMemcpyHostToDevice, s));
            GCAM_HIP_TRY(hipEventRecord(b->staged, s));
            float ms = 0.0f;
            gcam_status r = time_bm(b, i, c, reps, s, &ms);
            if (r != GCAM_OK) return r;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
            return GCAM_OK;
        };
        const int rows_set[3] = {1, 2, 4};
        for (int k = 0; k < 3 && st == GCAM_OK; k++) {
            if (b->radius > 6 && rows_set[k] == 4) continue;
            st = consider(BurstBmConfig{rows_set[k], 0, 0, 32});
        }
        if (st == GCAM_OK && best.rows == 2 && b->radius <= 6) st = consider(BurstBmConfig{2, 0, 0, 64});
        const int rows = best.rows, tile_w = best.tile_w;
        const int grids[3] = {b->cus, 2 * b->cus, 4 * b->cus};
        for (int k = 0; k < 3 && st == GCAM_OK; k++) st = consider(BurstBmConfig{rows, grids[k], 0, tile_w});
        if (st == GCAM_OK) st = consider(BurstBmConfig{rows, best.grid, 1, tile_w});
        if (st != GCAM_OK) return st;
        b->tuning.rows_per_thread[i] = best.rows;
        b->tuning.grid[i] = best.grid;
        b->tuning.schedule[i] = best.schedule;
        b->tuning.tile_width[i] = best.tile_w;
        b->tuning.level_ms[i] = best_ms;
    }
    b->tuning.tuned = 1;
    drop_graph(b);
    return GCAM_OK;
}

gcam_status capture(gcam_burst* b) {
    drop_graph(b);
    GCAM_HIP_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    gcam_status st = enqueue(b, b->cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (st != GCAM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return set_error(GCAM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    GCAM_HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return GCAM_OK;
}

gcam_status run(gcam_burst* b, const gcam_denoise_params* params, const gcam_image* frames,
                const gcam_image* gradients, const gcam_ycbcr_nlf* nlf, float lens, void* stream, bool force_tune) {
    gcam_status st = check_call(b, params, frames, gradients, nlf);
    if (st != GCAM_OK) return st;
    DeviceGuard guard(b->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    for (int f = 0; f < b->n; f++)
        if (gradients && gradients[f].data) {
            st = ensure_gradient(b->frames[f]);
            if (st != GCAM_OK) return st;
        }
    if (force_tune || !b->tuning.tuned) {
        st = autotune(b, params, frames, gradients, nlf, lens, s);
        if (st != GCAM_OK) return st;
    }
    st = stage(b, params, frames, gradients, nlf, lens, s);
    if (st != GCAM_OK) return st;
    if (!b->exec) {
        st = capture(b);
        if (st != GCAM_OK) return st;
    }
    GCAM_HIP_TRY(hipGraphLaunch(b->exec, s));
    return GCAM_OK;
}

}  // namespace

extern "C" {

gcam_status gcam_burst_create(int32_t n_frames, const int32_t* widths, const int32_t* heights, int32_t levels,
                              const gcam_config* config, gcam_burst** out) {
    if (!out) return set_error(GCAM_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (n_frames < 1 || n_frames > GCAM_MAX_BURST)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "n_frames must be 1.." + std::to_string(GCAM_MAX_BURST));
    if (!widths || !heights) return set_error(GCAM_ERR_INVALID_ARGUMENT, "widths and heights are required");
    gcam_config cfg{10, GCAM_PRECISION_FP16, 0, 0, {0, 0, 0, 0}};
    if (config) cfg = *config;
    if (cfg.precision != GCAM_PRECISION_FP16 && cfg.precision != GCAM_PRECISION_BF16)
        return set_error(GCAM_ERR_UNSUPPORTED, "the burst form runs the precision form: FP16 or BF16");
    if (cfg.bilateral != 0) return set_error(GCAM_ERR_UNSUPPORTED, "the burst form runs the PCA block matching");
    if (cfg.bm_rows_per_thread != 0)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "the burst picks rows per thread itself (gcam_burst_set_tuning)");
    auto* b = new gcam_burst();
    b->n = n_frames;
    b->levels = levels;
    b->radius = cfg.search_radius;
    b->precision = cfg.precision;
    auto fail = [&](gcam_status st) {
        gcam_burst_destroy(b);
        return st;
    };
    if (hipGetDevice(&b->device) != hipSuccess) return fail(set_error(GCAM_ERR_NO_DEVICE, "no HIP device"));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess && prop.multiProcessorCount > 0)
        b->cus = prop.multiProcessorCount;
    for (int f = 0; f < n_frames; f++) {
        gcam_pyramid* p = nullptr;
        const gcam_status st = gcam_pyramid_create(widths[f], heights[f], levels, &cfg, &p);
        if (st != GCAM_OK) {
            const std::string msg = gcam_last_error();
            return fail(set_error(st, "frame " + std::to_string(f) + ": " + msg));
        }
        b->frames.push_back(p);
    }
    if (hipMalloc((void**)&b->dev, sizeof(BurstTables)) != hipSuccess)
        return fail(set_error(GCAM_ERR_OUT_OF_MEMORY, "burst tables"));
    if (hipHostMalloc((void**)&b->host, sizeof(BurstTables), hipHostMallocDefault) != hipSuccess)
        return fail(set_error(GCAM_ERR_OUT_OF_MEMORY, "burst staging tables"));
    std::memset(b->host, 0, sizeof(BurstTables));
    b->host->dir = make_bm_dir_table();
    if (hipEventCreateWithFlags(&b->staged, hipEventDisableTiming) != hipSuccess)
        return fail(set_error(GCAM_ERR_HIP, "hipEventCreate"));
    if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess)
        return fail(set_error(GCAM_ERR_HIP, "hipStreamCreate"));
    *out = b;
    return GCAM_OK;
}

void gcam_burst_destroy(gcam_burst* b) {
    if (!b) return;
    DeviceGuard g(b->device);
    if (b->staged) (void)hipEventSynchronize(b->staged);
    drop_graph(b);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->staged) (void)hipEventDestroy(b->staged);
    if (b->host) (void)hipHostFree(b->host);
    if (b->dev) (void)hipFree(b->dev);
    for (gcam_pyramid* p : b->frames) gcam_pyramid_destroy(p);
    delete b;
}

gcam_status gcam_burst_denoise(gcam_burst* b, const gcam_denoise_params* params, const gcam_image* frames,
                               const gcam_image* gradients, const gcam_ycbcr_nlf* nlf, float lens_shading_correction,
                               void* stream) {
    return run(b, params, frames, gradients, nlf, lens_shading_correction, stream, false);
}

gcam_status gcam_burst_tune(gcam_burst* b, const gcam_denoise_params* params, const gcam_image* frames,
                            const gcam_image* gradients, const gcam_ycbcr_nlf* nlf, float lens_shading_correction,
                            void* stream) {
    gcam_status st = run(b, params, frames, gradients, nlf, lens_shading_correction, stream, true);
    if (st == GCAM_OK) GCAM_HIP_TRY(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
    return st;
}

gcam_status gcam_burst_output(const gcam_burst* b, int32_t frame, int32_t level, gcam_image* out) {
    if (!b || !out) return set_error(GCAM_ERR_INVALID_ARGUMENT, "burst and out are required");
    if (frame < 0 || frame >= b->n) return set_error(GCAM_ERR_INVALID_ARGUMENT, "frame out of range");
    if (level < 0 || level >= b->levels) return set_error(GCAM_ERR_INVALID_ARGUMENT, "level out of range");
    const gcam_pyramid* p = b->frames[frame];
    *out = gcam_image{p->den[level], p->W[level], p->H[level], p->W[level], GCAM_PIXEL_RGBA32F};
    return GCAM_OK;
}

gcam_status gcam_burst_get_tuning(const gcam_burst* b, gcam_burst_tuning* tuning) {
    if (!b || !tuning) return set_error(GCAM_ERR_INVALID_ARGUMENT, "burst and tuning are required");
    *tuning = b->tuning;
    return GCAM_OK;
}

gcam_status gcam_burst_set_tuning(gcam_burst* b, const gcam_burst_tuning* tuning) {
    if (!b || !tuning) return set_error(GCAM_ERR_INVALID_ARGUMENT, "burst and tuning are required");
    for (int i = 0; i < b->levels; i++) {
        const int r = tuning->rows_per_thread[i];
        if (r != 1 && r != 2 && r != 4) return set_error(GCAM_ERR_INVALID_ARGUMENT, "rows_per_thread must be 1, 2 or 4");
        if (b->radius > 6 && r == 4)
            return set_error(GCAM_ERR_INVALID_ARGUMENT, "search radius > 6 runs one or two rows per thread");
        if (tuning->grid[i] < 0) return set_error(GCAM_ERR_INVALID_ARGUMENT, "grid must be >= 0");
        if (tuning->schedule[i] != 0 && tuning->schedule[i] != 1)
            return set_error(GCAM_ERR_INVALID_ARGUMENT, "schedule must be 0 or 1");
        const int tw = tuning->tile_width[i];
        if (tw != 32 && tw != 64) return set_error(GCAM_ERR_INVALID_ARGUMENT, "tile_width must be 32 or 64");
        if (tw == 64 && (r != 2 || b->radius > 6))
            return set_error(GCAM_ERR_INVALID_ARGUMENT, "64-wide tiles run two rows per thread at radius <= 6");
    }
    DeviceGuard g(b->device);
    if (b->staged) (void)hipEventSynchronize(b->staged);
    b->tuning = *tuning;
    for (int i = b->levels; i < GCAM_MAX_LEVELS; i++) {
        b->tuning.rows_per_thread[i] = 0;
        b->tuning.grid[i] = 0;
        b->tuning.schedule[i] = 0;
        b->tuning.tile_width[i] = 0;
    }
    for (int i = 0; i < GCAM_MAX_LEVELS; i++) b->tuning.level_ms[i] = 0.0f;
    b->tuning.tuned = 1;
    drop_graph(b);
    return GCAM_OK;
}

}  // extern "C"
