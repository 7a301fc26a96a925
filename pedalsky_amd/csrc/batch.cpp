// Many scenes per launch: every scene's LUT chain runs as one launch per LUT
// kind (blockIdx.y = scene) and every wave tile of every scene is rendered
// by one tile launch, optionally recorded once into a hipGraph and replayed.
// Scene parameters live in pinned staging copied to the device by the
// launch sequence itself, so a replay picks up ps_batch_update's new values.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>
#include "ps_host.h"
#include "ps_launch.h"

using namespace psi;

static_assert(sizeof(ps_scene_desc) == 96, "ps_scene_desc layout");
static_assert(sizeof(ps_tuning) == 20, "ps_tuning layout");

struct ps_batch {
    ps_ctx* c = nullptr;
    int n = 0;
    std::vector<SceneIn> sc;
    std::vector<const float*> depth;
    std::vector<ps_image_out> out;
    std::vector<int> W, H;
    std::vector<uint8_t> overcast;     // 128*128 per scene (host copy)
    std::vector<char> own_overcast;    // scene brings its own mask
    uint8_t* d_overcast = nullptr;     // n * 128*128
    float4 *T = nullptr, *MS = nullptr, *SV = nullptr, *AP = nullptr;  // n LUT sets each
    LutParams *d_lp = nullptr, *h_lp = nullptr;
    RenderParams *d_rp = nullptr, *h_rp = nullptr;
    uint32_t* d_tiles = nullptr;
    int n_tiles = 0;
    size_t tiles_cap = 0;
    int tiles_w = 0, tiles_order = -1;
    unsigned* d_next = nullptr;
    ps_tuning tune{8, 0, 8, 1, 1};
    // step schedule: resolution class of each scene, the classes' sizes and
    // their primary / light step counts (0: each scene's own)
    std::vector<int> cls;
    int n_sizes = 0;
    int cls_w[PS_MAX_STEP_CLASSES]{}, cls_h[PS_MAX_STEP_CLASSES]{};
    int cls_cloud[PS_MAX_STEP_CLASSES]{}, cls_light[PS_MAX_STEP_CLASSES]{};
    int n_cu = 256;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr, t0 = nullptr, t1 = nullptr;
    bool in_flight = false;
};

namespace {

constexpr int kMaxBatchScenes = 4096;  // 12 bits of a tile entry
constexpr int kMaxBatchSide = 4096;    // 10-bit tile coordinates at 4x16 / 16x4 tiles

size_t lut_texels(const ps_lut_dims& d, int which) {
    switch (which) {
        case 0: return (size_t)d.transmittance_w * d.transmittance_h;
        case 1: return (size_t)d.multiscatter_w * d.multiscatter_h;
        case 2: return (size_t)d.skyview_w * d.skyview_h;
        default: return (size_t)d.ap_w * d.ap_h * d.ap_d;
    }
}

ps_status read_descs(ps_batch* b, const ps_scene_desc* s, int n, bool shapes_fixed) {
    if (!s || n < 1 || n > kMaxBatchScenes)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch: need 1 <= n <= 4096 scene descriptors");
    std::vector<SceneIn> sc(n);
    std::vector<int> W(n), H(n);
    for (int i = 0; i < n; ++i) {
        const ps_scene_desc& d = s[i];
        const std::string who = "ps_batch scene " + std::to_string(i);
        if (!d.frame || !d.atmosphere || !d.clouds)
            return fail(PS_ERR_INVALID_ARGUMENT, who + ": frame, atmosphere and clouds are required");
        ps_status st;
        if ((st = check_atmosphere(d.atmosphere, who.c_str()))) return st;
        if ((st = check_clouds(d.clouds, d.layers, d.n_layers, who.c_str()))) return st;
        const float vw = d.frame->viewport_size[0], vh = d.frame->viewport_size[1];
        if (!(vw >= 1.f && vh >= 1.f && vw <= kMaxBatchSide && vh <= kMaxBatchSide) || vw != std::floor(vw) ||
            vh != std::floor(vh))
            return fail(PS_ERR_INVALID_ARGUMENT, who + ": viewport must be integer sizes in [1, 4096]");
        if (!d.out.composite && !d.out.cloud_luminance && !d.out.cloud_transmittance)
            return fail(PS_ERR_INVALID_ARGUMENT, who + ": no output image");
        sc[i].frame = *d.frame;
        sc[i].atm = *d.atmosphere;
        sc[i].clouds = *d.clouds;
        sc[i].n_layers = d.n_layers;
        for (int l = 0; l < d.n_layers; ++l) {
            sc[i].layers[l] = d.layers[l];
            sc[i].wind[2 * l] = d.layer_wind_xz ? d.layer_wind_xz[2 * l] : 0.f;
            sc[i].wind[2 * l + 1] = d.layer_wind_xz ? d.layer_wind_xz[2 * l + 1] : 0.f;
        }
        W[i] = (int)vw;
        H[i] = (int)vh;
    }
    if (shapes_fixed) {
        if (n != b->n) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_update: scene count differs from the batch");
        for (int i = 0; i < n; ++i)
            if (W[i] != b->W[i] || H[i] != b->H[i])
                return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_update: a scene's resolution differs from the batch");
    }
    b->n = n;
    b->sc = std::move(sc);
    b->W = std::move(W);
    b->H = std::move(H);
    b->depth.resize(n);
    b->out.resize(n);
    b->overcast.assign((size_t)n * 128 * 128, 0);
    b->own_overcast.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        b->depth[i] = s[i].depth;
        b->out[i] = s[i].out;
        if (s[i].overcast_mask_128) {
            std::memcpy(&b->overcast[(size_t)i * 128 * 128], s[i].overcast_mask_128, 128 * 128);
            b->own_overcast[i] = 1;
        }
    }
    return PS_OK;
}

ps_status wait_idle(ps_batch* b) {
    if (b->in_flight) {
        PS_HIP(hipEventSynchronize(b->done));
        b->in_flight = false;
    }
    return PS_OK;
}

// Resolution classes: distinct viewport sizes, most pixels first (ties:
// wider first). More than PS_MAX_STEP_CLASSES sizes leave the schedule off.
void build_classes(ps_batch* b) {
    std::vector<std::pair<int, int>> sizes;
    for (int i = 0; i < b->n; ++i) {
        const std::pair<int, int> wh{b->W[i], b->H[i]};
        if (std::find(sizes.begin(), sizes.end(), wh) == sizes.end()) sizes.push_back(wh);
    }
    std::stable_sort(sizes.begin(), sizes.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& c) {
        const long pa = (long)a.first * a.second, pc = (long)c.first * c.second;
        return pa != pc ? pa > pc : a.first > c.first;
    });
    b->n_sizes = (int)sizes.size();
    b->cls.assign(b->n, -1);
    for (int k = 0; k < PS_MAX_STEP_CLASSES; ++k) b->cls_w[k] = b->cls_h[k] = b->cls_cloud[k] = b->cls_light[k] = 0;
    if (b->n_sizes > PS_MAX_STEP_CLASSES) return;
    for (int k = 0; k < b->n_sizes; ++k) {
        b->cls_w[k] = sizes[k].first;
        b->cls_h[k] = sizes[k].second;
    }
    for (int i = 0; i < b->n; ++i)
        for (int k = 0; k < b->n_sizes; ++k)
            if (b->W[i] == b->cls_w[k] && b->H[i] == b->cls_h[k]) b->cls[i] = k;
}

// A scene's march step counts under the schedule (its own where its class
// has none).
void scene_steps(const ps_batch* b, int i, uint32_t& cloud, uint32_t& light) {
    cloud = b->sc[i].clouds.cloud_steps;
    light = b->sc[i].clouds.light_steps;
    const int k = b->cls[i];
    if (k >= 0 && b->cls_cloud[k] > 0) cloud = (uint32_t)b->cls_cloud[k];
    if (k >= 0 && b->cls_light[k] > 0) light = (uint32_t)b->cls_light[k];
}

// Fold every scene into the pinned staging blocks the launch sequence copies.
ps_status stage_params(ps_batch* b) {
    const ps_lut_dims& d = b->c->dims;
    for (int i = 0; i < b->n; ++i) {
        SceneIn si = b->sc[i];
        scene_steps(b, i, si.clouds.cloud_steps, si.clouds.light_steps);
        fill_lut_params(si, d, b->h_lp[i]);
        LutSet L;
        L.T = b->T + i * lut_texels(d, 0);
        L.MS = b->MS + i * lut_texels(d, 1);
        L.SV = b->SV + i * lut_texels(d, 2);
        L.AP = b->AP + i * lut_texels(d, 3);
        ps_status s;
        if ((s = fill_render_params(b->c, si, L, b->depth[i], &b->out[i], b->h_rp[i]))) return s;
        if (b->own_overcast[i]) b->h_rp[i].overcast = b->d_overcast + (size_t)i * 128 * 128;
    }
    // per-scene masks: the batch is idle here (create, or update after wait_idle)
    PS_HIP(hipMemcpy(b->d_overcast, b->overcast.data(), b->overcast.size(), hipMemcpyHostToDevice));
    return PS_OK;
}

void drop_graph(ps_batch* b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Tile list for the current tuning: scene-major, tiles row-major; order 1
// puts larger scenes first so the small ones fill the tail of the launch.
ps_status build_tiles(ps_batch* b) {
    const int tw = b->tune.tile_w, th = 64 / tw;
    std::vector<int> idx(b->n);
    std::iota(idx.begin(), idx.end(), 0);
    if (b->tune.order == 1)
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int a, int c) { return (size_t)b->W[a] * b->H[a] > (size_t)b->W[c] * b->H[c]; });
    std::vector<uint32_t> tiles;
    for (int i : idx) {
        const int nx = (b->W[i] + tw - 1) / tw, ny = (b->H[i] + th - 1) / th;
        if (nx > 1024 || ny > 1024) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch: too many tiles along one axis");
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) tiles.push_back((uint32_t)i << 20 | (uint32_t)y << 10 | (uint32_t)x);
    }
    if (tiles.size() > b->tiles_cap) {
        dev_free(b->d_tiles);
        ps_status s;
        if ((s = dev_alloc(&b->d_tiles, tiles.size() * sizeof(uint32_t), "batch tiles"))) return s;
        b->tiles_cap = tiles.size();
    }
    PS_HIP(hipMemcpy(b->d_tiles, tiles.data(), tiles.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    b->n_tiles = (int)tiles.size();
    b->tiles_w = b->tune.tile_w;
    b->tiles_order = b->tune.order;
    return PS_OK;
}

// The launch sequence of one batch run (also what a capture records).
ps_status enqueue(ps_batch* b) {
    hipStream_t s = b->c->stream;
    PS_HIP(hipMemcpyAsync(b->d_lp, b->h_lp, sizeof(LutParams) * b->n, hipMemcpyHostToDevice, s));
    PS_HIP(hipMemcpyAsync(b->d_rp, b->h_rp, sizeof(RenderParams) * b->n, hipMemcpyHostToDevice, s));
    PS_HIP(ps_launch_luts(b->d_lp, b->h_lp[0], b->n, 15u, b->T, b->MS, b->SV, b->AP, s, nullptr));
    PS_HIP(ps_launch_render_tiles(b->d_rp, b->d_tiles, b->n_tiles, b->tune.tile_w, b->tune.persistent,
                                  b->n_cu * std::max(b->tune.workers_per_cu, 1), std::max(b->tune.grab, 1),
                                  b->d_next, s));
    return PS_OK;
}

ps_status launch_once(ps_batch* b) {
    ps_status s;
    if ((s = check_shared_inputs(b->c))) return s;
    // a batch renders full-resolution single frames: the context's
    // half-resolution and TAA modes, switched on after ps_batch_create, are
    // refused rather than ignored
    if (b->c->half_res || b->c->taa)
        return fail(PS_ERR_UNSUPPORTED, "batch launch: switch half resolution and TAA off first");
    if (b->exec) {
        PS_HIP(hipGraphLaunch(b->exec, b->c->stream));
    } else if ((s = enqueue(b))) {
        return s;
    }
    PS_HIP(hipEventRecord(b->done, b->c->stream));
    b->in_flight = true;
    return PS_OK;
}

ps_status check_tuning(const ps_tuning* t) {
    if (t->tile_w != 4 && t->tile_w != 8 && t->tile_w != 16)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_tuning.tile_w must be 4, 8 or 16");
    if (t->persistent != 0 && t->persistent != 1) return fail(PS_ERR_INVALID_ARGUMENT, "ps_tuning.persistent is 0 or 1");
    if (t->persistent && (t->workers_per_cu < 1 || t->workers_per_cu > 64))
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_tuning.workers_per_cu must be in [1, 64]");
    if (t->order != 0 && t->order != 1) return fail(PS_ERR_INVALID_ARGUMENT, "ps_tuning.order is 0 or 1");
    if (t->grab < 0 || t->grab > 256) return fail(PS_ERR_INVALID_ARGUMENT, "ps_tuning.grab must be in [0, 256]");
    return PS_OK;
}

}  // namespace

extern "C" {

ps_status ps_batch_create(ps_ctx* c, const ps_scene_desc* scenes, int32_t n, ps_batch** out) {
    if (!c || !out) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_create: NULL argument");
    *out = nullptr;
    if (c->half_res) return fail(PS_ERR_UNSUPPORTED, "ps_batch_create: batches render at full resolution");
    if (c->taa) return fail(PS_ERR_UNSUPPORTED, "ps_batch_create: batches render single frames (TAA off)");
    ps_batch* b = new ps_batch();
    b->c = c;
    auto bail = [&](ps_status s) {
        ps_batch_destroy(b);
        return s;
    };
    ps_status s;
    if ((s = read_descs(b, scenes, n, false))) return bail(s);
    const ps_lut_dims& d = c->dims;
    if ((s = dev_alloc(&b->T, sizeof(float4) * lut_texels(d, 0) * n, "batch transmittance"))) return bail(s);
    if ((s = dev_alloc(&b->MS, sizeof(float4) * lut_texels(d, 1) * n, "batch multiscatter"))) return bail(s);
    if ((s = dev_alloc(&b->SV, sizeof(float4) * lut_texels(d, 2) * n, "batch skyview"))) return bail(s);
    if ((s = dev_alloc(&b->AP, sizeof(float4) * lut_texels(d, 3) * n, "batch aerial perspective"))) return bail(s);
    if ((s = dev_alloc(&b->d_lp, sizeof(LutParams) * n, "batch LUT params"))) return bail(s);
    if ((s = dev_alloc(&b->d_rp, sizeof(RenderParams) * n, "batch render params"))) return bail(s);
    if ((s = dev_alloc(&b->d_next, sizeof(unsigned), "batch tile counter"))) return bail(s);
    if ((s = dev_alloc(&b->d_overcast, (size_t)n * 128 * 128, "batch overcast masks"))) return bail(s);
    hipError_t e = hipHostMalloc((void**)&b->h_lp, sizeof(LutParams) * n, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_rp, sizeof(RenderParams) * n, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&b->t0);
    if (e == hipSuccess) e = hipEventCreate(&b->t1);
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, c->device);
        if (e == hipSuccess) b->n_cu = std::max(prop.multiProcessorCount, 1);
    }
    if (e != hipSuccess) return bail(fail(PS_ERR_DEVICE, std::string("ps_batch_create: ") + hipGetErrorString(e)));
    build_classes(b);
    if ((s = stage_params(b))) return bail(s);
    if ((s = build_tiles(b))) return bail(s);
    *out = b;
    return PS_OK;
}

ps_status ps_batch_update(ps_batch* b, const ps_scene_desc* scenes, int32_t n) {
    if (!b) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_update: batch is NULL");
    ps_status s;
    if ((s = wait_idle(b))) return s;
    if ((s = read_descs(b, scenes, n, true))) return s;
    return stage_params(b);
}

ps_status ps_batch_set_tuning(ps_batch* b, const ps_tuning* t) {
    if (!b || !t) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_set_tuning: NULL argument");
    ps_status s;
    if ((s = check_tuning(t))) return s;
    if ((s = wait_idle(b))) return s;
    const bool retile = t->tile_w != b->tiles_w || t->order != b->tiles_order;
    const bool changed = std::memcmp(t, &b->tune, sizeof(ps_tuning)) != 0;
    b->tune = *t;
    if (changed) drop_graph(b);
    if (retile && (s = build_tiles(b))) return s;
    return PS_OK;
}

ps_status ps_batch_get_tuning(const ps_batch* b, ps_tuning* t) {
    if (!b || !t) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_get_tuning: NULL argument");
    *t = b->tune;
    return PS_OK;
}

ps_status ps_batch_run(ps_batch* b) {
    if (!b) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_run: batch is NULL");
    return launch_once(b);
}

ps_status ps_batch_capture(ps_batch* b) {
    if (!b) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_capture: batch is NULL");
    ps_status s;
    if ((s = check_shared_inputs(b->c))) return s;
    if ((s = wait_idle(b))) return s;
    drop_graph(b);
    hipStream_t st = b->c->stream;
    PS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    ps_status es = enqueue(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (es) {
        if (g) hipGraphDestroy(g);
        return es;
    }
    if (e != hipSuccess) return fail(PS_ERR_DEVICE, std::string("ps_batch_capture: ") + hipGetErrorString(e));
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        return fail(PS_ERR_DEVICE, std::string("ps_batch_capture: instantiate: ") + hipGetErrorString(e));
    }
    return PS_OK;
}

ps_status ps_batch_replay(ps_batch* b) {
    if (!b) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_replay: batch is NULL");
    if (!b->exec) return fail(PS_ERR_NOT_READY, "ps_batch_replay: no captured graph (call ps_batch_capture)");
    return launch_once(b);
}

ps_status ps_batch_time(ps_batch* b, int32_t reps, float* ms_per_run) {
    if (!b || !ms_per_run) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_time: NULL argument");
    if (reps < 1 || reps > 10000) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_time: reps must be in [1, 10000]");
    ps_status s;
    PS_HIP(hipEventRecord(b->t0, b->c->stream));
    for (int r = 0; r < reps; ++r)
        if ((s = launch_once(b))) return s;
    PS_HIP(hipEventRecord(b->t1, b->c->stream));
    PS_HIP(hipEventSynchronize(b->t1));
    b->in_flight = false;
    float ms = 0.f;
    PS_HIP(hipEventElapsedTime(&ms, b->t0, b->t1));
    *ms_per_run = ms / reps;
    return PS_OK;
}

ps_status ps_batch_autotune(ps_batch* b, int32_t reps, ps_tuning* best_out) {
    if (!b) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_autotune: batch is NULL");
    if (reps < 1 || reps > 100) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_autotune: reps must be in [1, 100]");
    // hardware dispatch at each tile shape and order; then persistent workers,
    // 20 per CU (the five waves per SIMD the tile kernel holds) and 16, with
    // schedule steps of 1, 2 and 4 list entries per counter increment (longer
    // steps leave uneven tails: 64 entries cost 44 % on the 64-scene
    // benchmark batch)
    const ps_tuning cands[] = {
        {8, 0, 8, 1, 1},    {8, 0, 8, 0, 1},    {16, 0, 8, 1, 1},  {4, 0, 8, 1, 1},   {16, 1, 20, 1, 1},
        {16, 1, 20, 1, 2},  {16, 1, 20, 1, 4},  {8, 1, 20, 1, 1},  {8, 1, 20, 1, 2},  {16, 1, 16, 1, 2},
    };
    const bool had_graph = b->exec != nullptr;
    ps_tuning best = b->tune;
    float best_ms = 1e30f;
    ps_status s;
    for (const ps_tuning& t : cands) {
        if ((s = ps_batch_set_tuning(b, &t))) return s;
        float ms = 0.f;
        if ((s = ps_batch_time(b, 1, &ms))) return s;  // warm-up
        if ((s = ps_batch_time(b, reps, &ms))) return s;
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    }
    if ((s = ps_batch_set_tuning(b, &best))) return s;
    if (had_graph && (s = ps_batch_capture(b))) return s;
    if (best_out) *best_out = best;
    return PS_OK;
}

ps_status ps_batch_get_step_schedule(const ps_batch* b, ps_step_schedule* out) {
    if (!b || !out) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_get_step_schedule: NULL argument");
    if (b->n_sizes > PS_MAX_STEP_CLASSES)
        return fail(PS_ERR_UNSUPPORTED, "ps_batch_get_step_schedule: the batch has more than 8 viewport sizes");
    std::memset(out, 0, sizeof(*out));
    out->n_classes = b->n_sizes;
    for (int k = 0; k < b->n_sizes; ++k) {
        out->width[k] = b->cls_w[k];
        out->height[k] = b->cls_h[k];
        out->cloud_steps[k] = b->cls_cloud[k];
        out->light_steps[k] = b->cls_light[k];
    }
    return PS_OK;
}

ps_status ps_batch_set_step_schedule(ps_batch* b, const ps_step_schedule* t) {
    if (!b || !t) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_set_step_schedule: NULL argument");
    if (b->n_sizes > PS_MAX_STEP_CLASSES)
        return fail(PS_ERR_UNSUPPORTED, "ps_batch_set_step_schedule: the batch has more than 8 viewport sizes");
    if (t->n_classes != b->n_sizes)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_set_step_schedule: class count differs from the batch's");
    for (int k = 0; k < b->n_sizes; ++k) {
        if (t->width[k] != b->cls_w[k] || t->height[k] != b->cls_h[k])
            return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_set_step_schedule: class " + std::to_string(k) +
                                                     " is not the batch's (sizes in ps_batch_get_step_schedule order)");
        if (t->cloud_steps[k] < 0 || t->light_steps[k] < 0)
            return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_set_step_schedule: negative step count");
    }
    ps_status s;
    if ((s = wait_idle(b))) return s;
    for (int k = 0; k < b->n_sizes; ++k) {
        // set_march_steps' clamps (lib.rs:709-712); 0 keeps the scenes' own
        b->cls_cloud[k] = t->cloud_steps[k] > 0 ? std::clamp(t->cloud_steps[k], 8, 256) : 0;
        b->cls_light[k] = t->light_steps[k] > 0 ? std::clamp(t->light_steps[k], 1, 16) : 0;
    }
    return stage_params(b);
}

ps_status ps_batch_autotune_steps(ps_batch* b, float budget_ms, int32_t reps, ps_step_schedule* best) {
    if (!b) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_autotune_steps: batch is NULL");
    if (!(budget_ms > 0.f)) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_autotune_steps: budget_ms must be > 0");
    if (reps < 1 || reps > 100) return fail(PS_ERR_INVALID_ARGUMENT, "ps_batch_autotune_steps: reps must be in [1, 100]");
    if (b->n_sizes > PS_MAX_STEP_CLASSES)
        return fail(PS_ERR_UNSUPPORTED, "ps_batch_autotune_steps: the batch has more than 8 viewport sizes");
    static const int kCloudRungs[] = {256, 224, 192, 160, 128, 112, 96, 80, 64, 56, 48, 40, 32, 24, 16, 12, 8};
    static const int kLightRungs[] = {16, 12, 10, 8, 6, 5, 4, 3, 2, 1};
    auto next_rung = [](const int* r, int n, int v) {
        for (int j = 0; j < n; ++j)
            if (r[j] < v) return r[j];
        return v;
    };
    ps_step_schedule t{};
    ps_status s;
    if ((s = ps_batch_get_step_schedule(b, &t))) return s;
    // start: every class at the largest of its scenes' own counts, clamped
    std::vector<double> px(b->n_sizes, 0.0);
    for (int k = 0; k < b->n_sizes; ++k) t.cloud_steps[k] = t.light_steps[k] = 0;
    for (int i = 0; i < b->n; ++i) {
        const int k = b->cls[i];
        px[k] += (double)b->W[i] * b->H[i];
        t.cloud_steps[k] = std::max(t.cloud_steps[k], (int)std::clamp(b->sc[i].clouds.cloud_steps, 8u, 256u));
        t.light_steps[k] = std::max(t.light_steps[k], (int)std::clamp(b->sc[i].clouds.light_steps, 1u, 16u));
    }
    float ms = 0.f;
    for (;;) {
        if ((s = ps_batch_set_step_schedule(b, &t))) return s;
        if ((s = ps_batch_time(b, 1, &ms))) return s;  // warm-up
        if ((s = ps_batch_time(b, reps, &ms))) return s;
        if (ms <= budget_ms) break;
        // the class with the most march work whose primary steps can drop
        int kc = -1;
        for (int k = 0; k < b->n_sizes; ++k)
            if (t.cloud_steps[k] > 8 && (kc < 0 || px[k] * t.cloud_steps[k] > px[kc] * t.cloud_steps[kc])) kc = k;
        if (kc >= 0) {
            t.cloud_steps[kc] = next_rung(kCloudRungs, 17, t.cloud_steps[kc]);
            continue;
        }
        int kl = -1;
        for (int k = 0; k < b->n_sizes; ++k)
            if (t.light_steps[k] > 1 && (kl < 0 || px[k] * t.light_steps[k] > px[kl] * t.light_steps[kl])) kl = k;
        if (kl < 0) break;  // everything at the minimum
        t.light_steps[kl] = next_rung(kLightRungs, 10, t.light_steps[kl]);
    }
    t.batch_ms = ms;
    if (best) *best = t;
    return PS_OK;
}

void ps_batch_destroy(ps_batch* b) {
    if (!b) return;
    if (b->in_flight) hipEventSynchronize(b->done);
    drop_graph(b);
    dev_free(b->T); dev_free(b->MS); dev_free(b->SV); dev_free(b->AP);
    dev_free(b->d_lp); dev_free(b->d_rp); dev_free(b->d_tiles); dev_free(b->d_next); dev_free(b->d_overcast);
    if (b->h_lp) hipHostFree(b->h_lp);
    if (b->h_rp) hipHostFree(b->h_rp);
    if (b->done) hipEventDestroy(b->done);
    if (b->t0) hipEventDestroy(b->t0);
    if (b->t1) hipEventDestroy(b->t1);
    delete b;
}

ps_status ps_render_batch(ps_ctx* c, const ps_scene_desc* scenes, int32_t n, const ps_tuning* tuning) {
    ps_batch* b = nullptr;
    ps_status s = ps_batch_create(c, scenes, n, &b);
    if (s) return s;
    if (tuning) s = ps_batch_set_tuning(b, tuning);
    if (!s) s = ps_batch_run(b);
    if (!s && hipEventSynchronize(b->done) != hipSuccess) s = fail(PS_ERR_DEVICE, "ps_render_batch: device error");
    b->in_flight = false;
    ps_batch_destroy(b);
    return s;
}

}  // extern "C"
