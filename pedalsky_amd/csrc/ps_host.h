// Host-side internals shared by the ABI translation units (capi.cpp for the
// single-scene path, batch.cpp for many scenes per launch): error reporting,
// device allocation helpers, the per-scene input state and the context.
#pragma once
#include <hip/hip_runtime.h>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/pedalsky_hip.h"
#include "ps_device.h"

namespace psi {

// Sets the thread-local message ps_last_error() returns; returns `code`.
ps_status fail(ps_status code, const std::string& msg);

// Blue-noise splat weights by squared texel distance (73 entries), shared by
// the host generator and the GPU synthesis (host_synthesis.cpp).
void blue_noise_weights(float* w73);

#define PS_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ::psi::fail(e_ == hipErrorOutOfMemory ? PS_ERR_OUT_OF_MEMORY : PS_ERR_DEVICE, \
                               std::string(#call) + ": " + hipGetErrorString(e_));             \
    } while (0)

template <class T>
ps_status dev_alloc(T** p, size_t bytes, const char* what) {
    if (*p) return PS_OK;
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(e == hipErrorOutOfMemory ? PS_ERR_OUT_OF_MEMORY : PS_ERR_DEVICE,
                    std::string("hipMalloc(") + what + "): " + hipGetErrorString(e));
    }
    return PS_OK;
}
template <class T>
void dev_free(T*& p) {
    if (p) hipFree((void*)p);
    p = nullptr;
}

// Everything one scene's passes read from the host: the reference's uniform
// blocks (FrameUniformsGpu, AtmosphereParams, CloudParamsGpu, CloudLayerGpu[])
// plus the per-layer wind the march samples.
struct SceneIn {
    ps_frame_uniforms frame{};
    ps_atmosphere_params atm{};
    ps_cloud_params clouds{};
    ps_cloud_layer layers[PS_MAX_CLOUD_LAYERS]{};
    float wind[2 * PS_MAX_CLOUD_LAYERS]{};
    int n_layers = 0;
    // differentiation extension: per-layer extinction scale (1 = reference)
    float kappa[PS_MAX_CLOUD_LAYERS] = {1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f};
};

// One scene's LUT set (device).
struct LutSet {
    float4 *T = nullptr, *MS = nullptr, *SV = nullptr, *AP = nullptr;
};

// A ring of device parameter blocks fed from pinned host staging: a new
// block per frame needs no host/device synchronisation; a slot is rewritten
// only after the event of its last use has completed. Identical consecutive
// blocks reuse the current slot.
template <class T>
struct ParamRing {
    static constexpr int kSlots = 8;
    T* d = nullptr;
    T* h = nullptr;
    hipEvent_t ev[kSlots]{};
    bool used[kSlots]{};
    int cur = -1;

    ps_status upload(const T& v, hipStream_t s, T** out) {
        if (!d) {
            PS_HIP(hipMalloc((void**)&d, sizeof(T) * kSlots));
            PS_HIP(hipHostMalloc((void**)&h, sizeof(T) * kSlots, hipHostMallocDefault));
            for (int i = 0; i < kSlots; ++i) PS_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
        }
        if (cur < 0 || std::memcmp(&v, &h[cur], sizeof(T)) != 0) {
            int slot = (cur + 1) % kSlots;
            if (used[slot]) PS_HIP(hipEventSynchronize(ev[slot]));
            std::memcpy(&h[slot], &v, sizeof(T));
            PS_HIP(hipMemcpyAsync(&d[slot], &h[slot], sizeof(T), hipMemcpyHostToDevice, s));
            cur = slot;
        }
        *out = &d[cur];
        return PS_OK;
    }
    // Mark the current slot as read by work queued on `s` up to now.
    ps_status mark_used(hipStream_t s) {
        PS_HIP(hipEventRecord(ev[cur], s));
        used[cur] = true;
        return PS_OK;
    }
    void release() {
        if (d) hipFree(d);
        if (h) {
            hipHostFree(h);
            for (int i = 0; i < kSlots; ++i) hipEventDestroy(ev[i]);
        }
        d = h = nullptr;
    }
};

}  // namespace psi

struct ps_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    ps_lut_dims dims{256, 64, 32, 32, 192, 108, 32, 32, 64};

    psi::SceneIn sc;
    bool have_atm = false, have_frame = false, have_clouds = false;

    psi::LutSet luts;
    uint32_t lut_baked = 0;  // LUTs baked since the last atmosphere / size change
    bool luts_valid = false;
    // Frame pipelining (ps_frame rebaking every LUT): frame k bakes a second
    // LUT set on a side stream while frame k-1 still renders from the
    // current one, then the sets swap. ev_render_issued marks the main
    // stream just before the last pipelined render: everything that could
    // read the set about to be overwritten was queued before it.
    psi::LutSet luts_back;
    hipStream_t lut_stream = nullptr;
    hipEvent_t ev_luts_done = nullptr, ev_render_issued = nullptr;
    bool render_issued_recorded = false;

    int noise_format = -1;
    void *base_store = nullptr, *detail_store = nullptr, *curl_store = nullptr;
    uint32_t* base_rl = nullptr;
    uint16_t* bbound = nullptr;
    float* detail_hf = nullptr;
    float2* curl2 = nullptr;
    bool have_noise = false;

    uint2* weather = nullptr;         // binary16 2x2 quads
    uint16_t* weather_synth = nullptr;  // the last synthesised 128x128 RGBA16F map
    int ww = 0, wh = 0;
    uint8_t* type_grid = nullptr;
    bool type_grid_active = false;
    uint8_t* overcast = nullptr;
    bool overcast_active = false;
    bool have_weather = false;
    uint8_t* blue = nullptr;
    bool have_blue = false;

    psi::ParamRing<RenderParams> render_ring;
    psi::ParamRing<LutParams> lut_ring;
    unsigned long long* d_counters = nullptr;
    std::vector<hipEvent_t> events;
    // in-loop timing (ps_timing_begin / ps_timing_end)
    bool timing = false;
    int timing_max = 0, timing_n = 0;

    // differentiation: caller-owned perturbation volume and its brick bounds
    const float* delta = nullptr;
    float2* delta_bnd = nullptr;
    // training forward state read by ps_backward
    int ck_len = 16;
    float* ck = nullptr;
    size_t ck_cap = 0;            // floats
    float* px_state = nullptr;
    size_t px_cap = 0;
    float* rec = nullptr;
    size_t rec_cap = 0;
    float* lrec = nullptr;        // light-march records of the training forward
    size_t lrec_cap = 0;
    // ps_count_work_tiles: per-tile start / end times of the counting render
    unsigned long long* d_tile_times = nullptr;
    size_t tile_times_cap = 0;    // floats
    bool want_tile_times = false;
    double* g_delta64 = nullptr;  // FP64 accumulator of the perturbation-volume gradient
    RenderParams train{};         // parameter block of the last training forward
    bool train_valid = false;
    // half-resolution cloud march: RT luminance then transmittance (float4)
    bool half_res = false;
    float* rt = nullptr;
    size_t rt_cap = 0;            // floats
    // presentation scratch (float4 images): the working HDR image when the
    // caller keeps none, the godrays RT and the three bloom levels
    float* post_hdr = nullptr;
    size_t post_hdr_cap = 0;      // floats
    float* post_rays = nullptr;
    size_t post_rays_cap = 0;
    float* post_bloom[3] = {nullptr, nullptr, nullptr};
    size_t post_bloom_cap[3] = {0, 0, 0};
    // cloud TAA: per view, two history slots of luminance and transmittance
    // (float4, slot-major: lum0, trn0, lum1, trn1) at the RT size
    struct TaaView {
        float* hist = nullptr;
        size_t cap = 0;           // floats
        bool history_valid = false;
        bool prev_on = false;
        int write_slot = 0;
    };
    bool taa = false;
    int active_view = 0;
    int taa_rt_w = 0, taa_rt_h = 0;   // RT size the histories hold
    TaaView taa_views[PS_MAX_VIEWS];

    // atmosphere differentiation: LUT and tangent-LUT sets for n parameter sets
    int atmo_cap = 0;
    ps_lut_dims atmo_dims{};
    LutParams* atmo_params = nullptr;     // device, atmo_cap blocks
    LutParams* atmo_params_h = nullptr;   // pinned staging
    int atmo_uploaded = 0;                // blocks of atmo_params staged from atmo_params_h
    bool atmo_coeffs_dirty = false;       // coefficients rewritten on the device since staging
    hipEvent_t atmo_ev = nullptr;         // last staging copy
    float4 *aT = nullptr, *aMS = nullptr, *aSV = nullptr, *aAP = nullptr;
    float *aTT = nullptr, *aMST = nullptr, *aSVT = nullptr, *aAPT = nullptr;
};

namespace psi {

void fill_lut_params(const SceneIn& sc, const ps_lut_dims& d, LutParams& P);
// Fold one scene into the render parameter block; shared inputs (noise,
// weather, blue noise) come from the context, the LUTs from `luts`.
ps_status fill_render_params(const ps_ctx* c, const SceneIn& sc, const LutSet& luts, const float* depth,
                             const ps_image_out* out, RenderParams& P);
// Input validation shared by ps_set_* and the batch descriptors.
ps_status check_atmosphere(const ps_atmosphere_params* a, const char* who);
ps_status check_clouds(const ps_cloud_params* p, const ps_cloud_layer* layers, int32_t n, const char* who);
// Shared inputs every render needs (noise, weather, blue noise).
ps_status check_shared_inputs(const ps_ctx* c);
// Everything a render needs: uniforms, shared inputs and baked LUTs.
ps_status check_render_ready(const ps_ctx* c);
// Per-GPU set-up of ps_render_noalloc's scratch array (noalloc.hip), once per
// device with that device current.
ps_status noalloc_init(int device);

}  // namespace psi
