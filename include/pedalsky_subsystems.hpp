// C++ host layer over the C ABI of pedalsky_hip.h: the reference's
// RenderSubsystem boundary (crates/ps-core/src/subsystem.rs:1-118) for the
// two subsystems this library replaces, AtmosphereSubsystem
// (crates/ps-atmosphere/src/lib.rs:454-658) and CloudsSubsystem
// (crates/ps-clouds/src/lib.rs:1100-1260), and the executor's pass order
// (flatten every subsystem's passes, stable-sort by PassStage). Names,
// argument meaning and error behaviour follow the reference: passes are
// registered once, `prepare` runs once per frame before any pass, and
// `dispatch_pass` runs one registered pass; an unknown pass id is a
// programming error (the reference panics; here std::logic_error), and a
// failing library call throws pedalsky::Error with the library's message
// (the reference's anyhow errors).
//
// Header-only, C++17, no GPU types: device buffers are the caller's
// (depth and output images as device pointers, as in ps_render).
#pragma once

#include <algorithm>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pedalsky_hip.h"

namespace pedalsky {

// A failed library call: the status and ps_last_error()'s message.
class Error : public std::runtime_error {
public:
    Error(ps_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    ps_status status;
};

inline void check(ps_status s, const char* call) {
    if (s != PS_OK) throw Error(s, std::string(call) + ": " + ps_last_error());
}

// One device context (the reference's GpuContext plus the AtmosphereLuts
// bundle both subsystems share: here the LUTs live in the context).
class Device {
public:
    explicit Device(int32_t device = 0) { check(ps_ctx_create(device, &ctx_), "ps_ctx_create"); }
    ~Device() {
        if (ctx_) ps_ctx_destroy(ctx_);
    }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    ps_ctx* get() const { return ctx_; }
    void synchronize() const { check(ps_ctx_synchronize(ctx_), "ps_ctx_synchronize"); }

private:
    ps_ctx* ctx_ = nullptr;
};

// subsystem.rs:15-31, in increasing dependency order.
enum class PassStage { Compute, SkyBackdrop, Opaque, Translucent, PostProcess, ToneMap, Overlay };

using PassId = uint32_t;

// subsystem.rs:51-59
struct PassDescriptor {
    const char* name;
    PassStage stage;
    PassId id;
};

// The parts of ps_core's PrepareContext these subsystems read: the frame
// uniforms every frame, the synthesised weather state (atmosphere, cloud
// layers and winds, the weather textures) tagged with WeatherState's
// revision counter, which bumps when synthesis reruns.
struct PrepareContext {
    const ps_frame_uniforms* frame = nullptr;
    uint64_t weather_revision = 0;
    const ps_atmosphere_params* atmosphere = nullptr;
    const ps_cloud_layer* cloud_layers = nullptr;
    int32_t n_cloud_layers = 0;
    const float* layer_wind_xz = nullptr;        // 2 per layer, or NULL (calm)
    const uint16_t* weather_f16 = nullptr;       // w x h x channels binary16 texels, or NULL (keep)
    int32_t weather_w = 0, weather_h = 0, weather_channels = 0;
    const uint8_t* cloud_type_grid_128 = nullptr;
    const uint8_t* overcast_mask_128 = nullptr;
};

// The parts of RenderContext a dispatch reads: the scene depth (device, W*H
// reverse-Z, or NULL: every ray reaches the far side) and the HDR targets
// (device images as ps_render takes them).
struct RenderContext {
    const float* depth = nullptr;
    ps_image_out out{};
};

// subsystem.rs:66-118
class RenderSubsystem {
public:
    virtual ~RenderSubsystem() = default;
    virtual const char* name() const = 0;
    virtual void prepare(PrepareContext& ctx) = 0;
    virtual std::vector<PassDescriptor> register_passes() const = 0;
    virtual void dispatch_pass(PassId id, const RenderContext& ctx) = 0;
};

// AtmosphereSubsystem (crates/ps-atmosphere/src/lib.rs:454-658): four LUT
// bakes in the Compute stage and the sky in SkyBackdrop. The transmittance
// and multi-scatter LUTs depend only on the atmosphere: they bake while the
// static-dirty bit is set (set on creation, on reconfigure and when the
// weather revision brings a new atmosphere) and the multi-scatter pass
// clears it; with multi-scattering off the multi-scatter LUT is zeroed
// instead. Sky-view and aerial perspective bake every frame.
class AtmosphereSubsystem : public RenderSubsystem {
public:
    enum : PassId { PASS_TRANSMITTANCE = 0, PASS_MULTISCATTER, PASS_SKYVIEW, PASS_AERIAL_PERSPECTIVE, PASS_SKY };

    explicit AtmosphereSubsystem(Device& dev) : dev_(dev) {}

    const char* name() const override { return "atmosphere"; }

    // Config.render.atmosphere.multi_scattering; marks the static LUTs dirty
    // (lib.rs:646-657).
    void reconfigure(bool multi_scattering) {
        multi_scattering_ = multi_scattering;
        static_dirty_ = true;
    }

    // The sky is drawn by the cloud march's fused composite when a clouds
    // subsystem shares the device; otherwise the sky pass renders it alone.
    void set_sky_drawn_by_clouds(bool v) { sky_by_clouds_ = v; }

    void prepare(PrepareContext& ctx) override {
        if (!ctx.frame) throw std::invalid_argument("atmosphere: PrepareContext.frame is NULL");
        check(ps_set_frame(dev_.get(), ctx.frame), "ps_set_frame");
        if (!have_revision_ || ctx.weather_revision != revision_) {
            if (!ctx.atmosphere) throw std::invalid_argument("atmosphere: PrepareContext.atmosphere is NULL");
            check(ps_set_atmosphere(dev_.get(), ctx.atmosphere), "ps_set_atmosphere");
            revision_ = ctx.weather_revision;
            have_revision_ = true;
            static_dirty_ = true;
        }
    }

    std::vector<PassDescriptor> register_passes() const override {
        return {{"atmosphere::transmittance", PassStage::Compute, PASS_TRANSMITTANCE},
                {"atmosphere::multiscatter", PassStage::Compute, PASS_MULTISCATTER},
                {"atmosphere::skyview", PassStage::Compute, PASS_SKYVIEW},
                {"atmosphere::aerialperspective", PassStage::Compute, PASS_AERIAL_PERSPECTIVE},
                {"atmosphere::sky", PassStage::SkyBackdrop, PASS_SKY}};
    }

    void dispatch_pass(PassId id, const RenderContext& ctx) override {
        switch (id) {
            case PASS_TRANSMITTANCE:
                if (static_dirty_) check(ps_bake_luts(dev_.get(), PS_LUT_TRANSMITTANCE), "ps_bake_luts");
                break;
            case PASS_MULTISCATTER:
                if (!static_dirty_) break;
                check(ps_bake_luts(dev_.get(), multi_scattering_ ? (uint32_t)PS_LUT_MULTISCATTER
                                                                 : (uint32_t)(PS_LUT_MULTISCATTER |
                                                                              PS_LUT_MULTISCATTER_DISABLED)),
                      "ps_bake_luts");
                static_dirty_ = false;
                break;
            case PASS_SKYVIEW:
                check(ps_bake_luts(dev_.get(), PS_LUT_SKYVIEW), "ps_bake_luts");
                break;
            case PASS_AERIAL_PERSPECTIVE:
                check(ps_bake_luts(dev_.get(), PS_LUT_AERIAL_PERSPECTIVE), "ps_bake_luts");
                break;
            case PASS_SKY:
                if (!sky_by_clouds_) {
                    // the sky alone: a march over no layers composites the sky pass
                    ps_cloud_params p;
                    ps_default_cloud_params(&p);
                    check(ps_set_clouds(dev_.get(), &p, nullptr, 0, nullptr), "ps_set_clouds");
                    check(ps_render(dev_.get(), ctx.depth, &ctx.out), "ps_render");
                }
                break;
            default:
                throw std::logic_error("atmosphere: unknown pass id " + std::to_string(id));
        }
    }

    bool static_dirty() const { return static_dirty_; }

private:
    Device& dev_;
    bool static_dirty_ = true;
    bool multi_scattering_ = true;
    bool sky_by_clouds_ = false;
    bool have_revision_ = false;
    uint64_t revision_ = 0;
};

// CloudsSubsystem (crates/ps-clouds/src/lib.rs:1100-1260): the march, TAA
// and composite passes in the Translucent stage. This library fuses the sky
// pass, the march, the TAA resolve (when on) and the dual-source composite
// into the march dispatch, so the TAA and composite passes have nothing left
// to do. The layers (and the weather textures) are uploaded when the weather
// revision changes; the cloud parameters every frame, as the reference
// uploads its params uniform.
class CloudsSubsystem : public RenderSubsystem {
public:
    enum : PassId { PASS_MARCH = 0, PASS_TAA, PASS_COMPOSITE };

    CloudsSubsystem(Device& dev, const ps_cloud_params& params) : dev_(dev), params_(params) {}

    const char* name() const override { return "clouds"; }

    // lib.rs:701-712: clamped to [8, 256] and [1, 16]; from the next frame on.
    void set_march_steps(uint32_t cloud, uint32_t light) {
        params_.cloud_steps = std::clamp(cloud, 8u, 256u);
        params_.light_steps = std::clamp(light, 1u, 16u);
    }
    // clouds.temporal_taa && !freeze_time (lib.rs:119-142)
    void set_temporal_taa(bool on) { check(ps_set_taa(dev_.get(), on ? 1 : 0), "ps_set_taa"); }
    // lib.rs:726-734
    void invalidate_taa_history() { check(ps_invalidate_taa_history(dev_.get()), "ps_invalidate_taa_history"); }
    // lib.rs:716-724
    void set_active_view(int32_t view) { check(ps_set_active_view(dev_.get(), view), "ps_set_active_view"); }
    const ps_cloud_params& params() const { return params_; }

    void prepare(PrepareContext& ctx) override {
        if (!have_revision_ || ctx.weather_revision != revision_) {
            layers_.assign(ctx.cloud_layers, ctx.cloud_layers + std::max(ctx.n_cloud_layers, 0));
            winds_.clear();
            if (ctx.layer_wind_xz) winds_.assign(ctx.layer_wind_xz, ctx.layer_wind_xz + 2 * layers_.size());
            if (ctx.weather_f16)
                check(ps_upload_weather(dev_.get(), ctx.weather_f16, ctx.weather_w, ctx.weather_h,
                                        ctx.weather_channels, ctx.cloud_type_grid_128, ctx.overcast_mask_128),
                      "ps_upload_weather");
            revision_ = ctx.weather_revision;
            have_revision_ = true;
        }
        check(ps_set_clouds(dev_.get(), &params_, layers_.empty() ? nullptr : layers_.data(), (int32_t)layers_.size(),
                            winds_.empty() ? nullptr : winds_.data()),
              "ps_set_clouds");
    }

    std::vector<PassDescriptor> register_passes() const override {
        return {{"clouds::march", PassStage::Translucent, PASS_MARCH},
                {"clouds::taa", PassStage::Translucent, PASS_TAA},
                {"clouds::composite", PassStage::Translucent, PASS_COMPOSITE}};
    }

    void dispatch_pass(PassId id, const RenderContext& ctx) override {
        switch (id) {
            case PASS_MARCH:
                check(ps_render(dev_.get(), ctx.depth, &ctx.out), "ps_render");
                break;
            case PASS_TAA:
            case PASS_COMPOSITE:
                break;  // done by the fused march dispatch
            default:
                throw std::logic_error("clouds: unknown pass id " + std::to_string(id));
        }
    }

private:
    Device& dev_;
    ps_cloud_params params_;
    std::vector<ps_cloud_layer> layers_;
    std::vector<float> winds_;
    bool have_revision_ = false;
    uint64_t revision_ = 0;
};

// The executor's frame: every subsystem's prepare in registration order,
// then every registered pass, stable-sorted by stage (registration order
// within a stage).
class Executor {
public:
    struct Entry {
        PassDescriptor pass;
        RenderSubsystem* owner;
    };

    void add(RenderSubsystem* s) {
        subsystems_.push_back(s);
        rebuild();
    }
    const std::vector<Entry>& passes() const { return passes_; }

    void frame(PrepareContext& prep, const RenderContext& rc) {
        for (RenderSubsystem* s : subsystems_) s->prepare(prep);
        for (const Entry& e : passes_) e.owner->dispatch_pass(e.pass.id, rc);
    }

private:
    void rebuild() {
        passes_.clear();
        for (RenderSubsystem* s : subsystems_)
            for (const PassDescriptor& p : s->register_passes()) passes_.push_back({p, s});
        std::stable_sort(passes_.begin(), passes_.end(),
                         [](const Entry& a, const Entry& b) { return a.pass.stage < b.pass.stage; });
    }
    std::vector<RenderSubsystem*> subsystems_;
    std::vector<Entry> passes_;
};

}  // namespace pedalsky
