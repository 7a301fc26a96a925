// Drives the C++ host layer (include/pedalsky_subsystems.hpp) the way the
// reference's executor drives its subsystems, for tests/test_host_subsystems.py.
//
//   subsystem_frame --selftest
//       pass ordering of the executor and the error contract, no device
//   subsystem_frame --scene FILE --out FILE [--sky-only]
//       loads a scene written by the test (uniform structs, weather texels,
//       blue noise, depth), renders two frames through the executor (the
//       second with the same weather revision: the static LUTs are not baked
//       again) and writes the second frame's composite (H x W x 4 float32)
//
// Scene file: "PSSCENE1", then frame (448 B), atmosphere (144 B), cloud
// params (112 B), int32 n_layers, layers (n x 48 B), winds (n x 2 float32),
// int32 noise seed, int32 noise format, lut dims (ps_lut_dims), int32
// weather w, h, channels, weather texels (uint16), type grid flag + 128x128
// uint8, overcast 128x128 uint8, blue noise 64x64 uint8, int32 W, H, int32
// depth flag, depth (W x H float32).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "pedalsky_subsystems.hpp"

namespace {

int fails = 0;
#define EXPECT(cond)                                                                  \
    do {                                                                              \
        if (!(cond)) {                                                                \
            std::fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__, __LINE__, #cond); \
            ++fails;                                                                  \
        }                                                                             \
    } while (0)

// a subsystem that records its calls (no device)
struct Probe : pedalsky::RenderSubsystem {
    Probe(const char* n, std::vector<pedalsky::PassDescriptor> p, std::vector<std::string>* log)
        : n_(n), p_(std::move(p)), log_(log) {}
    const char* name() const override { return n_; }
    void prepare(pedalsky::PrepareContext&) override { log_->push_back(std::string("prepare ") + n_); }
    std::vector<pedalsky::PassDescriptor> register_passes() const override { return p_; }
    void dispatch_pass(pedalsky::PassId id, const pedalsky::RenderContext&) override {
        for (auto& d : p_)
            if (d.id == id) {
                log_->push_back(d.name);
                return;
            }
        throw std::logic_error("probe: unknown pass id");
    }
    const char* n_;
    std::vector<pedalsky::PassDescriptor> p_;
    std::vector<std::string>* log_;
};

int selftest() {
    using pedalsky::PassStage;
    std::vector<std::string> log;
    Probe clouds("clouds", {{"clouds::march", PassStage::Translucent, 0}, {"clouds::composite", PassStage::Translucent, 2}},
                 &log);
    Probe atm("atmosphere", {{"atmosphere::transmittance", PassStage::Compute, 0},
                             {"atmosphere::sky", PassStage::SkyBackdrop, 4},
                             {"atmosphere::skyview", PassStage::Compute, 2}},
              &log);
    Probe tone("tonemap", {{"tonemap", PassStage::ToneMap, 0}, {"bloom", PassStage::PostProcess, 1}}, &log);
    pedalsky::Executor ex;
    ex.add(&clouds);  // registration order: clouds before atmosphere
    ex.add(&atm);
    ex.add(&tone);
    pedalsky::PrepareContext pc;
    pedalsky::RenderContext rc;
    ex.frame(pc, rc);
    const std::vector<std::string> want = {"prepare clouds", "prepare atmosphere", "prepare tonemap",
                                           "atmosphere::transmittance", "atmosphere::skyview", "atmosphere::sky",
                                           "clouds::march", "clouds::composite", "bloom", "tonemap"};
    EXPECT(log == want);
    // the error contract: a failing call throws Error with status and message
    try {
        pedalsky::check(ps_set_march_steps(nullptr, 64, 6), "ps_set_march_steps");
        EXPECT(false);
    } catch (const pedalsky::Error& e) {
        EXPECT(e.status == PS_ERR_INVALID_ARGUMENT);
        EXPECT(std::string(e.what()).find("ps_set_march_steps") != std::string::npos);
    }
    if (fails == 0) std::printf("selftest ok\n");
    return fails == 0 ? 0 : 1;
}

template <class T>
void rd(std::ifstream& f, T* p, size_t n = 1) {
    f.read(reinterpret_cast<char*>(p), sizeof(T) * n);
    if (!f) throw std::runtime_error("scene file truncated");
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

int render(const std::string& scene, const std::string& out_path, bool sky_only) {
    std::ifstream f(scene, std::ios::binary);
    char magic[8];
    rd(f, magic, 8);
    if (std::memcmp(magic, "PSSCENE1", 8) != 0) throw std::runtime_error("not a scene file");
    ps_frame_uniforms frame;
    ps_atmosphere_params atm;
    ps_cloud_params cp;
    rd(f, &frame);
    rd(f, &atm);
    rd(f, &cp);
    int32_t nl;
    rd(f, &nl);
    std::vector<ps_cloud_layer> layers(nl);
    std::vector<float> winds(2 * nl);
    if (nl) {
        rd(f, layers.data(), nl);
        rd(f, winds.data(), 2 * nl);
    }
    int32_t seed, fmt;
    rd(f, &seed);
    rd(f, &fmt);
    ps_lut_dims dims;
    rd(f, &dims);
    int32_t ww, wh, wc;
    rd(f, &ww);
    rd(f, &wh);
    rd(f, &wc);
    std::vector<uint16_t> weather((size_t)ww * wh * wc);
    rd(f, weather.data(), weather.size());
    int32_t has_grid;
    rd(f, &has_grid);
    std::vector<uint8_t> grid(128 * 128), overcast(128 * 128), blue(64 * 64);
    rd(f, grid.data(), grid.size());
    rd(f, overcast.data(), overcast.size());
    rd(f, blue.data(), blue.size());
    int32_t W, H, has_depth;
    rd(f, &W);
    rd(f, &H);
    rd(f, &has_depth);
    std::vector<float> depth;
    if (has_depth) {
        depth.resize((size_t)W * H);
        rd(f, depth.data(), depth.size());
    }

    // the factories' one-time setup (AtmosphereFactory / CloudsFactory::build)
    pedalsky::Device dev(0);
    pedalsky::check(ps_ctx_set_lut_dims(dev.get(), &dims), "ps_ctx_set_lut_dims");
    pedalsky::check(ps_generate_noise(dev.get(), (uint32_t)seed, fmt), "ps_generate_noise");
    pedalsky::check(ps_upload_blue_noise(dev.get(), blue.data()), "ps_upload_blue_noise");
    if (sky_only)  // without a clouds subsystem the weather textures (overcast mask) come with setup
        pedalsky::check(ps_upload_weather(dev.get(), weather.data(), ww, wh, wc, has_grid ? grid.data() : nullptr,
                                          overcast.data()),
                        "ps_upload_weather");
    float* d_depth = nullptr;
    float* d_out = nullptr;
    const size_t px = (size_t)W * H;
    if (has_depth) {
        hip_check(hipMalloc(&d_depth, px * sizeof(float)), "hipMalloc depth");
        hip_check(hipMemcpy(d_depth, depth.data(), px * sizeof(float), hipMemcpyHostToDevice), "upload depth");
    }
    hip_check(hipMalloc(&d_out, px * 4 * sizeof(float)), "hipMalloc out");

    pedalsky::AtmosphereSubsystem atmosphere(dev);
    pedalsky::CloudsSubsystem clouds(dev, cp);
    pedalsky::Executor ex;
    ex.add(&atmosphere);
    if (!sky_only) {
        ex.add(&clouds);
        atmosphere.set_sky_drawn_by_clouds(true);
    }
    // the executor's order: four LUT bakes, the sky, then the clouds' passes
    const char* want[] = {"atmosphere::transmittance", "atmosphere::multiscatter", "atmosphere::skyview",
                          "atmosphere::aerialperspective", "atmosphere::sky", "clouds::march", "clouds::taa",
                          "clouds::composite"};
    EXPECT(ex.passes().size() == (sky_only ? 5u : 8u));
    for (size_t i = 0; i < ex.passes().size(); ++i) EXPECT(std::strcmp(ex.passes()[i].pass.name, want[i]) == 0);

    pedalsky::PrepareContext pc;
    pc.frame = &frame;
    pc.weather_revision = 1;
    pc.atmosphere = &atm;
    pc.cloud_layers = layers.data();
    pc.n_cloud_layers = nl;
    pc.layer_wind_xz = nl ? winds.data() : nullptr;
    pc.weather_f16 = weather.data();
    pc.weather_w = ww;
    pc.weather_h = wh;
    pc.weather_channels = wc;
    pc.cloud_type_grid_128 = has_grid ? grid.data() : nullptr;
    pc.overcast_mask_128 = overcast.data();
    pedalsky::RenderContext rc;
    rc.depth = d_depth;
    rc.out.composite = d_out;
    rc.out.format = PS_PIXEL_F32;

    EXPECT(atmosphere.static_dirty());
    ex.frame(pc, rc);
    EXPECT(!atmosphere.static_dirty());  // latched off by the multi-scatter pass
    ex.frame(pc, rc);                    // same revision: static LUTs kept
    EXPECT(!atmosphere.static_dirty());
    dev.synchronize();

    // the clamps of set_march_steps and the unknown-pass contract
    pedalsky::CloudsSubsystem probe(dev, cp);
    probe.set_march_steps(1000, 0);
    EXPECT(probe.params().cloud_steps == 256 && probe.params().light_steps == 1);
    try {
        atmosphere.dispatch_pass(99, rc);
        EXPECT(false);
    } catch (const std::logic_error&) {
    }
    try {
        clouds.dispatch_pass(7, rc);
        EXPECT(false);
    } catch (const std::logic_error&) {
    }

    std::vector<float> host(px * 4);
    hip_check(hipMemcpy(host.data(), d_out, host.size() * sizeof(float), hipMemcpyDeviceToHost), "download");
    std::ofstream o(out_path, std::ios::binary);
    o.write(reinterpret_cast<const char*>(host.data()), host.size() * sizeof(float));
    if (d_depth) (void)hipFree(d_depth);
    (void)hipFree(d_out);
    if (fails == 0) std::printf("frame ok\n");
    return fails == 0 ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc >= 2 && std::string(argv[1]) == "--selftest") return selftest();
        std::string scene, out;
        bool sky_only = false;
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            if (a == "--scene" && i + 1 < argc) scene = argv[++i];
            else if (a == "--out" && i + 1 < argc) out = argv[++i];
            else if (a == "--sky-only") sky_only = true;
        }
        if (scene.empty() || out.empty()) {
            std::fprintf(stderr, "usage: %s --selftest | --scene FILE --out FILE [--sky-only]\n", argv[0]);
            return 2;
        }
        return render(scene, out, sky_only);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
