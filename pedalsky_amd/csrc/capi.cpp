// C ABI of the sky + cloud transport library: context lifecycle, uploads,
// per-frame parameter folding (host FP64) and kernel dispatch on one HIP
// stream. See include/pedalsky_hip.h for the contract of each entry point.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include "ps_host.h"
#include "ps_launch.h"

namespace psi {

namespace {
thread_local std::string g_last_error = "no error";
}

ps_status fail(ps_status code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

}  // namespace psi

using namespace psi;

static_assert(sizeof(ps_frame_uniforms) == 448, "FrameUniformsGpu is 448 B (frame_uniforms.rs:195-201)");
static_assert(sizeof(ps_atmosphere_params) == 144, "AtmosphereParams is 144 B (weather.rs:544-555)");
static_assert(sizeof(ps_cloud_params) == 112, "CloudParamsGpu is 112 B (params.rs:221-231)");
static_assert(sizeof(ps_cloud_layer) == 48, "CloudLayerGpu is 48 B (weather.rs:544-555)");
static_assert(sizeof(ps_surface_params) == 64, "SurfaceParams is 64 B (weather.rs:541-554)");
static_assert(offsetof(ps_surface_params, material) == 48, "SurfaceParams.material offset");
static_assert(offsetof(ps_frame_uniforms, inv_view_proj) == 192, "inv_view_proj offset");
static_assert(offsetof(ps_frame_uniforms, camera_position_world) == 256, "camera offset");
static_assert(offsetof(ps_frame_uniforms, sun_direction) == 288, "sun_direction offset");
static_assert(offsetof(ps_frame_uniforms, viewport_size) == 352, "viewport offset");
static_assert(offsetof(ps_frame_uniforms, time_seconds) == 368, "time offset");
static_assert(offsetof(ps_frame_uniforms, prev_view_proj) == 384, "prev_view_proj offset");
static_assert(offsetof(ps_atmosphere_params, rayleigh_scattering) == 16, "rayleigh offset");
static_assert(offsetof(ps_atmosphere_params, mie_g) == 64, "mie_g offset");
static_assert(offsetof(ps_atmosphere_params, ozone_absorption) == 80, "ozone offset");
static_assert(offsetof(ps_atmosphere_params, ozone_center_m) == 96, "ozone centre offset");
static_assert(offsetof(ps_atmosphere_params, ground_albedo) == 112, "albedo offset");
static_assert(offsetof(ps_cloud_params, sigma_a) == 16, "sigma_a offset");
static_assert(offsetof(ps_cloud_params, droplet_diameter_bias) == 28, "droplet bias offset");
static_assert(offsetof(ps_cloud_params, cone_light_sampling) == 32, "cone offset");
static_assert(offsetof(ps_cloud_params, powder_strength) == 48, "powder offset");
static_assert(offsetof(ps_cloud_params, ambient_strength) == 64, "ambient offset");
static_assert(offsetof(ps_cloud_params, light_steps) == 80, "light_steps offset");
static_assert(offsetof(ps_cloud_params, temporal_jitter) == 96, "temporal_jitter offset");
static_assert(offsetof(ps_cloud_layer, cloud_type) == 16, "cloud_type offset");
static_assert(offsetof(ps_cloud_layer, droplet_diameter_um) == 32, "droplet offset");



namespace {

size_t noise_texel_bytes(int format, int channels) { return (format == PS_NOISE_UNORM8 ? 1u : 2u) * channels; }

ps_status ensure_luts(ps_ctx* c) {
    const ps_lut_dims& d = c->dims;
    ps_status s;
    if ((s = dev_alloc(&c->luts.T, sizeof(float4) * d.transmittance_w * d.transmittance_h, "transmittance"))) return s;
    if ((s = dev_alloc(&c->luts.MS, sizeof(float4) * d.multiscatter_w * d.multiscatter_h, "multiscatter"))) return s;
    if ((s = dev_alloc(&c->luts.SV, sizeof(float4) * d.skyview_w * d.skyview_h, "skyview"))) return s;
    if ((s = dev_alloc(&c->luts.AP, sizeof(float4) * d.ap_w * d.ap_h * d.ap_d, "aerial perspective"))) return s;
    return PS_OK;
}

}  // namespace

void psi::fill_lut_params(const SceneIn& sc, const ps_lut_dims& d, LutParams& P) {
    const ps_atmosphere_params& a = sc.atm;
    P.a.R = a.planet_radius_m; P.a.top = a.atmosphere_top_m;
    P.a.hr = a.rayleigh_scale_height_m; P.a.hm = a.mie_scale_height_m;
    for (int i = 0; i < 3; ++i) {
        P.a.ray_s[i] = a.rayleigh_scattering[i];
        P.a.mie_s[i] = a.mie_scattering[i];
        P.a.mie_a[i] = a.mie_absorption[i];
        P.a.ozone[i] = a.ozone_absorption[i];
        P.a.albedo[i] = a.ground_albedo[i];
    }
    P.a.mie_g = a.mie_g; P.a.oz_c = a.ozone_center_m; P.a.oz_t = a.ozone_thickness_m;
    const ps_frame_uniforms& f = sc.frame;
    for (int i = 0; i < 16; ++i) P.f.inv_vp[i] = f.inv_view_proj[i];
    for (int i = 0; i < 3; ++i) {
        P.f.cam[i] = f.camera_position_world[i];
        P.f.sun[i] = f.sun_direction[i];
        P.f.E[i] = f.sun_illuminance[i];
    }
    P.tw = d.transmittance_w; P.th = d.transmittance_h;
    P.mw = d.multiscatter_w; P.mh = d.multiscatter_h;
    P.sw = d.skyview_w; P.sh = d.skyview_h;
    P.aw = d.ap_w; P.ah = d.ap_h; P.ad = d.ap_d;
    P.ms_disabled = 0;
}

namespace {

double smoothstep_d(double e0, double e1, double x) {
    double t = std::min(std::max((x - e0) / (e1 - e0), 0.0), 1.0);
    return t * t * (3.0 - 2.0 * t);
}

// Smallest common multiple of the periods p[0..n) that is a whole multiple
// of p[0] (n p[0], n <= 2^20) and at most 4096 km, to a relative 1e-9;
// 0 when there is none.
double common_period(const double* p, int n) {
    for (int k = 1; k <= (1 << 20); ++k) {
        const double L = k * p[0];
        if (L > 4.096e6) break;
        bool ok = true;
        for (int i = 1; i < n && ok; ++i) {
            const double q = L / p[i];
            ok = std::fabs(q - std::nearbyint(q)) <= 1e-9 * q;
        }
        if (ok) return L;
    }
    return 0.0;
}

}  // namespace

ps_status psi::fill_render_params(const ps_ctx* c, const SceneIn& sc, const LutSet& luts, const float* depth,
                                  const ps_image_out* out, RenderParams& P) {
    const ps_frame_uniforms& f = sc.frame;
    const ps_atmosphere_params& a = sc.atm;
    const ps_cloud_params& cp = sc.clouds;
    std::memset(&P, 0, sizeof(P));
    float vw = f.viewport_size[0], vh = f.viewport_size[1];
    if (!(vw >= 1.f && vh >= 1.f && vw <= 32768.f && vh <= 32768.f) || vw != std::floor(vw) || vh != std::floor(vh))
        return fail(PS_ERR_INVALID_ARGUMENT, "frame.viewport_size must hold integer width/height in [1, 32768]");
    P.W = (int)vw; P.H = (int)vh;
    P.full_W = P.W; P.full_H = P.H;
    P.rt_sx = 1.0; P.rt_sy = 1.0;
    P.inv_W = 1.f / vw; P.inv_H = 1.f / vh;
    for (int i = 0; i < 16; ++i) P.inv_vp[i] = f.inv_view_proj[i];
    const double R = a.planet_radius_m, top = a.atmosphere_top_m;
    double cam[3] = {f.camera_position_world[0], f.camera_position_world[1], f.camera_position_world[2]};
    double sun[3] = {f.sun_direction[0], f.sun_direction[1], f.sun_direction[2]};
    for (int i = 0; i < 3; ++i) {
        P.cam[i] = cam[i];
        P.sun_d[i] = sun[i];
        P.sun[i] = f.sun_direction[i];
        P.E[i] = f.sun_illuminance[i];
        P.cam_f[i] = f.camera_position_world[i];
    }
    P.sun_ang = f.sun_direction[3];
    P.cos_disk = std::cos((double)f.sun_direction[3]);
    P.toa = f.sun_illuminance[3];
    P.R = (float)R; P.top = (float)top;
    P.big_h = (float)std::sqrt(std::max(top * top - R * R, 0.0));
    P.top_minus_R = (float)(top - R);
    // camera radius and altitude without cancellation
    double oy = cam[1] + R;
    double r_cam = std::sqrt(cam[0] * cam[0] + oy * oy + cam[2] * cam[2]);
    double h_cam = (cam[0] * cam[0] + cam[2] * cam[2] + cam[1] * (cam[1] + 2.0 * R)) / (r_cam + R);
    P.r0 = r_cam;
    P.h0 = h_cam;
    P.cam_alt = (float)h_cam;
    double cam_r = std::max(r_cam, R + 1.0);
    P.cam_zha = (float)(PS_PI - std::asin(std::min(std::max(R / cam_r, 0.0), 1.0)));
    P.sun_az = (float)std::atan2(sun[0], sun[2]);
    {
        double hf = smoothstep_d(-0.09, 0.05, sun[1]);
        double saf = std::sqrt(std::min(std::max(sun[1], 0.05), 1.0)) * hf;
        P.overcast_lum = (float)((double)f.sun_illuminance[3] * saf * 0.25);
    }
    // clouds
    int lc = (int)std::min<uint32_t>(cp.cloud_layer_count, PS_MAX_CLOUD_LAYERS);
    lc = std::min(lc, sc.n_layers);
    P.n_layers = lc;
    P.steps = (int)cp.cloud_steps;
    P.light_steps = (int)cp.light_steps;
    P.octaves = (int)cp.multi_scatter_octaves;
    P.cone = cp.cone_light_sampling ? 1 : 0;
    P.temporal_jitter = cp.temporal_jitter ? 1 : 0;
    P.frame_index = f.frame_index;
    P.powder = cp.powder_strength;
    P.forward_bias = cp.forward_bias;
    P.ambient = cp.ambient_strength;
    P.curl_off_scale = cp.curl_strength * cp.detail_scale_m;
    P.inv_detail_scale = (float)(1.0 / std::max((double)cp.detail_scale_m, 1.0));
    P.inv_base_scale = (float)(1.0 / std::max((double)cp.base_scale_m, 1.0));
    P.inv_weather_scale = (float)(1.0 / std::max((double)cp.weather_scale_m, 1.0));
    {
        const double base = std::max((double)cp.base_scale_m, 1.0), det = std::max((double)cp.detail_scale_m, 1.0),
                     wea = std::max((double)cp.weather_scale_m, 1.0);
        const double xz[4] = {base, det, wea, 32000.0};  // 32 km: the type grid (128 cells of 250 m)
        P.wrap_xz = common_period(xz, 4);
        P.wrap_y = common_period(xz, 2);
    }
    for (int i = 0; i < 3; ++i) P.sigma_s[i] = cp.sigma_s[i];
    P.ms_a = cp.multi_scatter_a; P.ms_b = cp.multi_scatter_b; P.ms_c = cp.multi_scatter_c;
    P.droplet_bias = cp.droplet_diameter_bias;
    {
        double an = 1, bn = 1, cn = 1;
        for (int n = 0; n < PS_MAX_OCTAVES; ++n) {
            P.a_pow[n] = (float)an; P.b_pow[n] = (float)bn; P.c_pow[n] = (float)cn;
            an *= cp.multi_scatter_a; bn *= cp.multi_scatter_b; cn *= cp.multi_scatter_c;
        }
    }
    for (int i = 0; i < 3; ++i) P.light_rgb[i] = f.lightning_illuminance[i];
    P.light_radius = f.lightning_illuminance[3];
    P.light_x = f.lightning_origin_world[0];
    P.light_z = f.lightning_origin_world[2];
    const double drift = (double)cp.wind_drift_strength * (double)f.simulated_seconds;
    for (int l = 0; l < lc; ++l) {
        const ps_cloud_layer& L = sc.layers[l];
        LayerC& oc = P.layers[l];
        LayerF& o = oc.f;
        o.base = L.base_m; o.top = L.top_m;
        o.inv_thick = (float)(1.0 / std::max((double)L.top_m - (double)L.base_m, 1.0));
        o.coverage = L.coverage; o.density_scale = L.density_scale;
        double diurnal = 0.0;
        if (cp.diurnal_strength != 0.f) {
            uint32_t t = L.cloud_type;
            bool conv = t == 0 || t == 2 || t == 3 || t == 7;
            if (conv) diurnal = smoothstep_d(-0.1, 0.4, sun[1]) * cp.diurnal_strength;
        }
        o.esb = (float)(L.shape_bias * diurnal);
        o.ld = (float)std::min(std::max((double)cp.detail_strength + L.detail_bias * diurnal, 0.0), 1.0);
        o.anvil = L.anvil_bias;
        // chromatic_mie_modulation (cloud_march.wgsl:122-145) on sigma_s + sigma_a
        double t = std::min(std::max(1.0 - L.droplet_diameter_um / 20.0, 0.0), 1.0);
        double mod[3] = {1, 1, 1};
        if (t > 0.0) {
            double st = 0.25 * t * t;
            mod[0] = 1.0 - 0.5 * st; mod[2] = 1.0 + st;
        }
        for (int k = 0; k < 3; ++k) {
            const double st0 = ((double)cp.sigma_s[k] + cp.sigma_a[k]) * mod[k];
            oc.sigma_t0[k] = (float)st0;
            o.sigma_t[k] = (float)(st0 * (double)sc.kappa[l]);
        }
        double wx = sc.wind[2 * l], wz = sc.wind[2 * l + 1];
        o.wind_off_x = (float)(wx * drift);
        o.wind_off_z = (float)(wz * drift);
        o.skew_x = 0.f; o.skew_z = 0.f;
        if (cp.wind_skew_strength != 0.f) {
            double speed = std::sqrt(wx * wx + wz * wz);
            if (speed >= 0.1) {
                double thick = std::max((double)L.top_m - L.base_m, 1.0) * cp.wind_skew_strength;
                o.skew_x = (float)(wx / speed * thick);
                o.skew_z = (float)(wz / speed * thick);
            }
        }
        o.type = L.cloud_type;
        {
            static constexpr float ndf_rows[9][5] = PS_NDF_ROWS;
            const float* row = ndf_rows[L.cloud_type < 8u ? L.cloud_type : 8u];
            o.ndf_ia = row[0]; o.ndf_e0 = row[1]; o.ndf_ie = row[2]; o.ndf_a = row[3]; o.ndf_b = row[4];
        }
        oc.is_cb = L.cloud_type == 7u ? 1u : 0u;
        oc.droplet = L.droplet_diameter_um;
        oc.r_bottom = R + (double)L.base_m;
        oc.r_top = R + (double)L.top_m;
        double d = std::min(std::max((double)L.droplet_diameter_um * cp.droplet_diameter_bias, 5.0), 50.0);
        double xg = -0.0990567 / (d - 1.67154);
        oc.g_hg = (float)std::exp(xg);
        oc.omg_hg = (float)-std::expm1(xg);
        oc.g_d = (float)(std::exp(-2.20679 / (d + 3.91029)) - 0.428934);
        oc.alpha = (float)std::exp(3.62489 - 8.29288 / (d + 5.52825));
        oc.w_d = (float)std::exp(-0.599085 / (d - 0.641583) - 0.665888);
        oc.ice = (float)smoothstep_d(35.0, 50.0, d);
    }
    // resources
    const ps_lut_dims& dm = c->dims;
    P.T = luts.T; P.tw = dm.transmittance_w; P.th = dm.transmittance_h;
    P.SV = luts.SV; P.sw = dm.skyview_w; P.sh = dm.skyview_h;
    P.AP = luts.AP; P.aw = dm.ap_w; P.ah = dm.ap_h; P.ad = dm.ap_d;
    P.base_rl = c->base_rl; P.bbound = c->bbound; P.detail_hf = c->detail_hf; P.curl = c->curl2;
    base_fold_scales(c->noise_format, P.base_rs, P.base_ls);
    P.weather = c->weather; P.ww = c->ww; P.wh = c->wh;
    P.weather_pow2 = (c->ww & (c->ww - 1)) == 0 && (c->wh & (c->wh - 1)) == 0 ? 1 : 0;
    P.ww_log2 = 0;
    while ((1 << P.ww_log2) < c->ww) ++P.ww_log2;
    P.ww_f = (float)c->ww; P.wh_f = (float)c->wh;
    P.type_grid = c->type_grid; P.use_type_grid = c->type_grid_active ? 1 : 0;
    P.overcast = c->overcast_active ? c->overcast : nullptr;
    P.blue = c->blue;
    P.depth = depth; P.has_depth = depth ? 1 : 0;
    if (out) {
        P.out_comp = out->composite; P.out_lum = out->cloud_luminance; P.out_trans = out->cloud_transmittance;
        P.out_f16 = out->format == PS_PIXEL_F16 ? 1 : 0;
    }
    P.counters = c->d_counters;
    P.delta = c->delta;
    P.delta_bnd = c->delta_bnd;
    P.delta_su = (float)(128.0 / std::max((double)cp.weather_scale_m, 1.0));
    P.n_tiles_x = (P.W + 7) / 8;
    return PS_OK;
}

ps_status psi::check_shared_inputs(const ps_ctx* c) {
    if (!c->have_noise) return fail(PS_ERR_NOT_READY, "no noise volumes: call ps_generate_noise or ps_upload_noise");
    if (!c->have_weather) return fail(PS_ERR_NOT_READY, "no weather map: call ps_upload_weather");
    if (!c->have_blue) return fail(PS_ERR_NOT_READY, "no blue noise: call ps_upload_blue_noise");
    return PS_OK;
}

namespace {

ps_status check_ready(const ps_ctx* c, bool for_render) {
    if (!c->have_atm) return fail(PS_ERR_NOT_READY, "ps_set_atmosphere has not been called");
    if (!c->have_frame) return fail(PS_ERR_NOT_READY, "ps_set_frame has not been called");
    if (!for_render) return PS_OK;
    if (!c->have_clouds) return fail(PS_ERR_NOT_READY, "ps_set_clouds has not been called");
    ps_status s;
    if ((s = check_shared_inputs(c))) return s;
    if (!c->luts_valid) return fail(PS_ERR_NOT_READY, "LUTs not baked: call ps_bake_luts");
    return PS_OK;
}

}  // namespace

ps_status psi::check_render_ready(const ps_ctx* c) { return check_ready(c, true); }

namespace {

ps_status launch_luts(ps_ctx* c, uint32_t flags, hipEvent_t* ev, hipStream_t st) {
    ps_status s;
    if ((s = check_ready(c, false))) return s;
    if ((s = ensure_luts(c))) return s;
    LutParams P;
    fill_lut_params(c->sc, c->dims, P);
    uint32_t k = flags & 15u;
    if (flags & PS_LUT_MULTISCATTER_DISABLED) {
        P.ms_disabled = 1;
        k |= 2u;
    }
    LutParams* dP = nullptr;
    if ((s = c->lut_ring.upload(P, st, &dP))) return s;
    PS_HIP(ps_launch_luts(dP, P, 1, k, c->luts.T, c->luts.MS, c->luts.SV, c->luts.AP, st, ev));
    if ((s = c->lut_ring.mark_used(st))) return s;
    c->lut_baked |= (k & 15u);
    c->luts_valid = c->lut_baked == 15u;
    return PS_OK;
}

// Brick and volume bounds of the caller's perturbation volume (it may have
// changed since the last render: recomputed on every render that uses it).
ps_status prepare_delta(ps_ctx* c) {
    if (!c->delta) return PS_OK;
    ps_status s;
    if ((s = dev_alloc(&c->delta_bnd, sizeof(float2) * (32 * 32 * 32 + 1 + 128), "perturbation brick bounds"))) return s;
    PS_HIP(ps_launch_delta_bounds(c->delta, c->delta_bnd, c->stream));
    return PS_OK;
}

ps_status grow(float** p, size_t* cap, size_t floats, const char* what);

ps_status launch_render(ps_ctx* c, const float* depth, const ps_image_out* out, bool count) {
    ps_status s;
    if ((s = check_ready(c, true))) return s;
    if (!out && !count) return fail(PS_ERR_INVALID_ARGUMENT, "ps_render: out is NULL");
    if (count && c->half_res) return fail(PS_ERR_UNSUPPORTED, "ps_count_work: counts the full-resolution march only");
    if ((s = prepare_delta(c))) return s;
    RenderParams next;
    if ((s = fill_render_params(c, c->sc, c->luts, depth, out, next))) return s;
    RenderParams* dP = nullptr;
    // TAA bookkeeping of the active view (crates/ps-clouds/src/lib.rs:773-829):
    // the off -> on transition is tracked on every render, counted or not
    ps_ctx::TaaView& tv = c->taa_views[c->active_view];
    const bool taa = c->taa && !count;
    if (!count) {
        if (taa && !tv.prev_on) tv.history_valid = false;
        tv.prev_on = c->taa;
    }
    if (count && c->want_tile_times) next.tile_times = c->d_tile_times;
    if (!c->half_res && !taa) {
        // one kernel: sky pass, march and composite (0.7 % faster at 1280x720
        // than the RT march plus a composite kernel, with identical output)
        if ((s = c->render_ring.upload(next, c->stream, &dP))) return s;
        PS_HIP(ps_launch_render(dP, next.W, next.H, count, c->delta != nullptr, false, c->stream));
        return c->render_ring.mark_used(c->stream);
    }
    // the cloud RT: half resolution as the reference computes it (W * 0.5 in
    // FP32, truncated, at least 1) or the frame's own size; the march into
    // it, the TAA resolve, then the full-resolution composite
    const int rw = c->half_res ? std::max(1, (int)((float)next.W * 0.5f)) : next.W;
    const int rh = c->half_res ? std::max(1, (int)((float)next.H * 0.5f)) : next.H;
    const size_t texels = (size_t)rw * rh;
    if ((s = grow(&c->rt, &c->rt_cap, texels * 8, "cloud RT"))) return s;
    float4* rt_lum = (float4*)c->rt;
    float4* rt_trn = rt_lum + texels;
    RenderParams march = next;
    march.W = rw; march.H = rh;
    march.rt_sx = (double)next.W / rw; march.rt_sy = (double)next.H / rh;
    march.n_tiles_x = (rw + 7) / 8;
    march.out_comp = nullptr; march.out_lum = rt_lum; march.out_trans = rt_trn; march.out_f16 = 0;
    if ((s = c->render_ring.upload(march, c->stream, &dP))) return s;
    PS_HIP(ps_launch_render(dP, rw, rh, false, c->delta != nullptr, false, c->stream, true));
    if ((s = c->render_ring.mark_used(c->stream))) return s;
    const float4* res_lum = rt_lum;
    const float4* res_trn = rt_trn;
    if (taa) {
        // a new RT size discards every view's history (the reference rebuilds
        // its RT bundle and TAA state, lib.rs:300-312)
        if (rw != c->taa_rt_w || rh != c->taa_rt_h) {
            for (ps_ctx::TaaView& v : c->taa_views) v.history_valid = false;
            c->taa_rt_w = rw; c->taa_rt_h = rh;
        }
        if ((s = grow(&tv.hist, &tv.cap, texels * 16, "TAA history"))) return s;
        const int slot = tv.history_valid ? tv.write_slot ^ 1 : 0;
        tv.write_slot = slot;
        float4* h = (float4*)tv.hist;
        TaaParams tp{};
        tp.cur_lum = rt_lum; tp.cur_trn = rt_trn;
        tp.hist_lum = h + (size_t)(2 * (slot ^ 1)) * texels; tp.hist_trn = tp.hist_lum + texels;
        tp.out_lum = h + (size_t)(2 * slot) * texels; tp.out_trn = tp.out_lum + texels;
        tp.w = rw; tp.h = rh;
        tp.blend = tv.history_valid ? 1.f / 8.f : 1.f;
        tp.history_valid = tv.history_valid ? 1 : 0;
        const ps_frame_uniforms& f = c->sc.frame;
        for (int i = 0; i < 16; ++i) { tp.inv_vp[i] = f.inv_view_proj[i]; tp.prev_vp[i] = f.prev_view_proj[i]; }
        for (int i = 0; i < 3; ++i) tp.cam[i] = f.camera_position_world[i];
        PS_HIP(ps_launch_taa_resolve(tp, c->stream));
        tv.history_valid = true;
        res_lum = tp.out_lum; res_trn = tp.out_trn;
    }
    next.rt_lum = res_lum; next.rt_trn = res_trn; next.rt_W = rw; next.rt_H = rh;
    if ((s = c->render_ring.upload(next, c->stream, &dP))) return s;
    PS_HIP(ps_launch_composite_rt(dP, next.W, next.H, c->half_res, c->stream));
    return c->render_ring.mark_used(c->stream);
}

// Grow a context-owned device buffer to at least `floats` floats.
ps_status grow(float** p, size_t* cap, size_t floats, const char* what) {
    if (*cap >= floats && *p) return PS_OK;
    dev_free(*p);
    *cap = 0;
    ps_status s;
    if ((s = dev_alloc(p, floats * sizeof(float), what))) return s;
    *cap = floats;
    return PS_OK;
}

}  // namespace

ps_status psi::check_atmosphere(const ps_atmosphere_params* a, const char* who) {
    if (!(a->planet_radius_m > 0.f) || !(a->atmosphere_top_m > a->planet_radius_m))
        return fail(PS_ERR_INVALID_ARGUMENT, std::string(who) + ": need 0 < planet_radius_m < atmosphere_top_m");
    if (!(a->rayleigh_scale_height_m > 0.f) || !(a->mie_scale_height_m > 0.f))
        return fail(PS_ERR_INVALID_ARGUMENT, std::string(who) + ": scale heights must be positive");
    return PS_OK;
}

ps_status psi::check_clouds(const ps_cloud_params* p, const ps_cloud_layer* layers, int32_t n, const char* who) {
    if (n < 0 || n > PS_MAX_CLOUD_LAYERS)
        return fail(PS_ERR_INVALID_ARGUMENT, std::string(who) + ": n_layers must be in [0, 8]");
    if (n > 0 && !layers) return fail(PS_ERR_INVALID_ARGUMENT, std::string(who) + ": layers is NULL");
    if (p->cloud_steps < 1 || p->cloud_steps > 4096)
        return fail(PS_ERR_INVALID_ARGUMENT, std::string(who) + ": cloud_steps must be in [1, 4096]");
    if (p->light_steps < 1 || p->light_steps > 256)
        return fail(PS_ERR_INVALID_ARGUMENT, std::string(who) + ": light_steps must be in [1, 256]");
    if (p->multi_scatter_octaves > PS_MAX_OCTAVES)
        return fail(PS_ERR_UNSUPPORTED, std::string(who) + ": at most 8 multi-scatter octaves are supported");
    int i, j;
    if (n > 0 && ps_check_layers_overlap(layers, n, &i, &j)) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "%s: layers %d and %d overlap ([%g, %g] vs [%g, %g] m)", who, i, j,
                      layers[i].base_m, layers[i].top_m, layers[j].base_m, layers[j].top_m);
        return fail(PS_ERR_OVERLAPPING_LAYERS, buf);
    }
    return PS_OK;
}

extern "C" {

const char* ps_last_error(void) { return psi::g_last_error.c_str(); }
int32_t ps_abi_version(void) { return PS_ABI_VERSION; }

ps_status ps_ctx_create(int32_t device, ps_ctx** out_ctx) {
    if (!out_ctx) return fail(PS_ERR_INVALID_ARGUMENT, "ps_ctx_create: out_ctx is NULL");
    *out_ctx = nullptr;
    int n = 0;
    PS_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(PS_ERR_INVALID_ARGUMENT, "ps_ctx_create: device index out of range");
    PS_HIP(hipSetDevice(device));
    ps_ctx* c = new ps_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(PS_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    c->stream = c->own_stream;
    const ps_status ns = noalloc_init(device);
    if (ns != PS_OK) {
        hipStreamDestroy(c->own_stream);
        delete c;
        return ns;
    }
    e = hipMalloc((void**)&c->d_counters, 16 * sizeof(unsigned long long));
    if (e != hipSuccess) {
        hipStreamDestroy(c->own_stream);
        delete c;
        return fail(PS_ERR_OUT_OF_MEMORY, std::string("hipMalloc(counters): ") + hipGetErrorString(e));
    }
    *out_ctx = c;
    return PS_OK;
}

void ps_ctx_destroy(ps_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    if (c->lut_stream) hipStreamSynchronize(c->lut_stream);
    dev_free(c->luts.T); dev_free(c->luts.MS); dev_free(c->luts.SV); dev_free(c->luts.AP);
    dev_free(c->luts_back.T); dev_free(c->luts_back.MS); dev_free(c->luts_back.SV); dev_free(c->luts_back.AP);
    if (c->lut_stream) hipStreamDestroy(c->lut_stream);
    if (c->ev_luts_done) hipEventDestroy(c->ev_luts_done);
    if (c->ev_render_issued) hipEventDestroy(c->ev_render_issued);
    dev_free(c->base_store); dev_free(c->detail_store); dev_free(c->curl_store);
    dev_free(c->base_rl); dev_free(c->bbound); dev_free(c->detail_hf); dev_free(c->curl2);
    dev_free(c->weather); dev_free(c->weather_synth); dev_free(c->type_grid); dev_free(c->overcast); dev_free(c->blue);
    dev_free(c->d_counters);
    dev_free(c->d_tile_times);
    dev_free(c->delta_bnd); dev_free(c->rt); dev_free(c->post_hdr); dev_free(c->post_rays);
    for (float*& b : c->post_bloom) dev_free(b);
    for (ps_ctx::TaaView& v : c->taa_views) dev_free(v.hist);
    dev_free(c->ck); dev_free(c->px_state); dev_free(c->rec); dev_free(c->lrec); dev_free(c->g_delta64);
    dev_free(c->atmo_params); dev_free(c->aT); dev_free(c->aMS); dev_free(c->aSV); dev_free(c->aAP);
    dev_free(c->aTT); dev_free(c->aMST); dev_free(c->aSVT); dev_free(c->aAPT);
    if (c->atmo_params_h) hipHostFree(c->atmo_params_h);
    if (c->atmo_ev) hipEventDestroy(c->atmo_ev);
    c->render_ring.release();
    c->lut_ring.release();
    for (hipEvent_t e : c->events) hipEventDestroy(e);
    if (c->own_stream) hipStreamDestroy(c->own_stream);
    delete c;
}

ps_status ps_ctx_set_stream(ps_ctx* c, void* hip_stream) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_ctx_set_stream: ctx is NULL");
    c->stream = hip_stream ? (hipStream_t)hip_stream : c->own_stream;
    return PS_OK;
}

ps_status ps_ctx_synchronize(ps_ctx* c) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_ctx_synchronize: ctx is NULL");
    PS_HIP(hipStreamSynchronize(c->stream));
    return PS_OK;
}

ps_status ps_ctx_set_lut_dims(ps_ctx* c, const ps_lut_dims* d) {
    if (!c || !d) return fail(PS_ERR_INVALID_ARGUMENT, "ps_ctx_set_lut_dims: NULL argument");
    const int v[9] = {d->transmittance_w, d->transmittance_h, d->multiscatter_w, d->multiscatter_h,
                      d->skyview_w, d->skyview_h, d->ap_w, d->ap_h, d->ap_d};
    for (int i = 0; i < 9; ++i)
        if (v[i] < 1 || v[i] > 4096) return fail(PS_ERR_INVALID_ARGUMENT, "ps_ctx_set_lut_dims: sizes must be in [1, 4096]");
    if ((size_t)d->ap_w * d->ap_h * d->ap_d > (size_t)1 << 24)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_ctx_set_lut_dims: AP volume too large");
    PS_HIP(hipStreamSynchronize(c->stream));
    if (c->lut_stream) PS_HIP(hipStreamSynchronize(c->lut_stream));
    dev_free(c->luts.T); dev_free(c->luts.MS); dev_free(c->luts.SV); dev_free(c->luts.AP);
    psi::LutSet& b = c->luts_back;
    dev_free(b.T); dev_free(b.MS); dev_free(b.SV); dev_free(b.AP);
    c->dims = *d;
    c->lut_baked = 0;
    c->luts_valid = false;
    c->train_valid = false;  // its parameters point into the freed sets
    return ensure_luts(c);
}

ps_status ps_ctx_get_lut_dims(const ps_ctx* c, ps_lut_dims* out) {
    if (!c || !out) return fail(PS_ERR_INVALID_ARGUMENT, "ps_ctx_get_lut_dims: NULL argument");
    *out = c->dims;
    return PS_OK;
}

ps_status ps_set_atmosphere(ps_ctx* c, const ps_atmosphere_params* a) {
    if (!c || !a) return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_atmosphere: NULL argument");
    ps_status s;
    if ((s = check_atmosphere(a, "ps_set_atmosphere"))) return s;
    c->sc.atm = *a;
    c->have_atm = true;
    c->lut_baked = 0;
    c->luts_valid = false;
    return PS_OK;
}

ps_status ps_set_frame(ps_ctx* c, const ps_frame_uniforms* f) {
    if (!c || !f) return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_frame: NULL argument");
    c->sc.frame = *f;
    c->have_frame = true;
    return PS_OK;
}

ps_status ps_set_clouds(ps_ctx* c, const ps_cloud_params* p, const ps_cloud_layer* layers, int32_t n,
                        const float* wind) {
    if (!c || !p) return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_clouds: NULL argument");
    ps_status s;
    if ((s = check_clouds(p, layers, n, "ps_set_clouds"))) return s;
    c->sc.clouds = *p;
    c->sc.n_layers = n;
    for (int l = 0; l < n; ++l) {
        c->sc.layers[l] = layers[l];
        c->sc.wind[2 * l] = wind ? wind[2 * l] : 0.f;
        c->sc.wind[2 * l + 1] = wind ? wind[2 * l + 1] : 0.f;
    }
    c->have_clouds = true;
    return PS_OK;
}

ps_status ps_set_march_steps(ps_ctx* c, uint32_t cloud_steps, uint32_t light_steps) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_march_steps: ctx is NULL");
    if (!c->have_clouds) return fail(PS_ERR_NOT_READY, "ps_set_march_steps: call ps_set_clouds first");
    c->sc.clouds.cloud_steps = std::clamp(cloud_steps, 8u, 256u);
    c->sc.clouds.light_steps = std::clamp(light_steps, 1u, 16u);
    return PS_OK;
}

static ps_status alloc_noise(ps_ctx* c, int format) {
    if (c->noise_format != format) {
        PS_HIP(hipStreamSynchronize(c->stream));
        dev_free(c->base_store); dev_free(c->detail_store); dev_free(c->curl_store);
        c->noise_format = format;
        c->train_valid = false;  // its parameters point into the freed stores
    }
    ps_status s;
    if ((s = dev_alloc(&c->base_store, noise_texel_bytes(format, 4) * 128 * 128 * 128, "base noise"))) return s;
    if ((s = dev_alloc(&c->detail_store, noise_texel_bytes(format, 4) * 32 * 32 * 32, "detail noise"))) return s;
    if ((s = dev_alloc(&c->curl_store, noise_texel_bytes(format, 2) * 128 * 128, "curl noise"))) return s;
    if ((s = dev_alloc(&c->base_rl, 4 * sizeof(uint32_t) * 128 * 128 * 128, "base fold"))) return s;
    if ((s = dev_alloc(&c->bbound, sizeof(uint16_t) * PS_BN * PS_BN * PS_BN, "base brick bounds"))) return s;
    if ((s = dev_alloc(&c->detail_hf, 4 * sizeof(float) * 32 * 32 * 32, "detail fold"))) return s;
    if ((s = dev_alloc(&c->curl2, 2 * sizeof(float2) * 128 * 128, "curl fold"))) return s;
    return PS_OK;
}

ps_status ps_generate_noise(ps_ctx* c, uint32_t seed, int32_t format) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_generate_noise: ctx is NULL");
    if (format != PS_NOISE_UNORM8 && format != PS_NOISE_FP16)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_generate_noise: format must be PS_NOISE_UNORM8 or PS_NOISE_FP16");
    ps_status s;
    if ((s = alloc_noise(c, format))) return s;
    PS_HIP(ps_launch_noise_bake(seed, format, c->base_store, c->detail_store, c->curl_store, c->stream));
    PS_HIP(ps_launch_noise_fold(format, c->base_store, c->detail_store, c->curl_store, c->base_rl, c->detail_hf,
                                c->curl2, c->bbound, c->stream));
    c->have_noise = true;
    return PS_OK;
}

ps_status ps_upload_noise(ps_ctx* c, int32_t format, const void* base, const void* detail, const void* curl) {
    if (!c || !base || !detail || !curl) return fail(PS_ERR_INVALID_ARGUMENT, "ps_upload_noise: NULL argument");
    if (format != PS_NOISE_UNORM8 && format != PS_NOISE_FP16)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_upload_noise: format must be PS_NOISE_UNORM8 or PS_NOISE_FP16");
    ps_status s;
    if ((s = alloc_noise(c, format))) return s;
    PS_HIP(hipMemcpyAsync(c->base_store, base, noise_texel_bytes(format, 4) * 128 * 128 * 128, hipMemcpyHostToDevice,
                          c->stream));
    PS_HIP(hipMemcpyAsync(c->detail_store, detail, noise_texel_bytes(format, 4) * 32 * 32 * 32,
                          hipMemcpyHostToDevice, c->stream));
    PS_HIP(hipMemcpyAsync(c->curl_store, curl, noise_texel_bytes(format, 2) * 128 * 128, hipMemcpyHostToDevice,
                          c->stream));
    PS_HIP(ps_launch_noise_fold(format, c->base_store, c->detail_store, c->curl_store, c->base_rl, c->detail_hf,
                                c->curl2, c->bbound, c->stream));
    PS_HIP(hipStreamSynchronize(c->stream));
    c->have_noise = true;
    return PS_OK;
}

ps_status ps_read_noise(ps_ctx* c, int32_t which, void* host_out, size_t bytes) {
    if (!c || !host_out) return fail(PS_ERR_INVALID_ARGUMENT, "ps_read_noise: NULL argument");
    if (!c->have_noise) return fail(PS_ERR_NOT_READY, "ps_read_noise: no noise volumes");
    const void* src;
    size_t need;
    switch (which) {
        case PS_NOISE_BASE: src = c->base_store; need = noise_texel_bytes(c->noise_format, 4) * 128 * 128 * 128; break;
        case PS_NOISE_DETAIL: src = c->detail_store; need = noise_texel_bytes(c->noise_format, 4) * 32 * 32 * 32; break;
        case PS_NOISE_CURL: src = c->curl_store; need = noise_texel_bytes(c->noise_format, 2) * 128 * 128; break;
        default: return fail(PS_ERR_INVALID_ARGUMENT, "ps_read_noise: unknown volume");
    }
    if (bytes < need) return fail(PS_ERR_INVALID_ARGUMENT, "ps_read_noise: host buffer too small");
    PS_HIP(hipMemcpyAsync(host_out, src, need, hipMemcpyDeviceToHost, c->stream));
    PS_HIP(hipStreamSynchronize(c->stream));
    return PS_OK;
}

ps_status ps_upload_blue_noise(ps_ctx* c, const uint8_t* ranks) {
    if (!c || !ranks) return fail(PS_ERR_INVALID_ARGUMENT, "ps_upload_blue_noise: NULL argument");
    ps_status s;
    if ((s = dev_alloc(&c->blue, 64 * 64, "blue noise"))) return s;
    PS_HIP(hipMemcpyAsync(c->blue, ranks, 64 * 64, hipMemcpyHostToDevice, c->stream));
    PS_HIP(hipStreamSynchronize(c->stream));
    c->have_blue = true;
    return PS_OK;
}

ps_status ps_synthesise_blue_noise(ps_ctx* c) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_synthesise_blue_noise: NULL context");
    ps_status s;
    if ((s = dev_alloc(&c->blue, 64 * 64, "blue noise"))) return s;
    BlueNoiseWeights w;
    psi::blue_noise_weights(w.w);
    PS_HIP(ps_launch_blue_noise(w, c->blue, c->stream));
    c->have_blue = true;
    return PS_OK;
}

ps_status ps_read_blue_noise(ps_ctx* c, uint8_t* ranks_64x64) {
    if (!c || !ranks_64x64) return fail(PS_ERR_INVALID_ARGUMENT, "ps_read_blue_noise: NULL argument");
    if (!c->have_blue) return fail(PS_ERR_NOT_READY, "ps_read_blue_noise: no blue noise uploaded or synthesised");
    PS_HIP(hipMemcpyAsync(ranks_64x64, c->blue, 64 * 64, hipMemcpyDeviceToHost, c->stream));
    PS_HIP(hipStreamSynchronize(c->stream));
    return PS_OK;
}

ps_status ps_upload_weather(ps_ctx* c, const uint16_t* weather, int32_t w, int32_t h, int32_t channels,
                            const uint8_t* type_grid, const uint8_t* overcast) {
    if (!c || !weather) return fail(PS_ERR_INVALID_ARGUMENT, "ps_upload_weather: NULL argument");
    if (w < 1 || h < 1 || w > 8192 || h > 8192 || channels < 1 || channels > 4)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_upload_weather: need 1 <= w, h <= 8192 and 1 <= channels <= 4");
    PS_HIP(hipStreamSynchronize(c->stream));
    if (c->ww != w || c->wh != h) {
        dev_free(c->weather);
        c->train_valid = false;  // its parameters point into the freed map
    }
    ps_status s;
    if ((s = dev_alloc(&c->weather, sizeof(uint2) * w * h, "weather"))) return s;
    if ((s = dev_alloc(&c->type_grid, 128 * 128, "cloud type grid"))) return s;
    if ((s = dev_alloc(&c->overcast, 128 * 128, "overcast mask"))) return s;
    uint16_t* staging = nullptr;
    size_t bytes = sizeof(uint16_t) * (size_t)w * h * channels;
    PS_HIP(hipMalloc((void**)&staging, bytes));
    hipError_t e = hipMemcpy(staging, weather, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = ps_launch_weather_fold(staging, w, h, channels, c->weather, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(staging);
    if (e != hipSuccess) return fail(PS_ERR_DEVICE, std::string("ps_upload_weather: ") + hipGetErrorString(e));
    c->ww = w; c->wh = h;
    c->type_grid_active = false;
    if (type_grid) {
        for (int i = 0; i < 128 * 128; ++i)
            if (type_grid[i] != 255u) { c->type_grid_active = true; break; }
        PS_HIP(hipMemcpy(c->type_grid, type_grid, 128 * 128, hipMemcpyHostToDevice));
    }
    c->overcast_active = overcast != nullptr;
    if (overcast) PS_HIP(hipMemcpy(c->overcast, overcast, 128 * 128, hipMemcpyHostToDevice));
    c->have_weather = true;
    return PS_OK;
}

ps_status ps_synthesise_weather(ps_ctx* c, const ps_weather_synth* ws, const uint8_t* overcast) {
    if (!c || !ws) return fail(PS_ERR_INVALID_ARGUMENT, "ps_synthesise_weather: NULL argument");
    const bool grid = ws->coverage_grid != nullptr, types = ws->type_grid != nullptr;
    if ((grid || types) && (ws->grid_w < 1 || ws->grid_h < 1 || ws->grid_w > 8192 || ws->grid_h > 8192))
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_synthesise_weather: grid size must be 1..8192 per axis");
    if (grid && !(ws->grid_extent_m > 0.f && std::isfinite(ws->grid_extent_m)))
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_synthesise_weather: grid_extent_m must be positive");
    PS_HIP(hipStreamSynchronize(c->stream));
    const int n = 128 * 128;
    if (c->ww != 128 || c->wh != 128) {
        dev_free(c->weather);
        c->train_valid = false;
    }
    ps_status s;
    if ((s = dev_alloc(&c->weather, sizeof(uint2) * n, "weather"))) return s;
    if ((s = dev_alloc(&c->weather_synth, sizeof(uint16_t) * 4 * n, "synthesised weather"))) return s;
    if ((s = dev_alloc(&c->type_grid, n, "cloud type grid"))) return s;
    if ((s = dev_alloc(&c->overcast, n, "overcast mask"))) return s;
    // precipitation intensity over a 50 mm/h ceiling or ground wetness,
    // the larger (weather_map.rs:66-74)
    const float precip = std::fmin(std::fmax(ws->precip_mm_per_h / 50.f, 0.f), 1.f);
    const float wet = std::fmin(std::fmax(ws->ground_wetness, 0.f), 1.f);
    const float alpha = std::fmax(precip, wet);
    const size_t cells = grid || types ? (size_t)ws->grid_w * ws->grid_h : 0;
    float* d_grid = nullptr;
    uint8_t* d_types = nullptr;
    hipError_t e = hipSuccess;
    if (grid) e = hipMalloc((void**)&d_grid, sizeof(float) * cells);
    if (e == hipSuccess && grid) e = hipMemcpy(d_grid, ws->coverage_grid, sizeof(float) * cells, hipMemcpyHostToDevice);
    if (e == hipSuccess && types) e = hipMalloc((void**)&d_types, cells);
    if (e == hipSuccess && types) e = hipMemcpy(d_types, ws->type_grid, cells, hipMemcpyHostToDevice);
    if (e == hipSuccess && !types) e = hipMemsetAsync(c->type_grid, 0xff, n, c->stream);  // all sentinel
    if (e == hipSuccess)
        e = ps_launch_weather_synth(d_grid, ws->grid_w, ws->grid_h, ws->grid_extent_m, ws->has_layers ? 1.f : 0.f,
                                    alpha, d_types, c->weather_synth, c->type_grid, c->stream);
    if (e == hipSuccess) e = ps_launch_weather_fold(c->weather_synth, 128, 128, 4, c->weather, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (d_grid) hipFree(d_grid);
    if (d_types) hipFree(d_types);
    if (e != hipSuccess) return fail(PS_ERR_DEVICE, std::string("ps_synthesise_weather: ") + hipGetErrorString(e));
    c->ww = 128; c->wh = 128;
    // the march consults the grid only when some source cell is a valid type
    c->type_grid_active = false;
    for (size_t i = 0; types && i < cells; ++i)
        if (ws->type_grid[i] <= 7u) { c->type_grid_active = true; break; }
    c->overcast_active = overcast != nullptr;
    if (overcast) PS_HIP(hipMemcpy(c->overcast, overcast, n, hipMemcpyHostToDevice));
    c->have_weather = true;
    return PS_OK;
}

ps_status ps_read_weather(ps_ctx* c, uint16_t* rgba_f16, uint8_t* cloud_types) {
    if (!c || !rgba_f16) return fail(PS_ERR_INVALID_ARGUMENT, "ps_read_weather: NULL argument");
    if (!c->weather_synth) return fail(PS_ERR_NOT_READY, "ps_read_weather: no synthesised weather map");
    PS_HIP(hipMemcpyAsync(rgba_f16, c->weather_synth, sizeof(uint16_t) * 4 * 128 * 128, hipMemcpyDeviceToHost,
                          c->stream));
    if (cloud_types)
        PS_HIP(hipMemcpyAsync(cloud_types, c->type_grid, 128 * 128, hipMemcpyDeviceToHost, c->stream));
    PS_HIP(hipStreamSynchronize(c->stream));
    return PS_OK;
}

void ps_default_present_params(ps_present_params* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->ev100 = 14.f;                          // RenderConfig::default (config.rs:243-244)
    o->tonemap = PS_TONEMAP_ACES;
    o->ground_model = PS_GROUND_RADIANCE;
    o->godrays = 1;                           // GodraysTuning::default (config.rs:542-549)
    o->godrays_samples = 64;
    o->godrays_decay = 0.96f;
    o->godrays_intensity = 0.6f;
    o->godrays_bright_threshold = 5000.f;
    o->bloom = 1;                             // BloomTuning::default (config.rs:651-659)
    o->bloom_threshold_ev100 = 2.f;
    o->bloom_intensity = 1.f;
    o->bloom_knee_ev = 0.5f;
    ps_surface_params& f = o->surface;        // SurfaceParams::default (weather.rs:203-222)
    f.visibility_m = 30000.f;
    f.temperature_c = 17.f;
    f.dewpoint_c = 7.f;
    f.pressure_hpa = 1018.f;
    f.wind_dir_deg = 240.f;
    f.wind_speed_mps = 5.f;
    f.puddle_start = 0.6f;
}

ps_status ps_present(ps_ctx* c, const float* depth, const float* ground_lit, const ps_image_out* in,
                     const ps_present_params* pp, uint8_t* out_srgb8, float* out_hdr) {
    if (!c || !in || !pp) return fail(PS_ERR_INVALID_ARGUMENT, "ps_present: NULL argument");
    if (!in->composite) return fail(PS_ERR_INVALID_ARGUMENT, "ps_present: in->composite is NULL");
    if (in->format != PS_PIXEL_F32) return fail(PS_ERR_INVALID_ARGUMENT, "ps_present: images must be PS_PIXEL_F32");
    if (pp->ground_model != PS_GROUND_RADIANCE && pp->ground_model != PS_GROUND_PBR)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_present: ground_model must be PS_GROUND_RADIANCE or PS_GROUND_PBR");
    const bool pbr = pp->ground_model == PS_GROUND_PBR;
    if ((ground_lit || pbr) && (!depth || !in->cloud_luminance || !in->cloud_transmittance))
        return fail(PS_ERR_INVALID_ARGUMENT,
                    "ps_present: a shaded surface needs the depth buffer, cloud_luminance and cloud_transmittance");
    if (pp->tonemap != PS_TONEMAP_ACES && pp->tonemap != PS_TONEMAP_CLAMP)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_present: tonemap must be PS_TONEMAP_ACES or PS_TONEMAP_CLAMP");
    if (pp->target != PS_TARGET_SRGB8 && pp->target != PS_TARGET_UNORM8)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_present: target must be PS_TARGET_SRGB8 or PS_TARGET_UNORM8");
    if (!out_srgb8 && !out_hdr) return fail(PS_ERR_INVALID_ARGUMENT, "ps_present: no output");
    if (!std::isfinite(pp->ev100)) return fail(PS_ERR_INVALID_ARGUMENT, "ps_present: ev100 must be finite");
    if (pp->godrays && (pp->godrays_samples < 1 || pp->godrays_samples > 1024 || !std::isfinite(pp->godrays_decay) ||
                        !std::isfinite(pp->godrays_intensity) || !std::isfinite(pp->godrays_bright_threshold)))
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_present: godrays need 1..1024 samples and finite tunables");
    if (pp->bloom && (!std::isfinite(pp->bloom_threshold_ev100) || !std::isfinite(pp->bloom_intensity) ||
                      !std::isfinite(pp->bloom_knee_ev)))
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_present: bloom tunables must be finite");
    ps_status s;
    if ((s = check_ready(c, true))) return s;
    RenderParams next;
    if ((s = fill_render_params(c, c->sc, c->luts, depth, in, next))) return s;
    const int W = next.W, H = next.H;
    const ps_frame_uniforms& f = c->sc.frame;

    // godrays: the sun direction projected through view_proj as a direction
    // (w = 0), in f32 in glam's column order (ps-godrays lib.rs:358-386)
    bool rays_on = false;
    float sun_u = 0.f, sun_v = 0.f;
    if (pp->godrays && f.sun_direction[1] > 0.f) {
        const float* m = f.view_proj;
        const float sx = f.sun_direction[0], sy = f.sun_direction[1], sz = f.sun_direction[2];
        const float cx = ((m[0] * sx + m[4] * sy) + m[8] * sz) + m[12] * 0.f;
        const float cy = ((m[1] * sx + m[5] * sy) + m[9] * sz) + m[13] * 0.f;
        const float cw = ((m[3] * sx + m[7] * sy) + m[11] * sz) + m[15] * 0.f;
        if (cw > 1.0e-6f) {
            const float nx = cx / cw, ny = cy / cw;
            if (std::fabs(nx) <= 2.f && std::fabs(ny) <= 2.f) {
                rays_on = true;
                // radial.wgsl:71-74
                sun_u = nx * 0.5f + 0.5f;
                sun_v = 1.f - (ny * 0.5f + 0.5f);
            }
        }
    }
    const bool bloom_on = pp->bloom != 0;
    const bool post = rays_on || bloom_on;

    PresentArgs a{};
    a.composite = (const float4*)in->composite;
    a.luminance = (const float4*)in->cloud_luminance;
    a.transmittance = (const float4*)in->cloud_transmittance;
    a.ground = pbr ? nullptr : (const float4*)ground_lit;
    // tonemap.wgsl:50: exposure = 1 / (1.2 * 2^EV100)
    a.exposure = 1.f / (1.2f * std::pow(2.f, pp->ev100));
    a.mode = pp->tonemap | (pp->target == PS_TARGET_UNORM8 ? 2 : 0);
    a.pbr = pbr ? 1 : 0;
    if (pbr) {
        const ps_surface_params& sp = pp->surface;
        const float mf = sp.material;
        a.material = mf >= 0.f && mf < 4294967040.f ? (uint32_t)mf : 0u;
        for (int k = 0; k < 3; ++k) a.albedo_scale[k] = c->sc.atm.ground_albedo[k] / 0.18f;
        a.wetness = sp.ground_wetness;
        a.puddle_start = sp.puddle_start;
        a.puddle_coverage = sp.puddle_coverage;
        a.precip = sp.precip_intensity_mm_per_h;
        a.temperature = sp.temperature_c;
        a.snow_depth = sp.snow_depth_m;
        a.sim_seconds = f.simulated_seconds;
    }
    float4* work = (float4*)out_hdr;
    if (post && !work) {
        if ((s = grow(&c->post_hdr, &c->post_hdr_cap, (size_t)W * H * 4, "presentation HDR image"))) return s;
        work = (float4*)c->post_hdr;
    }
    a.hdr = work;
    a.ldr = post ? nullptr : (uint32_t*)out_srgb8;
    RenderParams* dP = nullptr;
    if ((s = c->render_ring.upload(next, c->stream, &dP))) return s;
    PS_HIP(ps_launch_present(dP, W, H, a, c->stream));
    if ((s = c->render_ring.mark_used(c->stream))) return s;
    if (!post) return PS_OK;

    PostArgs q{};
    q.hdr = work;
    q.W = W;
    q.H = H;
    q.godrays = rays_on ? 1 : 0;
    if (rays_on) {
        q.rw = std::max(W / 2, 1);    // HALF_RES_FACTOR 2 (ps-godrays lib.rs:435-438)
        q.rh = std::max(H / 2, 1);
        if ((s = grow(&c->post_rays, &c->post_rays_cap, (size_t)q.rw * q.rh * 4, "godrays image"))) return s;
        q.rays = (float4*)c->post_rays;
        q.sun_u = sun_u;
        q.sun_v = sun_v;
        q.samples = (int)pp->godrays_samples;
        q.decay = pp->godrays_decay;
        q.rays_intensity = pp->godrays_intensity;
        q.bright_threshold = pp->godrays_bright_threshold;
    }
    q.bloom = bloom_on ? 1 : 0;
    if (bloom_on) {
        // half / quarter / eighth resolution (ps-bloom lib.rs:439-467)
        int lw = std::max(W / 2, 1), lh = std::max(H / 2, 1);
        for (int l = 0; l < 3; ++l) {
            q.lw[l] = lw;
            q.lh[l] = lh;
            if ((s = grow(&c->post_bloom[l], &c->post_bloom_cap[l], (size_t)lw * lh * 4, "bloom level"))) return s;
            q.levels[l] = (float4*)c->post_bloom[l];
            lw = std::max(lw / 2, 1);
            lh = std::max(lh / 2, 1);
        }
        // threshold exp2(ev100 + threshold_ev100), knee knee_ev * ln 2
        // (ps-bloom lib.rs:482-484)
        q.bloom_threshold = std::exp2(pp->ev100 + pp->bloom_threshold_ev100);
        q.bloom_knee = std::max(pp->bloom_knee_ev * 0.693147180559945309f, 1e-6f);
        q.bloom_intensity = pp->bloom_intensity;
    }
    q.out_hdr = (float4*)out_hdr;
    q.ldr = (uint32_t*)out_srgb8;
    q.exposure = a.exposure;
    q.mode = a.mode;
    PS_HIP(ps_launch_post(q, c->stream));
    return PS_OK;
}

ps_status ps_exposure_stats(ps_ctx* c, const float* hdr, int32_t w, int32_t h, float* out) {
    if (!c || !hdr || !out) return fail(PS_ERR_INVALID_ARGUMENT, "ps_exposure_stats: NULL argument");
    if (w < 1 || h < 1) return fail(PS_ERR_INVALID_ARGUMENT, "ps_exposure_stats: empty image");
    PS_HIP(ps_launch_exposure_stats((const float4*)hdr, w, h, out, c->stream));
    return PS_OK;
}

float ps_ev100_from_exposure_stats(float log_lum_sum, float pixel_count) {
    // auto_exposure.rs:209-212: mid-grey 0.18 times the tone map's 1.2
    const float avg_log = log_lum_sum / pixel_count;
    const float avg_lum = std::max(std::exp2(avg_log), 1e-6f);
    return std::log2(avg_lum / 0.216f);
}

ps_status ps_bake_luts(ps_ctx* c, uint32_t flags) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_bake_luts: ctx is NULL");
    return launch_luts(c, flags, nullptr, c->stream);
}

ps_status ps_read_lut(ps_ctx* c, int32_t which, float* host_out, size_t floats) {
    if (!c || !host_out) return fail(PS_ERR_INVALID_ARGUMENT, "ps_read_lut: NULL argument");
    if (!c->luts.T) return fail(PS_ERR_NOT_READY, "ps_read_lut: LUTs not allocated");
    const ps_lut_dims& d = c->dims;
    const float4* src;
    size_t n;
    switch (which) {
        case PS_LUT_TRANSMITTANCE: src = c->luts.T; n = (size_t)d.transmittance_w * d.transmittance_h; break;
        case PS_LUT_MULTISCATTER: src = c->luts.MS; n = (size_t)d.multiscatter_w * d.multiscatter_h; break;
        case PS_LUT_SKYVIEW: src = c->luts.SV; n = (size_t)d.skyview_w * d.skyview_h; break;
        case PS_LUT_AERIAL_PERSPECTIVE: src = c->luts.AP; n = (size_t)d.ap_w * d.ap_h * d.ap_d; break;
        default: return fail(PS_ERR_INVALID_ARGUMENT, "ps_read_lut: unknown LUT");
    }
    if (floats < 4 * n) return fail(PS_ERR_INVALID_ARGUMENT, "ps_read_lut: host buffer too small");
    PS_HIP(hipMemcpyAsync(host_out, src, 4 * n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PS_HIP(hipStreamSynchronize(c->stream));
    return PS_OK;
}

ps_status ps_render(ps_ctx* c, const float* depth, const ps_image_out* out) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_render: ctx is NULL");
    return launch_render(c, depth, out, false);
}

// Whether ps_frame may bake into the second LUT set on the side stream:
// every LUT is rebaked (the second set then holds nothing stale), and the
// main stream is not being captured into a graph (the side stream's work
// would not be part of the capture).
static bool can_pipeline(ps_ctx* c, uint32_t flags) {
    if ((flags & 15u) != 15u) return false;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &st) != hipSuccess) return false;
    return st == hipStreamCaptureStatusNone;
}

static ps_status ensure_pipeline(ps_ctx* c) {
    if (!c->lut_stream) {
        // default priority, like every other stream of the library
        PS_HIP(hipStreamCreateWithFlags(&c->lut_stream, hipStreamNonBlocking));
        PS_HIP(hipEventCreateWithFlags(&c->ev_luts_done, hipEventDisableTiming));
        PS_HIP(hipEventCreateWithFlags(&c->ev_render_issued, hipEventDisableTiming));
    }
    const ps_lut_dims& d = c->dims;
    ps_status s;
    psi::LutSet& b = c->luts_back;
    if ((s = dev_alloc(&b.T, sizeof(float4) * d.transmittance_w * d.transmittance_h, "transmittance"))) return s;
    if ((s = dev_alloc(&b.MS, sizeof(float4) * d.multiscatter_w * d.multiscatter_h, "multiscatter"))) return s;
    if ((s = dev_alloc(&b.SV, sizeof(float4) * d.skyview_w * d.skyview_h, "skyview"))) return s;
    if ((s = dev_alloc(&b.AP, sizeof(float4) * d.ap_w * d.ap_h * d.ap_d, "aerial perspective"))) return s;
    return PS_OK;
}

// Per frame of in-loop timing, seven events: the starts of the four LUT
// bakes and their end (on the stream that bakes), then the render's start
// and end on the main stream.
static constexpr size_t kFrameEvents = 7;

// The LUT chain of a frame (ps_frame, ps_frame_train): pipelined into the
// second set when every LUT is rebaked, else baked on the main stream.
static ps_status bake_frame_luts(ps_ctx* c, uint32_t lut_flags, hipEvent_t* ev) {
    ps_status s;
    if (can_pipeline(c, lut_flags)) {
        // bake this frame's LUTs into the second set on the side stream
        // while the previous frame renders from the current one, then swap.
        // The set being overwritten was last read by work queued on the main
        // stream before the previous pipelined render (or never).
        if ((s = check_ready(c, false))) return s;
        if ((s = ensure_pipeline(c))) return s;
        std::swap(c->luts, c->luts_back);
        if (c->render_issued_recorded) PS_HIP(hipStreamWaitEvent(c->lut_stream, c->ev_render_issued, 0));
        if ((s = launch_luts(c, lut_flags, ev, c->lut_stream))) {
            std::swap(c->luts, c->luts_back);  // the set the last frame used stays current
            return s;
        }
        PS_HIP(hipEventRecord(c->ev_luts_done, c->lut_stream));
        PS_HIP(hipStreamWaitEvent(c->stream, c->ev_luts_done, 0));
        PS_HIP(hipEventRecord(c->ev_render_issued, c->stream));
        c->render_issued_recorded = true;
    } else if ((s = launch_luts(c, lut_flags, ev, c->stream))) {
        return s;
    }
    return PS_OK;
}

ps_status ps_frame(ps_ctx* c, uint32_t lut_flags, const float* depth, const ps_image_out* out) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_frame: ctx is NULL");
    ps_status s;
    hipEvent_t* ev = nullptr;
    if (c->timing && c->timing_n < c->timing_max) ev = &c->events[(size_t)c->timing_n * kFrameEvents];
    if ((s = bake_frame_luts(c, lut_flags, ev))) return s;
    if (ev) PS_HIP(hipEventRecord(ev[5], c->stream));
    if ((s = launch_render(c, depth, out, false))) return s;
    if (ev) {
        PS_HIP(hipEventRecord(ev[6], c->stream));
        c->timing_n++;
    }
    return PS_OK;
}

static ps_status ensure_events(ps_ctx* c, size_t need) {
    while (c->events.size() < need) {
        hipEvent_t e;
        PS_HIP(hipEventCreate(&e));
        c->events.push_back(e);
    }
    return PS_OK;
}

static ps_status reduce_timing(ps_ctx* c, int frames, ps_frame_timing* timing) {
    PS_HIP(hipStreamSynchronize(c->stream));
    double acc[5] = {0, 0, 0, 0, 0};
    for (int f = 0; f < frames; ++f) {
        hipEvent_t* ev = &c->events[(size_t)f * kFrameEvents];
        for (int k = 0; k < 5; ++k) {
            // LUT bakes ev[k] -> ev[k + 1]; the render ev[5] -> ev[6]
            float ms = 0.f;
            PS_HIP(hipEventElapsedTime(&ms, ev[k < 4 ? k : 5], ev[k < 4 ? k + 1 : 6]));
            acc[k] += ms;
        }
    }
    float total = 0.f;
    PS_HIP(hipEventElapsedTime(&total, c->events[0], c->events[(size_t)(frames - 1) * kFrameEvents + 6]));
    timing->transmittance_ms = (float)(acc[0] / frames);
    timing->multiscatter_ms = (float)(acc[1] / frames);
    timing->skyview_ms = (float)(acc[2] / frames);
    timing->aerial_perspective_ms = (float)(acc[3] / frames);
    timing->render_ms = (float)(acc[4] / frames);
    timing->frame_ms = total / frames;
    timing->frames = frames;
    timing->_reserved = 0;
    return PS_OK;
}

ps_status ps_timing_begin(ps_ctx* c, int32_t max_frames) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_timing_begin: ctx is NULL");
    if (max_frames < 1 || max_frames > 65536)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_timing_begin: max_frames must be in [1, 65536]");
    ps_status s;
    if ((s = ensure_events(c, (size_t)max_frames * kFrameEvents))) return s;
    c->timing = true;
    c->timing_max = max_frames;
    c->timing_n = 0;
    return PS_OK;
}

ps_status ps_timing_end(ps_ctx* c, ps_frame_timing* timing) {
    if (!c || !timing) return fail(PS_ERR_INVALID_ARGUMENT, "ps_timing_end: NULL argument");
    if (!c->timing) return fail(PS_ERR_NOT_READY, "ps_timing_end: timing was not started");
    c->timing = false;
    if (c->timing_n == 0) return fail(PS_ERR_NOT_READY, "ps_timing_end: no frame was recorded");
    return reduce_timing(c, c->timing_n, timing);
}

ps_status ps_time_frames(ps_ctx* c, int32_t frames, uint32_t lut_flags, const float* depth, const ps_image_out* out,
                         ps_frame_timing* timing) {
    if (!c || !timing) return fail(PS_ERR_INVALID_ARGUMENT, "ps_time_frames: NULL argument");
    if (frames < 1 || frames > 4096) return fail(PS_ERR_INVALID_ARGUMENT, "ps_time_frames: frames must be in [1, 4096]");
    ps_status s;
    if (c->timing) return fail(PS_ERR_NOT_READY, "ps_time_frames: in-loop timing is active");
    if ((s = ensure_events(c, (size_t)frames * kFrameEvents))) return s;
    for (int f = 0; f < frames; ++f) {
        hipEvent_t* ev = &c->events[(size_t)f * kFrameEvents];
        if ((s = launch_luts(c, lut_flags, ev, c->stream))) return s;
        PS_HIP(hipEventRecord(ev[5], c->stream));
        if ((s = launch_render(c, depth, out, false))) return s;
        PS_HIP(hipEventRecord(ev[6], c->stream));
    }
    return reduce_timing(c, frames, timing);
}

ps_status ps_count_work(ps_ctx* c, const float* depth, ps_work_counters* out) {
    return ps_count_work_tiles(c, depth, out, nullptr, 0);
}

ps_status ps_count_work_tiles(ps_ctx* c, const float* depth, ps_work_counters* out, uint64_t* tile_times,
                              int32_t capacity) {
    if (!c || !out) return fail(PS_ERR_INVALID_ARGUMENT, "ps_count_work: NULL argument");
    if (tile_times && c->have_frame) {
        const int W = (int)c->sc.frame.viewport_size[0], H = (int)c->sc.frame.viewport_size[1];
        const size_t tiles = (size_t)((W + 7) / 8) * ((H + 7) / 8);
        if (capacity < 0 || (size_t)capacity < 2 * tiles)
            return fail(PS_ERR_INVALID_ARGUMENT, "ps_count_work_tiles: tile_times needs 2 entries per 8x8 tile");
        ps_status s0;
        if ((s0 = grow((float**)&c->d_tile_times, &c->tile_times_cap, 4 * tiles, "tile times"))) return s0;
    }
    c->want_tile_times = tile_times != nullptr;
    PS_HIP(hipMemsetAsync(c->d_counters, 0, 16 * sizeof(unsigned long long), c->stream));
    ps_status s = launch_render(c, depth, nullptr, true);
    c->want_tile_times = false;
    if (s) return s;
    if (tile_times) {
        const int W = (int)c->sc.frame.viewport_size[0], H = (int)c->sc.frame.viewport_size[1];
        const size_t tiles = (size_t)((W + 7) / 8) * ((H + 7) / 8);
        PS_HIP(hipMemcpyAsync(tile_times, c->d_tile_times, 16 * tiles, hipMemcpyDeviceToHost, c->stream));
    }
    unsigned long long h[16];
    PS_HIP(hipMemcpyAsync(h, c->d_counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    PS_HIP(hipStreamSynchronize(c->stream));
    out->pixels = h[0]; out->layer_marches = h[1]; out->primary_steps = h[2]; out->covered_steps = h[3];
    out->eroded_steps = h[4]; out->dense_samples = h[5]; out->light_steps = h[6]; out->light_covered = h[7];
    out->octaves = h[8]; out->ap_taps = h[9];
    return PS_OK;
}

// ---- differentiation -----------------------------------------------------------------
ps_status ps_set_cloud_extensions(ps_ctx* c, const float* extinction_scale, int32_t n, const float* density_delta) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_cloud_extensions: NULL context");
    if (n < 0 || n > PS_MAX_CLOUD_LAYERS || (n > 0 && !extinction_scale))
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_cloud_extensions: need 0 <= n <= 8 extinction scales");
    for (int l = 0; l < PS_MAX_CLOUD_LAYERS; ++l) {
        const float k = l < n ? extinction_scale[l] : 1.f;
        if (!(k >= 0.f) || !std::isfinite(k))
            return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_cloud_extensions: extinction scales must be finite and >= 0");
        c->sc.kappa[l] = k;
    }
    c->delta = density_delta;
    c->train_valid = false;
    return PS_OK;
}

ps_status ps_set_checkpoint_interval(ps_ctx* c, int32_t steps) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_checkpoint_interval: NULL context");
    if (steps < 1 || steps > 32)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_checkpoint_interval: the interval must be in [1, 32] steps");
    c->ck_len = steps;
    c->train_valid = false;
    return PS_OK;
}

ps_status ps_set_half_res(ps_ctx* c, int32_t enabled) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_half_res: NULL context");
    c->half_res = enabled != 0;
    return PS_OK;
}

ps_status ps_set_taa(ps_ctx* c, int32_t enabled) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_taa: NULL context");
    c->taa = enabled != 0;
    return PS_OK;
}

ps_status ps_invalidate_taa_history(ps_ctx* c) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_invalidate_taa_history: NULL context");
    for (ps_ctx::TaaView& v : c->taa_views) v.history_valid = false;
    return PS_OK;
}

ps_status ps_set_active_view(ps_ctx* c, int32_t view) {
    if (!c) return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_active_view: NULL context");
    if (view < 0 || view >= PS_MAX_VIEWS) return fail(PS_ERR_INVALID_ARGUMENT, "ps_set_active_view: view must be 0 or 1");
    c->active_view = view;
    return PS_OK;
}

ps_status ps_render_train(ps_ctx* c, const float* depth, const ps_image_out* out) {
    if (!c || !out) return fail(PS_ERR_INVALID_ARGUMENT, "ps_render_train: NULL argument");
    if (c->half_res) return fail(PS_ERR_UNSUPPORTED, "ps_render_train: the training forward runs at full resolution");
    if (c->taa) return fail(PS_ERR_UNSUPPORTED, "ps_render_train: the training forward renders single frames (TAA off)");
    ps_status s;
    if ((s = check_ready(c, true))) return s;
    if ((s = prepare_delta(c))) return s;
    RenderParams P;
    if ((s = fill_render_params(c, c->sc, c->luts, depth, out, P))) return s;
    const size_t tiles = (size_t)((P.W + 7) / 8) * ((P.H + 7) / 8);
    P.ck_len = c->ck_len;
    P.ck_per_seg = (P.steps + c->ck_len - 1) / c->ck_len;
    P.ck_max = std::max(P.n_layers, 1) * P.ck_per_seg;
    if ((s = grow(&c->ck, &c->ck_cap, tiles * P.ck_max * PS_CK_STRIDE, "transmittance checkpoints"))) return s;
    if ((s = grow(&c->px_state, &c->px_cap, tiles * PX_FIELDS * 64, "forward totals"))) return s;
    if ((s = grow(&c->lrec, &c->lrec_cap, tiles * (size_t)PS_LREC_CAP * LR_FIELDS * 64, "light-march records")))
        return s;
    P.ck = c->ck;
    P.px_state = c->px_state;
    P.lrec = c->lrec;
    P.lrec_cap = PS_LREC_CAP;
    RenderParams* dP = nullptr;
    if ((s = c->render_ring.upload(P, c->stream, &dP))) return s;
    PS_HIP(ps_launch_render(dP, P.W, P.H, false, c->delta != nullptr, true, c->stream));
    if ((s = c->render_ring.mark_used(c->stream))) return s;
    c->train = P;
    c->train_valid = true;
    return PS_OK;
}

ps_status ps_frame_train(ps_ctx* c, uint32_t lut_flags, const float* depth, const ps_image_out* out) {
    if (!c || !out) return fail(PS_ERR_INVALID_ARGUMENT, "ps_frame_train: NULL argument");
    if (c->half_res) return fail(PS_ERR_UNSUPPORTED, "ps_frame_train: the training forward runs at full resolution");
    if (c->taa) return fail(PS_ERR_UNSUPPORTED, "ps_frame_train: the training forward renders single frames (TAA off)");
    ps_status s;
    if ((s = bake_frame_luts(c, lut_flags, nullptr))) return s;
    return ps_render_train(c, depth, out);
}

ps_status ps_backward(ps_ctx* c, const float* g_composite, const float* g_luminance, const float* g_transmittance,
                      const ps_cloud_grads* grads) {
    if (!c || !grads) return fail(PS_ERR_INVALID_ARGUMENT, "ps_backward: NULL argument");
    if (!c->train_valid)
        return fail(PS_ERR_NOT_READY, "ps_backward: no training forward to differentiate (call ps_render_train; "
                                      "changing the extensions, the checkpoint interval, the LUT sizes, the noise "
                                      "format or the weather-map size invalidates it)");
    if (!grads->layers && !grads->density_delta)
        return fail(PS_ERR_INVALID_ARGUMENT, "ps_backward: no gradient output requested");
    RenderParams P = c->train;
    if (grads->density_delta && !P.cone && P.light_steps > 32)
        return fail(PS_ERR_UNSUPPORTED, "ps_backward: perturbation-volume gradients support at most 32 light steps");
    const size_t tiles = (size_t)((P.W + 7) / 8) * ((P.H + 7) / 8);
    ps_status s;
    if ((s = grow(&c->rec, &c->rec_cap, tiles * (size_t)P.ck_len * 8 * 64, "backward records"))) return s;
    if (!grads->layers) {
        if ((s = dev_alloc(&c->d_counters, 16 * sizeof(unsigned long long), "counters"))) return s;
    }
    P.g_comp = g_composite;
    P.g_lum = g_luminance;
    P.g_trn = g_transmittance;
    // without a caller buffer for the per-layer sums, the (ignored) sums go
    // to the context's counter scratch
    P.g_layers = grads->layers ? grads->layers : (float*)c->d_counters;
    P.g_delta = grads->density_delta;
    if (P.g_delta && !c->g_delta64) {
        // the backward leaves the accumulator at zero: cleared once here
        if ((s = dev_alloc(&c->g_delta64, sizeof(double) * 128 * 128 * 128, "FP64 volume gradient"))) return s;
        PS_HIP(ps_launch_zero_f64(c->g_delta64, 128 * 128 * 128, c->stream));
    }
    P.g_delta64 = c->g_delta64;
    P.rec = c->rec;
    RenderParams* dP = nullptr;
    if ((s = c->render_ring.upload(P, c->stream, &dP))) return s;
    PS_HIP(ps_launch_render_backward(dP, P.W, P.H, P.delta != nullptr, P.g_delta, c->g_delta64, c->stream));
    if (c->lut_stream && P.T == c->luts_back.T) {
        // the training forward's LUT set has since become the set the next
        // pipelined frame bakes into: that bake must wait for this backward
        PS_HIP(hipEventRecord(c->ev_render_issued, c->stream));
        c->render_issued_recorded = true;
    }
    return c->render_ring.mark_used(c->stream);
}

// ---- atmosphere differentiation ------------------------------------------------------
}  // extern "C"

namespace {

// Scratch LUTs, tangent LUTs and parameter blocks for n sets at the context's
// LUT sizes.
ps_status ensure_atmo(ps_ctx* c, int n) {
    const ps_lut_dims& d = c->dims;
    if (n <= c->atmo_cap && std::memcmp(&d, &c->atmo_dims, sizeof(d)) == 0) return PS_OK;
    if (c->atmo_ev) PS_HIP(hipEventSynchronize(c->atmo_ev));
    dev_free(c->atmo_params); dev_free(c->aT); dev_free(c->aMS); dev_free(c->aSV); dev_free(c->aAP);
    dev_free(c->aTT); dev_free(c->aMST); dev_free(c->aSVT); dev_free(c->aAPT);
    if (c->atmo_params_h) hipHostFree(c->atmo_params_h);
    c->atmo_params_h = nullptr;
    c->atmo_cap = 0;
    c->atmo_uploaded = 0;
    const size_t nt = (size_t)d.transmittance_w * d.transmittance_h, nm = (size_t)d.multiscatter_w * d.multiscatter_h,
                 ns = (size_t)d.skyview_w * d.skyview_h, na = (size_t)d.ap_w * d.ap_h * d.ap_d;
    ps_status s;
    if ((s = dev_alloc(&c->atmo_params, sizeof(LutParams) * n, "atmosphere parameter blocks"))) return s;
    if ((s = dev_alloc(&c->aT, sizeof(float4) * nt * n, "transmittance LUTs"))) return s;
    if ((s = dev_alloc(&c->aMS, sizeof(float4) * nm * n, "multi-scatter LUTs"))) return s;
    if ((s = dev_alloc(&c->aSV, sizeof(float4) * ns * n, "sky-view LUTs"))) return s;
    if ((s = dev_alloc(&c->aAP, sizeof(float4) * na * n, "aerial-perspective LUTs"))) return s;
    if ((s = dev_alloc(&c->aTT, sizeof(float) * 18 * nt * n, "transmittance tangents"))) return s;
    if ((s = dev_alloc(&c->aMST, sizeof(float) * 18 * nm * n, "multi-scatter tangents"))) return s;
    if ((s = dev_alloc(&c->aSVT, sizeof(float) * 18 * ns * n, "sky-view tangents"))) return s;
    if ((s = dev_alloc(&c->aAPT, sizeof(float) * 36 * na * n, "aerial-perspective tangents"))) return s;
    PS_HIP(hipHostMalloc((void**)&c->atmo_params_h, sizeof(LutParams) * n, hipHostMallocDefault));
    if (!c->atmo_ev) PS_HIP(hipEventCreateWithFlags(&c->atmo_ev, hipEventDisableTiming));
    c->atmo_cap = n;
    c->atmo_dims = d;
    return PS_OK;
}

// The frame geometry of the sky + AP image (the quantities fill_render_params
// folds for the sky pass).
void sky_geometry(const psi::SceneIn& sc, SkyApLaunch& L) {
    const ps_frame_uniforms& f = sc.frame;
    const ps_atmosphere_params& a = sc.atm;
    L.W = (int)f.viewport_size[0];
    L.H = (int)f.viewport_size[1];
    for (int i = 0; i < 16; ++i) L.inv_vp[i] = f.inv_view_proj[i];
    const double R = a.planet_radius_m, top = a.atmosphere_top_m;
    for (int i = 0; i < 3; ++i) {
        L.cam[i] = f.camera_position_world[i];
        L.sun_d[i] = f.sun_direction[i];
        L.sun[i] = f.sun_direction[i];
        L.E[i] = f.sun_illuminance[i];
    }
    L.sun_ang = f.sun_direction[3];
    L.cos_disk = std::cos((double)f.sun_direction[3]);
    L.R = (float)R;
    L.top = (float)top;
    L.big_h = (float)std::sqrt(std::max(top * top - R * R, 0.0));
    L.top_minus_R = (float)(top - R);
    const double oy = L.cam[1] + R;
    const double r_cam = std::sqrt(L.cam[0] * L.cam[0] + oy * oy + L.cam[2] * L.cam[2]);
    const double h_cam = (L.cam[0] * L.cam[0] + L.cam[2] * L.cam[2] + L.cam[1] * (L.cam[1] + 2.0 * R)) / (r_cam + R);
    L.cam_alt = (float)h_cam;
    const double cam_r = std::max(r_cam, R + 1.0);
    L.cam_zha = (float)(PS_PI - std::asin(std::min(std::max(R / cam_r, 0.0), 1.0)));
    L.sun_az = (float)std::atan2(L.sun_d[0], L.sun_d[2]);
}

// Bake the n value LUT sets (coefficients from `coeffs`, or the context's
// atmosphere when null), optionally their tangents, and run the image pass.
ps_status sky_ap_run(ps_ctx* c, int n, const float* coeffs, const float* depth, float ground, const float* target,
                     const float* weight, float* out, float* image, bool grad) {
    ps_status s;
    if ((s = check_ready(c, false))) return s;
    const float vw = c->sc.frame.viewport_size[0], vh = c->sc.frame.viewport_size[1];
    if (!(vw >= 1.f && vh >= 1.f && vw <= 32768.f && vh <= 32768.f))
        return fail(PS_ERR_INVALID_ARGUMENT, "frame.viewport_size must hold the image size");
    if ((s = ensure_atmo(c, n))) return s;
    LutParams base;
    fill_lut_params(c->sc, c->dims, base);
    // The device blocks must equal `base` except for the 14 coefficients,
    // which ps_launch_set_coeffs rewrites when `coeffs` is given. Restage
    // when the base changed, when more blocks are needed, or when a call
    // without coefficients follows a coefficient overwrite.
    const bool staged = c->atmo_uploaded >= n && std::memcmp(&base, c->atmo_params_h, sizeof(base)) == 0;
    if (!staged || (!coeffs && c->atmo_coeffs_dirty)) {
        PS_HIP(hipEventSynchronize(c->atmo_ev));
        for (int i = 0; i < n; ++i) c->atmo_params_h[i] = base;
        PS_HIP(hipMemcpyAsync(c->atmo_params, c->atmo_params_h, sizeof(LutParams) * n, hipMemcpyHostToDevice,
                              c->stream));
        PS_HIP(hipEventRecord(c->atmo_ev, c->stream));
        c->atmo_uploaded = n;
        c->atmo_coeffs_dirty = false;
    }
    if (coeffs) {
        PS_HIP(ps_launch_set_coeffs(c->atmo_params, coeffs, n, c->stream));
        c->atmo_coeffs_dirty = true;
    }
    PS_HIP(ps_launch_luts(c->atmo_params, base, n, 15u, c->aT, c->aMS, c->aSV, c->aAP, c->stream, nullptr));
    if (grad)
        PS_HIP(ps_launch_lut_tangents(c->atmo_params, base, n, c->aT, c->aMS, c->aTT, c->aMST, c->aSVT, c->aAPT,
                                      c->stream));
    SkyApLaunch L{};
    sky_geometry(c->sc, L);
    L.lp = c->atmo_params;
    L.T = c->aT; L.SV = c->aSV; L.AP = c->aAP;
    L.TT = c->aTT; L.SVT = c->aSVT; L.APT = c->aAPT;
    L.depth = depth; L.target = target; L.weight = weight; L.out = out; L.image = image;
    L.ground = ground;
    L.want_grad = grad ? 1 : 0;
    PS_HIP(ps_launch_sky_ap_grad(L, n, c->stream));
    return PS_OK;
}

}  // namespace

extern "C" {

ps_status ps_sky_ap_image(ps_ctx* c, const float* depth, float ground_radiance, float* out) {
    if (!c || !out) return fail(PS_ERR_INVALID_ARGUMENT, "ps_sky_ap_image: NULL argument");
    return sky_ap_run(c, 1, nullptr, depth, ground_radiance, nullptr, nullptr, nullptr, out, false);
}

ps_status ps_atmosphere_grad(ps_ctx* c, int32_t n, const float* coeffs, const float* depth, float ground_radiance,
                             const float* target, const float* weight, float* out, float* image) {
    if (!c || !coeffs || !target || !out) return fail(PS_ERR_INVALID_ARGUMENT, "ps_atmosphere_grad: NULL argument");
    if (n < 1 || n > 4096) return fail(PS_ERR_INVALID_ARGUMENT, "ps_atmosphere_grad: n must be in [1, 4096]");
    return sky_ap_run(c, n, coeffs, depth, ground_radiance, target, weight, out, image, true);
}

}  // extern "C"
