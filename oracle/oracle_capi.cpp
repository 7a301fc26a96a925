// Oracle (test infrastructure, not product code): extern "C" entry points of
// the CPU restatement, loaded by tests/ through ctypes and by bench.py's CPU
// baseline. `or_*` run in double precision (the parity oracle); `orf_*` are
// the float instantiation of the same code (the CPU port of the reference
// used as the timing baseline).
#include <omp.h>
#include <vector>
#include <cstring>
#include "oracle_clouds.hpp"
#include "oracle_noise.hpp"
#include "oracle_taa.hpp"

using namespace oracle;

extern "C" {

typedef struct or_inputs {
    const ps_frame_uniforms* frame;
    const ps_atmosphere_params* atm;
    const ps_cloud_params* clouds;
    const ps_cloud_layer* layers;
    int32_t n_layers;
    const float* layer_wind_xz;  // n_layers * 2 or null
    const void* T; int32_t tw, th;           // LUTs, RGBA in the instantiation's scalar type
    const void* SV; int32_t sw, sh;
    const void* AP; int32_t aw, ah, ad;
    const float* base;           // 128^3 RGBA decoded
    const float* detail;         // 32^3 RGBA decoded
    const float* curl;           // 128^2 RG decoded
    const float* weather; int32_t ww, wh, wch;
    const uint8_t* type_grid;    // 128^2 or null
    const uint8_t* overcast;     // 128^2 or null
    const uint8_t* blue;         // 64^2
    const float* depth;          // W*H or null
} or_inputs;

// Differentiation extensions and parameter overrides in double precision
// (FP64 finite differences need perturbations finer than the float layer
// structs hold). Null arrays keep the inputs' values (kappa 1, delta 0).
// The tracked parameters seed the tangents of the dual instantiation:
// kind 0 density_scale[index], 1 extinction_scale[index], 2 coverage[index],
// 3 delta voxel `index` (x + 128 (h + 128 z)).
typedef struct or_diff {
    const double* density_scale;     // per layer or null
    const double* extinction_scale;  // per layer or null
    const double* coverage;          // per layer or null
    const double* delta;             // 128^3 or null
    int32_t n_params;
    const int32_t* param_kind;
    const int32_t* param_index;
} or_diff;

}  // extern "C"

namespace {

template <class R>
Scene<R> make_scene(const or_inputs* in) {
    Scene<R> s;
    s.f = frame_from<R>(*in->frame);
    s.a = atm_from<R>(*in->atm);
    s.c = clouds_from<R>(*in->clouds);
    s.n_layers = std::min<int>(in->n_layers, PS_MAX_CLOUD_LAYERS);
    s.rt_w = 0;
    s.rt_h = 0;
    for (int i = 0; i < s.n_layers; ++i) {
        s.layers[i] = layer_from<R>(in->layers[i]);
        if (in->layer_wind_xz) s.layer_wind[i] = V3<R>(in->layer_wind_xz[2 * i], 0, in->layer_wind_xz[2 * i + 1]);
        else s.layer_wind[i] = V3<R>(0);
    }
    typedef typename LutStore<R>::type L;
    s.T.data = static_cast<const L*>(in->T); s.T.w = in->tw; s.T.h = in->th; s.T.ch = 4;
    s.SV.data = static_cast<const L*>(in->SV); s.SV.w = in->sw; s.SV.h = in->sh; s.SV.ch = 4;
    s.AP.data = static_cast<const L*>(in->AP); s.AP.w = in->aw; s.AP.h = in->ah; s.AP.d = in->ad; s.AP.ch = 4;
    s.base.data = in->base; s.base.w = s.base.h = s.base.d = 128; s.base.ch = 4;
    s.detail.data = in->detail; s.detail.w = s.detail.h = s.detail.d = 32; s.detail.ch = 4;
    s.curl.data = in->curl; s.curl.w = s.curl.h = 128; s.curl.ch = 2;
    s.weather.data = in->weather; s.weather.w = in->ww; s.weather.h = in->wh; s.weather.ch = in->wch;
    s.type_grid = in->type_grid;
    s.overcast = in->overcast;
    s.blue = in->blue;
    s.depth = in->depth;
    for (int i = 0; i < PS_MAX_CLOUD_LAYERS; ++i) s.kappa[i] = R(1);
    s.delta = nullptr;
    s.delta_track = s.delta_track_slot = nullptr;
    s.n_delta_track = 0;
    return s;
}

// Overrides and tangent seeds of `d` on a scene built by make_scene.
template <class R>
void apply_diff(Scene<R>& s, const or_diff* d, std::vector<int>& track, std::vector<int>& slot) {
    if (!d) return;
    for (int l = 0; l < s.n_layers; ++l) {
        if (d->density_scale) s.layers[l].density_scale = R(d->density_scale[l]);
        if (d->extinction_scale) s.kappa[l] = R(d->extinction_scale[l]);
        if (d->coverage) s.layers[l].coverage = R(d->coverage[l]);
    }
    s.delta = d->delta;
    if constexpr (is_dual<R>::value) {
        for (int k = 0; k < d->n_params; ++k) {
            const int kind = d->param_kind[k], idx = d->param_index[k];
            if (kind == 0) s.layers[idx].density_scale.d[k] = 1.0;
            else if (kind == 1) s.kappa[idx].d[k] = 1.0;
            else if (kind == 2) s.layers[idx].coverage.d[k] = 1.0;
            else {
                track.push_back(idx);
                slot.push_back(k);
            }
        }
        s.delta_track = track.data();
        s.delta_track_slot = slot.data();
        s.n_delta_track = (int)track.size();
    }
}

template <class R>
void render_region(const or_inputs* in, int x0, int y0, int w, int h, R* comp, R* lum, R* trn,
                   uint64_t* counters, int nthreads, const or_diff* diff = nullptr) {
    Scene<R> s = make_scene<R>(in);
    std::vector<int> track, slot;
    apply_diff(s, diff, track, slot);
    const int n = w * h;
    Counters total;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        Counters local;
        Counters* cp = counters ? &local : nullptr;
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < n; ++i) {
            int ix = x0 + i % w, iy = y0 + i / w;
            R c4[4], l4[4], t4[4];
            render_pixel(s, ix, iy, c4, l4, t4, cp);
            for (int k = 0; k < 4; ++k) {
                if (comp) comp[size_t(i) * 4 + k] = c4[k];
                if (lum) lum[size_t(i) * 4 + k] = l4[k];
                if (trn) trn[size_t(i) * 4 + k] = t4[k];
            }
        }
        if (counters) {
#pragma omp critical
            {
                total.pixels += local.pixels; total.layer_marches += local.layer_marches;
                total.primary_steps += local.primary_steps; total.covered_steps += local.covered_steps;
                total.eroded_steps += local.eroded_steps; total.dense_samples += local.dense_samples;
                total.light_steps += local.light_steps; total.light_covered += local.light_covered;
                total.octaves += local.octaves; total.ap_taps += local.ap_taps;
            }
        }
    }
    if (counters) {
        uint64_t v[10] = {total.pixels, total.layer_marches, total.primary_steps, total.covered_steps,
                          total.eroded_steps, total.dense_samples, total.light_steps, total.light_covered,
                          total.octaves, total.ap_taps};
        std::memcpy(counters, v, sizeof(v));
    }
}

template <class R>
void bake_ms(const ps_atmosphere_params* atm, const R* T, int tw, int th, int W, int H, R* out, int nthreads) {
    Atm<R> a = atm_from<R>(*atm);
    Tex<R, R> t; t.data = T; t.w = tw; t.h = th; t.ch = 4;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < W * H; ++i) {
        int x = i % W, y = i / W;
        bake_multiscatter_texel(a, t, (R(x) + R(0.5)) / R(W), (R(y) + R(0.5)) / R(H), out + size_t(i) * 4);
    }
}

template <class R>
void bake_sv(const ps_frame_uniforms* fr, const ps_atmosphere_params* atm, const R* T, int tw, int th,
             const R* MS, int mw, int mh, int W, int H, R* out, int nthreads) {
    Atm<R> a = atm_from<R>(*atm);
    Frame<R> f = frame_from<R>(*fr);
    Tex<R, R> t; t.data = T; t.w = tw; t.h = th; t.ch = 4;
    Tex<R, R> m; m.data = MS; m.w = mw; m.h = mh; m.ch = 4;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < W * H; ++i) {
        int x = i % W, y = i / W;
        bake_skyview_texel(f, a, t, m, (R(x) + R(0.5)) / R(W), (R(y) + R(0.5)) / R(H), out + size_t(i) * 4);
    }
}

template <class R>
void bake_aerial(const ps_frame_uniforms* fr, const ps_atmosphere_params* atm, const R* T, int tw, int th,
                 const R* MS, int mw, int mh, int W, int H, int D, R* out, int nthreads) {
    Atm<R> a = atm_from<R>(*atm);
    Frame<R> f = frame_from<R>(*fr);
    Tex<R, R> t; t.data = T; t.w = tw; t.h = th; t.ch = 4;
    Tex<R, R> m; m.data = MS; m.w = mw; m.h = mh; m.ch = 4;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < W * H * D; ++i) {
        int x = i % W, y = (i / W) % H, z = i / (W * H);
        bake_ap_froxel(f, a, t, m, x, y, z, W, H, D, out + size_t(i) * 4);
    }
}

template <class R>
void bake_t(const ps_atmosphere_params* atm, int W, int H, R* out, int nthreads) {
    Atm<R> a = atm_from<R>(*atm);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            R u = (R(x) + R(0.5)) / R(W), v = (R(y) + R(0.5)) / R(H);
            V3<R> p, dir;
            transmittance_lut_uv_to_pos_dir(a, u, v, p, dir);
            V3<R> od = integrate_optical_depth(a, p, dir, 40);
            V3<R> tt = vexp(-od);
            R* o = out + (size_t(y) * W + x) * 4;
            o[0] = tt.x; o[1] = tt.y; o[2] = tt.z; o[3] = R(1);
        }
}

}  // namespace

extern "C" {

// ---- LUT chain -----------------------------------------------------------------
void or_bake_transmittance(const ps_atmosphere_params* atm, int W, int H, double* out, int nthreads) {
    bake_t<double>(atm, W, H, out, nthreads);
}
void or_bake_multiscatter(const ps_atmosphere_params* atm, const double* T, int tw, int th, int W, int H,
                          double* out, int nthreads) {
    bake_ms<double>(atm, T, tw, th, W, H, out, nthreads);
}
void or_bake_skyview(const ps_frame_uniforms* f, const ps_atmosphere_params* atm, const double* T, int tw, int th,
                     const double* MS, int mw, int mh, int W, int H, double* out, int nthreads) {
    bake_sv<double>(f, atm, T, tw, th, MS, mw, mh, W, H, out, nthreads);
}
void or_bake_ap(const ps_frame_uniforms* f, const ps_atmosphere_params* atm, const double* T, int tw, int th,
                const double* MS, int mw, int mh, int W, int H, int D, double* out, int nthreads) {
    bake_aerial<double>(f, atm, T, tw, th, MS, mw, mh, W, H, D, out, nthreads);
}
void orf_bake_transmittance(const ps_atmosphere_params* atm, int W, int H, float* out, int nthreads) {
    bake_t<float>(atm, W, H, out, nthreads);
}
void orf_bake_multiscatter(const ps_atmosphere_params* atm, const float* T, int tw, int th, int W, int H,
                           float* out, int nthreads) {
    bake_ms<float>(atm, T, tw, th, W, H, out, nthreads);
}
void orf_bake_skyview(const ps_frame_uniforms* f, const ps_atmosphere_params* atm, const float* T, int tw, int th,
                      const float* MS, int mw, int mh, int W, int H, float* out, int nthreads) {
    bake_sv<float>(f, atm, T, tw, th, MS, mw, mh, W, H, out, nthreads);
}
void orf_bake_ap(const ps_frame_uniforms* f, const ps_atmosphere_params* atm, const float* T, int tw, int th,
                 const float* MS, int mw, int mh, int W, int H, int D, float* out, int nthreads) {
    bake_aerial<float>(f, atm, T, tw, th, MS, mw, mh, W, H, D, out, nthreads);
}

// ---- per-pixel sky + cloud path ------------------------------------------------------
// Renders the w x h region at (x0, y0) of the frame; outputs are w*h RGBA.
// counters (10 x u64, ps_work_counters order) may be null.
void or_render(const or_inputs* in, int x0, int y0, int w, int h, double* comp, double* lum, double* trn,
               uint64_t* counters, int nthreads) {
    render_region<double>(in, x0, y0, w, h, comp, lum, trn, counters, nthreads);
}
void orf_render(const or_inputs* in, int x0, int y0, int w, int h, float* comp, float* lum, float* trn,
                uint64_t* counters, int nthreads) {
    render_region<float>(in, x0, y0, w, h, comp, lum, trn, counters, nthreads);
}
// The half-resolution path (crates/ps-clouds/src/lib.rs:748-760 RT size,
// pipeline.rs:277-319 march mapping, cloud_composite_halfres.wgsl): marches
// the RT texels the w x h region's Catmull-Rom taps read, then composites
// the region at full resolution. Outputs w*h*4 (any may be null).
void or_render_halfres(const or_inputs* in, int x0, int y0, int w, int h, double* comp, double* lum, double* trn,
                       int nthreads) {
    Scene<double> s = make_scene<double>(in);
    const double vw = s.f.vw, vh = s.f.vh;
    s.rt_w = std::max(1, int(float(vw) * 0.5f));
    s.rt_h = std::max(1, int(float(vh) * 0.5f));
    // RT texels read by the region: floor((x + 0.5) / vw * rt_w - 0.5) - 1 .. + 2
    auto lo = [](double p) { return std::max(int(std::floor(p - 0.5)) - 1, 0); };
    auto hi = [](double p, int n) { return std::min(int(std::floor(p - 0.5)) + 2, n - 1); };
    const int tx0 = lo((x0 + 0.5) / vw * s.rt_w), tx1 = hi((x0 + w - 0.5) / vw * s.rt_w, s.rt_w);
    const int ty0 = lo((y0 + 0.5) / vh * s.rt_h), ty1 = hi((y0 + h - 0.5) / vh * s.rt_h, s.rt_h);
    std::vector<double> rl(size_t(s.rt_w) * s.rt_h * 4, 0.0), rt(size_t(s.rt_w) * s.rt_h * 4, 0.0);
    const int tw = tx1 - tx0 + 1, th = ty1 - ty0 + 1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < tw * th; ++i) {
        const int ix = tx0 + i % tw, iy = ty0 + i / tw;
        V3<double> cl, ct;
        cloud_march_pixel(s, ix, iy, cl, ct, (Counters*)nullptr);
        const size_t o = (size_t(iy) * s.rt_w + ix) * 4;
        rl[o] = cl.x; rl[o + 1] = cl.y; rl[o + 2] = cl.z; rl[o + 3] = 1.0;
        rt[o] = ct.x; rt[o + 1] = ct.y; rt[o + 2] = ct.z; rt[o + 3] = 1.0;
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < w * h; ++i) {
        double c4[4], l4[4], t4[4];
        render_pixel_half(s, x0 + i % w, y0 + i / w, rl.data(), rt.data(), c4, l4, t4);
        for (int k = 0; k < 4; ++k) {
            if (comp) comp[size_t(i) * 4 + k] = c4[k];
            if (lum) lum[size_t(i) * 4 + k] = l4[k];
            if (trn) trn[size_t(i) * 4 + k] = t4[k];
        }
    }
}

// The cloud march into an RT of rw x rh texels covering the frame's viewport
// (rw = W: the full-resolution march; floor(W / 2): the half-resolution
// one), luminance and transmittance RGBA, rw*rh*4 each; `d` (or null) holds
// the differentiation extensions and overrides as in or_render_ext.
void or_march_rt(const or_inputs* in, const or_diff* d, int rw, int rh, double* lum, double* trn, int nthreads) {
    Scene<double> s = make_scene<double>(in);
    std::vector<int> track, slot;
    apply_diff(s, d, track, slot);
    s.rt_w = rw;
    s.rt_h = rh;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < rw * rh; ++i) {
        V3<double> cl, ct;
        cloud_march_pixel(s, i % rw, i / rw, cl, ct, (Counters*)nullptr);
        const size_t o = size_t(i) * 4;
        lum[o] = cl.x; lum[o + 1] = cl.y; lum[o + 2] = cl.z; lum[o + 3] = 1.0;
        trn[o] = ct.x; trn[o + 1] = ct.y; trn[o + 2] = ct.z; trn[o + 3] = 1.0;
    }
}

// The TAA resolve of a whole rw x rh RT (cloud_taa.wgsl) with the frame's
// inv_view_proj, camera position and prev_view_proj.
void or_taa_resolve(const or_inputs* in, int rw, int rh, const double* cur_l, const double* cur_t,
                    const double* hist_l, const double* hist_t, double blend, double history_valid, double* out_l,
                    double* out_t) {
    for (int y = 0; y < rh; ++y)
        for (int x = 0; x < rw; ++x) {
            const size_t o = (size_t(y) * rw + x) * 4;
            taa_resolve_texel(*in->frame, rw, rh, x, y, cur_l, cur_t, hist_l, hist_t, blend, history_valid,
                              out_l + o, out_t + o);
        }
}

// The full-resolution composite of an RT over a w x h region: `upsample` 1 =
// the half-resolution composite (Catmull-Rom, clamps), 0 = the plain
// composite of an RT of the frame's own size.
void or_composite_rt(const or_inputs* in, int rw, int rh, const double* rt_lum, const double* rt_trn, int upsample,
                     int x0, int y0, int w, int h, double* comp, double* lum, double* trn, int nthreads) {
    Scene<double> s = make_scene<double>(in);
    s.rt_w = rw;
    s.rt_h = rh;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < w * h; ++i) {
        double c4[4], l4[4], t4[4];
        if (upsample)
            render_pixel_half(s, x0 + i % w, y0 + i / w, rt_lum, rt_trn, c4, l4, t4);
        else
            render_pixel_rt(s, x0 + i % w, y0 + i / w, rt_lum, rt_trn, c4, l4, t4);
        for (int k = 0; k < 4; ++k) {
            if (comp) comp[size_t(i) * 4 + k] = c4[k];
            if (lum) lum[size_t(i) * 4 + k] = l4[k];
            if (trn) trn[size_t(i) * 4 + k] = t4[k];
        }
    }
}

// The two LUT-sampled light terms of a dense cloud sample at world position
// p (n points, p and the outputs n*3): the sun transmittance
// (cloud_march.wgsl:328-354) and the sky ambient before its strength and
// height gradient (cloud_march.wgsl:693-700).
void or_sun_transmittance_at(const or_inputs* in, const double* p, int n, double* out) {
    Scene<double> s = make_scene<double>(in);
    for (int i = 0; i < n; ++i) {
        V3<double> v = sun_transmittance_at(s, V3<double>(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
        out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
    }
}
void or_sky_ambient_at(const or_inputs* in, const double* p, int n, double* out) {
    Scene<double> s = make_scene<double>(in);
    for (int i = 0; i < n; ++i) {
        V3<double> v = sample_sky_ambient(s, V3<double>(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
        out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
    }
}

// Renders an arbitrary set of n pixels given as (x, y) pairs; outputs n*4.
void or_render_pixels(const or_inputs* in, const int32_t* xy, int n, double* comp, double* lum, double* trn,
                      int nthreads) {
    Scene<double> s = make_scene<double>(in);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        double c4[4], l4[4], t4[4];
        render_pixel(s, xy[2 * i], xy[2 * i + 1], c4, l4, t4, (Counters*)nullptr);
        for (int k = 0; k < 4; ++k) {
            if (comp) comp[size_t(i) * 4 + k] = c4[k];
            if (lum) lum[size_t(i) * 4 + k] = l4[k];
            if (trn) trn[size_t(i) * 4 + k] = t4[k];
        }
    }
}

// ---- differentiation (extensions + exact derivatives) ------------------------------------
// The FP64 render with the differentiation extensions and overrides of `d`.
void or_render_ext(const or_inputs* in, const or_diff* d, int x0, int y0, int w, int h, double* comp,
                   double* lum, double* trn, uint64_t* counters, int nthreads) {
    render_region<double>(in, x0, y0, w, h, comp, lum, trn, counters, nthreads, d);
}

constexpr int kDualN = 16;

// Vector-Jacobian product of the w x h region at (x0, y0): grad[k] = sum over
// pixels and RGB of g_comp * d comp / d theta_k + g_lum * d lum / d theta_k +
// g_trn * d trn / d theta_k for the (at most 16) tracked parameters of `d`,
// computed with the Dual<16> instantiation of the FP64 restatement. The
// cotangent images are w*h RGBA (alpha ignored), any of them may be null.
// near_tol > 0: *near_mask receives the tangent slots that a quantity
// within near_tol of one of the march's decision thresholds depends on (see
// Counters and note_near).
int or_vjp_clouds(const or_inputs* in, const or_diff* d, int x0, int y0, int w, int h, const double* g_comp,
                  const double* g_lum, const double* g_trn, double* grad, int nthreads, double near_tol,
                  uint32_t* near_mask) {
    typedef Dual<kDualN> D;
    if (!d || d->n_params < 0 || d->n_params > kDualN) return 1;
    Scene<D> s = make_scene<D>(in);
    std::vector<int> track, slot;
    apply_diff(s, d, track, slot);
    const int n = w * h, np = d->n_params;
    std::vector<double> total(np, 0.0);
    uint32_t mask = 0u;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        std::vector<double> part(np, 0.0);
        Counters cnt;
        cnt.near_tol = near_tol;
#pragma omp for schedule(dynamic, 4)
        for (int i = 0; i < n; ++i) {
            int ix = x0 + i % w, iy = y0 + i / w;
            D c4[4], l4[4], t4[4];
            render_pixel(s, ix, iy, c4, l4, t4, near_tol > 0.0 ? &cnt : (Counters*)nullptr);
            for (int c = 0; c < 3; ++c) {
                const size_t o = size_t(i) * 4 + c;
                for (int k = 0; k < np; ++k) {
                    double v = 0.0;
                    if (g_comp) v += g_comp[o] * c4[c].d[k];
                    if (g_lum) v += g_lum[o] * l4[c].d[k];
                    if (g_trn) v += g_trn[o] * t4[c].d[k];
                    part[k] += v;
                }
            }
        }
#pragma omp critical
        {
            for (int k = 0; k < np; ++k) total[k] += part[k];
            mask |= cnt.near_mask;
        }
    }
    for (int k = 0; k < np; ++k) grad[k] = total[k];
    if (near_mask) *near_mask = mask;
    return 0;
}

// ---- atmosphere differentiation: sky + AP image with FP64 coefficients ----------------------
// coeffs (14 doubles, or null = the struct's values): rayleigh_scattering[3],
// mie_scattering[3], mie_absorption[3], ozone_absorption[3],
// rayleigh_scale_height_m, mie_scale_height_m.
static void apply_coeffs(Atm<double>& a, const double* k) {
    if (!k) return;
    a.ray_s = V3<double>(k[0], k[1], k[2]);
    a.mie_s = V3<double>(k[3], k[4], k[5]);
    a.mie_a = V3<double>(k[6], k[7], k[8]);
    a.ozone = V3<double>(k[9], k[10], k[11]);
    a.hr = k[12];
    a.hm = k[13];
}

// Bakes the four LUTs in FP64 with the coefficients, then renders the w x h
// region at (x0, y0) of the sky + AP image (sky_ap_pixel) into out (w*h*3).
// Only frame, atm and depth of `in` are read. luts_out (optional, may be
// null): T, MS, SV, AP concatenated as RGBA doubles.
void or_sky_ap_image(const or_inputs* in, const double* coeffs, int tw, int th, int mw, int mh, int sw, int sh,
                     int aw, int ah, int ad, double ground_radiance, int x0, int y0, int w, int h, double* out,
                     int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    Atm<double> a = atm_from<double>(*in->atm);
    apply_coeffs(a, coeffs);
    Frame<double> f = frame_from<double>(*in->frame);
    std::vector<double> T(size_t(tw) * th * 4), MS(size_t(mw) * mh * 4), SV(size_t(sw) * sh * 4),
        AP(size_t(aw) * ah * ad * 4);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < tw * th; ++i) {
        int x = i % tw, y = i / tw;
        V3<double> p, dir;
        transmittance_lut_uv_to_pos_dir(a, (x + 0.5) / tw, (y + 0.5) / th, p, dir);
        V3<double> t = vexp(-integrate_optical_depth(a, p, dir, 40));
        double* o = &T[size_t(i) * 4];
        o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = 1.0;
    }
    Tex<double, double> tt; tt.data = T.data(); tt.w = tw; tt.h = th; tt.ch = 4;
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < mw * mh; ++i)
        bake_multiscatter_texel(a, tt, (i % mw + 0.5) / mw, (i / mw + 0.5) / mh, &MS[size_t(i) * 4]);
    Tex<double, double> mt; mt.data = MS.data(); mt.w = mw; mt.h = mh; mt.ch = 4;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < sw * sh; ++i)
        bake_skyview_texel(f, a, tt, mt, (i % sw + 0.5) / sw, (i / sw + 0.5) / sh, &SV[size_t(i) * 4]);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < aw * ah * ad; ++i)
        bake_ap_froxel(f, a, tt, mt, i % aw, (i / aw) % ah, i / (aw * ah), aw, ah, ad, &AP[size_t(i) * 4]);
    Scene<double> s{};
    s.f = f;
    s.a = a;
    s.depth = in->depth;
    s.T.data = T.data(); s.T.w = tw; s.T.h = th;
    s.SV.data = SV.data(); s.SV.w = sw; s.SV.h = sh;
    s.AP.data = AP.data(); s.AP.w = aw; s.AP.h = ah; s.AP.d = ad;
    s.n_layers = 0;
    s.overcast = nullptr;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < w * h; ++i) {
        V3<double> v = sky_ap_pixel(s, x0 + i % w, y0 + i / w, ground_radiance);
        out[size_t(i) * 3] = v.x; out[size_t(i) * 3 + 1] = v.y; out[size_t(i) * 3 + 2] = v.z;
    }
}

// ---- noise bakes (decoded float values before storage quantisation) ---------------------
void or_noise_base(uint32_t seed, int z0, int z1, double* out, int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const int per = 128 * 128;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < (z1 - z0) * per; ++i) {
        int z = z0 + i / per, y = (i % per) / 128, x = i % 128;
        base_texel<double>(x, y, z, seed, out + size_t(i) * 4);
    }
}
void or_noise_detail(uint32_t seed, double* out, int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < 32 * 32 * 32; ++i) {
        int z = i / 1024, y = (i % 1024) / 32, x = i % 32;
        detail_texel<double>(x, y, z, seed, out + size_t(i) * 4);
    }
}
void or_noise_curl(uint32_t seed, double* out, int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < 128 * 128; ++i) curl_texel<double>(i % 128, i / 128, seed, out + size_t(i) * 2);
}

// ---- single functions, for closed-form checks ------------------------------------------
double or_henyey_greenstein(double c, double g) { return henyey_greenstein(c, g); }
double or_draine_phase(double c, double g, double alpha) { return draine_phase(c, g, alpha); }
double or_phase_rayleigh(double c) { return phase_rayleigh(c); }
double or_phase_mie(double c, double g) { return phase_mie(c, g); }
double or_cloud_phase(double c, double g_scale, double droplet_um, double bias, uint32_t type, double h_norm) {
    Scene<double> s;
    s.c.droplet_bias = bias;
    Layer<double> L;
    L.droplet = droplet_um; L.type = type;
    PhaseConsts<double> k = cloud_phase_constants(s, L, h_norm);
    return cloud_phase_with(c, g_scale, k);
}
double or_ndf(double h, uint32_t type, double anvil) { return ndf(h, type, anvil); }
void or_transmittance_uv(const ps_atmosphere_params* atm, double r, double mu, double* uv) {
    Atm<double> a = atm_from<double>(*atm);
    V3<double> p(0, r, 0), d(std::sqrt(std::max(0.0, 1 - mu * mu)), mu, 0);
    transmittance_lut_uv(a, p, d, uv[0], uv[1]);
}
void or_transmittance_pos_dir(const ps_atmosphere_params* atm, double u, double v, double* r_mu) {
    Atm<double> a = atm_from<double>(*atm);
    V3<double> p, d;
    transmittance_lut_uv_to_pos_dir(a, u, v, p, d);
    r_mu[0] = p.y; r_mu[1] = d.y;
}
void or_optical_depth(const ps_atmosphere_params* atm, double r, double mu, int n_steps, double* od) {
    Atm<double> a = atm_from<double>(*atm);
    V3<double> p(0, r, 0), d(std::sqrt(std::max(0.0, 1 - mu * mu)), mu, 0);
    V3<double> o = integrate_optical_depth(a, p, d, n_steps);
    od[0] = o.x; od[1] = o.y; od[2] = o.z;
}
void or_skyview_uv(const ps_atmosphere_params* atm, const double* view_dir, const double* sun, double cam_r,
                   double* uv) {
    Atm<double> a = atm_from<double>(*atm);
    skyview_uv(a, V3<double>(view_dir[0], view_dir[1], view_dir[2]), V3<double>(sun[0], sun[1], sun[2]), cam_r,
               uv[0], uv[1]);
}
void or_sample3d_repeat(const float* data, int w, int h, int d, int ch, double u, double v, double ww, double* out) {
    Tex<double, float> t; t.data = data; t.w = w; t.h = h; t.d = d; t.ch = ch;
    double o[4];
    sample3d(t, u, v, ww, true, o);
    for (int c = 0; c < ch && c < 4; ++c) out[c] = o[c];
}
void or_sample2d(const float* data, int w, int h, int ch, double u, double v, int repeat, double* out) {
    Tex<double, float> t; t.data = data; t.w = w; t.h = h; t.ch = ch;
    double o[4];
    sample2d(t, u, v, repeat != 0, o);
    for (int c = 0; c < ch && c < 4; ++c) out[c] = o[c];
}
uint32_t or_hash3(uint32_t x, uint32_t y, uint32_t z, uint32_t seed) { return hash3(x, y, z, seed); }
// The noise primitives at an arbitrary point (for the lattice and
// brute-force checks of the bakes).
double or_perlin_tile(double x, double y, double z, uint32_t freq, uint32_t seed) {
    return perlin_tile<double>(V3<double>(x, y, z), freq, seed);
}
double or_worley_tile(double x, double y, double z, uint32_t freq, uint32_t ox, uint32_t oy, uint32_t oz,
                      uint32_t seed) {
    return worley_tile<double>(V3<double>(x, y, z), freq, ox, oy, oz, seed);
}
double or_perlin2(double x, double y, uint32_t freq, uint32_t seed) { return perlin2<double>(x, y, freq, seed); }

}  // extern "C"
