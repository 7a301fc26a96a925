// Oracle (test infrastructure, not product code): CPU restatement of the
// Hillaire-2020 atmosphere path of PedalSky, templated on the scalar type.
// The double instantiation is the FP64 parity oracle; the float instantiation
// is the CPU "port" timing baseline. Each function cites the WGSL it follows.
#pragma once
#include "oracle_math.hpp"
#include "../include/pedalsky_hip.h"

namespace oracle {

// shaders/common/atmosphere.wgsl:13-14
constexpr double kPI = 3.14159265358979;

template <class R>
struct Atm {
    R planet_r, top, hr, hm;
    V3<R> ray_s, mie_s, mie_a;
    R mie_g;
    V3<R> ozone;
    R oz_c, oz_t;
    V3<R> albedo;
};

template <class R>
Atm<R> atm_from(const ps_atmosphere_params& p) {
    Atm<R> a;
    a.planet_r = p.planet_radius_m; a.top = p.atmosphere_top_m;
    a.hr = p.rayleigh_scale_height_m; a.hm = p.mie_scale_height_m;
    a.ray_s = V3<R>(p.rayleigh_scattering[0], p.rayleigh_scattering[1], p.rayleigh_scattering[2]);
    a.mie_s = V3<R>(p.mie_scattering[0], p.mie_scattering[1], p.mie_scattering[2]);
    a.mie_a = V3<R>(p.mie_absorption[0], p.mie_absorption[1], p.mie_absorption[2]);
    a.mie_g = p.mie_g;
    a.ozone = V3<R>(p.ozone_absorption[0], p.ozone_absorption[1], p.ozone_absorption[2]);
    a.oz_c = p.ozone_center_m; a.oz_t = p.ozone_thickness_m;
    a.albedo = V3<R>(p.ground_albedo[0], p.ground_albedo[1], p.ground_albedo[2]);
    return a;
}

template <class R>
struct Frame {
    R inv_vp[16];
    V3<R> cam;
    V3<R> sun;
    R sun_ang;
    V3<R> E;
    R toa;
    R vw, vh;
    V3<R> light_rgb;
    R light_radius;
    V3<R> light_origin;
    R sim_s;
    uint32_t frame_index;
};

template <class R>
Frame<R> frame_from(const ps_frame_uniforms& f) {
    Frame<R> o;
    for (int i = 0; i < 16; ++i) o.inv_vp[i] = f.inv_view_proj[i];
    o.cam = V3<R>(f.camera_position_world[0], f.camera_position_world[1], f.camera_position_world[2]);
    o.sun = V3<R>(f.sun_direction[0], f.sun_direction[1], f.sun_direction[2]);
    o.sun_ang = f.sun_direction[3];
    o.E = V3<R>(f.sun_illuminance[0], f.sun_illuminance[1], f.sun_illuminance[2]);
    o.toa = f.sun_illuminance[3];
    o.vw = f.viewport_size[0]; o.vh = f.viewport_size[1];
    o.light_rgb = V3<R>(f.lightning_illuminance[0], f.lightning_illuminance[1], f.lightning_illuminance[2]);
    o.light_radius = f.lightning_illuminance[3];
    o.light_origin = V3<R>(f.lightning_origin_world[0], f.lightning_origin_world[1], f.lightning_origin_world[2]);
    o.sim_s = f.simulated_seconds;
    o.frame_index = f.frame_index;
    return o;
}

// Column-major mat4 * vec4 (WGSL `mat4x4<f32> * vec4<f32>`).
template <class R>
inline void mat4_mul(const R m[16], const R v[4], R out[4]) {
    for (int r = 0; r < 4; ++r)
        out[r] = m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1] + m[2 * 4 + r] * v[2] + m[3 * 4 + r] * v[3];
}

// ---- atmosphere.wgsl:23-61 density profiles ------------------------------
template <class R>
inline V3<R> atmosphere_density(const Atm<R>& a, R h) {
    R r = exp(-h / a.hr);
    R m = exp(-h / a.hm);
    R half_thickness = a.oz_t * R(0.5);
    R dz = fabs(h - a.oz_c);
    R o = std::max(R(0), R(1) - dz / std::max(half_thickness, R(1)));
    return V3<R>(r, m, o);
}
template <class R>
inline V3<R> extinction_at(const Atm<R>& a, R h) {
    V3<R> d = atmosphere_density(a, h);
    return a.ray_s * d.x + (a.mie_s + a.mie_a) * d.y + a.ozone * d.z;
}
template <class R>
inline V3<R> scattering_at(const Atm<R>& a, R h) {
    V3<R> d = atmosphere_density(a, h);
    return a.ray_s * d.x + a.mie_s * d.y;
}
template <class R>
inline void scattering_pair(const Atm<R>& a, R h, V3<R>& ray, V3<R>& mie) {
    V3<R> d = atmosphere_density(a, h);
    ray = a.ray_s * d.x;
    mie = a.mie_s * d.y;
}

// ---- atmosphere.wgsl:67-75 phase functions --------------------------------
template <class R>
inline R phase_rayleigh(R c) { return (R(3) / (R(16) * R(kPI))) * (R(1) + c * c); }
template <class R>
inline R phase_mie(R c, R g) {
    R g2 = g * g;
    R denom = R(1) + g2 - R(2) * g * c;
    return (R(1) - g2) / (R(4) * R(kPI) * pow(std::max(denom, R(1e-4)), R(1.5)));
}

// ---- atmosphere.wgsl:95-138 geometry ---------------------------------------
template <class R>
inline bool ray_sphere_intersect_origin(const V3<R>& ro, const V3<R>& rd, R r, R& t0, R& t1) {
    R b = dot(ro, rd);
    R c = dot(ro, ro) - r * r;
    R disc = b * b - c;
    if (disc < R(0)) return false;
    R sq = sqrt(disc);
    t0 = -b - sq; t1 = -b + sq;
    return true;
}
template <class R>
inline R distance_to_top_atmosphere(const Atm<R>& a, const V3<R>& p, const V3<R>& dir) {
    R t0, t1;
    if (!ray_sphere_intersect_origin(p, dir, a.top, t0, t1)) return R(0);
    return std::max(t1, R(0));
}
template <class R>
inline R distance_to_atmosphere_boundary(const Atm<R>& a, const V3<R>& p, const V3<R>& dir) {
    R tt0, tt1, tp0, tp1;
    bool hit_top = ray_sphere_intersect_origin(p, dir, a.top, tt0, tt1);
    bool hit_planet = ray_sphere_intersect_origin(p, dir, a.planet_r, tp0, tp1);
    if (!hit_top) return R(0);
    R t_max = tt1;
    if (hit_planet && tp0 > R(0)) t_max = std::min(t_max, tp0);
    return std::max(t_max, R(0));
}
template <class R>
inline V3<R> world_to_atmosphere_pos(const Atm<R>& a, const V3<R>& pw) {
    return pw + V3<R>(R(0), a.planet_r, R(0));
}

// ---- atmosphere.wgsl:146-179 transmittance LUT parametrisation --------------
template <class R>
inline void transmittance_lut_uv(const Atm<R>& a, const V3<R>& p, const V3<R>& dir, R& u, R& v) {
    R r = length(p);
    R H = sqrt(std::max(a.top * a.top - a.planet_r * a.planet_r, R(0)));
    R rho = sqrt(std::max(r * r - a.planet_r * a.planet_r, R(0)));
    R d = std::max(distance_to_top_atmosphere(a, p, dir), R(0));
    R d_min = a.top - r;
    R d_max = rho + H;
    u = clampr<R>((d - d_min) / std::max(d_max - d_min, R(1)), R(0), R(1));
    v = rho / std::max(H, R(1));
}
template <class R>
inline void transmittance_lut_uv_to_pos_dir(const Atm<R>& a, R u, R v, V3<R>& p, V3<R>& dir) {
    R H = sqrt(std::max(a.top * a.top - a.planet_r * a.planet_r, R(0)));
    R rho = v * H;
    R r = sqrt(rho * rho + a.planet_r * a.planet_r);
    R d_min = a.top - r;
    R d_max = rho + H;
    R d = d_min + u * (d_max - d_min);
    R mu = R(1);
    if (d != R(0)) mu = clampr<R>((H * H - rho * rho - d * d) / std::max(R(2) * r * d, R(1e-6)), R(-1), R(1));
    p = V3<R>(R(0), r, R(0));
    R sin_t = sqrt(std::max(R(1) - mu * mu, R(0)));
    dir = V3<R>(sin_t, mu, R(0));
}

// ---- atmosphere.wgsl:192-214 trapezoid optical depth -----------------------
template <class R>
inline V3<R> integrate_optical_depth(const Atm<R>& a, const V3<R>& p, const V3<R>& dir, int n_steps) {
    R t_max = distance_to_atmosphere_boundary(a, p, dir);
    if (t_max <= R(0)) return V3<R>(R(0));
    R dt = t_max / R(n_steps);
    R r0 = length(p);
    R cos_view = dot(p / std::max(r0, R(1)), dir);
    R h0 = r0 - a.planet_r;
    V3<R> od(R(0));
    V3<R> prev = extinction_at(a, h0);
    for (int i = 1; i <= n_steps; ++i) {
        R t = R(i) * dt;
        V3<R> pi = p + dir * t;
        R num = R(2) * t * r0 * cos_view + t * t;
        R r_delta = num / (std::max(r0, R(1)) + length(pi));
        R h = h0 + r_delta;
        V3<R> cur = extinction_at(a, h);
        od = od + (prev + cur) * (R(0.5) * dt);
        prev = cur;
    }
    return od;
}

// ---- atmosphere_lut_sampling.wgsl:6-22 ---------------------------------------
template <class R, class S>
inline V3<R> sample_transmittance_lut(const Atm<R>& a, const Tex<R, S>& T, const V3<R>& p, const V3<R>& dir) {
    R t0, t1;
    if (ray_sphere_intersect_origin(p, dir, a.planet_r, t0, t1) && t0 > R(0)) return V3<R>(R(0));
    R u, v;
    transmittance_lut_uv(a, p, dir, u, v);
    R s[4];
    sample2d(T, u, v, false, s);
    return V3<R>(s[0], s[1], s[2]);
}

// ---- transmittance.comp.wgsl:18-31 -------------------------------------------
template <class R>
void bake_transmittance(const Atm<R>& a, int W, int H, R* out) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            R u = (R(x) + R(0.5)) / R(W), v = (R(y) + R(0.5)) / R(H);
            V3<R> p, dir;
            transmittance_lut_uv_to_pos_dir(a, u, v, p, dir);
            V3<R> od = integrate_optical_depth(a, p, dir, 40);
            V3<R> t = vexp(-od);
            R* o = out + (size_t(y) * W + x) * 4;
            o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = R(1);
        }
}

// ---- multiscatter.comp.wgsl:58-65 Fibonacci sphere sample ---------------------
template <class R>
inline V3<R> sphere_sample(int i, int n_dirs) {
    R golden = R(0.6180339887498949);
    R z = R(1) - R(2) * (R(i) + R(0.5)) / R(n_dirs);
    R r = sqrt(std::max(R(1) - z * z, R(0)));
    R phi = R(2) * R(kPI) * fractr<R>(R(i) * golden);
    return V3<R>(r * cos(phi), r * sin(phi), z);
}

// ---- multiscatter.comp.wgsl:67-156 -------------------------------------------
template <class R, class S>
void bake_multiscatter_texel(const Atm<R>& a, const Tex<R, S>& T, R u, R v, R out[4]) {
    const int N_DIRS = 64, N_STEPS = 20;
    R sun_cos = clampr<R>(u * R(2) - R(1), R(-1), R(1));
    R sun_sin = sqrt(std::max(R(1) - sun_cos * sun_cos, R(0)));
    R h_norm = v;
    R r = a.planet_r + h_norm * (a.top - a.planet_r);
    V3<R> p0(R(0), r, R(0));
    V3<R> sun_dir(sun_sin, sun_cos, R(0));
    V3<R> l2(R(0)), fms(R(0));
    R h0 = r - a.planet_r;
    for (int d = 0; d < N_DIRS; ++d) {
        V3<R> view_dir = sphere_sample<R>(d, N_DIRS);
        R t_max = distance_to_atmosphere_boundary(a, p0, view_dir);
        if (t_max <= R(0)) continue;
        R dt = t_max / R(N_STEPS);
        R cos_view = dot(p0 / std::max(r, R(1)), view_dir);
        V3<R> trans(R(1)), l_dir(R(0)), f_dir(R(0));
        for (int s = 0; s < N_STEPS; ++s) {
            R t = (R(s) + R(0.5)) * dt;
            V3<R> p = p0 + view_dir * t;
            R num = R(2) * t * r * cos_view + t * t;
            R r_delta = num / (std::max(r, R(1)) + length(p));
            R h = h0 + r_delta;
            V3<R> sigma_t = extinction_at(a, h);
            V3<R> sigma_s = scattering_at(a, h);
            V3<R> sample_transmit = vexp(-sigma_t * dt);
            R r_p = r + r_delta;
            R cos_sun = dot(p, sun_dir) / std::max(r_p, R(1));
            V3<R> sun_vis(R(0));
            R ratio = a.planet_r / r_p;
            if (cos_sun > -sqrt(std::max(R(1) - ratio * ratio, R(0))))
                sun_vis = sample_transmittance_lut(a, T, p, sun_dir);
            V3<R> in_scatter = sigma_s * sun_vis;
            V3<R> safe = vmax(sigma_t, V3<R>(R(1e-7)));
            V3<R> s_int = (in_scatter - in_scatter * sample_transmit) / safe;
            l_dir = l_dir + trans * s_int;
            V3<R> s_int_iso = (sigma_s - sigma_s * sample_transmit) / safe;
            f_dir = f_dir + trans * s_int_iso;
            trans = trans * sample_transmit;
        }
        l2 = l2 + l_dir;
        fms = fms + f_dir;
    }
    R norm = (R(4) * R(kPI)) / R(N_DIRS);
    V3<R> L2 = l2 * norm * (R(1) / (R(4) * R(kPI)));
    V3<R> F = fms * norm * (R(1) / (R(4) * R(kPI)));
    V3<R> fc(clampr<R>(F.x, R(0), R(0.95)), clampr<R>(F.y, R(0), R(0.95)), clampr<R>(F.z, R(0), R(0.95)));
    V3<R> full = ORACLE_SLIP == 4 ? L2 : L2 / (V3<R>(R(1)) - fc);
    out[0] = full.x; out[1] = full.y; out[2] = full.z; out[3] = R(1);
}

template <class R, class S>
void bake_multiscatter(const Atm<R>& a, const Tex<R, S>& T, int W, int H, R* out) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            bake_multiscatter_texel(a, T, (R(x) + R(0.5)) / R(W), (R(y) + R(0.5)) / R(H),
                                    out + (size_t(y) * W + x) * 4);
}

// ---- skyview.comp.wgsl:46-174 ----------------------------------------------------
template <class R, class S>
void bake_skyview_texel(const Frame<R>& f, const Atm<R>& a, const Tex<R, S>& T, const Tex<R, S>& MS,
                        R u, R v, R out[4]) {
    const int N_STEPS = 32;
    V3<R> p = world_to_atmosphere_pos(a, f.cam);
    R r = length(p);
    if (r < a.planet_r + R(1)) {
        p = normalize(p) * (a.planet_r + R(1));
        r = a.planet_r + R(1);
    }
    R sin_horizon = clampr<R>(a.planet_r / r, R(0), R(1));
    R zha = R(kPI) - asin(sin_horizon);
    R vza;
    if (v < R(0.5)) {
        R coord = v * R(2);
        vza = coord * coord * zha;
    } else {
        R coord = (v - R(0.5)) * R(2);
        R below = R(kPI) - zha;
        vza = zha + coord * coord * below;
    }
    V3<R> sun_dir = f.sun;
    R sun_az = atan2(sun_dir.x, sun_dir.z);
    R azimuth = sun_az + u * R(2) * R(kPI);
    R cos_v = cos(vza), sin_v = sin(vza);
    V3<R> view_dir(sin_v * sin(azimuth), cos_v, sin_v * cos(azimuth));
    R cos_theta = dot(view_dir, sun_dir);
    R phase_r = phase_rayleigh(cos_theta);
    R phase_m = phase_mie(cos_theta, a.mie_g);
    R t_max = distance_to_atmosphere_boundary(a, p, view_dir);
    if (t_max <= R(0)) { out[0] = out[1] = out[2] = R(0); out[3] = R(1); return; }
    R dt = t_max / R(N_STEPS);
    R cos_view = dot(p / std::max(r, R(1)), view_dir);
    R h0 = r - a.planet_r;
    V3<R> lum(R(0)), trans(R(1));
    for (int s = 0; s < N_STEPS; ++s) {
        R t = (R(s) + R(0.5)) * dt;
        V3<R> pi = p + view_dir * t;
        R num = R(2) * t * r * cos_view + t * t;
        R r_delta = num / (std::max(r, R(1)) + length(pi));
        R h = h0 + r_delta;
        V3<R> sigma_t = extinction_at(a, h);
        V3<R> sample_transmit = vexp(-sigma_t * dt);
        V3<R> sun_vis = sample_transmittance_lut(a, T, pi, sun_dir);
        V3<R> sr, sm;
        scattering_pair(a, h, sr, sm);
        V3<R> sun_in = (sr * phase_r + sm * phase_m) * sun_vis;
        R r_p = r + r_delta;
        R h_norm = clampr<R>(h / std::max(a.top - a.planet_r, R(1)), R(0), R(1));
        R sun_cos = clampr<R>(dot(pi / std::max(r_p, R(1)), sun_dir), R(-1), R(1));
        R ms[4];
        sample2d(MS, sun_cos * R(0.5) + R(0.5), h_norm, false, ms);
        V3<R> multi = (ORACLE_SLIP == 11 ? sigma_t : sr + sm) * V3<R>(ms[0], ms[1], ms[2]);
        V3<R> in_scatter = sun_in + multi;
        V3<R> safe = vmax(sigma_t, V3<R>(R(1e-7)));
        V3<R> s_int = (in_scatter - in_scatter * sample_transmit) / safe;
        lum = lum + trans * s_int;
        trans = trans * sample_transmit;
    }
    R tp0, tp1;
    bool hit = ray_sphere_intersect_origin(p, view_dir, a.planet_r, tp0, tp1);
    if (hit && tp0 > R(0)) {
        V3<R> pg = p + view_dir * tp0;
        V3<R> ng = normalize(pg);
        V3<R> sv = sample_transmittance_lut(a, T, pg, sun_dir);
        R ndl = std::max(dot(ng, sun_dir), R(0));
        V3<R> bounce = a.albedo * (R(1) / R(3.14159265)) * ndl * sv;
        lum = lum + trans * bounce;
    }
    out[0] = lum.x; out[1] = lum.y; out[2] = lum.z; out[3] = R(1);
}

template <class R, class S>
void bake_skyview(const Frame<R>& f, const Atm<R>& a, const Tex<R, S>& T, const Tex<R, S>& MS, int W, int H, R* out) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            bake_skyview_texel(f, a, T, MS, (R(x) + R(0.5)) / R(W), (R(y) + R(0.5)) / R(H),
                               out + (size_t(y) * W + x) * 4);
}

// ---- aerialperspective.comp.wgsl:41-127 ------------------------------------------
constexpr double kApNear = 50.0, kApFar = 100000.0;

template <class R, class S>
void bake_ap_froxel(const Frame<R>& f, const Atm<R>& a, const Tex<R, S>& T, const Tex<R, S>& MS,
                    int gx, int gy, int gz, int W, int H, int D, R out[4]) {
    const int AP_STEPS = 16;
    R u = (R(gx) + R(0.5)) / R(W), v = (R(gy) + R(0.5)) / R(H);
    R ndc_x = u * R(2) - R(1), ndc_y = v * R(2) - R(1);
    R hv[4] = {ndc_x, ORACLE_SLIP == 12 ? ndc_y : -ndc_y, R(1), R(1)}, nh[4];
    mat4_mul(f.inv_vp, hv, nh);
    V3<R> near_p(nh[0] / nh[3], nh[1] / nh[3], nh[2] / nh[3]);
    V3<R> view_dir = normalize(near_p - f.cam);
    R z_norm = (R(gz) + R(0.5)) / R(D);
    R t_target = R(kApNear) * pow(R(kApFar) / R(kApNear), z_norm);
    V3<R> p0 = world_to_atmosphere_pos(a, f.cam);
    R r0 = length(p0);
    R cos_view = dot(p0 / std::max(r0, R(1)), view_dir);
    V3<R> sun_dir = f.sun;
    R cos_theta = dot(view_dir, sun_dir);
    R phase_r = phase_rayleigh(cos_theta);
    R phase_m = phase_mie(cos_theta, a.mie_g);
    R dt = t_target / R(AP_STEPS);
    V3<R> lum(R(0)), trans(R(1));
    for (int s = 0; s < AP_STEPS; ++s) {
        R t = (R(s) + R(0.5)) * dt;
        V3<R> pi = p0 + view_dir * t;
        R r0_safe = std::max(r0, R(1));
        R num = R(2) * t * r0 * cos_view + t * t;
        R r_delta = num / (r0_safe + length(pi));
        R h = (r0 - a.planet_r) + r_delta;
        if (h < R(0)) break;
        V3<R> sigma_t = extinction_at(a, h);
        V3<R> sample_transmit = vexp(-sigma_t * dt);
        V3<R> sun_vis = sample_transmittance_lut(a, T, pi, sun_dir);
        V3<R> sr, sm;
        scattering_pair(a, h, sr, sm);
        V3<R> sun_in = (sr * phase_r + sm * phase_m) * sun_vis;
        R r_p = r0 + r_delta;
        R h_norm = clampr<R>(h / std::max(a.top - a.planet_r, R(1)), R(0), R(1));
        R sun_cos = clampr<R>(dot(pi / std::max(r_p, R(1)), sun_dir), R(-1), R(1));
        R ms[4];
        sample2d(MS, sun_cos * R(0.5) + R(0.5), h_norm, false, ms);
        V3<R> multi = (sr + sm) * V3<R>(ms[0], ms[1], ms[2]);
        V3<R> in_scatter = ORACLE_SLIP == 5 ? sun_in + multi : (sun_in + multi) * f.E;
        V3<R> safe = vmax(sigma_t, V3<R>(R(1e-7)));
        V3<R> s_int = (in_scatter - in_scatter * sample_transmit) / safe;
        lum = lum + trans * s_int;
        trans = trans * sample_transmit;
    }
    R t_lum = dot(trans, V3<R>(R(0.2126), R(0.7152), R(0.0722)));
    out[0] = lum.x; out[1] = lum.y; out[2] = lum.z; out[3] = t_lum;
}

template <class R, class S>
void bake_ap(const Frame<R>& f, const Atm<R>& a, const Tex<R, S>& T, const Tex<R, S>& MS,
             int W, int H, int D, R* out) {
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
                bake_ap_froxel(f, a, T, MS, x, y, z, W, H, D, out + ((size_t(z) * H + y) * W + x) * 4);
}

}  // namespace oracle
