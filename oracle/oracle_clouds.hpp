// Oracle (test infrastructure, not product code): CPU restatement of the sky
// pass (shaders/atmosphere/sky.wgsl) and the volumetric cloud march +
// composite (shaders/clouds/cloud_march.wgsl, cloud_density_common.wgsl,
// cloud_composite.wgsl), templated on the scalar type.
#pragma once
#include "oracle_atmosphere.hpp"
#include "oracle_dual.hpp"

namespace oracle {

template <class R>
struct Clouds {
    V3<R> sigma_s, sigma_a;
    R droplet_bias;
    uint32_t cone;
    R forward_bias, detail_strength, curl_strength, powder;
    R ms_a, ms_b, ms_c, ambient;
    R base_scale, detail_scale, weather_scale;
    uint32_t light_steps, cloud_steps, octaves, layer_count, temporal_jitter;
    R drift, skew, diurnal;
};

template <class R>
Clouds<R> clouds_from(const ps_cloud_params& p) {
    Clouds<R> c;
    c.sigma_s = V3<R>(p.sigma_s[0], p.sigma_s[1], p.sigma_s[2]);
    c.sigma_a = V3<R>(p.sigma_a[0], p.sigma_a[1], p.sigma_a[2]);
    c.droplet_bias = p.droplet_diameter_bias;
    c.cone = p.cone_light_sampling;
    c.forward_bias = p.forward_bias; c.detail_strength = p.detail_strength; c.curl_strength = p.curl_strength;
    c.powder = p.powder_strength;
    c.ms_a = p.multi_scatter_a; c.ms_b = p.multi_scatter_b; c.ms_c = p.multi_scatter_c;
    c.ambient = p.ambient_strength;
    c.base_scale = p.base_scale_m; c.detail_scale = p.detail_scale_m; c.weather_scale = p.weather_scale_m;
    c.light_steps = p.light_steps; c.cloud_steps = p.cloud_steps; c.octaves = p.multi_scatter_octaves;
    c.layer_count = p.cloud_layer_count; c.temporal_jitter = p.temporal_jitter;
    c.drift = p.wind_drift_strength; c.skew = p.wind_skew_strength; c.diurnal = p.diurnal_strength;
    return c;
}

template <class R>
struct Layer {
    R base, top, coverage, density_scale;
    uint32_t type;
    R shape_bias, detail_bias, anvil_bias, droplet;
};

template <class R>
Layer<R> layer_from(const ps_cloud_layer& l) {
    Layer<R> o;
    o.base = l.base_m; o.top = l.top_m; o.coverage = l.coverage; o.density_scale = l.density_scale;
    o.type = l.cloud_type; o.shape_bias = l.shape_bias; o.detail_bias = l.detail_bias;
    o.anvil_bias = l.anvil_bias; o.droplet = l.droplet_diameter_um;
    return o;
}

// Everything one frame of the sky + cloud path reads. Noise and weather
// values are the decoded texel values (UNORM8 /255, or binary16 as float).
template <class R>
struct Scene {
    Frame<R> f;
    Atm<R> a;
    Clouds<R> c;
    Layer<R> layers[PS_MAX_CLOUD_LAYERS];
    V3<R> layer_wind[PS_MAX_CLOUD_LAYERS];  // (u, 0, v) wind m/s per layer (x, z used)
    int n_layers;
    // Half-resolution march (crates/ps-clouds/src/pipeline.rs:277-319): the
    // cloud RT size, 0 for the full-resolution march. An RT pixel's ray goes
    // through the viewport coordinates (x + 0.5) * vw / rt_w, its depth is
    // read at (int)(x * vw / rt_w), and its jitter at (x, y).
    int rt_w, rt_h;
    // Forward-identity extensions for differentiation (absent from the
    // reference; the neutral values reproduce it exactly): a per-layer
    // extinction multiplier on sigma_t, and an additive 128^3 density
    // perturbation (see delta_sample).
    R kappa[PS_MAX_CLOUD_LAYERS];
    const double* delta;                   // 128^3 or null (zero)
    // dual instantiation: delta voxels whose value seeds tangent slot k
    const int* delta_track;
    const int* delta_track_slot;
    int n_delta_track;
    Tex<R, typename LutStore<R>::type> T, SV, AP;  // LUTs (RGBA)
    Tex<R, float> base, detail, curl, weather;
    const uint8_t* type_grid;              // 128x128 or null (all sentinel)
    const uint8_t* overcast;               // 128x128 R8Unorm or null (zero)
    const uint8_t* blue;                   // 64x64 R8Unorm
    const float* depth;                    // W*H reverse-Z NDC or null
};

struct Counters {
    uint64_t pixels = 0, layer_marches = 0, primary_steps = 0, covered_steps = 0, eroded_steps = 0;
    uint64_t dense_samples = 0, light_steps = 0, light_covered = 0, octaves = 0, ap_taps = 0;
    // dual instantiation only: tangent slots that a quantity within near_tol
    // of one of the march's decision thresholds depends on (an FP32
    // evaluation may take the other side of that decision): the dense
    // threshold, the erosion branch, and the final clamp's kinks in the
    // primary and light-march densities (see note_near)
    double near_tol = 0.0;
    uint32_t near_mask = 0u;
};

// Dual instantiation: when x lies within cnt->near_tol of the decision
// threshold t, marks the tangent slots x depends on.
template <class R>
inline void note_near(Counters* cnt, const R& x, double t) {
    if constexpr (is_dual<R>::value) {
        if (cnt && cnt->near_tol > 0.0 && std::fabs(x.v - t) < cnt->near_tol)
            for (int k = 0; k < R::n; ++k)
                if (x.d[k] != 0.0) cnt->near_mask |= 1u << k;
    }
}

// ---- math.wgsl:7-9 -----------------------------------------------------------------
template <class R>
inline R remap(R v, R omin, R omax, R nmin, R nmax) {
    return nmin + (v - omin) * (nmax - nmin) / std::max(omax - omin, R(1e-5));
}

// ---- cloud_density_common.wgsl:159-201 -------------------------------------------
template <class R>
inline R ndf(R h, uint32_t t, R anvil_bias) {
    switch (t) {
        case 0: return smoothstep<R>(0, R(0.07), h) * smoothstep<R>(1, R(0.2), h);
        case 1: return smoothstep<R>(0, R(0.10), h) * (R(1) - smoothstep<R>(R(0.6), 1, h));
        case 2: return smoothstep<R>(0, R(0.15), h) * smoothstep<R>(1, R(0.4), h);
        case 3: return smoothstep<R>(0, R(0.20), h) * smoothstep<R>(1, R(0.3), h);
        case 4: return smoothstep<R>(0, R(0.30), h) * (R(1) - smoothstep<R>(R(0.7), 1, h));
        case 5: return smoothstep<R>(0, R(0.40), h) * (R(1) - smoothstep<R>(R(0.6), 1, h));
        case 6: return smoothstep<R>(0, R(0.50), h) * (R(1) - smoothstep<R>(R(0.8), 1, h));
        case 7: {
            R base = smoothstep<R>(0, R(0.05), h);
            R mid = smoothstep<R>(R(0.95), R(0.5), h);
            R anvil = smoothstep<R>(R(0.7), R(0.9), h) * R(1.5) * clampr<R>(anvil_bias, 0, 2);
            R mix_t = smoothstep<R>(R(0.65), R(0.8), h);
            return base * mixr(mid, anvil, mix_t);
        }
        default: return R(0);
    }
}

// ---- cloud_density_common.wgsl:247-293 (LayerWindCache) ---------------------------
template <class R>
struct Wind {
    R off_x, off_z, skew_x, skew_z, skew_thickness;
};
template <class R>
inline Wind<R> layer_wind_cache(const Scene<R>& s, int li) {
    const Layer<R>& L = s.layers[li];
    R wx = s.layer_wind[li].x, wz = s.layer_wind[li].z;
    R drift = s.c.drift * s.f.sim_s;
    Wind<R> w;
    w.off_x = wx * drift; w.off_z = wz * drift;
    w.skew_x = 0; w.skew_z = 0; w.skew_thickness = 0;
    if (s.c.skew != R(0)) {
        R speed = sqrt(wx * wx + wz * wz);
        if (speed >= R(0.1)) {
            w.skew_x = wx / speed; w.skew_z = wz / speed;
            w.skew_thickness = std::max(L.top - L.base, R(1)) * s.c.skew;
        }
    }
    return w;
}

// ---- cloud_density_common.wgsl:314-329 ----------------------------------------------
template <class R>
inline R diurnal_modulation(const Scene<R>& s, uint32_t t) {
    R strength = s.c.diurnal;
    if (strength == R(0)) return R(0);
    bool conv = t == 0 || t == 2 || t == 3 || t == 7;
    if (!conv) return R(0);
    return smoothstep<R>(R(-0.1), R(0.4), s.f.sun.y) * strength;
}

// ---- cloud_density_common.wgsl:343-361 ------------------------------------------------
template <class R>
inline uint32_t effective_cloud_type(const Scene<R>& s, R px, R pz, const Layer<R>& L) {
    if (!s.type_grid) return L.type;
    R half = R(16000);
    R u = (px + half) / R(32000), v = (pz + half) / R(32000);
    int tx_raw = int(floor(u * R(128))), ty_raw = int(floor(v * R(128)));
    int tx = ((tx_raw % 128) + 128) % 128, ty = ((ty_raw % 128) + 128) % 128;
    uint32_t t = s.type_grid[ty * 128 + tx];
    return t == 255u ? L.type : t;
}

// ---- density perturbation (extension; no reference counterpart) ------------------------
// delta is a 128^3 volume over the weather tile horizontally and the layer's
// normalised height vertically: texel (i, j, k) covers x = i, h = j, z = k
// (x fastest), sampled trilinearly at u = wx / weather_scale + 0.5 (Repeat),
// h (ClampToEdge), v = wz / weather_scale + 0.5 (Repeat), where (wx, wz) is
// the drift-offset position the weather map is read at. It is added to the
// shaped cloud value just before the final clamp, in both the primary and
// the light-march density, past the altitude and coverage gates.
constexpr int kDeltaN = 128;

template <class R>
inline R delta_texel(const Scene<R>& s, int idx) {
    R v = R(s.delta[idx]);
    if constexpr (is_dual<R>::value) {
        for (int k = 0; k < s.n_delta_track; ++k)
            if (s.delta_track[k] == idx) v.d[s.delta_track_slot[k]] = 1.0;
    }
    return v;
}

template <class R>
inline R delta_sample(const Scene<R>& s, R wx, R wz, R h) {
    if (!s.delta) return R(0);
    R wsc = std::max(s.c.weather_scale, R(1));
    R cx = (wx / wsc + R(0.5)) * R(kDeltaN) - R(0.5);
    R cy = h * R(kDeltaN) - R(0.5);
    R cz = (wz / wsc + R(0.5)) * R(kDeltaN) - R(0.5);
    R fx0 = floor(cx), fy0 = floor(cy), fz0 = floor(cz);
    R fx = cx - fx0, fy = cy - fy0, fz = cz - fz0;
    int ix = int(fx0), iy = int(fy0), iz = int(fz0);
    int x0 = wrap_index(ix, kDeltaN), x1 = wrap_index(ix + 1, kDeltaN);
    int y0 = clamp_index(iy, kDeltaN), y1 = clamp_index(iy + 1, kDeltaN);
    int z0 = wrap_index(iz, kDeltaN), z1 = wrap_index(iz + 1, kDeltaN);
    auto at = [&](int x, int y, int z) { return delta_texel(s, (z * kDeltaN + y) * kDeltaN + x); };
    R a0 = at(x0, y0, z0) + (at(x1, y0, z0) - at(x0, y0, z0)) * fx;
    R a1 = at(x0, y1, z0) + (at(x1, y1, z0) - at(x0, y1, z0)) * fx;
    R b0 = at(x0, y0, z1) + (at(x1, y0, z1) - at(x0, y0, z1)) * fx;
    R b1 = at(x0, y1, z1) + (at(x1, y1, z1) - at(x0, y1, z1)) * fx;
    R p0 = a0 + (a1 - a0) * fy, p1 = b0 + (b1 - b0) * fy;
    return p0 + (p1 - p0) * fz;
}

// ---- cloud_density_common.wgsl:376-415 ------------------------------------------------
template <class R>
inline R sample_density_light(const Scene<R>& s, const V3<R>& p, R p_alt, const Layer<R>& L, R weather_r,
                              const Wind<R>& w, uint32_t ctype, Counters* cnt) {
    R h = (p_alt - L.base) / std::max(L.top - L.base, R(1));
    if (h < R(0) || h > R(1)) return R(0);
    R eff = clampr<R>(weather_r * L.coverage, 0, 1);
    if (eff < R(0.01)) return R(0);
    if (cnt) cnt->light_covered++;
    R sx = w.skew_x * (h * w.skew_thickness), sz = w.skew_z * (h * w.skew_thickness);
    V3<R> pa = p - V3<R>(w.off_x, 0, w.off_z) - V3<R>(sx, 0, sz);
    R bs = std::max(s.c.base_scale, R(1));
    R base[4];
    sample3d(s.base, pa.x / bs, pa.y / bs, pa.z / bs, true, base);
    R diurnal = diurnal_modulation(s, L.type);
    R esb = L.shape_bias * diurnal;
    R lf = clampr<R>(base[1] * R(0.625) + base[2] * R(0.25) + base[3] * R(0.125) + esb * R(0.5), 0, 1);
    R base_cloud = remap<R>(base[0], -(R(1) - lf), 1, 0, 1);
    R profile = ndf(h, ctype, L.anvil_bias);
    R cloud = base_cloud * profile;
    R cov = eff;
    cloud = remap<R>(cloud, R(1) - cov, 1, 0, 1);
    cloud = cloud * cov;
    cloud = cloud + delta_sample(s, p.x - w.off_x, p.z - w.off_z, h);
    note_near(cnt, cloud, 0.0);
    note_near(cnt, cloud, 1.0);
    return clampr<R>(cloud, 0, 1) * L.density_scale;
}

// Upper tolerance of the primary march's altitude gate. With blue-noise
// jitter 255/255 the last sample of a layer sits exactly at the interval end
// t1 = t0 + steps * step, i.e. exactly on the top shell (h = 1) for rays
// leaving through the top; the literal `h > 1` test then decides it by
// rounding noise in any precision (the WGSL's FP32 included). h = 1 is
// inside in exact arithmetic, and only the cumulonimbus profile is non-zero
// there (anvil), so the gate accepts h <= 1 + 1e-5 (3 cm of a 3 km layer);
// the GPU kernel uses the same bound.
constexpr double kPrimaryGateTop = 1.0 + 1e-5;

// ---- cloud_march.wgsl:432-530 --------------------------------------------------------------
template <class R>
inline R sample_density(const Scene<R>& s, const V3<R>& p, R p_alt, const Layer<R>& L, R weather_r,
                        const Wind<R>& w, uint32_t ctype, Counters* cnt) {
    R h = (p_alt - L.base) / std::max(L.top - L.base, R(1));
    if (h < R(0) || h > R(kPrimaryGateTop)) return R(0);
    R eff = clampr<R>(weather_r * L.coverage, 0, 1);
    if (eff < R(0.01)) return R(0);
    if (cnt) cnt->covered_steps++;
    R sx = w.skew_x * (h * w.skew_thickness), sz = w.skew_z * (h * w.skew_thickness);
    V3<R> pa = p - V3<R>(w.off_x, 0, w.off_z) - V3<R>(sx, 0, sz);
    R bs = std::max(s.c.base_scale, R(1));
    R base[4];
    sample3d(s.base, pa.x / bs, pa.y / bs, pa.z / bs, true, base);
    R diurnal = diurnal_modulation(s, L.type);
    R esb = L.shape_bias * diurnal;
    R lf = clampr<R>(base[1] * R(0.625) + base[2] * R(0.25) + base[3] * R(0.125) + esb * R(0.5), 0, 1);
    R base_cloud = remap<R>(base[0], -(R(1) - lf), 1, 0, 1);
    R profile = ndf(h, ctype, L.anvil_bias);
    R cloud = base_cloud * profile;
    R cov = clampr<R>(weather_r * L.coverage, 0, 1);
    cloud = remap<R>(cloud, R(1) - cov, 1, 0, 1);
    cloud = cloud * cov;
    note_near(cnt, cloud, 0.0);
    if (cloud > R(0)) {
        if (cnt) cnt->eroded_steps++;
        R ds = std::max(s.c.detail_scale, R(1));
        R cu[4];
        sample2d(s.curl, pa.x / ds, pa.z / ds, true, cu);
        R cx = cu[0] * R(2) - R(1), cz = cu[1] * R(2) - R(1);
        V3<R> dp = pa + V3<R>(cx, 0, cz) * (s.c.curl_strength * s.c.detail_scale);
        R det[4];
        sample3d(s.detail, dp.x / ds, dp.y / ds, dp.z / ds, true, det);
        R hf = det[0] * R(0.625) + det[1] * R(0.25) + det[2] * R(0.125);
        R dmod = mixr(hf, R(1) - hf, clampr<R>(h * R(10), 0, 1));
        R edb = L.detail_bias * diurnal;
        R ld = clampr<R>(s.c.detail_strength + edb, 0, 1);
        cloud = remap<R>(cloud, dmod * ld, 1, 0, 1);
    }
    cloud = cloud + delta_sample(s, p.x - w.off_x, p.z - w.off_z, h);
    note_near(cnt, cloud, 1.0);
    return clampr<R>(cloud, 0, 1) * L.density_scale;
}

// ---- cloud_march.wgsl:360-365 ---------------------------------------------------------------
template <class R>
inline R altitude_from_entry(const Scene<R>& s, const V3<R>& pw, R r0, R cos_view, R h0, R t) {
    V3<R> pi = pw - V3<R>(0, -s.a.planet_r, 0);
    R num = R(2) * t * r0 * cos_view + t * t;
    R r_delta = num / (std::max(r0, R(1)) + length(pi));
    return h0 + r_delta;
}

// ---- cloud_march.wgsl:599-608 ----------------------------------------------------------------
template <class R>
inline V3<R> cone_kernel_offset(int i) {
    switch (i) {
        case 0: return V3<R>(R(0.38051305), R(0.92453449), R(-0.02111345));
        case 1: return V3<R>(R(-0.50625799), R(-0.03590792), R(-0.86163418));
        case 2: return V3<R>(R(-0.32509218), R(-0.94557439), R(0.01428793));
        case 3: return V3<R>(R(0.09026238), R(-0.27376545), R(0.95755165));
        case 4: return V3<R>(R(0.28128598), R(0.42443639), R(-0.86065785));
        default: return V3<R>(R(-0.16852403), R(0.14748697), R(0.97460106));
    }
}

// ---- cloud_march.wgsl:543-578 and 614-666 ----------------------------------------------------
template <class R>
inline R march_to_light(const Scene<R>& s, const V3<R>& p, R p_alt, const V3<R>& sun, const Layer<R>& L,
                        R weather_r, const Wind<R>& w, uint32_t ctype, Counters* cnt) {
    R cos_sun = std::max(sun.y, R(0.05));
    R dist_to_top = std::max((L.top - p_alt) / cos_sun, R(1));
    V3<R> centre(0, -s.a.planet_r, 0);
    R r0 = s.a.planet_r + p_alt;
    R cos_view = dot((p - centre) / std::max(r0, R(1)), sun);
    R h0 = p_alt;
    R od = 0;
    if (s.c.cone != 0u) {
        const int nf = 5;
        R step = dist_to_top / R(nf);
        for (int i = 0; i < nf; ++i) {
            R t = (R(i) + R(0.5)) * step;
            V3<R> off = cone_kernel_offset<R>(i) * (step * (R(i) + R(1)) / R(nf));
            V3<R> pi = p + sun * t + off;
            R alt = altitude_from_entry(s, pi, r0, cos_view, h0, t);
            if (cnt) cnt->light_steps++;
            od = od + sample_density_light(s, pi, alt, L, weather_r, w, ctype, cnt) * step;
        }
        R t_far = dist_to_top * R(3);
        V3<R> off_far = cone_kernel_offset<R>(5) * (step * R(3));
        V3<R> pf = p + sun * t_far + off_far;
        R alt_far = altitude_from_entry(s, pf, r0, cos_view, h0, t_far);
        if (cnt) cnt->light_steps++;
        od = od + sample_density_light(s, pf, alt_far, L, weather_r, w, ctype, cnt) * step * R(3);
        return od;
    }
    R step = dist_to_top / R(s.c.light_steps);
    for (uint32_t i = 0; i < s.c.light_steps; ++i) {
        R t = (R(i) + R(ORACLE_SLIP == 3 ? 1.0 : 0.5)) * step;
        V3<R> pi = p + sun * t;
        R alt = altitude_from_entry(s, pi, r0, cos_view, h0, t);
        if (cnt) cnt->light_steps++;
        od = od + sample_density_light(s, pi, alt, L, weather_r, w, ctype, cnt) * step;
    }
    return od;
}

// ---- cloud_march.wgsl:122-145 ----------------------------------------------------------------
template <class R>
inline V3<R> chromatic_mie_modulation(R d_um) {
    R t = clampr<R>(R(1) - d_um / R(20), 0, 1);
    if (t <= R(0)) return V3<R>(R(1));
    R strength = R(0.25) * t * t;
    return V3<R>(R(1) - R(0.5) * strength, R(1), R(1) + strength);
}

// ---- cloud_march.wgsl:151-168 ----------------------------------------------------------------
template <class R>
inline R henyey_greenstein(R c, R g) {
    R g2 = g * g;
    R denom = R(1) + g2 - R(2) * g * c;
    return (R(1) - g2) / (R(4) * R(kPI) * pow(std::max(denom, R(1e-4)), R(1.5)));
}
template <class R>
inline R draine_phase(R c, R g, R alpha) {
    R g2 = g * g;
    R denom = R(1) + g2 - R(2) * g * c;
    R num = (R(1) - g2) * (R(1) + alpha * c * c);
    R den = R(4) * (R(1) + alpha * (R(1) + R(2) * g2) / R(3)) * R(kPI) * pow(std::max(denom, R(1e-4)), R(1.5));
    return num / std::max(den, R(1e-6));
}

// ---- cloud_march.wgsl:229-296 ----------------------------------------------------------------
template <class R>
struct PhaseConsts { R g_hg, g_d, alpha, w_d, ice; };

template <class R>
inline R ice_halo_lobe(R c, R peak_cos, R width_rad) {
    R sigma_cos = width_rad * sqrt(std::max(R(1) - peak_cos * peak_cos, R(1e-4)));
    R dx = (c - peak_cos) / std::max(sigma_cos, R(1e-4));
    return exp(R(-0.5) * dx * dx);
}
template <class R>
inline PhaseConsts<R> cloud_phase_constants(const Scene<R>& s, const Layer<R>& L, R h_norm) {
    R d = L.droplet;
    if (L.type == 7u) {
        R mt = smoothstep<R>(R(0.60), R(0.85), clampr<R>(h_norm, 0, 1));
        d = mixr(L.droplet, R(50), mt);
    }
    d = clampr<R>(d * s.c.droplet_bias, R(5), R(50));
    PhaseConsts<R> c;
    c.g_hg = exp(R(-0.0990567) / (d - R(1.67154)));
    c.g_d = exp(R(-2.20679) / (d + R(3.91029))) - R(0.428934);
    c.alpha = exp(R(3.62489) - R(8.29288) / (d + R(5.52825)));
    c.w_d = exp(R(-0.599085) / (d - R(0.641583)) - R(0.665888));
    c.ice = smoothstep<R>(R(35), R(50), d);
    return c;
}
template <class R>
inline R cloud_phase_with(R c, R g_scale, const PhaseConsts<R>& k) {
    R p_hg = henyey_greenstein(c, k.g_hg * g_scale);
    R p_d = draine_phase(c, k.g_d * g_scale, k.alpha);
    R p = (R(1) - k.w_d) * p_hg + k.w_d * p_d;
    if (k.ice > R(0)) {
        // radians(0.5 deg)
        R width = (R(0.5) * R(kPI) / R(180)) / std::max(g_scale, R(0.1));
        R h22 = ice_halo_lobe(c, R(0.92718), width);
        R h46 = ice_halo_lobe(c, R(0.69466), width);
        p = p + k.ice * (R(0.08) * h22 + R(0.03) * h46);
    }
    return p;
}

// ---- cloud_march.wgsl:328-354 ----------------------------------------------------------------
template <class R>
inline V3<R> sun_transmittance_at(const Scene<R>& s, const V3<R>& p) {
    V3<R> sun = s.f.sun;
    V3<R> pa = p - V3<R>(0, -s.a.planet_r, 0);
    R r = std::max(length(pa), s.a.planet_r + R(1));
    V3<R> up = pa / r;
    R mu = clampr<R>(dot(up, sun), -1, 1);
    R ratio = s.a.planet_r / r;
    R sin_h = -sqrt(std::max(R(1) - ratio * ratio, R(0)));
    if (mu <= sin_h) return V3<R>(0);
    R top = s.a.top, Rp = s.a.planet_r;
    R big_h = sqrt(std::max(top * top - Rp * Rp, R(0)));
    R rho = sqrt(std::max(r * r - Rp * Rp, R(0)));
    R disc = r * r * (mu * mu - R(1)) + top * top;
    R d = std::max(-r * mu + sqrt(std::max(disc, R(0))), R(0));
    R d_min = top - r, d_max = rho + big_h;
    R xm = clampr<R>((d - d_min) / std::max(d_max - d_min, R(1)), 0, 1);
    R xr = rho / std::max(big_h, R(1));
    R t[4];
    if (ORACLE_SLIP == 15) sample2d(s.T, xr, xm, false, t);
    else sample2d(s.T, xm, xr, false, t);
    return V3<R>(t[0], t[1], t[2]);
}

// ---- cloud_march.wgsl:672-700 and sky.wgsl:73-102 (identical mapping) -------------------------
template <class R>
inline void skyview_uv(const Atm<R>& a, const V3<R>& view_dir, const V3<R>& sun, R cam_r, R& u, R& v) {
    R cos_v = clampr<R>(view_dir.y, -1, 1);
    R vza = acos(cos_v);
    R sin_h = clampr<R>(a.planet_r / cam_r, 0, 1);
    R zha = R(kPI) - asin(sin_h);
    if (vza < zha) {
        R coord = sqrt(clampr<R>(vza / std::max(zha, R(1e-6)), 0, 1));
        v = coord * R(0.5);
    } else {
        R below = std::max(R(kPI) - zha, R(1e-6));
        R coord = sqrt(clampr<R>((vza - zha) / below, 0, 1));
        v = R(0.5) + coord * R(0.5);
    }
    R view_az = atan2(view_dir.x, view_dir.z);
    R sun_az = atan2(sun.x, sun.z);
    R du = (view_az - sun_az) / (R(2) * R(kPI));
    u = du - floor(du);
}

template <class R>
inline V3<R> sample_sky_ambient(const Scene<R>& s, const V3<R>& p) {
    V3<R> c(0, -s.a.planet_r, 0);
    V3<R> up = normalize(ORACLE_SLIP == 16 ? p : p - c);
    R cam_r = length(p - c);
    R u, v;
    skyview_uv(s.a, up, s.f.sun, cam_r, u, v);
    R sv[4];
    sample2d(s.SV, u, v, false, sv);
    return V3<R>(sv[0], sv[1], sv[2]) * s.f.E;
}

// ---- cloud_march.wgsl:370-416 -----------------------------------------------------------------
template <class R>
inline bool ray_sphere_intersect(const V3<R>& ro, const V3<R>& rd, const V3<R>& c, R radius, R& t0, R& t1) {
    V3<R> oc = ro - c;
    R b = dot(oc, rd);
    R cc = dot(oc, oc) - radius * radius;
    R disc = b * b - cc;
    if (disc < R(0)) return false;
    R sq = sqrt(disc);
    t0 = -b - sq; t1 = -b + sq;
    return true;
}
template <class R>
inline bool ray_layer_intersect(const Scene<R>& s, const V3<R>& ro, const V3<R>& rd, const Layer<R>& L,
                                R& t0, R& t1) {
    V3<R> c(0, -s.a.planet_r, 0);
    R rb = s.a.planet_r + L.base, rt = s.a.planet_r + L.top;
    R tt0, tt1, tb0, tb1;
    bool hit_top = ray_sphere_intersect(ro, rd, c, rt, tt0, tt1);
    bool hit_bot = ray_sphere_intersect(ro, rd, c, rb, tb0, tb1);
    if (!hit_top) return false;
    R r_cam = length(ro - c);
    R ta, tb;
    if (r_cam < rb) {
        if (!hit_bot) return false;
        ta = std::max(tb1, R(0));
        tb = std::max(tt1, R(0));
    } else if (r_cam > rt) {
        ta = std::max(tt0, R(0));
        if (hit_bot && tb0 > R(0)) tb = std::max(tb0, R(0));
        else tb = std::max(tt1, R(0));
    } else {
        ta = R(0);
        if (hit_bot && tb0 > R(0)) tb = std::min(std::max(tt1, R(0)), tb0);
        else tb = std::max(tt1, R(0));
    }
    t0 = ta; t1 = tb;
    return tb > ta;
}

// ---- sky.wgsl:104-157 ------------------------------------------------------------------------------
template <class R>
inline R sample_overcast_at_view(const Scene<R>& s, const V3<R>& vd) {
    V3<R> cam = s.f.cam;
    R sx = cam.x, sz = cam.z;
    if (vd.y > R(1e-3)) {
        R dh = std::max(R(1500) - cam.y, R(0));
        R t = dh / vd.y;
        sx = cam.x + vd.x * t; sz = cam.z + vd.z * t;
    }
    R mask = 0;
    if (s.overcast) {
        Tex<R, uint8_t> m;
        m.data = s.overcast; m.w = 128; m.h = 128; m.d = 1; m.ch = 1;
        R u = clampr<R>((sx + R(16000)) / R(32000), 0, 1), v = clampr<R>((sz + R(16000)) / R(32000), 0, 1);
        R o[4];
        sample2d(m, u, v, false, o);
        mask = o[0] / R(255);
    }
    return R(1) - exp(R(-10) * mask);
}
template <class R>
inline V3<R> overcast_diffuse_luminance(const Scene<R>& s) {
    R hf = smoothstep<R>(R(-0.09), R(0.05), s.f.sun.y);
    R saf = sqrt(clampr<R>(s.f.sun.y, R(0.05), R(1))) * hf;
    R irr = s.f.toa * saf * R(0.25);
    return V3<R>(irr);
}

// Sky pass fragment (sky.wgsl:159-220) for the pixel centre (px, py).
template <class R>
inline V3<R> sky_pass_pixel(const Scene<R>& s, R ndc_x, R ndc_y) {
    R hv[4] = {ndc_x, ndc_y, R(1), R(1)}, nh[4];
    mat4_mul(s.f.inv_vp, hv, nh);
    V3<R> near_p(nh[0] / nh[3], nh[1] / nh[3], nh[2] / nh[3]);
    V3<R> vd = normalize(near_p - s.f.cam);
    V3<R> sun = s.f.sun;
    V3<R> cam_atm = world_to_atmosphere_pos(s.a, s.f.cam);
    R cam_r = std::max(length(cam_atm), s.a.planet_r + R(1));
    R u, v;
    skyview_uv(s.a, vd, sun, cam_r, u, v);
    R sv[4];
    sample2d(s.SV, u, v, false, sv);
    V3<R> clear = V3<R>(sv[0], sv[1], sv[2]) * s.f.E;
    V3<R> lum = clear;
    if (vd.y > R(0)) {
        R cb = sample_overcast_at_view(s, vd);
        lum = mixv(clear, overcast_diffuse_luminance(s), cb);
    }
    R cos_disk = cos(s.f.sun_ang);
    R cvs = dot(vd, sun);
    if (cvs > cos_disk) {
        R theta = acos(clampr<R>(cvs, -1, 1));
        R tn = theta / std::max(s.f.sun_ang, R(1e-6));
        R limb = R(1) - R(ORACLE_SLIP == 10 ? 0.4 : 0.6) * (R(1) - sqrt(std::max(R(1) - tn * tn, R(0))));
        R omega = R(kPI) * s.f.sun_ang * s.f.sun_ang;
        V3<R> ld = s.f.E * limb / std::max(omega, R(1e-6));
        V3<R> svis = sample_transmittance_lut(s.a, s.T, cam_atm, sun);
        lum = lum + ld * svis;
    }
    return vmax(lum, V3<R>(0));
}

// Cloud march fragment (cloud_march.wgsl:731-1054) for pixel (ix, iy).
template <class R>
inline void cloud_march_pixel(const Scene<R>& s, int ix, int iy, V3<R>& out_lum, V3<R>& out_trans, Counters* cnt) {
    // viewport / RT scale (1 at full resolution)
    const double scx = s.rt_w > 0 ? value_of(s.f.vw) / s.rt_w : 1.0;
    const double scy = s.rt_h > 0 ? value_of(s.f.vh) / s.rt_h : 1.0;
    R fx = (R(ix) + R(0.5)) * R(scx), fy = (R(iy) + R(0.5)) * R(scy);
    R ndc_x = (fx / s.f.vw) * R(2) - R(1), ndc_y = R(1) - (fy / s.f.vh) * R(2);
    R hv[4] = {ndc_x, ndc_y, R(1), R(1)}, nh[4];
    mat4_mul(s.f.inv_vp, hv, nh);
    V3<R> near_p(nh[0] / nh[3], nh[1] / nh[3], nh[2] / nh[3]);
    V3<R> ro = s.f.cam;
    V3<R> rd = normalize(near_p - ro);
    V3<R> sun = s.f.sun;
    R cos_theta = dot(rd, sun);
    R fb_dir = smoothstep<R>(0, 1, std::max(cos_theta, R(0)));

    int sxj = ix & 63, syj = iy & 63;
    if (s.c.temporal_jitter != 0u) {
        uint32_t phase = s.f.frame_index & 15u;
        int ox = int((phase * 11u + 7u) & 63u), oy = int((phase * 23u + 13u) & 63u);
        sxj = (sxj ^ ox) & 63; syj = (syj ^ oy) & 63;
    }
    R jitter = R(s.blue[syj * 64 + sxj]) / R(255);

    R depth_t_max = R(1e30);
    if (s.depth) {
        const int dx = int(ix * scx), dy = int(iy * scy);
        R dn = s.depth[size_t(dy) * size_t(value_of(s.f.vw)) + dx];
        if (dn > R(0)) {
            R dv[4] = {ndc_x, ndc_y, dn, R(1)}, wh[4];
            mat4_mul(s.f.inv_vp, dv, wh);
            V3<R> wp(wh[0] / wh[3], wh[1] / wh[3], wh[2] / wh[3]);
            depth_t_max = std::max(length(wp - ro), R(0));
        }
    }

    struct Hit { int idx; R t0, t1; int hit; };
    Hit hits[PS_MAX_CLOUD_LAYERS];
    int lc = std::min<int>(int(s.c.layer_count), PS_MAX_CLOUD_LAYERS);
    lc = std::min(lc, s.n_layers);
    for (int l = 0; l < lc; ++l) {
        R t0 = 0, t1 = 0;
        bool ok = ray_layer_intersect(s, ro, rd, s.layers[l], t0, t1);
        int flag = 0;
        if (ok) {
            t1 = std::min(t1, depth_t_max);
            if (t1 > t0) flag = 1;
        }
        hits[l] = Hit{l, t0, t1, flag};
    }
    for (int i = 1; i < lc; ++i) {
        int j = i;
        while (j > 0 && !(hits[j - 1].t0 <= hits[j].t0)) {
            std::swap(hits[j], hits[j - 1]);
            --j;
        }
    }

    V3<R> lum(0), trans(1);
    V3<R> sigma_t_base = s.c.sigma_s + s.c.sigma_a;
    R tw = 0, twn = 0;
    V3<R> c(0, -s.a.planet_r, 0);
    const V3<R> lw(R(0.2126), R(0.7152), R(0.0722));

    for (int i = 0; i < lc; ++i) {
        if (max3(trans) < R(0.01)) break;
        const Hit& lh = hits[i];
        if (lh.hit == 0) continue;
        const Layer<R>& L = s.layers[lh.idx];
        V3<R> sigma_t = sigma_t_base * chromatic_mie_modulation(L.droplet) * s.kappa[lh.idx];
        R segment = lh.t1 - lh.t0;
        if (segment <= R(0)) continue;
        if (cnt) cnt->layer_marches++;
        R step = segment / R(s.c.cloud_steps);
        R t = lh.t0 + jitter * step;
        R r0 = length(ro - c);
        R cos_view = dot((ro - c) / std::max(r0, R(1)), rd);
        R h0 = r0 - s.a.planet_r;
        Wind<R> wind = layer_wind_cache(s, lh.idx);
        for (uint32_t st = 0; st < s.c.cloud_steps; ++st) {
            if (cnt) cnt->primary_steps++;
            V3<R> p = ro + rd * t;
            R alt = altitude_from_entry(s, p, r0, cos_view, h0, t);
            R wx = p.x - wind.off_x, wz = p.z - wind.off_z;
            R wsc = std::max(s.c.weather_scale, R(1));
            R wsample[4];
            sample2d(s.weather, wx / wsc + R(0.5), wz / wsc + R(0.5), true, wsample);
            uint32_t ctype = effective_cloud_type(s, wx, wz, L);
            R density = sample_density(s, p, alt, L, wsample[0], wind, ctype, cnt);
            note_near(cnt, density, 1e-3);
            if (density > R(1e-3)) {
                if (cnt) cnt->dense_samples++;
                R od = march_to_light(s, p, alt, sun, L, wsample[0], wind, ctype, cnt);
                V3<R> beer = vexp(-sigma_t * od);
                V3<R> powder = V3<R>(1) - vexp(sigma_t * od * R(ORACLE_SLIP == 2 ? -1 : -2));
                V3<R> bp = beer * powder * R(2);
                V3<R> energy = ORACLE_SLIP == 14 ? mixv(bp, beer, s.c.powder) : mixv(beer, bp, s.c.powder);
                R phase_h = clampr<R>((alt - L.base) / std::max(L.top - L.base, R(1)), 0, 1);
                PhaseConsts<R> pc = cloud_phase_constants(s, L, phase_h);
                V3<R> sun_tr = sun_transmittance_at(s, p);
                V3<R> sun_in(0);
                R a = 1, b = 1, cc = 1;
                for (uint32_t n = 0; n < s.c.octaves; ++n) {
                    if (cnt) cnt->octaves++;
                    R ph = cloud_phase_with(cos_theta, cc, pc);
                    sun_in = sun_in + s.f.E * sun_tr * (a * ph) * vexp(-sigma_t * od * ((ORACLE_SLIP == 1 && n > 0) ? s.c.ms_b : b));
                    a = a * s.c.ms_a; b = b * s.c.ms_b; cc = cc * s.c.ms_c;
                }
                if (s.c.forward_bias > R(0)) sun_in = sun_in * (R(1) + s.c.forward_bias * fb_dir);
                R h_norm = clampr<R>((alt - L.base) / std::max(L.top - L.base, R(1)), 0, 1);
                V3<R> amb = sample_sky_ambient(s, p) * s.c.ambient * (ORACLE_SLIP == 6 ? mixr(R(1), R(0.3), h_norm) : mixr(R(0.3), R(1), h_norm));
                V3<R> lightning(0);
                R lr = s.f.light_radius;
                if (lr > R(0)) {
                    R dx = p.x - s.f.light_origin.x, dz = p.z - s.f.light_origin.z;
                    R d = sqrt(dx * dx + dz * dz);
                    R fo = clampr<R>(R(1) - d / lr, 0, 1);
                    R wgt = fo * fo * (R(3) - R(2) * fo);
                    lightning = s.f.light_rgb * wgt;
                }
                V3<R> in_scatter = s.c.sigma_s * density * (sun_in * energy + amb + lightning);
                V3<R> local_sigma = vmax(sigma_t * density, V3<R>(R(1e-6)));
                V3<R> sample_t = vexp(-local_sigma * step);
                V3<R> s_int = (in_scatter - in_scatter * sample_t) / local_sigma;
                V3<R> step_lum = trans * s_int;
                R sw = dot(step_lum, lw);
                lum = lum + step_lum;
                tw = tw + sw * t;
                twn = twn + sw;
                trans = trans * sample_t;
            }
            if (max3(trans) < R(0.01)) break;
            t = t + step;
        }
    }

    R t_lum = dot(trans, lw);
    R opacity = R(1) - t_lum;
    R t_cloud = tw / std::max(twn, R(1e-6));
    V3<R> lo = lum;
    if (opacity > R(1e-4)) {
        if (cnt) cnt->ap_taps++;
        R d_safe = std::max(t_cloud, R(50));
        R ap_z = clampr<R>(log(d_safe / R(50)) / log(R(100000) / R(50)), 0, 1);
        R ap[4];
        sample3d(s.AP, fx / s.f.vw, fy / s.f.vh, ap_z, false, ap);
        lo = lum * ap[3] + V3<R>(ap[0], ap[1], ap[2]) * opacity;
    }
    out_lum = lo;
    out_trans = trans;
}

// Sky + aerial-perspective image (the atmosphere calibration image; no
// reference counterpart as a pass): the sky pass (sky.wgsl:159-220) where the
// view ray meets no opaque surface, and where the depth buffer holds one,
// that surface's radiance attenuated and in-scattered by the AP froxels the
// way pbr.wgsl:653-655 applies them (L = L_ground * ap.a + ap.rgb, sampled at
// the surface distance with the mapping of atmosphere_lut_sampling.wgsl:33-38).
// The ground BRDF is out of scope: the surface radiance is a constant.
template <class R>
inline V3<R> sky_ap_pixel(const Scene<R>& s, int ix, int iy, R ground_radiance) {
    R fx = R(ix) + R(0.5), fy = R(iy) + R(0.5);
    R ndc_x = (fx / s.f.vw) * R(2) - R(1), ndc_y = R(1) - (fy / s.f.vh) * R(2);
    if (s.depth) {
        R dn = s.depth[size_t(iy) * size_t(value_of(s.f.vw)) + ix];
        if (dn > R(0)) {
            R dv[4] = {ndc_x, ndc_y, dn, R(1)}, wh[4];
            mat4_mul(s.f.inv_vp, dv, wh);
            V3<R> wp(wh[0] / wh[3], wh[1] / wh[3], wh[2] / wh[3]);
            R d = length(wp - s.f.cam);
            R d_safe = std::max(d, R(50));
            R ap_z = clampr<R>(log(d_safe / R(50)) / log(R(100000) / R(50)), 0, 1);
            R ap[4];
            sample3d(s.AP, fx / s.f.vw, fy / s.f.vh, ap_z, false, ap);
            return V3<R>(ground_radiance * ap[3] + ap[0], ground_radiance * ap[3] + ap[1],
                         ground_radiance * ap[3] + ap[2]);
        }
    }
    return sky_pass_pixel(s, ndc_x, ndc_y);
}

// One output pixel: sky pass, cloud march, dual-source composite
// (cloud_composite.wgsl:47-57 with blend dst = src0 + dst * src1).
template <class R>
inline void render_pixel(const Scene<R>& s, int ix, int iy, R comp[4], R lum[4], R trn[4], Counters* cnt) {
    if (cnt) cnt->pixels++;
    R fx = R(ix) + R(0.5), fy = R(iy) + R(0.5);
    R ndc_x = (fx / s.f.vw) * R(2) - R(1), ndc_y = R(1) - (fy / s.f.vh) * R(2);
    V3<R> sky = sky_pass_pixel(s, ndc_x, ndc_y);
    V3<R> cl, ct;
    cloud_march_pixel(s, ix, iy, cl, ct, cnt);
    V3<R> out = cl + sky * ct;
    comp[0] = out.x; comp[1] = out.y; comp[2] = out.z; comp[3] = R(1);
    lum[0] = cl.x; lum[1] = cl.y; lum[2] = cl.z; lum[3] = R(1);
    trn[0] = ct.x; trn[1] = ct.y; trn[2] = ct.z; trn[3] = R(1);
}

// ---- cloud_composite_halfres.wgsl:56-126 ----------------------------------------------
// The Catmull-Rom (B = 0, C = 0.5) reconstruction of a w x h RT at uv,
// ClampToEdge: the 16-tap kernel the reference's 9 offset bilinear taps
// reproduce. `rt` holds w*h RGBA texels.
inline void catmull_rom_rt(const double* rt, int w, int h, double u, double v, double out[4]) {
    auto weights = [](double f, double wt[4]) {
        wt[0] = f * (-0.5 + f * (1.0 - 0.5 * f));
        wt[1] = 1.0 + f * f * (-2.5 + 1.5 * f);
        wt[2] = f * (0.5 + f * (2.0 - 1.5 * f));
        wt[3] = f * f * (-0.5 + 0.5 * f);
    };
    const double sx = u * w, sy = v * h;
    const double bx = std::floor(sx - 0.5), by = std::floor(sy - 0.5);
    double wx[4], wy[4];
    weights(sx - (bx + 0.5), wx);
    weights(sy - (by + 0.5), wy);
    for (int k = 0; k < 4; ++k) out[k] = 0.0;
    for (int j = 0; j < 4; ++j) {
        const int yy = std::min(std::max(int(by) - 1 + j, 0), h - 1);
        for (int i = 0; i < 4; ++i) {
            const int xx = std::min(std::max(int(bx) - 1 + i, 0), w - 1);
            for (int k = 0; k < 4; ++k) out[k] += rt[(size_t(yy) * w + xx) * 4 + k] * wx[i] * wy[j];
        }
    }
}

// One output pixel of the half-resolution path: the full-resolution sky pass
// and the upsampled RT (luminance clamped >= 0, transmittance to [0, 1]),
// composited as render_pixel does.
inline void render_pixel_half(const Scene<double>& s, int ix, int iy, const double* rt_lum, const double* rt_trn,
                              double comp[4], double lum[4], double trn[4]) {
    double fx = ix + 0.5, fy = iy + 0.5;
    double ndc_x = (fx / s.f.vw) * 2.0 - 1.0, ndc_y = 1.0 - (fy / s.f.vh) * 2.0;
    V3<double> sky = sky_pass_pixel(s, ndc_x, ndc_y);
    double l[4], t[4];
    catmull_rom_rt(rt_lum, s.rt_w, s.rt_h, fx / s.f.vw, fy / s.f.vh, l);
    catmull_rom_rt(rt_trn, s.rt_w, s.rt_h, fx / s.f.vw, fy / s.f.vh, t);
    for (int k = 0; k < 3; ++k) {
        l[k] = std::max(l[k], 0.0);
        t[k] = std::min(std::max(t[k], 0.0), 1.0);
    }
    comp[0] = l[0] + sky.x * t[0]; comp[1] = l[1] + sky.y * t[1]; comp[2] = l[2] + sky.z * t[2]; comp[3] = 1.0;
    lum[0] = l[0]; lum[1] = l[1]; lum[2] = l[2]; lum[3] = 1.0;
    trn[0] = t[0]; trn[1] = t[1]; trn[2] = t[2]; trn[3] = 1.0;
}

// One output pixel of the full-resolution composite over a cloud RT of the
// frame's own size (cloud_composite.wgsl: the RT texel at the pixel, added
// as lum + sky * T by the dual-source blend, no clamps): the path the TAA
// resolve feeds at full resolution.
inline void render_pixel_rt(const Scene<double>& s, int ix, int iy, const double* rt_lum, const double* rt_trn,
                            double comp[4], double lum[4], double trn[4]) {
    double fx = ix + 0.5, fy = iy + 0.5;
    double ndc_x = (fx / s.f.vw) * 2.0 - 1.0, ndc_y = 1.0 - (fy / s.f.vh) * 2.0;
    V3<double> sky = sky_pass_pixel(s, ndc_x, ndc_y);
    const size_t o = (size_t(iy) * int(s.f.vw) + ix) * 4;
    const double* l = rt_lum + o;
    const double* t = rt_trn + o;
    comp[0] = l[0] + sky.x * t[0]; comp[1] = l[1] + sky.y * t[1]; comp[2] = l[2] + sky.z * t[2]; comp[3] = 1.0;
    lum[0] = l[0]; lum[1] = l[1]; lum[2] = l[2]; lum[3] = 1.0;
    trn[0] = t[0]; trn[1] = t[1]; trn[2] = t[2]; trn[3] = 1.0;
}

}  // namespace oracle
