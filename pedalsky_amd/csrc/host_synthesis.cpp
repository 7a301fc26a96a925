// Host-side (CPU) pieces of the path that the reference computes on the CPU
// before the GPU passes: parameter defaults, the scene -> CloudLayerGpu
// synthesis, the per-layer wind tap, the top-down overcast mask value and the
// blue-noise rank map. All arithmetic is f32 like the Rust originals.
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <vector>
#include <limits>
#include "../../include/pedalsky_hip.h"

namespace {

// IEEE binary16 round-to-nearest-even, as half::f16::from_f32.
uint16_t f32_to_f16_bits(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xffu) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u : 0u));
    int e = (int)exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        uint32_t shift = (uint32_t)(14 - e);
        uint32_t half_man = man >> shift;
        uint32_t rem = man & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1u);
        if (rem > halfway || (rem == halfway && (half_man & 1u))) half_man++;
        return (uint16_t)(sign | half_man);
    }
    uint32_t half_man = man >> 13;
    uint32_t rem = man & 0x1fffu;
    uint32_t out = sign | ((uint32_t)e << 10) | half_man;
    if (rem > 0x1000u || (rem == 0x1000u && (half_man & 1u))) out++;
    return (uint16_t)out;
}
float f16_bits_to_f32(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1fu;
    uint32_t man = h & 0x3ffu;
    uint32_t out;
    if (exp == 0) {
        if (man == 0) out = sign;
        else {
            int e = -1;
            do { man <<= 1; ++e; } while (!(man & 0x400u));
            out = sign | ((uint32_t)(127 - 15 - e) << 23) | ((man & 0x3ffu) << 13);
        }
    } else if (exp == 31) {
        out = sign | 0x7f800000u | (man << 13);
    } else {
        out = sign | ((exp + 127 - 15) << 23) | (man << 13);
    }
    float f;
    std::memcpy(&f, &out, 4);
    return f;
}

// ps-synthesis/src/ndf.rs:11-17 (CPU smoothstep with the degenerate-edge guard)
float cpu_smoothstep(float e0, float e1, float x) {
    if (std::fabs(e1 - e0) < std::numeric_limits<float>::epsilon()) return x < e0 ? 0.f : 1.f;
    float t = std::fmin(std::fmax((x - e0) / (e1 - e0), 0.f), 1.f);
    return t * t * (3.f - 2.f * t);
}
// ps-synthesis/src/ndf.rs:20-58 (keeps the per-type multipliers of the CPU mirror)
float cpu_ndf(float h, uint32_t t) {
    switch (t) {
        case 0: return cpu_smoothstep(0.f, 0.07f, h) * cpu_smoothstep(1.f, 0.2f, h);
        case 1: return cpu_smoothstep(0.f, 0.10f, h) * (1.f - cpu_smoothstep(0.6f, 1.f, h));
        case 2: return cpu_smoothstep(0.f, 0.15f, h) * cpu_smoothstep(1.f, 0.4f, h);
        case 3: return cpu_smoothstep(0.f, 0.20f, h) * cpu_smoothstep(1.f, 0.3f, h) * 0.8f;
        case 4: return cpu_smoothstep(0.f, 0.30f, h) * (1.f - cpu_smoothstep(0.7f, 1.f, h)) * 0.6f;
        case 5: return cpu_smoothstep(0.f, 0.40f, h) * (1.f - cpu_smoothstep(0.6f, 1.f, h)) * 0.4f;
        case 6: return cpu_smoothstep(0.f, 0.50f, h) * (1.f - cpu_smoothstep(0.8f, 1.f, h)) * 0.3f;
        case 7: {
            float base = cpu_smoothstep(0.f, 0.05f, h);
            float mid = cpu_smoothstep(0.95f, 0.5f, h);
            float anvil = cpu_smoothstep(0.7f, 0.9f, h) * 1.5f;
            float mt = cpu_smoothstep(0.65f, 0.8f, h);
            return base * (mid + (anvil - mid) * mt);
        }
        default: return 0.f;
    }
}

// ps-synthesis/src/wind_field.rs:167-178
void horizontal_wind(float surface_speed, float dir_to_rad, float alt_m, float& u, float& v) {
    const float Z_REF_M = 10.f, ALPHA = 0.143f, TOP_M = 12000.f;
    const float EKMAN_MAX_RAD = 30.0f * 3.14159265358979323846f / 180.0f;
    float speed = surface_speed * std::pow(std::fmax(alt_m, Z_REF_M) / Z_REF_M, ALPHA);
    float veer = std::pow(std::fmin(std::fmax(alt_m / TOP_M, 0.f), 1.f), 0.7f);
    float veered = dir_to_rad + veer * EKMAN_MAX_RAD;
    u = speed * std::sin(veered);
    v = speed * -std::cos(veered);
}

}  // namespace

namespace psi {

// The blue-noise splat weight by squared texel distance r2 = 0..72:
// exp(-r2 / (2 sigma^2)) with sigma = 1.9, in f32 as noise.rs:655-675
// evaluates it.
void blue_noise_weights(float* w73) {
    const float sigma = 1.9f;
    const float inv_two_sigma2 = 1.0f / (2.0f * sigma * sigma);
    for (int r2 = 0; r2 < 73; ++r2) w73[r2] = std::exp(-(float)r2 * inv_two_sigma2);
}

}  // namespace psi

extern "C" {

void ps_default_atmosphere(ps_atmosphere_params* o) {
    std::memset(o, 0, sizeof(*o));
    o->planet_radius_m = 6360000.f;
    o->atmosphere_top_m = 6460000.f;
    o->rayleigh_scale_height_m = 8000.f;
    o->mie_scale_height_m = 1200.f;
    o->rayleigh_scattering[0] = 5.802e-6f; o->rayleigh_scattering[1] = 13.558e-6f; o->rayleigh_scattering[2] = 33.100e-6f;
    for (int i = 0; i < 3; ++i) { o->mie_scattering[i] = 3.996e-6f; o->mie_absorption[i] = 4.4e-6f; }
    o->mie_g = 0.8f;
    o->ozone_absorption[0] = 0.650e-6f; o->ozone_absorption[1] = 1.881e-6f; o->ozone_absorption[2] = 0.085e-6f;
    o->ozone_center_m = 25000.f;
    o->ozone_thickness_m = 30000.f;
    for (int i = 0; i < 3; ++i) o->ground_albedo[i] = 0.18f;
}

void ps_default_cloud_params(ps_cloud_params* o) {
    std::memset(o, 0, sizeof(*o));
    o->sigma_s[0] = 0.030f; o->sigma_s[1] = 0.040f; o->sigma_s[2] = 0.060f;
    o->droplet_diameter_bias = 1.f;
    o->cone_light_sampling = 0;
    o->forward_bias = 0.f;
    o->detail_strength = 0.05f;
    o->curl_strength = 0.1f;
    o->powder_strength = 1.f;
    o->multi_scatter_a = 0.5f; o->multi_scatter_b = 0.5f; o->multi_scatter_c = 0.5f;
    o->ambient_strength = 1.f;
    o->base_scale_m = 4500.f;
    o->detail_scale_m = 800.f;
    o->weather_scale_m = 32000.f;
    o->light_steps = 6;
    o->cloud_steps = 192;
    o->multi_scatter_octaves = 4;
    o->cloud_layer_count = 0;
    o->temporal_jitter = 0;
    o->wind_drift_strength = 1.f;
    o->wind_skew_strength = 0.5f;
    o->diurnal_strength = 1.f;
}

float ps_remap_coverage(float scene_coverage) {
    const float VISIBLE_LOW = 0.40f, VISIBLE_HIGH = 0.60f, KNEE_END = 0.10f;
    float t = std::fmin(std::fmax(scene_coverage, 0.f), 1.f);
    if (t <= 0.f) return 0.f;
    if (t < KNEE_END) {
        float s = t / KNEE_END;
        float sm = s * s * (3.f - 2.f * s);
        return sm * VISIBLE_LOW;
    }
    float band_t = std::fmin(std::fmax((t - KNEE_END) / (1.f - KNEE_END), 0.f), 1.f);
    return VISIBLE_LOW + band_t * (VISIBLE_HIGH - VISIBLE_LOW);
}

float ps_default_density_scale(uint32_t t) {
    switch (t) {
        case 0: case 1: case 2: return 1.0f;
        case 3: return 0.85f;
        case 4: return 0.7f;
        case 5: return 0.55f;
        case 6: return 0.4f;
        case 7: return 1.4f;
        default: return 1.0f;
    }
}

float ps_default_droplet_diameter_um(uint32_t t) {
    switch (t) {
        case 0: return 20.f;
        case 1: case 2: return 16.f;
        case 3: return 14.f;
        case 4: return 30.f;
        case 5: case 6: return 50.f;
        case 7: return 20.f;
        default: return 20.f;
    }
}

void ps_synthesise_cloud_layer(uint32_t cloud_type, float base_m, float top_m, float scene_coverage,
                               float density_scale, float shape_bias, float detail_bias, float anvil_bias,
                               float droplet_diameter_um, ps_cloud_layer* out) {
    std::memset(out, 0, sizeof(*out));
    out->base_m = base_m;
    out->top_m = top_m;
    out->coverage = ps_remap_coverage(scene_coverage);
    out->density_scale = std::isnan(density_scale) ? ps_default_density_scale(cloud_type) : density_scale;
    out->cloud_type = cloud_type;
    out->shape_bias = shape_bias;
    out->detail_bias = detail_bias;
    out->anvil_bias = std::isnan(anvil_bias) ? (cloud_type == 7 ? 1.f : 0.f) : anvil_bias;
    out->droplet_diameter_um =
        std::isnan(droplet_diameter_um) ? ps_default_droplet_diameter_um(cloud_type) : droplet_diameter_um;
}

int32_t ps_check_layers_overlap(const ps_cloud_layer* layers, int32_t n, int32_t* out_i, int32_t* out_j) {
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            const ps_cloud_layer& a = layers[i];
            const ps_cloud_layer& b = layers[j];
            if (a.base_m < b.top_m && b.base_m < a.top_m) {
                if (out_i) *out_i = i;
                if (out_j) *out_j = j;
                return 1;
            }
        }
    return 0;
}

void ps_layer_wind(float surface_speed_mps, float wind_dir_deg, float base_m, float top_m, float out_xz[2]) {
    // WindField::synthesise (wind_field.rs:71-115): 16 altitude levels of
    // f16 (u, v); layer_wind_mps (cloud_density_common.wgsl:264-271) taps the
    // texture at (0.5, clamp(mid/12000, 0, 0.9999), 0.5) with linear
    // filtering and Repeat addressing. The profile does not vary in x or z,
    // so only the y interpolation remains.
    const int SY = 16;
    const float TOP_M = 12000.f;
    float dir_to_rad = (wind_dir_deg + 180.f) * (3.14159265358979323846f / 180.f);
    float lu[SY], lv[SY];
    for (int y = 0; y < SY; ++y) {
        float alt = ((float)y + 0.5f) / (float)SY * TOP_M;
        float u, v;
        horizontal_wind(surface_speed_mps, dir_to_rad, alt, u, v);
        lu[y] = f16_bits_to_f32(f32_to_f16_bits(u));
        lv[y] = f16_bits_to_f32(f32_to_f16_bits(v));
    }
    float mid = (base_m + top_m) * 0.5f;
    float uy = std::fmin(std::fmax(mid / TOP_M, 0.f), 0.9999f);
    float cy = uy * (float)SY - 0.5f;
    float fl = std::floor(cy);
    float fy = cy - fl;
    int iy = (int)fl;
    int y0 = ((iy % SY) + SY) % SY, y1 = (((iy + 1) % SY) + SY) % SY;
    out_xz[0] = lu[y0] + (lu[y1] - lu[y0]) * fy;
    out_xz[1] = lv[y0] + (lv[y1] - lv[y0]) * fy;
}

// One layer's column density at a coverage (density_mask.rs:155-166): the
// mean NDF over 100 m height steps x thickness in km x coverage x density.
static float layer_column(const ps_cloud_layer& L, float coverage) {
    float thickness = std::fmax(L.top_m - L.base_m, 1.f);
    uint32_t n_steps = (uint32_t)std::fmax(std::ceil(thickness / 100.f), 1.f);
    float dh = 1.f / (float)n_steps;
    float sum = 0.f;
    for (uint32_t k = 0; k < n_steps; ++k) sum += cpu_ndf(((float)k + 0.5f) * dh, L.cloud_type);
    float mean = sum / (float)n_steps;
    return mean * (thickness / 1000.f) * coverage * L.density_scale;
}

// Column total / 10 in [0, 1] as the mask's u8 (density_mask.rs:64, 134-151).
static uint8_t mask_u8(float total) {
    float col = std::fmin(std::fmax(total / 10.f, 0.f), 1.f);
    float v = std::fmin(std::fmax(col * 255.f, 0.f), 255.f);
    return (uint8_t)v;
}

uint8_t ps_overcast_mask_value(const ps_cloud_layer* layers, int32_t n) {
    // density_mask.rs:134-140 (uniform case, each layer's own coverage)
    float total = 0.f;
    for (int i = 0; i < n; ++i) total += layer_column(layers[i], layers[i].coverage);
    return mask_u8(total);
}

void ps_overcast_mask(const ps_cloud_layer* layers, int32_t n, const float* grid, int32_t grid_w, int32_t grid_h,
                      float grid_extent_m, uint8_t* out) {
    // TopDownMask::synthesise (density_mask.rs:47-83): with a coverage grid,
    // texel (x, y) of the 128^2 mask over 32 km takes the grid's bilinear,
    // edge-clamped coverage at its centre (LoadedCoverageGrid::sample_world,
    // coverage_grid.rs:31-61) for every layer; without one, the uniform value.
    const int size = 128;
    if (!grid || grid_w < 1 || grid_h < 1) {
        std::memset(out, ps_overcast_mask_value(layers, n), size * size);
        return;
    }
    const float extent = 32000.f, half = extent * 0.5f, ghalf = grid_extent_m * 0.5f;
    for (int y = 0; y < size; ++y)
        for (int x = 0; x < size; ++x) {
            const float gx = ((float)x + 0.5f) / (float)size * extent - half;
            const float gy = ((float)y + 0.5f) / (float)size * extent - half;
            const float u = (gx + ghalf) / grid_extent_m, v = (gy + ghalf) / grid_extent_m;
            const float fx = std::fmin(std::fmax(u, 0.f), 1.f) * ((float)grid_w - 1.f);
            const float fy = std::fmin(std::fmax(v, 0.f), 1.f) * ((float)grid_h - 1.f);
            const uint32_t x0 = (uint32_t)std::floor(fx), y0 = (uint32_t)std::floor(fy);
            const uint32_t x1 = std::min(x0 + 1u, (uint32_t)grid_w - 1u), y1 = std::min(y0 + 1u, (uint32_t)grid_h - 1u);
            const float tx = fx - (float)x0, ty = fy - (float)y0;
            const float v00 = grid[y0 * grid_w + x0], v10 = grid[y0 * grid_w + x1];
            const float v01 = grid[y1 * grid_w + x0], v11 = grid[y1 * grid_w + x1];
            const float bottom = v00 + (v10 - v00) * tx, top = v01 + (v11 - v01) * tx;
            const float cov = std::fmin(std::fmax(bottom + (top - bottom) * ty, 0.f), 1.f);
            float total = 0.f;
            for (int i = 0; i < n; ++i) total += layer_column(layers[i], cov);
            out[y * size + x] = mask_u8(total);
        }
}

void ps_generate_blue_noise(uint8_t* out) {
    // generate_blue_noise(64), crates/ps-clouds/src/noise.rs:646-704
    const int size = 64;
    const int n = size * size;
    const int radius = 6;
    float w_r2[73];
    psi::blue_noise_weights(w_r2);
    std::vector<uint8_t> taken(n, 0);
    std::vector<uint32_t> ranks(n, 0);
    std::vector<float> energy(n, 0.f);
    for (int placement = 0; placement < n; ++placement) {
        int best = 0;
        float best_e = std::numeric_limits<float>::infinity();
        for (int i = 0; i < n; ++i)
            if (!taken[i] && energy[i] < best_e) { best_e = energy[i]; best = i; }
        taken[best] = 1;
        ranks[best] = (uint32_t)placement;
        int cx = best % size, cy = best / size;
        for (int dy = -radius; dy <= radius; ++dy)
            for (int dx = -radius; dx <= radius; ++dx) {
                int mx = ((cx + dx) % size + size) % size, my = ((cy + dy) % size + size) % size;
                energy[my * size + mx] += w_r2[dx * dx + dy * dy];
            }
    }
    const float max_rank = (float)(n - 1);
    for (int i = 0; i < n; ++i) {
        float v = std::round(((float)ranks[i] / max_rank) * 255.f);
        v = std::fmin(std::fmax(v, 0.f), 255.f);
        out[i] = (uint8_t)v;
    }
}

}  // extern "C"
