// Fused per-pixel sky pass + volumetric cloud march + AP + composite for
// gfx950 (shaders/atmosphere/sky.wgsl:159-220, shaders/clouds/cloud_march.wgsl:731-1054,
// shaders/clouds/cloud_composite.wgsl:47-57 with the dual-source blend of
// crates/ps-clouds/src/pipeline.rs:469-484).
//
// Design
//  * One lane per pixel; each wave64 owns an 8x8 pixel tile so its rays stay
//    coherent (neighbouring gathers hit the same cache lines) and its lanes
//    leave the march loop together. A block is one wave: per-tile cost
//    varies by orders of magnitude (clear sky vs. horizon cloud), and
//    single-wave blocks let the dispatcher balance the CUs at the finest
//    grain (a four-wave block holds its slots until its slowest wave ends).
//  * Per-pixel ray setup is FP64: the view ray from inv_view_proj, the
//    ray/shell intersections (|o-c|^2 - r^2 at planet scale), the sun-disk
//    test and the per-layer phase coefficients. The step loop is FP32.
//  * Noise volumes are pre-folded: the base volume stores (R, 0.625G +
//    0.25B + 0.125A) so the trilinear fetch interpolates two channels, the
//    detail volume stores its FBM sum, the curl map stores 2*rg - 1.
//    (Trilinear filtering is linear, so folding the channel weights before
//    or after the fetch is the same function.)
//  * Numerically stable restatements of FP32-fragile reference formulas:
//    altitudes from h instead of |p| - R, vza via atan2 instead of acos near
//    the zenith, 1 - exp(-x) via a series for thin steps.
#include "ps_device.h"
#include "ps_launch.h"

namespace {

struct Cnt {
    uint32_t layer_marches = 0, primary = 0, covered = 0, eroded = 0, dense = 0, light = 0, light_cov = 0,
             octaves = 0, ap = 0;
};

// The 16-bit channels are interpolated as integers-in-float and scaled by
// 1/65535 once after the trilinear blend (filtering is linear).
__device__ __forceinline__ float2 unpack_rl(uint32_t v) {
    return make_float2((float)(v & 0xffffu), (float)(v >> 16));
}

// Trilinear Repeat fetch of the folded base volume (128^3, bricked, 2 x u16).
__device__ __forceinline__ float2 base_fetch(const uint32_t* __restrict__ v, float cx, float cy, float cz) {
    float fx0 = floorf(cx), fy0 = floorf(cy), fz0 = floorf(cz);
    float fx = cx - fx0, fy = cy - fy0, fz = cz - fz0;
    int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
    uint32_t x0 = brick_x(ix & 127), x1 = brick_x((ix + 1) & 127);
    uint32_t y0 = brick_y(iy & 127), y1 = brick_y((iy + 1) & 127);
    uint32_t z0 = brick_z(iz & 127), z1 = brick_z((iz + 1) & 127);
    float2 c000 = unpack_rl(ld_g<uint32_t>(v, (z0 + y0 + x0) * 4u)), c100 = unpack_rl(ld_g<uint32_t>(v, (z0 + y0 + x1) * 4u));
    float2 c010 = unpack_rl(ld_g<uint32_t>(v, (z0 + y1 + x0) * 4u)), c110 = unpack_rl(ld_g<uint32_t>(v, (z0 + y1 + x1) * 4u));
    float2 c001 = unpack_rl(ld_g<uint32_t>(v, (z1 + y0 + x0) * 4u)), c101 = unpack_rl(ld_g<uint32_t>(v, (z1 + y0 + x1) * 4u));
    float2 c011 = unpack_rl(ld_g<uint32_t>(v, (z1 + y1 + x0) * 4u)), c111 = unpack_rl(ld_g<uint32_t>(v, (z1 + y1 + x1) * 4u));
    float2 r;
    {
        float a0 = c000.x + (c100.x - c000.x) * fx, a1 = c010.x + (c110.x - c010.x) * fx;
        float b0 = c001.x + (c101.x - c001.x) * fx, b1 = c011.x + (c111.x - c011.x) * fx;
        float p0 = a0 + (a1 - a0) * fy, p1 = b0 + (b1 - b0) * fy;
        r.x = p0 + (p1 - p0) * fz;
    }
    {
        float a0 = c000.y + (c100.y - c000.y) * fx, a1 = c010.y + (c110.y - c010.y) * fx;
        float b0 = c001.y + (c101.y - c001.y) * fx, b1 = c011.y + (c111.y - c011.y) * fx;
        float p0 = a0 + (a1 - a0) * fy, p1 = b0 + (b1 - b0) * fy;
        r.y = p0 + (p1 - p0) * fz;
    }
    r.x *= 1.f / 65535.f;
    r.y *= 1.f / 65535.f;
    return r;
}

// Trilinear Repeat fetch of the folded detail volume (32^3 float).
__device__ __forceinline__ float detail_fetch(const float* __restrict__ v, float cx, float cy, float cz) {
    float fx0 = floorf(cx), fy0 = floorf(cy), fz0 = floorf(cz);
    float fx = cx - fx0, fy = cy - fy0, fz = cz - fz0;
    int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
    uint32_t x0 = ix & 31, x1 = (ix + 1) & 31;
    uint32_t y0 = (iy & 31) << 5, y1 = ((iy + 1) & 31) << 5;
    uint32_t z0 = (iz & 31) << 10, z1 = ((iz + 1) & 31) << 10;
    float c000 = ld_g<float>(v, (z0 + y0 + x0) * 4u), c100 = ld_g<float>(v, (z0 + y0 + x1) * 4u);
    float c010 = ld_g<float>(v, (z0 + y1 + x0) * 4u), c110 = ld_g<float>(v, (z0 + y1 + x1) * 4u);
    float c001 = ld_g<float>(v, (z1 + y0 + x0) * 4u), c101 = ld_g<float>(v, (z1 + y0 + x1) * 4u);
    float c011 = ld_g<float>(v, (z1 + y1 + x0) * 4u), c111 = ld_g<float>(v, (z1 + y1 + x1) * 4u);
    float a0 = c000 + (c100 - c000) * fx, a1 = c010 + (c110 - c010) * fx;
    float b0 = c001 + (c101 - c001) * fx, b1 = c011 + (c111 - c011) * fx;
    float p0 = a0 + (a1 - a0) * fy, p1 = b0 + (b1 - b0) * fy;
    return p0 + (p1 - p0) * fz;
}

// Bilinear Repeat fetch of the curl map (128^2 float2).
__device__ __forceinline__ float2 curl_fetch(const float2* __restrict__ v, float cx, float cy) {
    float fx0 = floorf(cx), fy0 = floorf(cy);
    float fx = cx - fx0, fy = cy - fy0;
    int ix = (int)fx0, iy = (int)fy0;
    uint32_t x0 = ix & 127, x1 = (ix + 1) & 127, y0 = (iy & 127) << 7, y1 = ((iy + 1) & 127) << 7;
    float2 a = ld_g<float2>(v, (y0 + x0) * 8u), b = ld_g<float2>(v, (y0 + x1) * 8u);
    float2 c = ld_g<float2>(v, (y1 + x0) * 8u), d = ld_g<float2>(v, (y1 + x1) * 8u);
    float2 r;
    float t0 = a.x + (b.x - a.x) * fx, t1 = c.x + (d.x - c.x) * fx;
    r.x = t0 + (t1 - t0) * fy;
    t0 = a.y + (b.y - a.y) * fx; t1 = c.y + (d.y - c.y) * fx;
    r.y = t0 + (t1 - t0) * fy;
    return r;
}

// Byte offset and weights of a bilinear Repeat tap of the weather plane at
// uv. The plane is stored as 2x2 texel quads, so one 8-byte load returns
// all four taps. Power-of-two sizes wrap by masking; other sizes wrap u, v
// into [0, 1) first (ix, iy in [-1, n-1]; the clamps only guard non-finite
// coordinates).
__device__ __forceinline__ uint32_t weather_tap(const RenderParams& P, float u, float v, float& fx, float& fy) {
    if (P.weather_pow2) {
        const float cx = u * P.ww_f - 0.5f, cy = v * P.wh_f - 0.5f;
        const float fx0 = floorf(cx), fy0 = floorf(cy);
        fx = cx - fx0; fy = cy - fy0;
        const uint32_t x0 = (uint32_t)(int)fx0 & (uint32_t)(P.ww - 1), y0 = (uint32_t)(int)fy0 & (uint32_t)(P.wh - 1);
        return ((y0 << P.ww_log2) | x0) * 8u;
    }
    u = u - floorf(u);
    v = v - floorf(v);
    const float cx = u * P.ww_f - 0.5f, cy = v * P.wh_f - 0.5f;
    const float fx0 = floorf(cx), fy0 = floorf(cy);
    fx = cx - fx0; fy = cy - fy0;
    const int ix = (int)fx0, iy = (int)fy0;
    const int x0 = clampi(ix < 0 ? P.ww - 1 : ix, P.ww), y0 = clampi(iy < 0 ? P.wh - 1 : iy, P.wh);
    return (uint32_t)(y0 * P.ww + x0) * 8u;
}

typedef _Float16 ps_half2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float weather_blend(uint2 q, float fx, float fy) {
    const ps_half2 r0 = __builtin_bit_cast(ps_half2, q.x), r1 = __builtin_bit_cast(ps_half2, q.y);
    const float a = (float)r0.x, b = (float)r0.y, c = (float)r1.x, d = (float)r1.y;
    const float t0 = a + (b - a) * fx, t1 = c + (d - c) * fx;
    return t0 + (t1 - t0) * fy;
}

// Vertical density profiles (cloud_density_common.wgsl:49-91). WGSL
// smoothstep(e0, e1, x) = s(clamp((x - e0) / (e1 - e0), 0, 1)) with s(t) =
// t^2 (3 - 2t), reversed edges included. Types 0..6 are a rising edge
// smoothstep(0, a, h) times either smoothstep(1, b, h) or 1 - smoothstep(c,
// 1, h); both are a + b s(clamp((h - e0) ie)) with (e0, ie, a, b) from the
// table. Type 7 (cumulonimbus) has a mid/anvil blend; other values give 0.
struct NdfC { float ia, e0, ie, a, b; };

__device__ __forceinline__ NdfC ndf_coef(uint32_t t) {
    constexpr float k[9][5] = PS_NDF_ROWS;
    const uint32_t i = t < 8u ? t : 8u;
    return NdfC{k[i][0], k[i][1], k[i][2], k[i][3], k[i][4]};
}

__device__ __forceinline__ float ndf_cb(float h, float anvil) {
    const float base = s3(clampf(h * (1.f / 0.05f), 0.f, 1.f));
    const float mid = s3(clampf((h - 0.95f) * (1.f / (0.5f - 0.95f)), 0.f, 1.f));
    const float anv = s3(clampf((h - 0.7f) * (1.f / (0.9f - 0.7f)), 0.f, 1.f)) * 1.5f * clampf(anvil, 0.f, 2.f);
    const float mt = s3(clampf((h - 0.65f) * (1.f / (0.8f - 0.65f)), 0.f, 1.f));
    return base * (mid + (anv - mid) * mt);
}

__device__ __forceinline__ float ndf_eval(const NdfC& c, uint32_t t, float h, float anvil) {
    if (t == 7u) return ndf_cb(h, anvil);
    return s3(clampf(h * c.ia, 0.f, 1.f)) * (c.a + c.b * s3(clampf((h - c.e0) * c.ie, 0.f, 1.f)));
}

__device__ __forceinline__ NdfC layer_ndf(const LayerF& L) { return NdfC{L.ndf_ia, L.ndf_e0, L.ndf_ie, L.ndf_a, L.ndf_b}; }

// Base shape through the coverage remap (shared by sample_density and
// sample_density_light). Returns the pre-erosion cloud value and writes the
// advected position. `cov` >= 0.01 here, so remap(c, 1-cov, 1, 0, 1) * cov
// equals c - (1 - cov) exactly.
//
// Empty-space skipping. base_cloud = (r + 1 - lf) / (2 - lf) is increasing in
// r and non-increasing in lf, so base_cloud <= (r_max + 1) / 2 for any LF-FBM
// value. When (r_max + 1)/2 * ndf <= (1 - cov) the shape is <= 0 and the
// density is exactly 0 whatever the fetch would return; the bound is checked
// first with r_max = 1 (no memory access), then with the brick's r_max. The
// 1e-5 margin covers FP32 rounding of the skipped arithmetic, so skipping
// never changes a result. Returns -1 when skipped.
//
// `rm` is the brick maximum when the caller has already loaded it (negative:
// load it here).
__device__ __forceinline__ float shape(const RenderParams& P, const LayerF& L, float cx, float cy, float cz,
                                       float ndfv, float cov, float rm) {
    const float thr = (1.f - cov) * (1.f - 1e-5f);
    if (ndfv <= thr) return -1.f;
    if (rm < 0.f) {
        const uint32_t bx = ((int)floorf(cx) >> 2) & 31, by = ((int)floorf(cy) >> 2) & 31,
                       bz = ((int)floorf(cz) >> 2) & 31;
        rm = (float)ld_g<uint8_t>(P.rmax, (bz << 10) | (by << 5) | bx) * (1.f / 255.f);
    }
    if ((rm + 1.f) * 0.5f * ndfv <= thr) return -1.f;
    float2 b = base_fetch(P.base_rl, cx, cy, cz);
    float lf = clampf(b.y + L.esb * 0.5f, 0.f, 1.f);
    float base_cloud = (b.x + (1.f - lf)) * frcp(fmaxf(2.f - lf, 1e-5f));
    return base_cloud * ndfv - (1.f - cov);
}

// Advected position of a sample (cloud_march.wgsl:448-458) and its texel
// coordinates in the base volume.
__device__ __forceinline__ void advect(const RenderParams& P, const LayerF& L, float px, float py, float pz, float h,
                                       float& ax, float& ay, float& az, float& cx, float& cy, float& cz) {
    ax = px - L.wind_off_x - L.skew_x * h; ay = py; az = pz - L.wind_off_z - L.skew_z * h;
    const float s = P.inv_base_scale * 128.f;
    cx = ax * s - 0.5f; cy = ay * s - 0.5f; cz = az * s - 0.5f;
}

// cloud_density_common.wgsl:266-305
__device__ __forceinline__ float density_light(const RenderParams& P, const LayerF& L, const NdfC& nc, float px,
                                               float py, float pz, float alt, float weather_r, uint32_t ctype,
                                               bool& covered) {
    float h = (alt - L.base) * L.inv_thick;
    covered = false;
    if (h < 0.f || h > 1.f) return 0.f;
    float cov = clampf(weather_r * L.coverage, 0.f, 1.f);
    if (cov < 0.01f) return 0.f;
    covered = true;
    float ax, ay, az, cx, cy, cz;
    advect(P, L, px, py, pz, h, ax, ay, az, cx, cy, cz);
    float cloud = shape(P, L, cx, cy, cz, ndf_eval(nc, ctype, h, L.anvil), cov, -1.f);
    return clampf(cloud, 0.f, 1.f) * L.density_scale;
}

// Detail erosion of a positive base shape (cloud_march.wgsl:502-527).
__device__ __forceinline__ float erode(const RenderParams& P, const LayerF& L, float cloud, float ax, float ay,
                                      float az, float h) {
    const float ds = P.inv_detail_scale;
    const float2 cu = curl_fetch(P.curl, ax * ds * 128.f - 0.5f, az * ds * 128.f - 0.5f);
    const float dx = ax + cu.x * P.curl_off_scale, dz = az + cu.y * P.curl_off_scale;
    const float hf = detail_fetch(P.detail_hf, dx * ds * 32.f - 0.5f, ay * ds * 32.f - 0.5f, dz * ds * 32.f - 0.5f);
    const float dmod = hf + ((1.f - hf) - hf) * clampf(h * 10.f, 0.f, 1.f);
    const float e = dmod * L.ld;
    return (cloud - e) * frcp(fmaxf(1.f - e, 1e-5f));
}

// cloud_march.wgsl:543-578 (straight) and 599-666 (cone), FP32.
__device__ __forceinline__ float march_light(const RenderParams& P, const LayerF& L, const NdfC& nc, float px,
                                             float py, float pz, float alt, float weather_r, uint32_t ctype, Cnt& cnt) {
    const float sx = P.sun[0], sy = P.sun[1], sz = P.sun[2];
    float cos_sun = fmaxf(sy, 0.05f);
    float dist = fmaxf((L.top - alt) / cos_sun, 1.f);
    float r0 = P.R + alt;
    float qy = py + P.R;
    const float r0c = fmaxf(r0, 1.f);
    float cv = (px * sx + qy * sy + pz * sz) * frcp(r0c);
    float k = 2.f * r0 * cv;
    float od = 0.f;
    bool cov;
    if (P.cone) {
        const float off[6][3] = {{0.38051305f, 0.92453449f, -0.02111345f}, {-0.50625799f, -0.03590792f, -0.86163418f},
                                 {-0.32509218f, -0.94557439f, 0.01428793f}, {0.09026238f, -0.27376545f, 0.95755165f},
                                 {0.28128598f, 0.42443639f, -0.86065785f}, {-0.16852403f, 0.14748697f, 0.97460106f}};
        float step = dist / 5.f;
        for (int i = 0; i < 6; ++i) {
            float t, scale, w;
            if (i < 5) { t = (i + 0.5f) * step; scale = step * (i + 1.f) / 5.f; w = step; }
            else { t = dist * 3.f; scale = step * 3.f; w = step * 3.f; }
            float qx = px + sx * t + off[i][0] * scale;
            float qyy = py + sy * t + off[i][1] * scale;
            float qz = pz + sz * t + off[i][2] * scale;
            float yy = qyy + P.R;
            float rr = fsqrt(qx * qx + yy * yy + qz * qz);
            float a = alt + t * (k + t) * frcp(r0c + rr);
            float d = density_light(P, L, nc, qx, qyy, qz, a, weather_r, ctype, cov);
            cnt.light++;
            cnt.light_cov += cov ? 1u : 0u;
            od += d * w;
        }
        return od;
    }
    float step = dist / (float)P.light_steps;
    for (int i = 0; i < P.light_steps; ++i) {
        float t = (i + 0.5f) * step;
        float qx = px + sx * t, qyy = py + sy * t, qz = pz + sz * t;
        float yy = qyy + P.R;
        float rr = fsqrt(qx * qx + yy * yy + qz * qz);
        float a = alt + t * (k + t) * frcp(r0c + rr);
        float d = density_light(P, L, nc, qx, qyy, qz, a, weather_r, ctype, cov);
        cnt.light++;
        cnt.light_cov += cov ? 1u : 0u;
        od += d * step;
    }
    return od;
}

// Transmittance-LUT coordinates for a point at altitude h whose local up
// makes cosine mu with the sun (cloud_march.wgsl:328-354), computed from h so
// r^2 - R^2 never cancels. Returns false inside the Earth's shadow.
__device__ __forceinline__ bool trans_uv_from_h(const RenderParams& P, float h, float mu, float& u, float& v) {
    float r = P.R + h;
    float rho = sqrtf(fmaxf(h * (2.f * P.R + h), 0.f));
    if (mu <= -rho / r) return false;
    float dtop = P.top_minus_R - h;                  // top - r
    float disc = dtop * (P.top + r) + (r * mu) * (r * mu);
    float sq = sqrtf(fmaxf(disc, 0.f));
    float d = mu >= 0.f ? (dtop * (P.top + r)) / fmaxf(r * mu + sq, 1e-30f) : -r * mu + sq;
    d = fmaxf(d, 0.f);
    float d_min = dtop, d_max = rho + P.big_h;
    u = clampf((d - d_min) / fmaxf(d_max - d_min, 1.f), 0.f, 1.f);
    v = rho / fmaxf(P.big_h, 1.f);
    return true;
}

// Sky-view LUT coordinates (sky.wgsl:73-102, cloud_march.wgsl:672-691) for the
// direction (dx, dy, dz) (any length); zha is the viewer's zenith-horizon angle.
// vza = atan2(|xz|, y) equals acos(y / |d|) without the cancellation near
// the zenith.
__device__ __forceinline__ void skyview_uv_f(const RenderParams& P, float dx, float dy, float dz, float zha,
                                             float& u, float& v) {
    float rxz = sqrtf(dx * dx + dz * dz);
    float vza = atan2f(rxz, dy);
    if (vza < zha) {
        v = sqrtf(clampf(vza / fmaxf(zha, 1e-6f), 0.f, 1.f)) * 0.5f;
    } else {
        float below = fmaxf((float)PS_PI - zha, 1e-6f);
        v = 0.5f + sqrtf(clampf((vza - zha) / below, 0.f, 1.f)) * 0.5f;
    }
    float du = (atan2f(dx, dz) - P.sun_az) * (float)(1.0 / (2.0 * PS_PI));
    u = du - floorf(du);
}

// One octave of the approximate-Mie cloud phase (cloud_march.wgsl:151-168,
// 278-296). The HG / Draine denominator 1 + g^2 - 2 g cos is evaluated as
// (1 - g)^2 + 2 g (1 - cos) and 1 - g^2 as (1 - g)(1 + g): for water droplets
// g ~ 0.995, where the textbook form loses ~4 digits in FP32. omg0 = 1 - g_hg
// comes from the host in FP64; omc = 1 - cos from the FP64 ray setup.
__device__ __forceinline__ float phase_octave_f(float omc, float c, float g_scale, float g_hg, float omg0,
                                                float g_d, float alpha, float w_d, float ice) {
    float g = g_hg * g_scale;
    float omg = g_scale == 1.f ? omg0 : 1.f - g;
    float den = fmaxf(omg * omg + 2.f * g * omc, 1e-4f);
    float p_hg = omg * (1.f + g) / (4.f * (float)PS_PI * den * sqrtf(den));
    float gd = g_d * g_scale;
    float omgd = 1.f - gd;
    float dend = fmaxf(omgd * omgd + 2.f * gd * omc, 1e-4f);
    float num = omgd * (1.f + gd) * (1.f + alpha * c * c);
    float dd = 4.f * (1.f + alpha * (1.f + 2.f * gd * gd) / 3.f) * (float)PS_PI * dend * sqrtf(dend);
    float p = (1.f - w_d) * p_hg + w_d * (num / fmaxf(dd, 1e-6f));
    if (ice > 0.f) {
        float width = (float)(0.5 * PS_PI / 180.0) / fmaxf(g_scale, 0.1f);
        float s22 = width * 0.37460022f, s46 = width * 0.71933556f;  // sqrt(1 - cos^2) at the halo angles
        float d22 = (c - 0.92718f) / fmaxf(s22, 1e-4f), d46 = (c - 0.69466f) / fmaxf(s46, 1e-4f);
        p += ice * (0.08f * __expf(-0.5f * d22 * d22) + 0.03f * __expf(-0.5f * d46 * d46));
    }
    return p;
}

// ---------------------------------------------------------------------------
// Deferred dense-sample shading.
//
// Only luminance depends on the light march; transmittance, the early exit
// and the step schedule depend on density alone. The primary march therefore
// appends each dense sample (density > 1e-3) to a per-wave LDS queue with
// the transmittance it saw, and whenever 64 samples are queued the whole
// wave shades them together: light march, Beer-Powder, multi-octave phase,
// sun transmittance and sky ambient, one sample per lane. Contributions go
// to per-pixel accumulators in LDS. A wave never idles lanes on the light
// march because only some of its pixels hit cloud.
// ---------------------------------------------------------------------------
constexpr int kQCap = 128;        // >= 64 + 63: a full flush plus one step's arrivals
constexpr int kQFields = 13;
enum { QF_PX = 0, QF_PY, QF_PZ, QF_ALT, QF_WR, QF_DENS, QF_TR, QF_TG, QF_TB, QF_T, QF_STEP, QF_OMC, QF_META };

struct WaveQueue {
    float f[kQFields][kQCap];
    float acc[5][64];             // per-pixel luminance rgb, sum(w t), sum(w)
};

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Shade one queued dense sample (cloud_march.wgsl:886-993) and add its
// luminance and depth-weight contributions to the owning pixel.
__device__ __forceinline__ void shade_record(const RenderParams& P, WaveQueue& Q, int j, Cnt& cnt) {
    const float px = Q.f[QF_PX][j], py = Q.f[QF_PY][j], pz = Q.f[QF_PZ][j];
    const float alt = Q.f[QF_ALT][j], wr = Q.f[QF_WR][j], density = Q.f[QF_DENS][j];
    const float Tr = Q.f[QF_TR][j], Tg = Q.f[QF_TG][j], Tb = Q.f[QF_TB][j];
    const float t = Q.f[QF_T][j], step = Q.f[QF_STEP][j], omc = Q.f[QF_OMC][j];
    const uint32_t meta = __float_as_uint(Q.f[QF_META][j]);
    const int owner = (int)(meta & 63u), li = (int)((meta >> 6) & 7u);
    const uint32_t ctype = (meta >> 9) & 7u;
    const LayerC& LC = P.layers[li];
    const LayerF L = LC.f;
    cnt.dense++;
    const float od = march_light(P, L, ndf_coef(ctype), px, py, pz, alt, wr, ctype, cnt);
    const float sig0 = L.sigma_t[0], sig1 = L.sigma_t[1], sig2 = L.sigma_t[2];
    // Beer-Powder (cloud_march.wgsl:888-892); exp(-2x) = exp(-x)^2
    const float e0r = __expf(-sig0 * od), e0g = __expf(-sig1 * od), e0b = __expf(-sig2 * od);
    const float pw = P.powder;
    const float enr = e0r + (2.f * e0r * (1.f - e0r * e0r) - e0r) * pw;
    const float eng = e0g + (2.f * e0g * (1.f - e0g * e0g) - e0g) * pw;
    const float enb = e0b + (2.f * e0b * (1.f - e0b * e0b) - e0b) * pw;
    const float hn = clampf((alt - L.base) * L.inv_thick, 0.f, 1.f);
    const float qy = py + P.R;
    // sun transmittance (cloud_march.wgsl:328-354)
    float str = 0.f, stg = 0.f, stb = 0.f;
    {
        const float hh = fmaxf(alt, 1.f);
        const float mu = clampf((px * P.sun[0] + qy * P.sun[1] + pz * P.sun[2]) / (P.R + hh), -1.f, 1.f);
        float tu, tv;
        if (trans_uv_from_h(P, hh, mu, tu, tv)) {
            const float4 tt = lut2d(P.T, P.tw, P.th, tu, tv);
            str = tt.x; stg = tt.y; stb = tt.z;
        }
    }
    // phase constants (cloud_march.wgsl:256-272); cumulonimbus blends to ice with height
    float g_hg = LC.g_hg, omg0 = LC.omg_hg, g_d = LC.g_d, alpha = LC.alpha, w_d = LC.w_d, ice = LC.ice;
    if (LC.is_cb) {
        const float mt = smoothstepf(0.60f, 0.85f, hn);
        float d = LC.droplet + (50.f - LC.droplet) * mt;
        d = clampf(d * P.droplet_bias, 5.f, 50.f);
        const float xg = -0.0990567f / (d - 1.67154f);
        g_hg = __expf(xg);
        omg0 = -expm1f(xg);
        g_d = __expf(-2.20679f / (d + 3.91029f)) - 0.428934f;
        alpha = __expf(3.62489f - 8.29288f / (d + 5.52825f));
        w_d = __expf(-0.599085f / (d - 0.641583f) - 0.665888f);
        ice = smoothstepf(35.f, 50.f, d);
    }
    // multi-octave sun term (cloud_march.wgsl:915-927)
    const float cosv = 1.f - omc;
    float sir = 0.f, sig_ = 0.f, sib = 0.f;
    const int noct = P.octaves;
    for (int n = 0; n < noct; ++n) {
        const float ph = P.a_pow[n] * phase_octave_f(omc, cosv, P.c_pow[n], g_hg, omg0, g_d, alpha, w_d, ice);
        if (n == 0) {
            sir += ph * e0r; sig_ += ph * e0g; sib += ph * e0b;
        } else {
            const float bn = P.b_pow[n];
            sir += ph * __expf(-sig0 * od * bn);
            sig_ += ph * __expf(-sig1 * od * bn);
            sib += ph * __expf(-sig2 * od * bn);
        }
    }
    cnt.octaves += (uint32_t)noct;
    sir *= P.E[0] * str; sig_ *= P.E[1] * stg; sib *= P.E[2] * stb;
    if (P.forward_bias > 0.f) {
        const float fbm = 1.f + P.forward_bias * smoothstepf(0.f, 1.f, fmaxf(cosv, 0.f));
        sir *= fbm; sig_ *= fbm; sib *= fbm;
    }
    // sky ambient at the local up (cloud_march.wgsl:693-700, 951-955)
    float ar, ag, ab;
    {
        const float rho = sqrtf(fmaxf(alt * (2.f * P.R + alt), 0.f));
        const float zha = (float)PS_PI - atan2f(P.R, rho);
        float u, v;
        skyview_uv_f(P, px, qy, pz, zha, u, v);
        const float4 sv = lut2d(P.SV, P.sw, P.sh, u, v);
        const float g = P.ambient * (0.3f + 0.7f * hn);
        ar = sv.x * P.E[0] * g; ag = sv.y * P.E[1] * g; ab = sv.z * P.E[2] * g;
    }
    // lightning (cloud_march.wgsl:968-977)
    if (P.light_radius > 0.f) {
        const float lx = px - P.light_x, lz = pz - P.light_z;
        const float fo = clampf(1.f - sqrtf(lx * lx + lz * lz) / P.light_radius, 0.f, 1.f);
        const float wgt = fo * fo * (3.f - 2.f * fo);
        ar += P.light_rgb[0] * wgt; ag += P.light_rgb[1] * wgt; ab += P.light_rgb[2] * wgt;
    }
    // energy-conserving step (cloud_march.wgsl:979-993)
    const float inr = P.sigma_s[0] * density * (sir * enr + ar);
    const float ing = P.sigma_s[1] * density * (sig_ * eng + ag);
    const float inb = P.sigma_s[2] * density * (sib * enb + ab);
    const float lsr = fmaxf(density * sig0, 1e-6f), lsg = fmaxf(density * sig1, 1e-6f),
                lsb = fmaxf(density * sig2, 1e-6f);
    const float slr = Tr * (inr * one_minus_exp_neg(lsr * step) / lsr);
    const float slg = Tg * (ing * one_minus_exp_neg(lsg * step) / lsg);
    const float slb = Tb * (inb * one_minus_exp_neg(lsb * step) / lsb);
    const float sw = slr * 0.2126f + slg * 0.7152f + slb * 0.0722f;
    atomicAdd(&Q.acc[0][owner], slr);
    atomicAdd(&Q.acc[1][owner], slg);
    atomicAdd(&Q.acc[2][owner], slb);
    atomicAdd(&Q.acc[3][owner], sw * t);
    atomicAdd(&Q.acc[4][owner], sw);
}

// Shade the first n (<= 64) queued samples with the whole wave, then move
// the entries [64, qn) to the front. Called with all 64 lanes active.
__device__ __forceinline__ void flush_queue(const RenderParams& P, WaveQueue& Q, int lane, int n, int qn, Cnt& cnt) {
    wave_sync_lds();
    if (lane < n) shade_record(P, Q, lane, cnt);
    wave_sync_lds();
    const int rest = qn - n;
    if (rest > 0) {
        float v[kQFields];
        if (lane < rest) {
#pragma unroll
            for (int k = 0; k < kQFields; ++k) v[k] = Q.f[k][64 + lane];
        }
        wave_sync_lds();
        if (lane < rest) {
#pragma unroll
            for (int k = 0; k < kQFields; ++k) Q.f[k][lane] = v[k];
        }
        wave_sync_lds();
    }
}

// Ray / layer-shell interval (cloud_march.wgsl:370-416) in FP64 with
// |oc|^2 - r^2 as (|oc| - r)(|oc| + r), clipped to the scene depth
// (cloud_march.wgsl:777-806). Returns false when the interval is empty.
__device__ __forceinline__ bool layer_interval(const LayerC& L, double r_cam, double b_, double depth_t_max,
                                              double& ta, double& tb) {
    const double ct = (r_cam - L.r_top) * (r_cam + L.r_top), cbt = (r_cam - L.r_bottom) * (r_cam + L.r_bottom);
    const double disc_t = b_ * b_ - ct, disc_b = b_ * b_ - cbt;
    const bool hit_top = disc_t >= 0.0, hit_bot = disc_b >= 0.0;
    if (!hit_top) return false;
    const double sqt = sqrt(disc_t), sqb = hit_bot ? sqrt(disc_b) : 0.0;
    const double tt0 = -b_ - sqt, tt1 = -b_ + sqt, tb0 = -b_ - sqb, tb1 = -b_ + sqb;
    if (r_cam < L.r_bottom) {
        if (!hit_bot) return false;
        ta = fmax(tb1, 0.0);
        tb = fmax(tt1, 0.0);
    } else if (r_cam > L.r_top) {
        ta = fmax(tt0, 0.0);
        tb = (hit_bot && tb0 > 0.0) ? fmax(tb0, 0.0) : fmax(tt1, 0.0);
    } else {
        ta = 0.0;
        tb = (hit_bot && tb0 > 0.0) ? fmin(fmax(tt1, 0.0), tb0) : fmax(tt1, 0.0);
    }
    if (!(tb > ta)) return false;
    tb = fmin(tb, depth_t_max);
    return tb > ta;
}

// One wave tile of TW x (64 / TW) pixels: everything from the view ray to
// the written output pixels. Called by one full wave.
template <bool COUNT, int TW>
__device__ __forceinline__ void render_tile(const RenderParams& P, int tile_x, int tile_y, WaveQueue& Q) {
    static_assert(TW == 4 || TW == 8 || TW == 16, "wave tiles are 4x16, 8x8 or 16x4 pixels");
    const int lane = threadIdx.x;
    const int x = tile_x * TW + (lane % TW);
    const int y = tile_y * (64 / TW) + lane / TW;
    // Lanes outside the image stay in the wave (they help shade queued
    // samples) but march nothing and write nothing.
    const bool valid = x < P.W && y < P.H;
    const int xc = min(x, P.W - 1), yc = min(y, P.H - 1);
    Cnt cnt;
#pragma unroll
    for (int k = 0; k < 5; ++k) Q.acc[k][lane] = 0.f;

    // ---- FP64 view ray (math.wgsl:23-33 / sky.wgsl:166-168) -------------------------
    const double fx = xc + 0.5, fy = yc + 0.5;
    const double ndx = fx / P.W * 2.0 - 1.0, ndy = 1.0 - fy / P.H * 2.0;
    const double* m = P.inv_vp;
    double nh0 = m[0] * ndx + m[4] * ndy + m[8] + m[12];
    double nh1 = m[1] * ndx + m[5] * ndy + m[9] + m[13];
    double nh2 = m[2] * ndx + m[6] * ndy + m[10] + m[14];
    double nh3 = m[3] * ndx + m[7] * ndy + m[11] + m[15];
    double dxd = nh0 / nh3 - P.cam[0], dyd = nh1 / nh3 - P.cam[1], dzd = nh2 / nh3 - P.cam[2];
    double il = 1.0 / sqrt(dxd * dxd + dyd * dyd + dzd * dzd);
    dxd *= il; dyd *= il; dzd *= il;
    const float dx = (float)dxd, dy = (float)dyd, dz = (float)dzd;
    // cos(theta) against the sun direction as uploaded (not renormalised, as
    // the reference uses it); 1 - cos in FP64 keeps the forward HG peak exact.
    const double cos_vs = dxd * P.sun_d[0] + dyd * P.sun_d[1] + dzd * P.sun_d[2];
    const float omcf = (float)(1.0 - cos_vs);

    // ---- sky pass (sky.wgsl:159-220) ------------------------------------------------------
    float3 sky;
    {
        float u, v;
        skyview_uv_f(P, dx, dy, dz, P.cam_zha, u, v);
        float4 s = lut2d(P.SV, P.sw, P.sh, u, v);
        float cr = s.x * P.E[0], cg = s.y * P.E[1], cb = s.z * P.E[2];
        if (dy > 0.f) {
            float sxz0 = P.cam_f[0], sxz1 = P.cam_f[2];
            if (dy > 1e-3f) {
                float dh = fmaxf(1500.f - P.cam_f[1], 0.f);
                float t = dh / dy;
                sxz0 = P.cam_f[0] + dx * t;
                sxz1 = P.cam_f[2] + dz * t;
            }
            float mu_ = clampf((sxz0 + 16000.f) / 32000.f, 0.f, 1.f), mv = clampf((sxz1 + 16000.f) / 32000.f, 0.f, 1.f);
            float mask = 0.f;
            if (P.overcast) {
                float cx = mu_ * 128.f - 0.5f, cy = mv * 128.f - 0.5f;
                float fx0 = floorf(cx), fy0 = floorf(cy);
                float wx = cx - fx0, wy = cy - fy0;
                int ix = (int)fx0, iy = (int)fy0;
                int x0 = clampi(ix, 128), x1 = clampi(ix + 1, 128), y0 = clampi(iy, 128), y1 = clampi(iy + 1, 128);
                float a = P.overcast[y0 * 128 + x0], b = P.overcast[y0 * 128 + x1];
                float c = P.overcast[y1 * 128 + x0], d = P.overcast[y1 * 128 + x1];
                float t0 = a + (b - a) * wx, t1 = c + (d - c) * wx;
                mask = (t0 + (t1 - t0) * wy) * (1.f / 255.f);
            }
            float blocking = 1.f - __expf(-10.f * mask);
            cr = cr + (P.overcast_lum - cr) * blocking;
            cg = cg + (P.overcast_lum - cg) * blocking;
            cb = cb + (P.overcast_lum - cb) * blocking;
        }
        if (cos_vs > P.cos_disk) {
            double theta = acos(clampd(cos_vs, -1.0, 1.0));
            double tn = theta / fmax(P.sun_ang, 1e-6);
            double limb = 1.0 - 0.6 * (1.0 - sqrt(fmax(1.0 - tn * tn, 0.0)));
            double omega = PS_PI * P.sun_ang * P.sun_ang;
            double sc = limb / fmax(omega, 1e-6);
            float mu_c = (float)(((P.cam[0]) * P.sun_d[0] + (P.cam[1] + P.R) * P.sun_d[1] + (P.cam[2]) * P.sun_d[2]) /
                                 (P.R + (double)P.cam_alt));
            float tu, tv;
            float3 vis = make_float3(0.f, 0.f, 0.f);
            if (trans_uv_from_h(P, P.cam_alt, mu_c, tu, tv)) {
                float4 tt = lut2d(P.T, P.tw, P.th, tu, tv);
                vis = make_float3(tt.x, tt.y, tt.z);
            }
            cr += (float)(P.E[0] * sc) * vis.x;
            cg += (float)(P.E[1] * sc) * vis.y;
            cb += (float)(P.E[2] * sc) * vis.z;
        }
        sky = make_float3(fmaxf(cr, 0.f), fmaxf(cg, 0.f), fmaxf(cb, 0.f));
    }

    // ---- blue-noise jitter (cloud_march.wgsl:745-756) ------------------------------------
    int jx = xc & 63, jy = yc & 63;
    if (P.temporal_jitter) {
        uint32_t ph = P.frame_index & 15u;
        jx = (jx ^ (int)((ph * 11u + 7u) & 63u)) & 63;
        jy = (jy ^ (int)((ph * 23u + 13u) & 63u)) & 63;
    }
    const float jitter = (float)P.blue[jy * 64 + jx] * (1.f / 255.f);

    // ---- depth clip (cloud_march.wgsl:764-775) --------------------------------------------
    double depth_t_max = 1e30;
    if (P.has_depth) {
        float dn = P.depth[(size_t)yc * P.W + xc];
        if (dn > 0.f) {
            double w0 = m[0] * ndx + m[4] * ndy + m[8] * dn + m[12];
            double w1 = m[1] * ndx + m[5] * ndy + m[9] * dn + m[13];
            double w2 = m[2] * ndx + m[6] * ndy + m[10] * dn + m[14];
            double w3 = m[3] * ndx + m[7] * ndy + m[11] * dn + m[15];
            double qx = w0 / w3 - P.cam[0], qy = w1 / w3 - P.cam[1], qz = w2 / w3 - P.cam[2];
            depth_t_max = fmax(sqrt(qx * qx + qy * qy + qz * qz), 0.0);
        }
    }

    // ---- per-layer intervals in FP64 (cloud_march.wgsl:370-416, 777-806) ---------------
    // Layers hit are sorted by entry distance (stable: equal keys keep layer
    // order). Only the order is kept (3 bits per entry); each interval is
    // recomputed when the march reaches it, which keeps the per-lane state
    // small.
    const double ocx = P.cam[0], ocy = P.cam[1] + P.R, ocz = P.cam[2];
    const double b_ = ocx * dxd + ocy * dyd + ocz * dzd;
    const double r_cam = P.r0;
    uint32_t order = 0;
    int nseg = 0;
    if (valid) {
        double key[PS_MAX_LAYERS];
        const int lc = P.n_layers;
        for (int l = 0; l < lc; ++l) {
            double ta, tb;
            if (!layer_interval(P.layers[l], r_cam, b_, depth_t_max, ta, tb)) continue;
            int j = nseg++;
            while (j > 0 && key[j - 1] > ta) {
                key[j] = key[j - 1];
                j--;
            }
            // insert l at rank j: shift ranks >= j up by one 3-bit slot
            const uint32_t lo = order & ((1u << (3 * j)) - 1u);
            order = lo | ((order & ~((1u << (3 * j)) - 1u)) << 3) | ((uint32_t)l << (3 * j));
            key[j] = ta;
        }
    }

    // ---- wave-uniform primary march ---------------------------------------------------
    const float r0c = fmaxf((float)P.r0, 1.f);
    const float kcam = (float)(2.0 * P.r0 * b_ / fmax(P.r0, 1.0));
    const float h0f = (float)P.h0;
    const float camx = P.cam_f[0], camy = P.cam_f[1], camz = P.cam_f[2];
    const float wscale = P.inv_weather_scale, nscale = P.inv_base_scale * 128.f;
    float tr = 1.f, tg = 1.f, tbb = 1.f;
    int cur = 0, s = 0, li = 0;
    float tstart = 0.f, step = 0.f;
    LayerF L;
    bool active = nseg > 0;
    // enter the march's next layer interval (the cur-th in entry order)
    auto enter = [&](int k) {
        li = (int)((order >> (3 * k)) & 7u);
        const LayerC& LC = P.layers[li];
        L = LC.f;
        double ta, tb;
        layer_interval(LC, r_cam, b_, depth_t_max, ta, tb);
        const double stepd = (tb - ta) / P.steps;
        tstart = (float)(ta + jitter * stepd);
        step = (float)stepd;
        s = 0;
    };
    if (active) enter(0);
    int qn = 0;  // queued samples (wave-uniform)
    while (__ballot(active) != 0ull) {
        bool dense = false;
        float t = 0.f, px = 0.f, py = 0.f, pz = 0.f, alt = 0.f, wr = 0.f, density = 0.f;
        uint32_t ctype = 0u;
        if (active) {
            cnt.primary++;
            t = tstart + (float)s * step;
            px = camx + dx * t; py = camy + dy * t; pz = camz + dz * t;
            const float qy = py + P.R;
            const float rr = fsqrt(px * px + qy * qy + pz * pz);
            alt = h0f + t * (kcam + t) * frcp(r0c + rr);
            const float h = (alt - L.base) * L.inv_thick;
            // the weather quad and the brick maximum depend on the position
            // only: both loads are issued before either is used
            const float wx = px - L.wind_off_x, wz = pz - L.wind_off_z;
            float wfx, wfy;
            const uint32_t woff = weather_tap(P, wx * wscale + 0.5f, wz * wscale + 0.5f, wfx, wfy);
            const float ax = wx - L.skew_x * h, ay = py, az = wz - L.skew_z * h;
            const float cx = ax * nscale - 0.5f, cy = ay * nscale - 0.5f, cz = az * nscale - 0.5f;
            const uint32_t roff = ((((int)floorf(cz) >> 2) & 31) << 10) | ((((int)floorf(cy) >> 2) & 31) << 5) |
                                  (((int)floorf(cx) >> 2) & 31);
            const uint2 wq = ld_g<uint2>(P.weather, woff);
            const uint32_t rmb = ld_g<uint8_t>(P.rmax, roff);
            wr = weather_blend(wq, wfx, wfy);
            ctype = L.type;
            NdfC nc = layer_ndf(L);
            if (P.use_type_grid) {
                float gu = (wx + 16000.f) / 32000.f, gv = (wz + 16000.f) / 32000.f;
                const uint32_t gx = (uint32_t)(int)floorf(gu * 128.f) & 127u, gy = (uint32_t)(int)floorf(gv * 128.f) & 127u;
                const uint32_t g = ld_g<uint8_t>(P.type_grid, gy * 128u + gx);
                if (g != 255u) { ctype = g; nc = ndf_coef(g); }
            }
            // sample_density (cloud_march.wgsl:432-530). The upper gate has a
            // 1e-5 tolerance: with blue-noise jitter 255/255 the last sample
            // lies exactly on the top shell (h = 1 in exact arithmetic), which
            // the literal `h > 1` test would decide by rounding; only the
            // cumulonimbus anvil is non-zero there.
            int stage = 0;
            if (h >= 0.f && h <= 1.f + 1e-5f) {
                const float cov = clampf(wr * L.coverage, 0.f, 1.f);
                if (cov >= 0.01f) {
                    stage = 1;
                    const float ndfv = ndf_eval(nc, ctype, h, L.anvil);
                    const float thr = (1.f - cov) * (1.f - 1e-5f);
                    const float rm = (float)rmb * (1.f / 255.f);
                    float cloud = -1.f;
                    if (ndfv > thr && (rm + 1.f) * 0.5f * ndfv > thr) {
                        const float ds = P.inv_detail_scale;
                        const float2 cu = curl_fetch(P.curl, ax * ds * 128.f - 0.5f, az * ds * 128.f - 0.5f);
                        const float2 b = base_fetch(P.base_rl, cx, cy, cz);
                        const float lf = clampf(b.y + L.esb * 0.5f, 0.f, 1.f);
                        const float base_cloud = (b.x + (1.f - lf)) * frcp(fmaxf(2.f - lf, 1e-5f));
                        cloud = base_cloud * ndfv - (1.f - cov);
                        if (cloud > 0.f) {
                            stage = 2;
                            const float ddx = ax + cu.x * P.curl_off_scale, ddz = az + cu.y * P.curl_off_scale;
                            const float hf = detail_fetch(P.detail_hf, ddx * ds * 32.f - 0.5f, ay * ds * 32.f - 0.5f,
                                                          ddz * ds * 32.f - 0.5f);
                            const float dmod = hf + ((1.f - hf) - hf) * clampf(h * 10.f, 0.f, 1.f);
                            const float e = dmod * L.ld;
                            cloud = (cloud - e) * frcp(fmaxf(1.f - e, 1e-5f));
                        }
                    }
                    density = clampf(cloud, 0.f, 1.f) * L.density_scale;
                }
            }
            cnt.covered += stage >= 1 ? 1u : 0u;
            cnt.eroded += stage >= 2 ? 1u : 0u;
            dense = density > 1e-3f;
        }
        const unsigned long long bal = __ballot(dense);
        if (bal != 0ull) {
            if (dense) {
                const int pos = qn + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                Q.f[QF_PX][pos] = px; Q.f[QF_PY][pos] = py; Q.f[QF_PZ][pos] = pz;
                Q.f[QF_ALT][pos] = alt; Q.f[QF_WR][pos] = wr; Q.f[QF_DENS][pos] = density;
                Q.f[QF_TR][pos] = tr; Q.f[QF_TG][pos] = tg; Q.f[QF_TB][pos] = tbb;
                Q.f[QF_T][pos] = t; Q.f[QF_STEP][pos] = step; Q.f[QF_OMC][pos] = omcf;
                Q.f[QF_META][pos] = __uint_as_float((uint32_t)lane | ((uint32_t)li << 6) | (ctype << 9));
            }
            qn += __popcll(bal);
        }
        if (active) {
            if (dense) {
                tr *= __expf(-fmaxf(density * L.sigma_t[0], 1e-6f) * step);
                tg *= __expf(-fmaxf(density * L.sigma_t[1], 1e-6f) * step);
                tbb *= __expf(-fmaxf(density * L.sigma_t[2], 1e-6f) * step);
            }
            ++s;
            if (fmaxf(tr, fmaxf(tg, tbb)) < 0.01f) {
                active = false;
            } else if (s >= P.steps) {
                ++cur;
                if (cur < nseg) enter(cur);
                else active = false;
            }
        }
        if (qn >= 64) {
            flush_queue(P, Q, lane, 64, qn, cnt);
            qn -= 64;
        }
    }
    if (qn > 0) flush_queue(P, Q, lane, qn, qn, cnt);
    wave_sync_lds();
    cnt.layer_marches = (uint32_t)nseg;

    // ---- aerial perspective on the cloud luminance (cloud_march.wgsl:1016-1048) ----
    const float lr = Q.acc[0][lane], lg = Q.acc[1][lane], lb = Q.acc[2][lane];
    const float tw = Q.acc[3][lane], twn = Q.acc[4][lane];
    const float t_lum = tr * 0.2126f + tg * 0.7152f + tbb * 0.0722f;
    const float opacity = 1.f - t_lum;
    float olr = lr, olg = lg, olb = lb;
    if (valid && opacity > 1e-4f) {
        cnt.ap++;
        const float tc = tw / fmaxf(twn, 1e-6f);
        const float ds = fmaxf(tc, 50.f);
        const float apz = clampf(__logf(ds / 50.f) * (float)(1.0 / 7.6009024595420822), 0.f, 1.f);
        const float u = (float)fx * P.inv_W, v = (float)fy * P.inv_H;
        // trilinear ClampToEdge on the froxel volume
        const float cx = u * P.aw - 0.5f, cy = v * P.ah - 0.5f, cz = apz * P.ad - 0.5f;
        const float fx0 = floorf(cx), fy0 = floorf(cy), fz0 = floorf(cz);
        const float wx = cx - fx0, wy = cy - fy0, wz = cz - fz0;
        const int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
        const int x0 = clampi(ix, P.aw), x1 = clampi(ix + 1, P.aw), y0 = clampi(iy, P.ah), y1 = clampi(iy + 1, P.ah),
                  z0 = clampi(iz, P.ad), z1 = clampi(iz + 1, P.ad);
        const int sxy = P.aw * P.ah;
        const float4 c000 = P.AP[z0 * sxy + y0 * P.aw + x0], c100 = P.AP[z0 * sxy + y0 * P.aw + x1];
        const float4 c010 = P.AP[z0 * sxy + y1 * P.aw + x0], c110 = P.AP[z0 * sxy + y1 * P.aw + x1];
        const float4 c001 = P.AP[z1 * sxy + y0 * P.aw + x0], c101 = P.AP[z1 * sxy + y0 * P.aw + x1];
        const float4 c011 = P.AP[z1 * sxy + y1 * P.aw + x0], c111 = P.AP[z1 * sxy + y1 * P.aw + x1];
        float ap[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float v000 = ((const float*)&c000)[k], v100 = ((const float*)&c100)[k];
            const float v010 = ((const float*)&c010)[k], v110 = ((const float*)&c110)[k];
            const float v001 = ((const float*)&c001)[k], v101 = ((const float*)&c101)[k];
            const float v011 = ((const float*)&c011)[k], v111 = ((const float*)&c111)[k];
            const float a0 = v000 + (v100 - v000) * wx, a1 = v010 + (v110 - v010) * wx;
            const float b0 = v001 + (v101 - v001) * wx, b1 = v011 + (v111 - v011) * wx;
            const float p0 = a0 + (a1 - a0) * wy, p1 = b0 + (b1 - b0) * wy;
            ap[k] = p0 + (p1 - p0) * wz;
        }
        olr = lr * ap[3] + ap[0] * opacity;
        olg = lg * ap[3] + ap[1] * opacity;
        olb = lb * ap[3] + ap[2] * opacity;
    }

    // ---- outputs: composite dst = lum + sky * T ------------------------------------------------
    if (valid) {
        const size_t o = (size_t)y * P.W + x;
        const float4 comp = make_float4(olr + sky.x * tr, olg + sky.y * tg, olb + sky.z * tbb, 1.f);
        const float4 lum4 = make_float4(olr, olg, olb, 1.f);
        const float4 tr4 = make_float4(tr, tg, tbb, 1.f);
        if (P.out_f16) {
            if (P.out_comp) {
                __half2* q = (__half2*)P.out_comp + 2 * o;
                q[0] = __floats2half2_rn(comp.x, comp.y); q[1] = __floats2half2_rn(comp.z, comp.w);
            }
            if (P.out_lum) {
                __half2* q = (__half2*)P.out_lum + 2 * o;
                q[0] = __floats2half2_rn(lum4.x, lum4.y); q[1] = __floats2half2_rn(lum4.z, lum4.w);
            }
            if (P.out_trans) {
                __half2* q = (__half2*)P.out_trans + 2 * o;
                q[0] = __floats2half2_rn(tr4.x, tr4.y); q[1] = __floats2half2_rn(tr4.z, tr4.w);
            }
        } else {
            if (P.out_comp) ((float4*)P.out_comp)[o] = comp;
            if (P.out_lum) ((float4*)P.out_lum)[o] = lum4;
            if (P.out_trans) ((float4*)P.out_trans)[o] = tr4;
        }
    }

    if (COUNT) {
        // wave sums first (all 64 lanes are converged here), then one
        // atomic per counter per wave
        uint32_t v[10] = {valid ? 1u : 0u, cnt.layer_marches, cnt.primary, cnt.covered, cnt.eroded,
                          cnt.dense,       cnt.light,         cnt.light_cov, cnt.octaves, cnt.ap};
#pragma unroll
        for (int k = 0; k < 10; ++k)
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) v[k] += __shfl_xor(v[k], o);
        if (lane == 0) {
            unsigned long long* c = P.counters;
#pragma unroll
            for (int k = 0; k < 10; ++k) atomicAdd(c + k, (unsigned long long)v[k]);
        }
    }
}

template <bool COUNT>
__global__ __launch_bounds__(64) void k_render(const RenderParams* __restrict__ Pp) {
    __shared__ WaveQueue Qs;
    render_tile<COUNT, 8>(*Pp, blockIdx.x, blockIdx.y, Qs);
}

// Tiles of many scenes in one launch. Each entry of `tiles` is
// scene << 20 | tile_y << 10 | tile_x. Persistent: every one-wave block
// pulls the next entry from the counter until the list is exhausted (the
// exit test is wave-uniform and the counter only grows, so every wave
// leaves). Otherwise block b renders entry b.
template <int TW, bool PERSISTENT>
__global__ __launch_bounds__(64) void k_render_tiles(const RenderParams* __restrict__ Ps,
                                                     const uint32_t* __restrict__ tiles, int n_tiles,
                                                     unsigned* __restrict__ next) {
    __shared__ WaveQueue Qs;
    if (PERSISTENT) {
        for (;;) {
            // readfirstlane keeps the tile index, the entry and hence the
            // scene's parameter block in scalar registers
            unsigned t = 0;
            if (threadIdx.x == 0) t = atomicAdd(next, 1u);
            t = __builtin_amdgcn_readfirstlane(t);
            if (t >= (unsigned)n_tiles) break;
            const uint32_t e = __builtin_amdgcn_readfirstlane(tiles[t]);
            render_tile<false, TW>(Ps[e >> 20], (int)(e & 1023u), (int)((e >> 10) & 1023u), Qs);
        }
    } else {
        const uint32_t e = __builtin_amdgcn_readfirstlane(tiles[blockIdx.x]);
        render_tile<false, TW>(Ps[e >> 20], (int)(e & 1023u), (int)((e >> 10) & 1023u), Qs);
    }
}

__global__ void k_zero_counter(unsigned* p) { *p = 0u; }

}  // namespace

hipError_t ps_launch_render(const RenderParams* d_params, int W, int H, bool count, hipStream_t s) {
    dim3 grid((W + 7) / 8, (H + 7) / 8), block(64);
    if (count) hipLaunchKernelGGL(k_render<true>, grid, block, 0, s, d_params);
    else hipLaunchKernelGGL(k_render<false>, grid, block, 0, s, d_params);
    return hipGetLastError();
}

hipError_t ps_launch_render_tiles(const RenderParams* d_params, const uint32_t* tiles, int n_tiles, int tile_w,
                                  int persistent, int blocks, unsigned* next, hipStream_t s) {
    if (n_tiles <= 0) return hipSuccess;
    if (persistent) {
        hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, s, next);
        const dim3 g(blocks), b(64);
        if (tile_w == 4) hipLaunchKernelGGL((k_render_tiles<4, true>), g, b, 0, s, d_params, tiles, n_tiles, next);
        else if (tile_w == 16) hipLaunchKernelGGL((k_render_tiles<16, true>), g, b, 0, s, d_params, tiles, n_tiles, next);
        else hipLaunchKernelGGL((k_render_tiles<8, true>), g, b, 0, s, d_params, tiles, n_tiles, next);
    } else {
        const dim3 g(n_tiles), b(64);
        if (tile_w == 4) hipLaunchKernelGGL((k_render_tiles<4, false>), g, b, 0, s, d_params, tiles, n_tiles, next);
        else if (tile_w == 16) hipLaunchKernelGGL((k_render_tiles<16, false>), g, b, 0, s, d_params, tiles, n_tiles, next);
        else hipLaunchKernelGGL((k_render_tiles<8, false>), g, b, 0, s, d_params, tiles, n_tiles, next);
    }
    return hipGetLastError();
}
