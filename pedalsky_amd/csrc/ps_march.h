// Device helpers of the cloud march shared by the forward render kernel
// (render_kernel.hip) and its adjoint (adjoint_kernel.hip): noise and
// weather fetches, the vertical profiles, the density function, the light
// march, LUT parametrisations, the cloud phase and the ray / layer-shell
// interval. Included by exactly those translation units (internal linkage).
#pragma once
#include "ps_device.h"

namespace {
struct Cnt {
    uint32_t layer_marches = 0, primary = 0, covered = 0, eroded = 0, dense = 0, light = 0, light_cov = 0,
             octaves = 0, ap = 0;
};

// The 16-bit channels are interpolated as integers-in-float and scaled once
// after the trilinear blend (filtering is linear).
__device__ __forceinline__ float2 unpack_rl(uint32_t v) {
    return make_float2((float)(v & 0xffffu), (float)(v >> 16));
}

// Trilinear Repeat fetch of the folded base volume (128^3, bricked, 2 x u16
// fixed point; rs and ls turn the blended R and LF integers into values).
// Each texel holds its (y, z) neighbourhood, so the eight corners are two
// 16-byte loads (the x0 and x1 texels).
__device__ __forceinline__ float2 base_fetch(const uint32_t* __restrict__ v, float cx, float cy, float cz, float rs,
                                             float ls) {
    float fx0 = floorf(cx), fy0 = floorf(cy), fz0 = floorf(cz);
    float fx = cx - fx0, fy = cy - fy0, fz = cz - fz0;
    int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
    const uint32_t yz = brick_y(iy & 127) + brick_z(iz & 127);
    const uint4 q0 = ld_g<uint4>(v, (brick_x(ix & 127) + yz) * 16u);
    const uint4 q1 = ld_g<uint4>(v, (brick_x((ix + 1) & 127) + yz) * 16u);
    float2 c000 = unpack_rl(q0.x), c100 = unpack_rl(q1.x), c010 = unpack_rl(q0.y), c110 = unpack_rl(q1.y);
    float2 c001 = unpack_rl(q0.z), c101 = unpack_rl(q1.z), c011 = unpack_rl(q0.w), c111 = unpack_rl(q1.w);
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
    r.x *= rs;
    r.y *= ls;
    return r;
}

// Trilinear Repeat fetch of the folded detail volume (32^3; each texel holds
// its (y, z) neighbourhood: two 16-byte loads).
__device__ __forceinline__ float detail_fetch(const float* __restrict__ v, float cx, float cy, float cz) {
    float fx0 = floorf(cx), fy0 = floorf(cy), fz0 = floorf(cz);
    float fx = cx - fx0, fy = cy - fy0, fz = cz - fz0;
    int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
    const uint32_t yz = ((uint32_t)(iy & 31) << 5) + ((uint32_t)(iz & 31) << 10);
    const float4 q0 = ld_g<float4>(v, (yz + (uint32_t)(ix & 31)) * 16u);
    const float4 q1 = ld_g<float4>(v, (yz + (uint32_t)((ix + 1) & 31)) * 16u);
    float c000 = q0.x, c100 = q1.x, c010 = q0.y, c110 = q1.y;
    float c001 = q0.z, c101 = q1.z, c011 = q0.w, c111 = q1.w;
    float a0 = c000 + (c100 - c000) * fx, a1 = c010 + (c110 - c010) * fx;
    float b0 = c001 + (c101 - c001) * fx, b1 = c011 + (c111 - c011) * fx;
    float p0 = a0 + (a1 - a0) * fy, p1 = b0 + (b1 - b0) * fy;
    return p0 + (p1 - p0) * fz;
}

// Bilinear Repeat fetch of the curl map (128^2 float2; each texel holds its
// y neighbour: two 16-byte loads).
__device__ __forceinline__ float2 curl_fetch(const float2* __restrict__ v, float cx, float cy) {
    float fx0 = floorf(cx), fy0 = floorf(cy);
    float fx = cx - fx0, fy = cy - fy0;
    int ix = (int)fx0, iy = (int)fy0;
    const uint32_t y = (uint32_t)(iy & 127) << 7;
    const float4 q0 = ld_g<float4>(v, (y + (uint32_t)(ix & 127)) * 16u);
    const float4 q1 = ld_g<float4>(v, (y + (uint32_t)((ix + 1) & 127)) * 16u);
    const float2 a = make_float2(q0.x, q0.y), b = make_float2(q1.x, q1.y);
    const float2 c = make_float2(q0.z, q0.w), d = make_float2(q1.z, q1.w);
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

// ---------------------------------------------------------------------------
// Density perturbation volume (differentiation extension). Texel coordinates
// of a sample at drift-offset position (wx, wz) and normalised height h are
// (wx * su + 63.5, 128 h - 0.5, wz * su + 63.5) with su = 128 / weather_scale.
// ---------------------------------------------------------------------------
struct DeltaTap {
    uint32_t x0, x1, y0, y1, z0, z1;  // element offsets per axis (summed)
    float fx, fy, fz;
};

__device__ __forceinline__ DeltaTap delta_tap(float cx, float cy, float cz) {
    const float fx0 = floorf(cx), fy0 = floorf(cy), fz0 = floorf(cz);
    DeltaTap d;
    d.fx = cx - fx0; d.fy = cy - fy0; d.fz = cz - fz0;
    const int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
    d.x0 = (uint32_t)ix & 127u; d.x1 = (uint32_t)(ix + 1) & 127u;
    d.y0 = (uint32_t)clampi(iy, 128) << 7; d.y1 = (uint32_t)clampi(iy + 1, 128) << 7;
    d.z0 = ((uint32_t)iz & 127u) << 14; d.z1 = ((uint32_t)(iz + 1) & 127u) << 14;
    return d;
}

// Bounds (max, min) of the perturbation around a tap: over the 4^3 brick
// holding its lower corner and the brick's +1 halo (all eight corners).
__device__ __forceinline__ float2 delta_bound(const RenderParams& P, float cx, float cy, float cz) {
    const uint32_t bx = ((uint32_t)(int)floorf(cx) & 127u) >> 2, bz = ((uint32_t)(int)floorf(cz) & 127u) >> 2;
    const uint32_t by = (uint32_t)clampi((int)floorf(cy), 128) >> 2;
    return ld_g<float2>(P.delta_bnd, ((bz << 10) | (by << 5) | bx) * 8u);
}

// Maximum of the perturbation over the whole volume. Written before the
// launch and constant during it: read through the constant address space,
// which makes it one scalar load the compiler may hoist.
__device__ __forceinline__ float delta_gmax(const RenderParams& P) {
    return ((const __attribute__((address_space(4))) float2*)(P.delta_bnd))[32 * 32 * 32].x;
}

__device__ __forceinline__ float delta_fetch(const float* __restrict__ v, const DeltaTap& d) {
    const float c000 = ld_g<float>(v, (d.z0 + d.y0 + d.x0) * 4u), c100 = ld_g<float>(v, (d.z0 + d.y0 + d.x1) * 4u);
    const float c010 = ld_g<float>(v, (d.z0 + d.y1 + d.x0) * 4u), c110 = ld_g<float>(v, (d.z0 + d.y1 + d.x1) * 4u);
    const float c001 = ld_g<float>(v, (d.z1 + d.y0 + d.x0) * 4u), c101 = ld_g<float>(v, (d.z1 + d.y0 + d.x1) * 4u);
    const float c011 = ld_g<float>(v, (d.z1 + d.y1 + d.x0) * 4u), c111 = ld_g<float>(v, (d.z1 + d.y1 + d.x1) * 4u);
    const float a0 = c000 + (c100 - c000) * d.fx, a1 = c010 + (c110 - c010) * d.fx;
    const float b0 = c001 + (c101 - c001) * d.fx, b1 = c011 + (c111 - c011) * d.fx;
    const float p0 = a0 + (a1 - a0) * d.fy, p1 = b0 + (b1 - b0) * d.fy;
    return p0 + (p1 - p0) * d.fz;
}

// Perturbation at a tap whose brick bounds are b: a brick whose eight-corner
// range is one value needs no fetch (the trilinear blend of equal corners is
// that value exactly).
__device__ __forceinline__ float delta_value(const RenderParams& P, float2 b, float cx, float cy, float cz) {
    return b.x == b.y ? b.x : delta_fetch(P.delta, delta_tap(cx, cy, cz));
}

__device__ __forceinline__ void delta_coords(const RenderParams& P, float wx, float wz, float h, float& cx, float& cy,
                                             float& cz) {
    cx = wx * P.delta_su + 63.5f;
    cy = h * 128.f - 0.5f;
    cz = wz * P.delta_su + 63.5f;
}

// Brick of the empty-space bound that holds the lower corner of a base tap.
__device__ __forceinline__ uint32_t bbound_index(float cx, float cy, float cz) {
    const uint32_t bx = ((uint32_t)(int)floorf(cx) >> PS_BB_LOG2) & (PS_BN - 1);
    const uint32_t by = ((uint32_t)(int)floorf(cy) >> PS_BB_LOG2) & (PS_BN - 1);
    const uint32_t bz = ((uint32_t)(int)floorf(cz) >> PS_BB_LOG2) & (PS_BN - 1);
    return (bz * PS_BN + by) * PS_BN + bx;
}

// False when no tap in the brick can make base_cloud * ndfv exceed thr.
// base_cloud = (R + 1 - lf) / (2 - lf) grows with R and (R <= 1) falls with
// lf = clamp(LF + esb / 2, 0, 1), so over the brick it is at most
// (Rmax + 1 - lm) / (2 - lm), lm = clamp(LFmin + esb / 2, 0, 1). The test is
// multiplied through by 2 - lm >= 1. An absolute 1e-5 on the bound covers the
// rounding of the interpolated R and lf (a few ulp of O(1) values); thr
// itself carries the relative 1e-5 that covers the products.
__device__ __forceinline__ bool brick_may_exceed(uint32_t bb, float hesb, float ndfv, float thr) {
    const float rm = (float)(bb & 0xffu) * (1.f / 255.f);
    const float lm = clampf((float)(bb >> 8) * (1.f / 255.f) + hesb, 0.f, 1.f);
    return ((rm + 1.f) - lm + 1e-5f) * ndfv > thr * (2.f - lm);
}

// Base shape through the coverage remap (shared by sample_density and
// sample_density_light): the pre-erosion cloud value. `cov` >= 0.01 here, so
// remap(c, 1-cov, 1, 0, 1) * cov equals c - (1 - cov) exactly. Returns -1
// without touching the base volume when ndf alone, or the brick bound
// (brick_may_exceed), already keeps the shape <= 0: the density is then 0
// whatever the fetch would return.
template <bool DELTA = false>
__device__ __forceinline__ float shape(const RenderParams& P, const LayerF& L, float cx, float cy, float cz,
                                       float ndfv, float cov, float dcx = 0.f, float dcy = 0.f, float dcz = 0.f) {
    // with the perturbation volume the bound is shape + max(delta) <= 0
    // (an absolute 1e-5 margin: every term is O(1)), with the volume's
    // maximum first and then the brick's (no smaller, so the first test
    // skips a subset of what the second would)
    float thr;
    float2 db = make_float2(0.f, 0.f);
    if constexpr (DELTA) {
        thr = (1.f - cov - delta_gmax(P)) - 1e-5f;
        if (ndfv <= thr) return -1.f;
        db = delta_bound(P, dcx, dcy, dcz);
        thr = (1.f - cov - db.x) - 1e-5f;
    } else {
        thr = (1.f - cov) * (1.f - 1e-5f);
    }
    if (ndfv <= thr) return -1.f;
    if (!brick_may_exceed(ld_g<uint16_t>(P.bbound, bbound_index(cx, cy, cz) * 2u), L.esb * 0.5f, ndfv, thr))
        return -1.f;
    float2 b = base_fetch(P.base_rl, cx, cy, cz, P.base_rs, P.base_ls);
    float lf = clampf(b.y + L.esb * 0.5f, 0.f, 1.f);
    float base_cloud = (b.x + (1.f - lf)) * frcp(fmaxf(2.f - lf, 1e-5f));
    float v = base_cloud * ndfv - (1.f - cov);
    if constexpr (DELTA) v += delta_value(P, db, dcx, dcy, dcz);
    return v;
}

// Advected position of a sample (cloud_march.wgsl:448-458) and its texel
// coordinates in the base volume.
__device__ __forceinline__ void advect(const RenderParams& P, const LayerF& L, float px, float py, float pz, float h,
                                       float& ax, float& ay, float& az, float& cx, float& cy, float& cz) {
    ax = px - L.wind_off_x - L.skew_x * h; ay = py; az = pz - L.wind_off_z - L.skew_z * h;
    const float s = P.inv_base_scale * 128.f;
    cx = ax * s - 0.5f; cy = ay * s - 0.5f; cz = az * s - 0.5f;
}

// cloud_density_common.wgsl:266-305 (with DELTA: plus the perturbation
// volume before the clamp)
template <bool DELTA = false>
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
    float dcx = 0.f, dcy = 0.f, dcz = 0.f;
    if constexpr (DELTA) delta_coords(P, px - L.wind_off_x, pz - L.wind_off_z, h, dcx, dcy, dcz);
    float cloud = shape<DELTA>(P, L, cx, cy, cz, ndf_eval(nc, ctype, h, L.anvil), cov, dcx, dcy, dcz);
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

// Geometry of the march toward the sun (cloud_march.wgsl:543-578 straight,
// 599-666 cone) with the arithmetic of march_light, so sample positions and
// altitudes match the forward bit for bit.
struct LightWalk {
    float dist, k, r0c, step;
    int n;
};

__device__ __forceinline__ LightWalk light_walk(const RenderParams& P, const LayerF& L, float px, float py, float pz,
                                                float alt) {
    const float sx = P.sun[0], sy = P.sun[1], sz = P.sun[2];
    LightWalk w;
    float cos_sun = fmaxf(sy, 0.05f);
    w.dist = fmaxf((L.top - alt) / cos_sun, 1.f);
    float r0 = P.R + alt;
    float qy = py + P.R;
    w.r0c = fmaxf(r0, 1.f);
    float cv = (px * sx + qy * sy + pz * sz) * frcp(w.r0c);
    w.k = 2.f * r0 * cv;
    if (P.cone) {
        w.step = w.dist / 5.f;
        w.n = 6;
    } else {
        w.step = w.dist / (float)P.light_steps;
        w.n = P.light_steps;
    }
    return w;
}

// Light sample i: lookup-frame position (from the sample's lookup position
// q0), altitude (from its world position p) and quadrature weight.
__device__ __forceinline__ void light_sample(const RenderParams& P, const LightWalk& w, int i, float qx0, float qy0,
                                             float qz0, float px, float py, float pz, float alt, float& qx,
                                             float& qyy, float& qz, float& a, float& wt) {
    const float sx = P.sun[0], sy = P.sun[1], sz = P.sun[2];
    if (P.cone) {
        const float off[6][3] = {{0.38051305f, 0.92453449f, -0.02111345f}, {-0.50625799f, -0.03590792f, -0.86163418f},
                                 {-0.32509218f, -0.94557439f, 0.01428793f}, {0.09026238f, -0.27376545f, 0.95755165f},
                                 {0.28128598f, 0.42443639f, -0.86065785f}, {-0.16852403f, 0.14748697f, 0.97460106f}};
        float t, scale;
        if (i < 5) { t = (i + 0.5f) * w.step; scale = w.step * (i + 1.f) / 5.f; wt = w.step; }
        else { t = w.dist * 3.f; scale = w.step * 3.f; wt = w.step * 3.f; }
        const float ox = sx * t + off[i][0] * scale, oy = sy * t + off[i][1] * scale, oz = sz * t + off[i][2] * scale;
        float wx = px + ox, yy = py + oy + P.R, wz = pz + oz;
        float rr = fsqrt(wx * wx + yy * yy + wz * wz);
        a = alt + t * (w.k + t) * frcp(w.r0c + rr);
        qx = qx0 + ox; qyy = qy0 + oy; qz = qz0 + oz;
        return;
    }
    float t = (i + 0.5f) * w.step;
    float wx = px + sx * t, yy = py + sy * t + P.R, wz = pz + sz * t;
    float rr = fsqrt(wx * wx + yy * yy + wz * wz);
    a = alt + t * (w.k + t) * frcp(w.r0c + rr);
    qx = qx0 + sx * t; qyy = qy0 + sy * t; qz = qz0 + sz * t;
    wt = w.step;
}

// The march toward the sun: returns the optical depth and, with PARTIALS,
// d od / d density_scale, d od / d coverage and the linear-range mask of
// light_adj. `cnt` (or null) counts the light samples and covered ones.
template <bool DELTA, bool PARTIALS>
__device__ __forceinline__ float light_march_impl(const RenderParams& P, const LayerF& L, const NdfC& nc, float qx0,
                                                  float qy0, float qz0, float px, float py, float pz, float alt,
                                                  float wr, uint32_t ctype, Cnt* cnt, float& dod_ds, float& dod_cov,
                                                  uint32_t& lin) {
    const LightWalk w = light_walk(P, L, px, py, pz, alt);
    const float wc = wr * L.coverage;
    const float cov = clampf(wc, 0.f, 1.f);
    const float dcov = (wc > 0.f && wc < 1.f) ? wr : 0.f;
    float od = 0.f;
    dod_ds = 0.f;
    dod_cov = 0.f;
    lin = 0u;
    for (int i = 0; i < w.n; ++i) {
        float qx, qyy, qz, a, wt;
        light_sample(P, w, i, qx0, qy0, qz0, px, py, pz, alt, qx, qyy, qz, a, wt);
        const float h = (a - L.base) * L.inv_thick;
        float d = 0.f;
        if (cnt) cnt->light++;
        if (!(h < 0.f || h > 1.f) && cov >= 0.01f) {
            if (cnt) cnt->light_cov++;
            float ax, ay, az, cx, cy, cz;
            advect(P, L, qx, qyy, qz, h, ax, ay, az, cx, cy, cz);
            float dcx = 0.f, dcy = 0.f, dcz = 0.f;
            if constexpr (DELTA) delta_coords(P, qx - L.wind_off_x, qz - L.wind_off_z, h, dcx, dcy, dcz);
            const float x = shape<DELTA>(P, L, cx, cy, cz, ndf_eval(nc, ctype, h, L.anvil), cov, dcx, dcy, dcz);
            const float c = clampf(x, 0.f, 1.f);
            d = c * L.density_scale;
            if constexpr (PARTIALS) {
                dod_ds += c * wt;
                if (x > 0.f && x < 1.f) {
                    dod_cov += L.density_scale * dcov * wt;
                    if (i < 32) lin |= 1u << i;
                }
            }
        }
        od += d * wt;
    }
    return od;
}

// The light march's optical depth (cloud_march.wgsl:543-578 straight,
// 599-666 cone) for the forward, counting its work. (px, py, pz) is the
// sample's world position (its radius and local up), (qx0, qy0, qz0) the same
// point in the segment's lookup frame (the density fetches).
template <bool DELTA = false>
__device__ __forceinline__ float march_light(const RenderParams& P, const LayerF& L, const NdfC& nc, float qx0,
                                             float qy0, float qz0, float px, float py, float pz, float alt,
                                             float weather_r, uint32_t ctype, Cnt& cnt) {
    float a, b;
    uint32_t c;
    return light_march_impl<DELTA, false>(P, L, nc, qx0, qy0, qz0, px, py, pz, alt, weather_r, ctype, &cnt, a, b, c);
}

// The same optical depth (the same arithmetic, so the training forward, the
// plain forward and the backward agree bit for bit) with its partials in
// density_scale and coverage; bit i of `lin` is set when light sample i
// sits inside the density clamp's linear range (0 < x < 1), where the
// perturbation volume receives d od / d delta = density_scale * weight.
template <bool DELTA>
__device__ __forceinline__ float light_adj(const RenderParams& P, const LayerF& L, const NdfC& nc, float qx0,
                                           float qy0, float qz0, float px, float py, float pz, float alt, float wr,
                                           uint32_t ctype, float& dod_ds, float& dod_cov, uint32_t& lin) {
    return light_march_impl<DELTA, true>(P, L, nc, qx0, qy0, qz0, px, py, pz, alt, wr, ctype, nullptr, dod_ds, dod_cov,
                                         lin);
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


// ---------------------------------------------------------------------------
// Per-pixel setup shared by the forward render and its adjoint.
// ---------------------------------------------------------------------------
struct PixelRay {
    double fx, fy, ndx, ndy;          // pixel centre and its NDC
    double dxd, dyd, dzd;             // unit view ray (FP64)
    float dx, dy, dz;
    double cos_vs;                    // cos(view, sun) against the sun direction as uploaded
    float omcf;                       // 1 - cos_vs from FP64 (keeps the forward HG peak exact)
    double b;                         // (cam - planet centre) . ray
};

// FP64 view ray (math.wgsl:23-33 / sky.wgsl:166-168).
__device__ __forceinline__ PixelRay pixel_ray(const RenderParams& P, int xc, int yc) {
    PixelRay r;
    // viewport coordinates of the pixel centre (x2 for a half-resolution RT)
    r.fx = (xc + 0.5) * P.rt_sx; r.fy = (yc + 0.5) * P.rt_sy;
    r.ndx = r.fx / P.full_W * 2.0 - 1.0; r.ndy = 1.0 - r.fy / P.full_H * 2.0;
    const double* m = P.inv_vp;
    const double ndx = r.ndx, ndy = r.ndy;
    double nh0 = m[0] * ndx + m[4] * ndy + m[8] + m[12];
    double nh1 = m[1] * ndx + m[5] * ndy + m[9] + m[13];
    double nh2 = m[2] * ndx + m[6] * ndy + m[10] + m[14];
    double nh3 = m[3] * ndx + m[7] * ndy + m[11] + m[15];
    double dxd = nh0 / nh3 - P.cam[0], dyd = nh1 / nh3 - P.cam[1], dzd = nh2 / nh3 - P.cam[2];
    double il = 1.0 / sqrt(dxd * dxd + dyd * dyd + dzd * dzd);
    dxd *= il; dyd *= il; dzd *= il;
    r.dxd = dxd; r.dyd = dyd; r.dzd = dzd;
    r.dx = (float)dxd; r.dy = (float)dyd; r.dz = (float)dzd;
    // cos(theta) against the sun direction as uploaded (not renormalised, as
    // the reference uses it)
    r.cos_vs = dxd * P.sun_d[0] + dyd * P.sun_d[1] + dzd * P.sun_d[2];
    r.omcf = (float)(1.0 - r.cos_vs);
    r.b = P.cam[0] * dxd + (P.cam[1] + P.R) * dyd + P.cam[2] * dzd;
    return r;
}

// ---------------------------------------------------------------------------
// One layer segment of a pixel's primary march (cloud_march.wgsl:808-1000),
// restated relative to the segment. Sample s sits at ray distance t = t0 +
// s step, t0 = ta + jitter step. Its density lookups use o + d (s step),
// where o is the first sample's position minus whole common periods of every
// periodic lookup (the wrap_* periods): the same texels as the world
// position, with coordinates that stay below a few hundred km. Its altitude
// is h_e + q / (r_e + sqrt(r_e^2 + q)), q = tl (k_e + tl), tl = s step, from
// the first sample's altitude h_e, radius r_e = R + h_e and k_e = 2 (p_e - c).d
// (|p - c|^2 = r_e^2 + tl k_e + tl^2), the reference's altitude_from_entry
// (cloud_march.wgsl:360-365) anchored at the segment instead of the camera.
// With world positions from the camera, FP32 rounding at a far-side shell
// (t ~ 1.3e4 km, rays through the planet when no depth buffer clips them)
// is ~1 m in both the lookups and the altitude; here it is ~1 cm.
// ---------------------------------------------------------------------------
struct SegGeom {
    float ox, oy, oz;                 // lookup-frame position of the first sample
    float he, ke;                     // its altitude and 2 (p - c).d (radius r_e = R + h_e)
    float t0;                         // its ray distance
    float step;
};

__device__ __forceinline__ double wrap_period(double v, double period) {
    return period > 0.0 ? v - period * rint(v / period) : v;
}

// Segment geometry of layer LC for a ray (FP64). False when the ray misses
// the layer's interval.
__device__ __forceinline__ bool seg_geom(const RenderParams& P, const LayerC& LC, const PixelRay& ray,
                                         double depth_t_max, float jitter, SegGeom& g) {
    double ta, tb;
    if (!layer_interval(LC, P.r0, ray.b, depth_t_max, ta, tb)) return false;
    const double stepd = (tb - ta) / P.steps;
    const double t0 = ta + (double)jitter * stepd;
    const double r2 = P.r0 * P.r0 + t0 * (2.0 * ray.b + t0);
    const double re = sqrt(fmax(r2, 0.0));
    g.he = (float)(re - (double)P.R);
    g.ke = (float)(2.0 * (ray.b + t0));
    g.ox = (float)wrap_period(P.cam[0] + ray.dxd * t0, P.wrap_xz);
    g.oy = (float)wrap_period(P.cam[1] + ray.dyd * t0, P.wrap_y);
    g.oz = (float)wrap_period(P.cam[2] + ray.dzd * t0, P.wrap_xz);
    g.t0 = (float)t0;
    g.step = (float)stepd;
    return true;
}

// Sample at tl = s step of a segment: lookup position, altitude. The radius
// enters only the denominator, where its FP32 rounding (R + h_e, 0.5 m) is a
// relative 1e-7.
__device__ __forceinline__ void seg_sample(const RenderParams& P, const SegGeom& g, float dx, float dy, float dz,
                                           float tl, float& qx, float& qy, float& qz, float& alt) {
    qx = g.ox + dx * tl; qy = g.oy + dy * tl; qz = g.oz + dz * tl;
    const float qq = tl * (g.ke + tl);
    const float re = P.R + g.he;
    alt = g.he + qq * frcp(re + fsqrt(fmaf(re, re, qq)));
}

// Sky pass (sky.wgsl:159-220): sky-view LUT, overcast blend, sun disk.
__device__ __forceinline__ float3 sky_pass(const RenderParams& P, const PixelRay& ray) {
    const float dx = ray.dx, dy = ray.dy, dz = ray.dz;
    const double cos_vs = ray.cos_vs;
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

    return sky;
}

// Blue-noise jitter (cloud_march.wgsl:745-756).
__device__ __forceinline__ float pixel_jitter(const RenderParams& P, int xc, int yc) {
    int jx = xc & 63, jy = yc & 63;
    if (P.temporal_jitter) {
        uint32_t ph = P.frame_index & 15u;
        jx = (jx ^ (int)((ph * 11u + 7u) & 63u)) & 63;
        jy = (jy ^ (int)((ph * 23u + 13u) & 63u)) & 63;
    }
    return (float)P.blue[jy * 64 + jx] * (1.f / 255.f);
}

// Depth clip (cloud_march.wgsl:764-775): view distance of the opaque surface.
__device__ __forceinline__ double pixel_depth_tmax(const RenderParams& P, const PixelRay& ray, int xc, int yc) {
    double depth_t_max = 1e30;
    if (P.has_depth) {
        const double* m = P.inv_vp;
        const double ndx = ray.ndx, ndy = ray.ndy;
        // the full-resolution depth at the RT pixel's scaled coordinates
        const int dx = (int)(xc * P.rt_sx), dy = (int)(yc * P.rt_sy);
        float dn = P.depth[(size_t)dy * P.full_W + dx];
        if (dn > 0.f) {
            double w0 = m[0] * ndx + m[4] * ndy + m[8] * dn + m[12];
            double w1 = m[1] * ndx + m[5] * ndy + m[9] * dn + m[13];
            double w2 = m[2] * ndx + m[6] * ndy + m[10] * dn + m[14];
            double w3 = m[3] * ndx + m[7] * ndy + m[11] * dn + m[15];
            double qx = w0 / w3 - P.cam[0], qy = w1 / w3 - P.cam[1], qz = w2 / w3 - P.cam[2];
            depth_t_max = fmax(sqrt(qx * qx + qy * qy + qz * qz), 0.0);
        }
    }
    return depth_t_max;
}

// Per-layer intervals in FP64 (cloud_march.wgsl:370-416, 777-806). Layers
// hit are sorted by entry distance (stable: equal keys keep layer order).
// Only the order is kept (3 bits per entry); each interval is recomputed
// when the march reaches it, which keeps the per-lane state small.
__device__ __forceinline__ uint32_t layer_order(const RenderParams& P, const PixelRay& ray, double depth_t_max,
                                                int& nseg) {
    uint32_t order = 0;
    nseg = 0;
    double key[PS_MAX_LAYERS];
    const int lc = P.n_layers;
    for (int l = 0; l < lc; ++l) {
        double ta, tb;
        if (!layer_interval(P.layers[l], P.r0, ray.b, depth_t_max, ta, tb)) continue;
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
    return order;
}

}  // namespace
