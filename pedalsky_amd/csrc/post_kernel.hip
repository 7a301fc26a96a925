// Presentation of a rendered frame for gfx950: the ground pass evaluated on
// the depth buffer's surface, aerial perspective over it, the dual-source
// cloud composite, godrays, bloom, and the tone map to the sRGB swapchain the
// reference reads its goldens from.
//
//  * Surface pixels (depth buffer holds a surface): `lit` is the caller's
//    radiance or the ground pass (shaders/ground/pbr.wgsl fs_main, 511-657)
//    on a flat ground (normal +y) at the pixel's surface position; then
//    L = lit * ap.a + ap.rgb, the AP froxel tap at the pixel's viewport uv
//    and the surface distance mapped as atmosphere_lut_sampling.wgsl:33-38
//    (pbr.wgsl:653-657); then hdr = cloud_luminance + L * cloud_transmittance
//    (cloud_composite.wgsl:47-57). Sky pixels keep the render's composite.
//  * Godrays (shaders/godrays/radial.wgsl:57-92, composite.wgsl:39-46):
//    at half resolution, the mean over N samples toward the sun of the
//    bright-passed HDR weighted by decay^i, then added to the HDR image
//    through a bilinear upsample times the intensity.
//  * Bloom (shaders/bloom/*.wgsl, crates/ps-bloom/src/lib.rs:511-797): a
//    soft-knee bright pass into a half-resolution level, 13-tap downsamples
//    to quarter and eighth resolution, 9-tap tent upsamples added back level
//    by level, and the half-resolution level added to the HDR image times the
//    intensity.
//  * Tone map (shaders/postprocess/tonemap.wgsl:37-58): exposure
//    1 / (1.2 * 2^EV100), ACES filmic (Narkowicz fit) or clamp, then the sRGB
//    encode of the *Srgb swapchain, rounded to 8 bits.
// Every image stays FP32 (the reference's are binary16 render targets) and
// every image tap is a bilinear ClampToEdge lookup at texel centres, the
// sampler the reference binds. One thread per output pixel: each pass reads
// and writes a few images once (HBM-bound); the radial blur's N taps per
// pixel converge on the sun and hit in L2.
#include "ps_device.h"
#include "ps_launch.h"

#include "ps_march.h"

namespace {

// Trilinear ClampToEdge tap of the AP LUT (RGBA32F, aw x ah x ad) at uvw.
__device__ __forceinline__ float4 ap_tap(const RenderParams& P, float u, float v, float w) {
    const float cx = u * P.aw - 0.5f, cy = v * P.ah - 0.5f, cz = w * P.ad - 0.5f;
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
    auto tri = [&](float v000, float v100, float v010, float v110, float v001, float v101, float v011,
                   float v111) {
        const float a0 = v000 + (v100 - v000) * wx, a1 = v010 + (v110 - v010) * wx;
        const float b0 = v001 + (v101 - v001) * wx, b1 = v011 + (v111 - v011) * wx;
        const float p0 = a0 + (a1 - a0) * wy, p1 = b0 + (b1 - b0) * wy;
        return p0 + (p1 - p0) * wz;
    };
    return make_float4(tri(c000.x, c100.x, c010.x, c110.x, c001.x, c101.x, c011.x, c111.x),
                       tri(c000.y, c100.y, c010.y, c110.y, c001.y, c101.y, c011.y, c111.y),
                       tri(c000.z, c100.z, c010.z, c110.z, c001.z, c101.z, c011.z, c111.z),
                       tri(c000.w, c100.w, c010.w, c110.w, c001.w, c101.w, c011.w, c111.w));
}

// ACES filmic, Narkowicz's fit (tonemap.wgsl:37-44), saturated.
__device__ __forceinline__ float aces(float x) {
    return fminf(fmaxf((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.f), 1.f);
}

// Linear -> sRGB 8-bit (the encode an Rgba8UnormSrgb target applies on
// store), rounded to nearest.
__device__ __forceinline__ uint32_t srgb8(float c) {
    c = fminf(fmaxf(c, 0.f), 1.f);
    const float e = c <= 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.f / 2.4f) - 0.055f;
    return (uint32_t)fminf(floorf(e * 255.f + 0.5f), 255.f);
}

// Linear -> 8-bit UNORM as an Rgba8Unorm target stores it, rounded to
// nearest.
__device__ __forceinline__ uint32_t unorm8(float c) {
    return (uint32_t)floorf(fminf(fmaxf(c, 0.f), 1.f) * 255.f + 0.5f);
}

// mode bit 0: clamp instead of ACES; bit 1: Rgba8Unorm target instead of
// the sRGB encode.
__device__ __forceinline__ uint32_t tonemap_rgba8(float4 hdr, float exposure, int mode) {
    float r = hdr.x * exposure, g = hdr.y * exposure, b = hdr.z * exposure;
    if ((mode & 1) == 0) {
        r = aces(r); g = aces(g); b = aces(b);
    }
    if (mode & 2) return unorm8(r) | (unorm8(g) << 8) | (unorm8(b) << 16) | (255u << 24);
    return srgb8(r) | (srgb8(g) << 8) | (srgb8(b) << 16) | (255u << 24);
}

// ---------------------------------------------------------------------------
// The ground pass (shaders/ground/pbr.wgsl)
// ---------------------------------------------------------------------------
constexpr float kGrPi = 3.14159265358979f;

// hash21 (pbr.wgsl:94-99): i32 coordinates reinterpreted as u32.
__device__ __forceinline__ uint32_t gr_hash21(int x, int y) {
    uint32_t h = (uint32_t)x * 0x27d4eb2du + (uint32_t)y * 0x165667b1u;
    h = (h ^ (h >> 15)) * 0x2c1b3c6du;
    h = (h ^ (h >> 12)) * 0x297a2d39u;
    return h ^ (h >> 15);
}

// Cell hash of the 5 m Voronoi tiling at ground position (x, z)
// (pbr.wgsl:101-131). FP32 with no contraction and correctly rounded
// divisions, so the nearest feature point, a discrete choice, is the one the
// FP32 statement of the shader picks.
__device__ uint32_t voronoi_cell(float px, float pz) {
#pragma clang fp contract(off)
    const float sx = __fdiv_rn(px, 5.f), sz = __fdiv_rn(pz, 5.f);
    const float flx = floorf(sx), flz = floorf(sz);
    const int cx = (int)flx, cz = (int)flz;
    const float frx = sx - flx, frz = sz - flz;
    float min_d2 = 4.f;
    int bx = 0, bz = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int nx = cx + dx, nz = cz + dy;
            const uint32_t h0 = gr_hash21(nx, nz), h1 = gr_hash21(nx + 7919, nz + 1);
            const float jx = __fdiv_rn((float)(h0 & 0xffffu), 65535.f), jz = __fdiv_rn((float)(h1 & 0xffffu), 65535.f);
            const float ex = ((float)dx + jx) - frx, ez = ((float)dy + jz) - frz;
            const float d2 = ex * ex + ez * ez;
            if (d2 < min_d2) {
                min_d2 = d2;
                bx = nx;
                bz = nz;
            }
        }
    return gr_hash21(bx, bz);
}

// Per-material tables (pbr.wgsl:145-194): 3-entry Voronoi palettes, dry
// roughness, dielectric F0. Index: grass, bare soil, tarmac, sand, water edge.
__constant__ float kPalette[5][3][3] = {
    {{0.10f, 0.20f, 0.06f}, {0.12f, 0.24f, 0.08f}, {0.08f, 0.16f, 0.05f}},
    {{0.30f, 0.22f, 0.14f}, {0.36f, 0.27f, 0.18f}, {0.24f, 0.18f, 0.11f}},
    {{0.05f, 0.05f, 0.06f}, {0.07f, 0.07f, 0.08f}, {0.04f, 0.04f, 0.05f}},
    {{0.55f, 0.48f, 0.36f}, {0.62f, 0.55f, 0.41f}, {0.50f, 0.43f, 0.32f}},
    {{0.08f, 0.18f, 0.28f}, {0.06f, 0.15f, 0.24f}, {0.10f, 0.20f, 0.30f}}};
__constant__ float kRoughness[5] = {0.85f, 0.88f, 0.92f, 0.78f, 0.30f};
__constant__ float kF0[5] = {0.04f, 0.04f, 0.04f, 0.05f, 0.08f};

__device__ __forceinline__ float fade5(float t) { return t * t * t * (t * (t * 6.f - 15.f) + 10.f); }
// WGSL mix(a, b, t) = a (1 - t) + b t
__device__ __forceinline__ float wmix(float a, float b, float t) { return a * (1.f - t) + b * t; }

// value_noise (pbr.wgsl:200-209): quintic-faded lattice of 16-bit hashes.
__device__ float gr_value_noise(float x, float y) {
    const float flx = floorf(x), fly = floorf(y);
    const int cx = (int)flx, cy = (int)fly;
    const float ux = fade5(x - flx), uy = fade5(y - fly);
    const float s = 1.f / 65535.f;
    const float h00 = (float)(gr_hash21(cx, cy) & 0xffffu) * s, h10 = (float)(gr_hash21(cx + 1, cy) & 0xffffu) * s;
    const float h01 = (float)(gr_hash21(cx, cy + 1) & 0xffffu) * s;
    const float h11 = (float)(gr_hash21(cx + 1, cy + 1) & 0xffffu) * s;
    return wmix(wmix(h00, h10, ux), wmix(h01, h11, ux), uy);
}

// Cook-Torrance GGX / Smith-Schlick / Fresnel-Schlick + Lambert
// (pbr.wgsl:215-271); returns the BRDF (not times n.l).
__device__ float3 gr_brdf(float3 n, float3 v, float3 l, float3 albedo, float rough, float f0) {
    float hx = v.x + l.x, hy = v.y + l.y, hz = v.z + l.z;
    const float hl = rsqrtf(hx * hx + hy * hy + hz * hz);
    hx *= hl; hy *= hl; hz *= hl;
    const float ndv = fmaxf(n.x * v.x + n.y * v.y + n.z * v.z, 0.f);
    const float ndl = fmaxf(n.x * l.x + n.y * l.y + n.z * l.z, 0.f);
    const float ndh = fmaxf(n.x * hx + n.y * hy + n.z * hz, 0.f);
    const float vdh = fmaxf(v.x * hx + v.y * hy + v.z * hz, 0.f);
    const float a = rough * rough, a2 = a * a;
    const float den = ndh * ndh * (a2 - 1.f) + 1.f;
    const float d = a2 / (kGrPi * den * den + 1e-6f);
    const float rr = rough + 1.f, k = rr * rr / 8.f;
    const float gv = fmaxf(ndv, 1e-4f), gl = fmaxf(ndl, 1e-4f);
    const float g = (gv / (gv * (1.f - k) + k)) * (gl / (gl * (1.f - k) + k));
    const float x = fminf(fmaxf(1.f - vdh, 0.f), 1.f), x5 = x * x * x * x * x;
    const float f = f0 + (1.f - f0) * x5;
    const float spec = (d * g) * f / fmaxf(4.f * ndv * ndl, 1e-4f);
    const float kd = 1.f - f;
    return make_float3(kd * albedo.x / kGrPi + spec, kd * albedo.y / kGrPi + spec, kd * albedo.z / kGrPi + spec);
}

// halton2 (pbr.wgsl:279-289): radical inverse of i in `base`.
__device__ float gr_halton(uint32_t i, uint32_t base) {
    float f = 1.f, r = 0.f;
    const float fb = (float)base;
    while (i != 0u) {
        f = f / fb;
        r = r + f * (float)(i % base);
        i = i / base;
    }
    return r;
}

// Ripple normal perturbation in the ground plane (pbr.wgsl:295-351): two
// spawns per 1 m cell per second over the 3x3 neighbourhood, damped
// sinusoidal rings expanding at 0.5 m/s for 0.6 s.
__device__ float2 gr_ripples(float px, float pz, float time_s, float intensity) {
    float2 dn = make_float2(0.f, 0.f);
    if (intensity <= 0.f) return dn;
    const int cx = (int)floorf(px), cz = (int)floorf(pz);
    const float t_floor = floorf(time_s);
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int gx = cx + dx, gz = cz + dy;
            const uint32_t h = gr_hash21(gx, gz);
            for (uint32_t spawn = 0; spawn < 2u; ++spawn) {
                const float phase = gr_halton(h ^ (spawn * 0x9e3779b9u), 2u);
                float spawn_t = t_floor + phase;
                if (spawn_t > time_s) spawn_t = spawn_t - 1.f;
                const float age = time_s - spawn_t;
                if (age < 0.f || age > 0.6f) continue;
                const float ox = gr_halton(h, 3u), oz = gr_halton(h ^ 0xa5a5a5a5u, 5u);
                const float tx = px - ((float)gx + ox), tz = pz - ((float)gz + oz);
                const float r = sqrtf(tx * tx + tz * tz);
                if (r < 1e-4f) continue;
                const float dr = r - age * 0.5f;
                if (dr > 0.f || dr < -2.f * 0.5f * 0.6f) continue;
                const float damp = fminf(fmaxf(1.f - age / 0.6f, 0.f), 1.f);
                const float amp = damp * damp * 0.4f * intensity;
                const float dh = sinf(dr * 6.f * 2.f * kGrPi / 0.5f) * amp;
                dn.x += tx / r * dh;
                dn.y += tz / r * dh;
            }
        }
    return dn;
}

// Overcast blocking at ground position (x, z) (pbr.wgsl:393-426): the
// top-down R8 mask over the 32 km square, bilinear ClampToEdge, through
// 1 - exp(-10 mask).
__device__ float gr_overcast_blocking(const RenderParams& P, float x, float z) {
    if (!P.overcast) return 0.f;
    const float cx = (x / 32000.f + 0.5f) * 128.f - 0.5f, cy = (z / 32000.f + 0.5f) * 128.f - 0.5f;
    const float fx0 = floorf(cx), fy0 = floorf(cy);
    const float wx = cx - fx0, wy = cy - fy0;
    const int ix = (int)fx0, iy = (int)fy0;
    const int x0 = clampi(ix, 128), x1 = clampi(ix + 1, 128), y0 = clampi(iy, 128), y1 = clampi(iy + 1, 128);
    const float a = P.overcast[y0 * 128 + x0], b = P.overcast[y0 * 128 + x1];
    const float c = P.overcast[y1 * 128 + x0], d = P.overcast[y1 * 128 + x1];
    const float t0 = a + (b - a) * wx, t1 = c + (d - c) * wx;
    const float mask = (t0 + (t1 - t0) * wy) * (1.f / 255.f);
    return 1.f - __expf(-10.f * mask);
}

// Sky irradiance seen along `dir` from a point at radius cam_r with
// zenith-horizon angle zha (sample_sky_at, pbr.wgsl:478-497): the sky-view
// LUT times the sun illuminance, blended toward the overcast irradiance.
__device__ float3 gr_sky_at(const RenderParams& P, float3 dir, float zha, float blocking) {
    float u, v;
    skyview_uv_f(P, dir.x, dir.y, dir.z, zha, u, v);
    const float4 s = lut2d(P.SV, P.sw, P.sh, u, v);
    const float o = P.overcast_lum;
    return make_float3(wmix(s.x * P.E[0], o, blocking), wmix(s.y * P.E[1], o, blocking),
                       wmix(s.z * P.E[2], o, blocking));
}

// The ground pass's `lit` at world position p (FP64 from the depth buffer),
// viewed along the unit ray (dx, dy, dz) (pbr.wgsl:511-650).
__device__ float3 ground_lit(const RenderParams& P, const PresentArgs& A, double pxd, double pyd, double pzd,
                             float3 ray) {
    const float px = (float)pxd, pz = (float)pzd;
    const float3 n = make_float3(0.f, 1.f, 0.f);
    const float3 v = make_float3(-ray.x, -ray.y, -ray.z);
    const float3 l = make_float3(P.sun[0], P.sun[1], P.sun[2]);
    // material: Voronoi albedo tinted by the world ground albedo, roughness, F0
    const uint32_t m = A.material <= 4u ? A.material : 0u;
    const uint32_t bucket = voronoi_cell(px, pz) % 3u;
    float3 albedo = make_float3(kPalette[m][bucket][0] * A.albedo_scale[0], kPalette[m][bucket][1] * A.albedo_scale[1],
                                kPalette[m][bucket][2] * A.albedo_scale[2]);
    const float wet = fminf(fmaxf(A.wetness, 0.f), 1.f);
    // Lagarde wet darkening and smoothing (pbr.wgsl:358-362)
    const float ex = wmix(1.f, 3.f, wet);
    albedo = make_float3(powf(albedo.x, ex), powf(albedo.y, ex), powf(albedo.z, ex));
    const float rough0 = kRoughness[m];
    const float rough = wmix(rough0, fmaxf(rough0 * 0.5f, 0.05f), wet);
    const float f0 = kF0[m];
    // cloud cover over the point, sun transmittance from the point (with the
    // Earth's shadow), the point's radius and zenith-horizon angle
    const float blocking = gr_overcast_blocking(P, px, pz);
    const double ay = pyd + (double)P.R;
    const double r2 = pxd * pxd + ay * ay + pzd * pzd;
    const double ra = sqrt(r2);
    const float h = (float)((pxd * pxd + pzd * pzd + pyd * (pyd + 2.0 * (double)P.R)) / (ra + (double)P.R));
    const float mu = (float)((pxd * P.sun_d[0] + ay * P.sun_d[1] + pzd * P.sun_d[2]) / fmax(ra, 1.0));
    float3 sun_t = make_float3(0.f, 0.f, 0.f);
    float tu, tv;
    if (trans_uv_from_h(P, h, mu, tu, tv)) {
        const float4 t = lut2d(P.T, P.tw, P.th, tu, tv);
        sun_t = make_float3(t.x, t.y, t.z);
    }
    const float zha = (float)(PS_PI - asin(fmin(fmax((double)P.R / ra, 0.0), 1.0)));
    const float ndl = fmaxf(l.y, 0.f);
    const float through = 1.f - blocking;
    const float3 dE = make_float3(P.E[0] * sun_t.x * ndl * through, P.E[1] * sun_t.y * ndl * through,
                                  P.E[2] * sun_t.z * ndl * through);
    const float3 fd = gr_brdf(n, v, l, albedo, rough, f0);
    const float3 sky = gr_sky_at(P, n, zha, blocking);
    float3 lit = make_float3(fd.x * dE.x + albedo.x / kGrPi * sky.x, fd.y * dE.y + albedo.y / kGrPi * sky.y,
                             fd.z * dE.z + albedo.z / kGrPi * sky.z);
    // puddles (pbr.wgsl:581-618)
    const float cov_gate = fminf(fmaxf(A.puddle_coverage, 0.f), 1.f);
    if (wet > A.puddle_start) {
        const float pt = fminf(fmaxf((wet - A.puddle_start) / fmaxf(1.f - A.puddle_start, 1e-3f), 0.f), 1.f);
        const float pmask = gr_value_noise(px * 0.15f, pz * 0.15f) * cov_gate * pt;   // n.y = 1 >= 0.95
        if (pmask > 0.01f) {
            float3 nw = n;
            if (wet > 0.5f && A.precip > 0.f) {
                const float2 dn = gr_ripples(px, pz, A.sim_seconds, fminf(fmaxf(A.precip / 25.f, 0.f), 1.f));
                const float il = rsqrtf(dn.x * dn.x + 1.f + dn.y * dn.y);
                nw = make_float3(-dn.x * il, il, -dn.y * il);
            }
            const float3 fw = gr_brdf(nw, v, l, make_float3(0.f, 0.f, 0.f), 0.05f, 0.02f);
            // reflect(-v, nw)
            const float dvn = -(v.x * nw.x + v.y * nw.y + v.z * nw.z);
            const float3 rf = make_float3(-v.x - 2.f * dvn * nw.x, -v.y - 2.f * dvn * nw.y, -v.z - 2.f * dvn * nw.z);
            const float3 srefl = gr_sky_at(P, rf, zha, blocking);
            const float nv = fmaxf(nw.x * v.x + nw.y * v.y + nw.z * v.z, 1e-4f);
            const float xv = fminf(fmaxf(1.f - nv, 0.f), 1.f);
            const float fv = 0.02f + 0.98f * (xv * xv * xv * xv * xv);
            lit.x = wmix(lit.x, fw.x * dE.x + fv * srefl.x, pmask);
            lit.y = wmix(lit.y, fw.y * dE.y + fv * srefl.y, pmask);
            lit.z = wmix(lit.z, fw.z * dE.z + fv * srefl.z, pmask);
        }
    }
    // snow (pbr.wgsl:621-642)
    if (A.temperature < 0.5f && A.snow_depth > 0.f) {
        const float inv_puddle = 1.f - gr_value_noise(px * 0.15f, pz * 0.15f) * cov_gate;
        const float amount = fminf(fmaxf(A.snow_depth / 0.05f, 0.f), 1.f) * inv_puddle;
        if (amount > 0.01f) {
            const float3 fs = gr_brdf(n, v, l, make_float3(0.9f, 0.9f, 0.9f), 0.85f, f0);
            const float sa = 0.9f / kGrPi;
            lit.x = wmix(lit.x, fs.x * dE.x + sa * sky.x, amount);
            lit.y = wmix(lit.y, fs.y * dE.y + sa * sky.y, amount);
            lit.z = wmix(lit.z, fs.z * dE.z + sa * sky.z, amount);
        }
    }
    return lit;
}

__global__ __launch_bounds__(256) void k_present(const RenderParams* __restrict__ Pp, PresentArgs A) {
    const RenderParams& P = *Pp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.W * P.H) return;
    const int x = i % P.W, y = i / P.W;
    float4 hdr = A.composite[i];
    if ((A.ground || A.pbr) && P.has_depth && P.depth[i] > 0.f) {
        const PixelRay ray = pixel_ray(P, x, y);
        const double dd = pixel_depth_tmax(P, ray, x, y);
        float3 g;
        if (A.pbr) {
            g = ground_lit(P, A, P.cam[0] + ray.dxd * dd, P.cam[1] + ray.dyd * dd, P.cam[2] + ray.dzd * dd,
                           make_float3(ray.dx, ray.dy, ray.dz));
        } else {
            const float4 gl = A.ground[i];
            g = make_float3(gl.x, gl.y, gl.z);
        }
        const float ds = fmaxf((float)dd, 50.f);
        const float apz = clampf(__logf(ds / 50.f) * (float)(1.0 / 7.6009024595420822), 0.f, 1.f);
        const float4 ap = ap_tap(P, ((float)x + 0.5f) * P.inv_W, ((float)y + 0.5f) * P.inv_H, apz);
        const float4 lum = A.luminance[i], tr = A.transmittance[i];
        hdr.x = lum.x + (g.x * ap.w + ap.x) * tr.x;
        hdr.y = lum.y + (g.y * ap.w + ap.y) * tr.y;
        hdr.z = lum.z + (g.z * ap.w + ap.z) * tr.z;
        hdr.w = 1.f;
    }
    if (A.hdr) A.hdr[i] = hdr;
    if (A.ldr) A.ldr[i] = tonemap_rgba8(hdr, A.exposure, A.mode);
}

// ---------------------------------------------------------------------------
// Godrays and bloom. `img_tap`: bilinear ClampToEdge of a w x h image at uv.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float4 img_tap(const float4* __restrict__ img, int w, int h, float u, float v) {
    return lut2d(img, w, h, u, v);
}

// Radial accumulation (radial.wgsl:57-92) at half-resolution pixel i.
__global__ __launch_bounds__(256) void k_godrays_radial(const float4* __restrict__ hdr, int W, int H,
                                                        float4* __restrict__ rays, int rw, int rh, float sun_u,
                                                        float sun_v, int n, float decay, float threshold) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rw * rh) return;
    const float u0 = ((float)(i % rw) + 0.5f) / (float)rw, v0 = ((float)(i / rw) + 0.5f) / (float)rh;
    const float du = (sun_u - u0) / (float)n, dv = (sun_v - v0) / (float)n;
    const float inv_t = 1.f / fmaxf(threshold, 1.f);
    float ar = 0.f, ag = 0.f, ab = 0.f, wgt = 1.f;
    for (int s = 0; s < n; ++s) {
        const float4 c = img_tap(hdr, W, H, fmaf((float)s, du, u0), fmaf((float)s, dv, v0));
        // bright_pass (radial.wgsl:51-55)
        const float lum = 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
        const float mask = fminf(fmaxf((lum - threshold) * inv_t, 0.f), 1.f) * wgt;
        ar += c.x * mask; ag += c.y * mask; ab += c.z * mask;
        wgt *= decay;
    }
    const float inv_n = 1.f / (float)n;
    rays[i] = make_float4(ar * inv_n, ag * inv_n, ab * inv_n, 1.f);
}

// hdr += rays(uv) * intensity (godrays composite.wgsl:39-46, additive blend).
__global__ __launch_bounds__(256) void k_godrays_composite(float4* __restrict__ hdr, int W, int H,
                                                           const float4* __restrict__ rays, int rw, int rh,
                                                           float intensity) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    const float4 r = img_tap(rays, rw, rh, ((float)(i % W) + 0.5f) / (float)W, ((float)(i / W) + 0.5f) / (float)H);
    float4 c = hdr[i];
    c.x += r.x * intensity; c.y += r.y * intensity; c.z += r.z * intensity;
    hdr[i] = c;
}

// Soft-knee bright pass into the half-resolution level (bright_pass.wgsl:29-45).
__global__ __launch_bounds__(256) void k_bloom_bright(const float4* __restrict__ hdr, int W, int H,
                                                      float4* __restrict__ dst, int dw, int dh, float threshold,
                                                      float knee) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dw * dh) return;
    const float4 c = img_tap(hdr, W, H, ((float)(i % dw) + 0.5f) / (float)dw, ((float)(i / dw) + 0.5f) / (float)dh);
    const float lum = 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
    const float lo = fmaxf(threshold - knee, 1e-6f);
    const float t = fminf(fmaxf((lum - lo) / (threshold - lo), 0.f), 1.f);
    const float wgt = t * t * (3.f - 2.f * t);
    dst[i] = make_float4(c.x * wgt, c.y * wgt, c.z * wgt, wgt);
}

// 13-tap downsample of src (sw x sh) into dst (downsample.wgsl:30-54).
__global__ __launch_bounds__(256) void k_bloom_down(const float4* __restrict__ src, int sw, int sh,
                                                    float4* __restrict__ dst, int dw, int dh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dw * dh) return;
    const float u = ((float)(i % dw) + 0.5f) / (float)dw, v = ((float)(i / dw) + 0.5f) / (float)dh;
    const float tx = 1.f / (float)sw, ty = 1.f / (float)sh;
    auto tap = [&](float ox, float oy) { return img_tap(src, sw, sh, u + ox * tx, v + oy * ty); };
    const float4 ce = tap(0.f, 0.f);
    const float4 a = tap(-1.f, -1.f), b = tap(1.f, -1.f), c = tap(-1.f, 1.f), d = tap(1.f, 1.f);
    const float4 e = tap(-2.f, 0.f), f = tap(2.f, 0.f), g = tap(0.f, -2.f), h = tap(0.f, 2.f);
    const float4 j = tap(-2.f, -2.f), k = tap(2.f, -2.f), l = tap(-2.f, 2.f), m = tap(2.f, 2.f);
    const float wi = 0.5f / 4.f, wo = 0.125f / 4.f;
    float4 r;
    r.x = ce.x * 0.125f + (a.x + b.x + c.x + d.x) * wi + (e.x + f.x + g.x + h.x) * wo + (j.x + k.x + l.x + m.x) * wo;
    r.y = ce.y * 0.125f + (a.y + b.y + c.y + d.y) * wi + (e.y + f.y + g.y + h.y) * wo + (j.y + k.y + l.y + m.y) * wo;
    r.z = ce.z * 0.125f + (a.z + b.z + c.z + d.z) * wi + (e.z + f.z + g.z + h.z) * wo + (j.z + k.z + l.z + m.z) * wo;
    r.w = 1.f;
    dst[i] = r;
}

// dst (dw x dh) += 9-tap tent upsample of the smaller src (upsample.wgsl:
// 29-50, additive blend; scale 1).
__global__ __launch_bounds__(256) void k_bloom_up(const float4* __restrict__ src, int sw, int sh,
                                                  float4* __restrict__ dst, int dw, int dh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dw * dh) return;
    const float u = ((float)(i % dw) + 0.5f) / (float)dw, v = ((float)(i / dw) + 0.5f) / (float)dh;
    const float tx = 1.f / (float)sw, ty = 1.f / (float)sh;
    auto tap = [&](float ox, float oy) { return img_tap(src, sw, sh, u + ox * tx, v + oy * ty); };
    const float4 a = tap(-1.f, -1.f), b = tap(0.f, -1.f), c = tap(1.f, -1.f), d = tap(-1.f, 0.f), e = tap(0.f, 0.f);
    const float4 f = tap(1.f, 0.f), g = tap(-1.f, 1.f), h = tap(0.f, 1.f), k = tap(1.f, 1.f);
    float4 o = dst[i];
    o.x += (a.x + c.x + g.x + k.x) * (1.f / 16.f) + (b.x + d.x + f.x + h.x) * (2.f / 16.f) + e.x * (4.f / 16.f);
    o.y += (a.y + c.y + g.y + k.y) * (1.f / 16.f) + (b.y + d.y + f.y + h.y) * (2.f / 16.f) + e.y * (4.f / 16.f);
    o.z += (a.z + c.z + g.z + k.z) * (1.f / 16.f) + (b.z + d.z + f.z + h.z) * (2.f / 16.f) + e.z * (4.f / 16.f);
    o.w += 1.f;
    dst[i] = o;
}

// Last pass: hdr + bloom(uv) * intensity (bloom composite.wgsl:29-34), the
// HDR copy, the tone map.
__global__ __launch_bounds__(256) void k_post_final(PostArgs A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.W * A.H) return;
    float4 c = A.hdr[i];
    if (A.bloom) {
        const float4 b = img_tap(A.levels[0], A.lw[0], A.lh[0], ((float)(i % A.W) + 0.5f) / (float)A.W,
                                 ((float)(i / A.W) + 0.5f) / (float)A.H);
        c.x += b.x * A.bloom_intensity; c.y += b.y * A.bloom_intensity; c.z += b.z * A.bloom_intensity;
        c.w += 1.f;
    }
    if (A.out_hdr) A.out_hdr[i] = c;
    if (A.ldr) A.ldr[i] = tonemap_rgba8(c, A.exposure, A.mode);
}

// Auto-exposure statistics (auto_exposure.comp.wgsl:31-73): 256 threads,
// thread (tx, ty) sums log2 luminance over the pixels (4 tx + 64 a, 4 ty + 64 b)
// row by row; the partial sums then reduce over the block.
__global__ __launch_bounds__(256) void k_exposure_stats(const float4* __restrict__ hdr, int W, int H,
                                                        float* __restrict__ out) {
    __shared__ float ps[256], pc[256];
    const int t = threadIdx.x, tx = t % 16, ty = t / 16;
    float sum = 0.f, cnt = 0.f;
    for (int y = ty * 4; y < H; y += 64)
        for (int x = tx * 4; x < W; x += 64) {
            const float4 c = hdr[(size_t)y * W + x];
            sum += __log2f(fmaxf(0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z, 1e-6f));
            cnt += 1.f;
        }
    ps[t] = sum;
    pc[t] = cnt;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            ps[t] += ps[t + s];
            pc[t] += pc[t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = ps[0];
        out[1] = pc[0];
    }
}

inline dim3 grid_for(int n) { return dim3((n + 255) / 256); }

}  // namespace

hipError_t ps_launch_present(const RenderParams* d_params, int W, int H, const PresentArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_present, grid_for(W * H), dim3(256), 0, s, d_params, a);
    return hipGetLastError();
}

hipError_t ps_launch_post(const PostArgs& a, hipStream_t s) {
    if (a.godrays) {
        hipLaunchKernelGGL(k_godrays_radial, grid_for(a.rw * a.rh), dim3(256), 0, s, a.hdr, a.W, a.H, a.rays, a.rw,
                           a.rh, a.sun_u, a.sun_v, a.samples, a.decay, a.bright_threshold);
        hipLaunchKernelGGL(k_godrays_composite, grid_for(a.W * a.H), dim3(256), 0, s, a.hdr, a.W, a.H, a.rays, a.rw,
                           a.rh, a.rays_intensity);
    }
    if (a.bloom) {
        hipLaunchKernelGGL(k_bloom_bright, grid_for(a.lw[0] * a.lh[0]), dim3(256), 0, s, a.hdr, a.W, a.H,
                           a.levels[0], a.lw[0], a.lh[0], a.bloom_threshold, a.bloom_knee);
        for (int l = 1; l < 3; ++l)
            hipLaunchKernelGGL(k_bloom_down, grid_for(a.lw[l] * a.lh[l]), dim3(256), 0, s, a.levels[l - 1],
                               a.lw[l - 1], a.lh[l - 1], a.levels[l], a.lw[l], a.lh[l]);
        for (int l = 2; l >= 1; --l)
            hipLaunchKernelGGL(k_bloom_up, grid_for(a.lw[l - 1] * a.lh[l - 1]), dim3(256), 0, s, a.levels[l],
                               a.lw[l], a.lh[l], a.levels[l - 1], a.lw[l - 1], a.lh[l - 1]);
    }
    hipLaunchKernelGGL(k_post_final, grid_for(a.W * a.H), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t ps_launch_exposure_stats(const float4* hdr, int W, int H, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_exposure_stats, dim3(1), dim3(256), 0, s, hdr, W, H, out);
    return hipGetLastError();
}
