// Cloud noise inputs for gfx950: the tileable Perlin-Worley / Worley-FBM /
// curl bakes of shaders/clouds/noise_{base,detail,curl}.comp.wgsl (integer
// hash exact, float work in FP32 as in the WGSL), stored as UNORM8 or FP16,
// plus the fold kernels that turn the stored volumes into the layouts the
// march reads. A seed is mixed into the hash as seed * 0x85ebca6b; seed 0 is
// the reference's fixed hash.
#include "ps_device.h"
#include "ps_launch.h"

namespace {

__device__ __forceinline__ uint32_t hash3(uint32_t x0, uint32_t y0, uint32_t z0, uint32_t seed) {
    uint32_t x = x0 * 0x27d4eb2du + y0 * 0x165667b1u + z0 * 0x9e3779b9u + seed * 0x85ebca6bu;
    x = (x ^ (x >> 15u)) * 0x2c1b3c6du;
    x = (x ^ (x >> 12u)) * 0x297a2d39u;
    return x ^ (x >> 15u);
}
__device__ __forceinline__ uint32_t hash2(uint32_t x0, uint32_t y0, uint32_t seed) {
    uint32_t x = x0 * 0x27d4eb2du + y0 * 0x165667b1u + seed * 0x85ebca6bu;
    x = (x ^ (x >> 15u)) * 0x2c1b3c6du;
    x = (x ^ (x >> 12u)) * 0x297a2d39u;
    return x ^ (x >> 15u);
}
__device__ __forceinline__ float u16f(uint32_t h) { return (float)(h & 0xffffu) / 65535.f; }
__device__ __forceinline__ int wrapf(int c, int f) { return ((c % f) + f) % f; }

// noise_base.comp.wgsl:45-75 / noise_detail.comp.wgsl:170-199
__device__ float worley_tile(float px, float py, float pz, int freq, uint32_t ox, uint32_t oy, uint32_t oz,
                             uint32_t seed) {
    float f = (float)freq;
    float sx = px * f, sy = py * f, sz = pz * f;
    float flx = floorf(sx), fly = floorf(sy), flz = floorf(sz);
    int cx = (int)flx, cy = (int)fly, cz = (int)flz;
    float frx = sx - flx, fry = sy - fly, frz = sz - flz;
    float min_d = 1.f;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                uint32_t wx = (uint32_t)wrapf(cx + dx, freq) + ox;
                uint32_t wy = (uint32_t)wrapf(cy + dy, freq) + oy;
                uint32_t wz = (uint32_t)wrapf(cz + dz, freq) + oz;
                float jx = u16f(hash3(wx, wy, wz, seed)), jy = u16f(hash3(wx + 1u, wy, wz, seed)),
                      jz = u16f(hash3(wx, wy + 1u, wz, seed));
                float ax = (float)dx + jx - frx, ay = (float)dy + jy - fry, az = (float)dz + jz - frz;
                float d = sqrtf(ax * ax + ay * ay + az * az);
                min_d = fminf(min_d, d);
            }
    return clampf(1.f - min_d, 0.f, 1.f);
}
__device__ float worley_fbm(float px, float py, float pz, int freq, uint32_t ox, uint32_t oy, uint32_t oz,
                            uint32_t seed) {
    return worley_tile(px, py, pz, freq, ox, oy, oz, seed) * 0.625f +
           worley_tile(px, py, pz, freq * 2, ox, oy, oz, seed) * 0.25f +
           worley_tile(px, py, pz, freq * 4, ox, oy, oz, seed) * 0.125f;
}
__device__ __forceinline__ float fade(float t) { return t * t * t * (t * (t * 6.f - 15.f) + 10.f); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a + (b - a) * t; }

// noise_base.comp.wgsl:81-120
__device__ float perlin_tile(float px, float py, float pz, int freq, uint32_t seed) {
    float f = (float)freq;
    float sx = px * f, sy = py * f, sz = pz * f;
    float flx = floorf(sx), fly = floorf(sy), flz = floorf(sz);
    int cx = (int)flx, cy = (int)fly, cz = (int)flz;
    float fx = sx - flx, fy = sy - fly, fz = sz - flz;
    float ux = fade(fx), uy = fade(fy), uz = fade(fz);
    float dots[8];
    for (int iz = 0; iz < 2; ++iz)
        for (int iy = 0; iy < 2; ++iy)
            for (int ix = 0; ix < 2; ++ix) {
                uint32_t wx = (uint32_t)wrapf(cx + ix, freq), wy = (uint32_t)wrapf(cy + iy, freq),
                         wz = (uint32_t)wrapf(cz + iz, freq);
                float gx = u16f(hash3(wx, wy, wz, seed)) * 2.f - 1.f;
                float gy = u16f(hash3(wx + 1u, wy, wz, seed)) * 2.f - 1.f;
                float gz = u16f(hash3(wx, wy + 1u, wz, seed)) * 2.f - 1.f;
                float il = 1.f / sqrtf(gx * gx + gy * gy + gz * gz);
                dots[ix + iy * 2 + iz * 4] = gx * il * (fx - ix) + gy * il * (fy - iy) + gz * il * (fz - iz);
            }
    float x00 = mixf(dots[0], dots[1], ux), x10 = mixf(dots[2], dots[3], ux);
    float x01 = mixf(dots[4], dots[5], ux), x11 = mixf(dots[6], dots[7], ux);
    float y0 = mixf(x00, x10, uy), y1 = mixf(x01, x11, uy);
    return mixf(y0, y1, uz) * 0.5f + 0.5f;
}

__device__ __forceinline__ uint8_t to_unorm8(float v) { return (uint8_t)rintf(clampf(v, 0.f, 1.f) * 255.f); }

template <typename Store>
__device__ __forceinline__ void store4(void* out, size_t i, int format, float a, float b, float c, float d) {
    if (format == 0) {
        uchar4 q = make_uchar4(to_unorm8(a), to_unorm8(b), to_unorm8(c), to_unorm8(d));
        ((uchar4*)out)[i] = q;
    } else {
        __half2* q = (__half2*)out + 2 * i;
        q[0] = __floats2half2_rn(a, b);
        q[1] = __floats2half2_rn(c, d);
    }
}

// noise_base.comp.wgsl:122-139
__global__ __launch_bounds__(64) void k_noise_base(uint32_t seed, int format, void* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 128 * 128 * 128) return;
    int x = i & 127, y = (i >> 7) & 127, z = i >> 14;
    float px = (x + 0.5f) / 128.f, py = (y + 0.5f) / 128.f, pz = (z + 0.5f) / 128.f;
    float perlin = perlin_tile(px, py, pz, 4, seed) * 0.5f + perlin_tile(px, py, pz, 8, seed) * 0.25f +
                   perlin_tile(px, py, pz, 16, seed) * 0.125f;
    float wl = worley_fbm(px, py, pz, 4, 0, 0, 0, seed);
    float pw = clampf((perlin - (wl - 1.f)) / fmaxf(1.f - (wl - 1.f), 1e-5f), 0.f, 1.f);
    float g = worley_fbm(px, py, pz, 2, 0, 0, 0, seed);
    float b = worley_fbm(px, py, pz, 8, 0, 0, 0, seed);
    float a = worley_fbm(px, py, pz, 14, 0, 0, 0, seed);
    store4<int>(out, (size_t)i, format, pw, g, b, a);
}

// noise_detail.comp.wgsl:201-211
__global__ __launch_bounds__(64) void k_noise_detail(uint32_t seed, int format, void* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 32 * 32 * 32) return;
    int x = i & 31, y = (i >> 5) & 31, z = i >> 10;
    float px = (x + 0.5f) / 32.f, py = (y + 0.5f) / 32.f, pz = (z + 0.5f) / 32.f;
    float r = worley_fbm(px, py, pz, 2, 73u, 19u, 41u, seed);
    float g = worley_fbm(px, py, pz, 8, 73u, 19u, 41u, seed);
    float b = worley_fbm(px, py, pz, 16, 73u, 19u, 41u, seed);
    store4<int>(out, (size_t)i, format, r, g, b, 0.f);
}

// noise_curl.comp.wgsl:230-285
__device__ float perlin2(float px, float py, int freq, uint32_t seed) {
    float f = (float)freq;
    float sx = px * f, sy = py * f;
    float flx = floorf(sx), fly = floorf(sy);
    int cx = (int)flx, cy = (int)fly;
    float fx = sx - flx, fy = sy - fly;
    float ux = fade(fx), uy = fade(fy);
    float d[4];
    for (int k = 0; k < 4; ++k) {
        int ox = k & 1, oy = k >> 1;
        uint32_t wx = (uint32_t)wrapf(cx + ox, freq), wy = (uint32_t)wrapf(cy + oy, freq);
        float a = u16f(hash2(wx, wy, seed)) * 2.f - 1.f;
        float b = u16f(hash2(wx + 1u, wy, seed)) * 2.f - 1.f;
        float il = 1.f / sqrtf(a * a + b * b);
        d[k] = a * il * (fx - ox) + b * il * (fy - oy);
    }
    float x0 = mixf(d[0], d[1], ux), x1 = mixf(d[2], d[3], ux);
    return mixf(x0, x1, uy);
}
__device__ __forceinline__ float potential(float px, float py, uint32_t seed) {
    return perlin2(px, py, 4, seed) + 0.5f * perlin2(px, py, 8, seed);
}
__global__ __launch_bounds__(64) void k_noise_curl(uint32_t seed, int format, void* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 128 * 128) return;
    int x = i & 127, y = i >> 7;
    float px = (x + 0.5f) / 128.f, py = (y + 0.5f) / 128.f, eps = 1.f / 128.f;
    float dpdx = (potential(px + eps, py, seed) - potential(px - eps, py, seed)) / (2.f * eps);
    float dpdy = (potential(px, py + eps, seed) - potential(px, py - eps, seed)) / (2.f * eps);
    float r = clampf(dpdy * 0.5f + 0.5f, 0.f, 1.f), g = clampf(-dpdx * 0.5f + 0.5f, 0.f, 1.f);
    if (format == 0) {
        ((uchar2*)out)[i] = make_uchar2(to_unorm8(r), to_unorm8(g));
    } else {
        ((__half2*)out)[i] = __floats2half2_rn(r, g);
    }
}

__device__ __forceinline__ void load4(const void* src, size_t i, int format, float& a, float& b, float& c, float& d) {
    if (format == 0) {
        uchar4 q = ((const uchar4*)src)[i];
        a = (float)q.x / 255.f; b = (float)q.y / 255.f; c = (float)q.z / 255.f; d = (float)q.w / 255.f;
    } else {
        const __half2* q = (const __half2*)src + 2 * i;
        float2 lo = __half22float2(q[0]), hi = __half22float2(q[1]);
        a = lo.x; b = lo.y; c = hi.x; d = hi.y;
    }
}

// R and the low-frequency FBM sum the march reads, as 16-bit fixed point,
// with scales that hold the stored values exactly where 16 bits can: UNORM8
// input R = k/255 as 257 k (scale 65535) and LF = (5 kG + 2 kB + kA) / 2040 as
// its numerator (scale 2040); binary16 input R as R * 32768 (exact for
// R >= 1/32) and LF as LF * 65535 (step 1.5e-5). base_fold_scales gives the
// scales the march multiplies by.
__device__ __forceinline__ uint32_t base_q(const void* base, int x, int y, int z, int format) {
    const size_t i = (size_t)x + ((size_t)y << 7) + ((size_t)z << 14);
    uint32_t qr, ql;
    if (format == 0) {
        const uchar4 q = ((const uchar4*)base)[i];
        qr = (uint32_t)q.x * 257u;
        ql = 5u * q.y + 2u * q.z + q.w;
    } else {
        float r, g, b, a;
        load4(base, i, format, r, g, b, a);
        const float lf = g * 0.625f + b * 0.25f + a * 0.125f;
        qr = (uint32_t)rintf(clampf(r, 0.f, 1.f) * 32768.f);
        ql = (uint32_t)rintf(clampf(lf, 0.f, 1.f) * 65535.f);
    }
    return qr | (ql << 16);
}

__device__ __forceinline__ float detail_h(const void* detail, int x, int y, int z, int format) {
    float r, g, b, a;
    load4(detail, (size_t)x + ((size_t)y << 5) + ((size_t)z << 10), format, r, g, b, a);
    return r * 0.625f + g * 0.25f + b * 0.125f;
}

__device__ __forceinline__ float2 curl_v(const void* curl, int x, int y, int format) {
    const int i = x + (y << 7);
    float r, g;
    if (format == 0) {
        uchar2 q = ((const uchar2*)curl)[i];
        r = (float)q.x / 255.f; g = (float)q.y / 255.f;
    } else {
        float2 v = __half22float2(((const __half2*)curl)[i]);
        r = v.x; g = v.y;
    }
    return make_float2(r * 2.f - 1.f, g * 2.f - 1.f);
}

// Folds the baked volumes into the layouts the march reads. Every texel
// carries the neighbours a filtered tap with that texel as its lower corner
// also reads (wrapped): base and detail (x, y, z), (x, y+1, z), (x, y, z+1),
// (x, y+1, z+1); curl (x, y), (x, y+1). A tap is then one 16-byte load per
// x corner.
__global__ __launch_bounds__(256) void k_fold(int format, const void* base, const void* detail, const void* curl,
                                              uint4* base_rl, float4* detail_hf, float4* curl2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 128 * 128 * 128) {
        const int x = i & 127, y = (i >> 7) & 127, z = i >> 14;
        const int y1 = (y + 1) & 127, z1 = (z + 1) & 127;
        base_rl[brick_x(x) + brick_y(y) + brick_z(z)] =
            make_uint4(base_q(base, x, y, z, format), base_q(base, x, y1, z, format), base_q(base, x, y, z1, format),
                       base_q(base, x, y1, z1, format));
    }
    if (i < 32 * 32 * 32) {
        const int x = i & 31, y = (i >> 5) & 31, z = i >> 10;
        const int y1 = (y + 1) & 31, z1 = (z + 1) & 31;
        detail_hf[i] = make_float4(detail_h(detail, x, y, z, format), detail_h(detail, x, y1, z, format),
                                   detail_h(detail, x, y, z1, format), detail_h(detail, x, y1, z1, format));
    }
    if (i < 128 * 128) {
        const int x = i & 127, y = i >> 7;
        const float2 a = curl_v(curl, x, y, format), b = curl_v(curl, x, (y + 1) & 127, format);
        curl2[i] = make_float4(a.x, a.y, b.x, b.y);
    }
}

// Conservative per-brick bounds for empty-space skipping, over texels
// [kb, kb + k] on each axis (the brick of k = PS_BB texels plus the +1
// neighbour a trilinear tap whose lower corner lies in the brick also reads):
// max R rounded up and min LF rounded down, each to a multiple of 1/255, from
// the first (own) entry of each folded texel. r_full and l_full are the
// fixed-point values of 1.0.
__global__ __launch_bounds__(64) void k_bbound(const uint32_t* __restrict__ base_rl, uint32_t r_full, uint32_t l_full,
                                               uint16_t* bbound) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= PS_BN * PS_BN * PS_BN) return;
    int bx = b % PS_BN, by = (b / PS_BN) % PS_BN, bz = b / (PS_BN * PS_BN);
    uint32_t mr = 0, ml = 0xffffu;
    for (int k = 0; k <= PS_BB; ++k)
        for (int j = 0; j <= PS_BB; ++j)
            for (int i = 0; i <= PS_BB; ++i) {
                int x = (bx * PS_BB + i) & 127, y = (by * PS_BB + j) & 127, z = (bz * PS_BB + k) & 127;
                const uint32_t v = base_rl[(brick_x(x) + brick_y(y) + brick_z(z)) * 4u];
                mr = max(mr, v & 0xffffu);
                ml = min(ml, v >> 16);
            }
    // fixed point (R = mr / r_full, LF = ml / l_full) -> u8:
    // ceil(mr * 255 / r_full) and floor(ml * 255 / l_full)
    const uint32_t qr = min((mr * 255u + r_full - 1u) / r_full, 255u);
    const uint32_t ql = min(ml * 255u / l_full, 255u);
    bbound[b] = (uint16_t)(qr | (ql << 8));
}

// Texel (x, y) of the output holds channel 0 of (x, y), (x+1, y) in .x and
// of (x, y+1), (x+1, y+1) in .y (low half first), neighbours wrapped: the
// four taps of a bilinear Repeat fetch whose lower-left texel is (x, y).
__global__ __launch_bounds__(256) void k_weather_fold(const uint16_t* src, int w, int h, int channels, uint2* dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const int x = i % w, y = i / w;
    const int x1 = x + 1 == w ? 0 : x + 1, y1 = y + 1 == h ? 0 : y + 1;
    auto c0 = [&](int xx, int yy) { return (uint32_t)src[((size_t)yy * w + xx) * channels]; };
    dst[i] = make_uint2(c0(x, y) | (c0(x1, y) << 16), c0(x, y1) | (c0(x1, y1) << 16));
}

// ---- weather-map synthesis (weather_map.rs:46-170, coverage_grid.rs:31-61,
// cloud_type_grid.rs:41-92) ------------------------------------------------
// FP32 arithmetic in the reference's order with no contraction (Rust does
// not fuse a * b + c), so every texel is bit-identical.
constexpr int kWeatherSize = 128;
constexpr float kWeatherExtent = 32000.f;

__device__ __forceinline__ float wm_hash2(int x, int y) {
    // hash2 (weather_map.rs:160-167)
    uint32_t h = (uint32_t)x * 2654435761u;
    h ^= (uint32_t)y * 2246822519u;
    h ^= h >> 16;
    h *= 2654435761u;
    h ^= h >> 13;
    return fminf(fmaxf((float)h / 4294967296.f, 0.f), 1.f);
}

__device__ float wm_value_noise(float x, float y) {
#pragma clang fp contract(off)
    // value_noise_2d (weather_map.rs:143-158): smoothstep-blended lattice hash
    const float xi = floorf(x), yi = floorf(y);
    const float xf = x - xi, yf = y - yi;
    const int ix = (int)xi, iy = (int)yi;
    const float h00 = wm_hash2(ix, iy), h10 = wm_hash2(ix + 1, iy);
    const float h01 = wm_hash2(ix, iy + 1), h11 = wm_hash2(ix + 1, iy + 1);
    const float sx = xf * xf * (3.f - 2.f * xf);
    const float sy = yf * yf * (3.f - 2.f * yf);
    const float bottom = h00 + (h10 - h00) * sx;
    const float top = h01 + (h11 - h01) * sx;
    return bottom + (top - bottom) * sy;
}

__device__ float wm_grid_sample(const float* __restrict__ g, int gw, int gh, float extent, float x, float z) {
#pragma clang fp contract(off)
    // LoadedCoverageGrid::sample_world / sample (coverage_grid.rs:31-61):
    // edge-clamped bilinear over the source grid spanning [-extent/2, extent/2]
    const float half = extent * 0.5f;
    const float u = (x + half) / extent, v = (z + half) / extent;
    const float fx = fminf(fmaxf(u, 0.f), 1.f) * ((float)gw - 1.f);
    const float fy = fminf(fmaxf(v, 0.f), 1.f) * ((float)gh - 1.f);
    const uint32_t x0 = (uint32_t)floorf(fx), y0 = (uint32_t)floorf(fy);
    const uint32_t x1 = min(x0 + 1u, (uint32_t)gw - 1u), y1 = min(y0 + 1u, (uint32_t)gh - 1u);
    const float tx = fx - (float)x0, ty = fy - (float)y0;
    const float v00 = g[y0 * gw + x0], v10 = g[y0 * gw + x1], v01 = g[y1 * gw + x0], v11 = g[y1 * gw + x1];
    const float bottom = v00 + (v10 - v00) * tx;
    const float top = v01 + (v11 - v01) * tx;
    return fminf(fmaxf(bottom + (top - bottom) * ty, 0.f), 1.f);
}

// One texel of the 128x128 RGBA16F weather map: R the coverage gate (the
// sampled coverage grid, or `gate`), G reserved 0, B the cloud-base offset
// noise in [-1, 1], A the precipitation / wetness scalar.
__global__ __launch_bounds__(256) void k_weather_synth(const float* __restrict__ grid, int gw, int gh, float extent,
                                                       float gate, float alpha, uint16_t* __restrict__ out) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kWeatherSize * kWeatherSize) return;
    const int x = i % kWeatherSize, y = i / kWeatherSize;
    float coverage = gate;
    if (grid) {
        // pixel_to_world (weather_map.rs:134-139)
        const float half = kWeatherExtent * 0.5f;
        const float fx = ((float)x + 0.5f) / (float)kWeatherSize, fy = ((float)y + 0.5f) / (float)kWeatherSize;
        coverage = wm_grid_sample(grid, gw, gh, extent, fx * kWeatherExtent - half, fy * kWeatherExtent - half);
    }
    const float bias = wm_value_noise((float)x * 0.05f, (float)y * 0.05f) * 2.f - 1.f;
    ushort4 o;
    o.x = __half_as_ushort(__float2half_rn(coverage));
    o.y = 0;
    o.z = __half_as_ushort(__float2half_rn(bias));
    o.w = __half_as_ushort(__float2half_rn(alpha));
    ((ushort4*)out)[i] = o;
}

// Nearest resample of a source cloud-type grid onto 128x128; values above 7
// become the sentinel 255 (cloud_type_grid.rs:71-86).
__global__ __launch_bounds__(256) void k_type_grid_synth(const uint8_t* __restrict__ src, int gw, int gh,
                                                         uint8_t* __restrict__ out) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kWeatherSize * kWeatherSize) return;
    const int x = i % kWeatherSize, y = i / kWeatherSize;
    const uint32_t sx = min((uint32_t)((float)x / (float)kWeatherSize * (float)gw), (uint32_t)gw - 1u);
    const uint32_t sy = min((uint32_t)((float)y / (float)kWeatherSize * (float)gh), (uint32_t)gh - 1u);
    const uint8_t v = src[sy * gw + sx];
    out[i] = v <= 7 ? v : (uint8_t)255;
}

// ---- blue-noise rank map (noise.rs:646-704) ---------------------------------
// The reference's greedy placement: 4096 times, take the untaken texel of
// lowest energy (the lowest index among equals), give it the next rank, and
// add a Gaussian of radius 6 around it to the energy field on the 64x64
// torus. One workgroup of 1024 threads; thread t keeps texels 4t..4t+3 (one
// row segment) with their energies and ranks in registers. Each placement is
// a 64-bit minimum over (energy bits, index): energies are non-negative
// sums, so their IEEE bits order like the values. A texel gains at most one
// term per placement (the splat is 13 texels wide on a 64-texel torus), in
// placement order, and the weights come from the host, so the energies, and
// therefore the ranks, are bit-identical to the sequential generator.
constexpr int kBlueSize = 64, kBlueCells = kBlueSize * kBlueSize, kBlueRadius = 6;

__global__ __launch_bounds__(1024) void k_blue_noise(BlueNoiseWeights wts, uint8_t* __restrict__ out) {
    __shared__ unsigned long long wave_min[2][16];
    __shared__ float w_r2[kBlueWeights];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t < kBlueWeights) w_r2[t] = wts.w[t];
    const int base = t * 4, y = base / kBlueSize, x0 = base % kBlueSize;
    float e[4] = {0.f, 0.f, 0.f, 0.f};
    int rank[4] = {-1, -1, -1, -1};
    __syncthreads();
    for (int p = 0; p < kBlueCells; ++p) {
        unsigned long long key = ~0ull;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (rank[k] < 0)
                key = min(key, ((unsigned long long)__float_as_uint(e[k]) << 32) | (unsigned)(base + k));
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) key = min(key, (unsigned long long)__shfl_xor(key, off));
        if (lane == 0) wave_min[p & 1][wave] = key;
        // one barrier per placement: the next placement writes the other
        // buffer, and the one after waits at this placement's successor's
        // barrier until every thread has read this one
        __syncthreads();
        unsigned long long best = wave_min[p & 1][0];
#pragma unroll
        for (int i = 1; i < 16; ++i) best = min(best, wave_min[p & 1][i]);
        const int b = (int)(uint32_t)best, cx = b % kBlueSize, cy = b / kBlueSize;
        // torus offset of this row from the placed texel, in [-32, 31]
        const int dy = ((y - cy + kBlueSize / 2) & (kBlueSize - 1)) - kBlueSize / 2;
        if (dy >= -kBlueRadius && dy <= kBlueRadius) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int dx = ((x0 + k - cx + kBlueSize / 2) & (kBlueSize - 1)) - kBlueSize / 2;
                if (dx >= -kBlueRadius && dx <= kBlueRadius) e[k] += w_r2[dx * dx + dy * dy];
            }
        }
        if (b >= base && b < base + 4) rank[b - base] = p;
    }
    // rank -> 0..255: (r / (n - 1)) * 255 rounded half away from zero, clamped
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float v = roundf(__fdiv_rn((float)rank[k], (float)(kBlueCells - 1)) * 255.f);
        out[base + k] = (uint8_t)fminf(fmaxf(v, 0.f), 255.f);
    }
}

}  // namespace

hipError_t ps_launch_blue_noise(const BlueNoiseWeights& w, uint8_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_blue_noise, dim3(1), dim3(1024), 0, s, w, out);
    return hipGetLastError();
}

hipError_t ps_launch_weather_synth(const float* grid, int gw, int gh, float extent, float gate, float alpha,
                                   const uint8_t* type_src, uint16_t* out_rgba, uint8_t* out_types, hipStream_t s) {
    const int n = kWeatherSize * kWeatherSize;
    hipLaunchKernelGGL(k_weather_synth, dim3((n + 255) / 256), dim3(256), 0, s, grid, gw, gh, extent, gate, alpha,
                       out_rgba);
    if (type_src)
        hipLaunchKernelGGL(k_type_grid_synth, dim3((n + 255) / 256), dim3(256), 0, s, type_src, gw, gh, out_types);
    return hipGetLastError();
}

hipError_t ps_launch_noise_bake(uint32_t seed, int format, void* base, void* detail, void* curl, hipStream_t s) {
    hipLaunchKernelGGL(k_noise_base, dim3(128 * 128 * 128 / 64), dim3(64), 0, s, seed, format, base);
    hipLaunchKernelGGL(k_noise_detail, dim3(32 * 32 * 32 / 64), dim3(64), 0, s, seed, format, detail);
    hipLaunchKernelGGL(k_noise_curl, dim3(128 * 128 / 64), dim3(64), 0, s, seed, format, curl);
    return hipGetLastError();
}

hipError_t ps_launch_noise_fold(int format, const void* base, const void* detail, const void* curl,
                                uint32_t* base_rl, float* detail_hf, float2* curl2, uint16_t* bbound,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_fold, dim3(128 * 128 * 128 / 256), dim3(256), 0, s, format, base, detail, curl,
                       (uint4*)base_rl, (float4*)detail_hf, (float4*)curl2);
    hipLaunchKernelGGL(k_bbound, dim3((PS_BN * PS_BN * PS_BN + 63) / 64), dim3(64), 0, s, base_rl,
                       format == 0 ? 65535u : 32768u, format == 0 ? 2040u : 65535u, bbound);
    return hipGetLastError();
}

hipError_t ps_launch_weather_fold(const uint16_t* src, int w, int h, int channels, uint2* dst, hipStream_t s) {
    int n = w * h;
    hipLaunchKernelGGL(k_weather_fold, dim3((n + 255) / 256), dim3(256), 0, s, src, w, h, channels, dst);
    return hipGetLastError();
}
