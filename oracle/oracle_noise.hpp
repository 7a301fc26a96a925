// Oracle (test infrastructure, not product code): CPU restatement of the
// cloud noise bakes (shaders/clouds/noise_base.comp.wgsl,
// noise_detail.comp.wgsl, noise_curl.comp.wgsl). The integer hash is exact;
// the float work runs in the template scalar type. `seed` is mixed into the
// hash input as seed * 0x85ebca6b, so seed 0 is the reference's unseeded hash.
#pragma once
#include "oracle_math.hpp"

namespace oracle {

// noise_base.comp.wgsl:25-30 (hash3), noise_curl.comp.wgsl:12-17 (hash2)
inline uint32_t hash3(uint32_t x0, uint32_t y0, uint32_t z0, uint32_t seed) {
    uint32_t x = x0 * 0x27d4eb2du + y0 * 0x165667b1u + z0 * 0x9e3779b9u + seed * 0x85ebca6bu;
    x = (x ^ (x >> 15u)) * 0x2c1b3c6du;
    x = (x ^ (x >> 12u)) * 0x297a2d39u;
    return x ^ (x >> 15u);
}
inline uint32_t hash2(uint32_t x0, uint32_t y0, uint32_t seed) {
    uint32_t x = x0 * 0x27d4eb2du + y0 * 0x165667b1u + seed * 0x85ebca6bu;
    x = (x ^ (x >> 15u)) * 0x2c1b3c6du;
    x = (x ^ (x >> 12u)) * 0x297a2d39u;
    return x ^ (x >> 15u);
}

// noise_base.comp.wgsl:32-41
template <class R>
inline V3<R> rand3(uint32_t x, uint32_t y, uint32_t z, uint32_t seed) {
    uint32_t h0 = hash3(x, y, z, seed), h1 = hash3(x + 1u, y, z, seed), h2 = hash3(x, y + 1u, z, seed);
    return V3<R>(R(h0 & 0xffffu) / R(65535), R(h1 & 0xffffu) / R(65535), R(h2 & 0xffffu) / R(65535));
}

inline int wrap_freq(int c, int f) { return ((c % f) + f) % f; }

// noise_base.comp.wgsl:45-68 and noise_detail.comp.wgsl:31-53 (hash offset added after the wrap)
template <class R>
inline R worley_tile(const V3<R>& p, uint32_t freq, uint32_t ox, uint32_t oy, uint32_t oz, uint32_t seed) {
    R f = R(freq);
    V3<R> sc = p * f;
    V3<R> fl(std::floor(sc.x), std::floor(sc.y), std::floor(sc.z));
    int cx = int(fl.x), cy = int(fl.y), cz = int(fl.z);
    V3<R> fr = sc - fl;
    R min_d = R(1);
    for (int dz = -1; dz <= (ORACLE_SLIP == 8 ? 0 : 1); ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                uint32_t wx = uint32_t(wrap_freq(cx + dx, int(freq)));
                uint32_t wy = uint32_t(wrap_freq(cy + dy, int(freq)));
                uint32_t wz = uint32_t(wrap_freq(cz + dz, int(freq)));
                V3<R> j = ORACLE_SLIP == 9
                              ? rand3<R>(uint32_t(wrap_freq(cx + dx + int(ox), int(freq))),
                                         uint32_t(wrap_freq(cy + dy + int(oy), int(freq))),
                                         uint32_t(wrap_freq(cz + dz + int(oz), int(freq))), seed)
                              : rand3<R>(wx + ox, wy + oy, wz + oz, seed);
                V3<R> cc = V3<R>(R(dx), R(dy), R(dz)) + j;
                R d = length(cc - fr);
                min_d = std::min(min_d, d);
            }
    return clampr<R>(R(1) - min_d, R(0), R(1));
}
template <class R>
inline R worley_fbm(const V3<R>& p, uint32_t freq, uint32_t ox, uint32_t oy, uint32_t oz, uint32_t seed) {
    R f1 = worley_tile(p, freq, ox, oy, oz, seed);
    R f2 = worley_tile(p, freq * 2u, ox, oy, oz, seed);
    R f3 = worley_tile(p, freq * 4u, ox, oy, oz, seed);
    return f1 * R(0.625) + f2 * R(0.25) + f3 * R(0.125);
}

template <class R>
inline R fade(R t) { return t * t * t * (t * (t * R(6) - R(15)) + R(10)); }

// noise_base.comp.wgsl:81-120
template <class R>
inline V3<R> perlin_grad_tile(int cx, int cy, int cz, uint32_t freq, uint32_t seed) {
    uint32_t wx = uint32_t(wrap_freq(cx, int(freq))), wy = uint32_t(wrap_freq(cy, int(freq))),
             wz = uint32_t(wrap_freq(cz, int(freq)));
    V3<R> r = rand3<R>(wx, wy, wz, seed) * R(2) - V3<R>(R(1));
    return normalize(r);
}
template <class R>
inline R perlin_tile(const V3<R>& p, uint32_t freq, uint32_t seed) {
    V3<R> sc = p * R(freq);
    V3<R> fl(std::floor(sc.x), std::floor(sc.y), std::floor(sc.z));
    int cx = int(fl.x), cy = int(fl.y), cz = int(fl.z);
    V3<R> f = sc - fl;
    V3<R> u(fade(f.x), fade(f.y), fade(f.z));
    R dots[8];
    for (int iz = 0; iz < 2; ++iz)
        for (int iy = 0; iy < 2; ++iy)
            for (int ix = 0; ix < 2; ++ix) {
                V3<R> g = perlin_grad_tile<R>(cx + ix, cy + iy, cz + iz, freq, seed);
                V3<R> d = f - V3<R>(R(ix), R(iy), R(iz));
                dots[ORACLE_SLIP == 7 ? iy + ix * 2 + iz * 4 : ix + iy * 2 + iz * 4] = dot(g, d);
            }
    R x00 = mixr(dots[0], dots[1], u.x), x10 = mixr(dots[2], dots[3], u.x);
    R x01 = mixr(dots[4], dots[5], u.x), x11 = mixr(dots[6], dots[7], u.x);
    R y0 = mixr(x00, x10, u.y), y1 = mixr(x01, x11, u.y);
    return mixr(y0, y1, u.z) * R(0.5) + R(0.5);
}
template <class R>
inline R perlin_fbm(const V3<R>& p, uint32_t bf, uint32_t seed) {
    return perlin_tile(p, bf, seed) * R(0.5) + perlin_tile(p, bf * 2u, seed) * R(0.25) +
           perlin_tile(p, bf * 4u, seed) * R(0.125);
}

// noise_base.comp.wgsl:122-139: RGBA for texel (x, y, z) of the 128^3 volume.
template <class R>
inline void base_texel(int x, int y, int z, uint32_t seed, R out[4]) {
    const R N = R(128);
    V3<R> p((R(x) + R(0.5)) / N, (R(y) + R(0.5)) / N, (R(z) + R(0.5)) / N);
    R perlin = perlin_fbm(p, 4u, seed);
    R wl = worley_fbm(p, 4u, 0, 0, 0, seed);
    // remap(perlin, wl - 1, 1, 0, 1) with the noise shader's own remap (:21-23)
    R pw = R(0) + (perlin - (wl - R(1))) * (R(1) - R(0)) / std::max(R(1) - (wl - R(1)), R(1e-5));
    out[0] = clampr<R>(pw, R(0), R(1));
    out[1] = worley_fbm(p, ORACLE_SLIP == 17 ? 4u : 2u, 0, 0, 0, seed);
    out[2] = worley_fbm(p, 8u, 0, 0, 0, seed);
    out[3] = worley_fbm(p, 14u, 0, 0, 0, seed);
}

// noise_detail.comp.wgsl:62-72
template <class R>
inline void detail_texel(int x, int y, int z, uint32_t seed, R out[4]) {
    const R N = R(32);
    V3<R> p((R(x) + R(0.5)) / N, (R(y) + R(0.5)) / N, (R(z) + R(0.5)) / N);
    out[0] = worley_fbm(p, 2u, 73u, 19u, 41u, seed);
    out[1] = worley_fbm(p, 8u, 73u, 19u, 41u, seed);
    out[2] = worley_fbm(p, 16u, 73u, 19u, 41u, seed);
    out[3] = R(0);
}

// noise_curl.comp.wgsl:19-74
template <class R>
inline void rand2(uint32_t x, uint32_t y, uint32_t seed, R& a, R& b) {
    uint32_t h0 = hash2(x, y, seed), h1 = hash2(x + 1u, y, seed);
    a = R(h0 & 0xffffu) / R(65535) * R(2) - R(1);
    b = R(h1 & 0xffffu) / R(65535) * R(2) - R(1);
}
template <class R>
inline void perlin_grad2(int cx, int cy, uint32_t freq, uint32_t seed, R& gx, R& gy) {
    uint32_t wx = uint32_t(wrap_freq(cx, int(freq))), wy = uint32_t(wrap_freq(cy, int(freq)));
    R a, b;
    rand2<R>(wx, wy, seed, a, b);
    R l = std::sqrt(a * a + b * b);
    gx = a / l; gy = b / l;
}
template <class R>
inline R perlin2(R px, R py, uint32_t freq, uint32_t seed) {
    R sx = px * R(freq), sy = py * R(freq);
    R flx = std::floor(sx), fly = std::floor(sy);
    int cx = int(flx), cy = int(fly);
    R fx = sx - flx, fy = sy - fly;
    R ux = fade(fx), uy = fade(fy);
    R g00x, g00y, g10x, g10y, g01x, g01y, g11x, g11y;
    perlin_grad2<R>(cx, cy, freq, seed, g00x, g00y);
    perlin_grad2<R>(cx + 1, cy, freq, seed, g10x, g10y);
    perlin_grad2<R>(cx, cy + 1, freq, seed, g01x, g01y);
    perlin_grad2<R>(cx + 1, cy + 1, freq, seed, g11x, g11y);
    R d00 = g00x * fx + g00y * fy;
    R d10 = g10x * (fx - R(1)) + g10y * fy;
    R d01 = g01x * fx + g01y * (fy - R(1));
    R d11 = g11x * (fx - R(1)) + g11y * (fy - R(1));
    R x0 = mixr(d00, d10, ux), x1 = mixr(d01, d11, ux);
    return mixr(x0, x1, uy);
}
template <class R>
inline R potential(R px, R py, uint32_t seed) {
    return perlin2(px, py, 4u, seed) + R(0.5) * perlin2(px, py, 8u, seed);
}
template <class R>
inline void curl_texel(int x, int y, uint32_t seed, R out[2]) {
    const R N = R(128);
    R px = (R(x) + R(0.5)) / N, py = (R(y) + R(0.5)) / N;
    R eps = R(1) / N;
    R dpdx = (potential(px + eps, py, seed) - potential(px - eps, py, seed)) / (R(2) * eps);
    R dpdy = (potential(px, py + eps, seed) - potential(px, py - eps, seed)) / (R(2) * eps);
    if (ORACLE_SLIP == 13) std::swap(dpdx, dpdy);
    out[0] = clampr<R>(dpdy * R(0.5) + R(0.5), R(0), R(1));
    out[1] = clampr<R>(-dpdx * R(0.5) + R(0.5), R(0), R(1));
}

}  // namespace oracle
