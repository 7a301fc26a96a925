// Oracle (test infrastructure, not product code): scalar/vector helpers and a
// software texture sampler used by the CPU restatement of PedalSky's WGSL sky
// + cloud path. Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline use anything under oracle/.
//
// WGSL built-ins are restated with their specification formulas:
//   smoothstep(e0, e1, x) = t*t*(3-2t), t = clamp((x-e0)/(e1-e0), 0, 1)
//   mix(a, b, t) = a + (b-a)*t      (WGSL spec: a*(1-t) + b*t)
//   fract(x) = x - floor(x)
// Texture sampling follows the WebGPU sampling rules: texel-centred linear
// filtering, coordinate = uv*N - 0.5, with Repeat or ClampToEdge addressing
// (crates/ps-clouds/src/noise.rs:178-197, crates/ps-core/src/atmosphere_luts.rs:130-139).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstddef>
#include <algorithm>

// Planted mistakes for the mutation checks of the oracle's pins
// (tests/test_oracle_pins.py builds this library with -DORACLE_SLIP=k, k > 0,
// and requires the pin of the function that carries slip k to fail). The
// default 0 is the restatement itself; every ORACLE_SLIP test below is a
// compile-time constant that is false in the checker build.
#ifndef ORACLE_SLIP
#define ORACLE_SLIP 0
#endif

namespace oracle {

// Math calls are unqualified so that the forward-mode dual numbers of
// oracle_dual.hpp (the oracle's exact derivatives) find their own overloads;
// for float and double these are the std:: functions.
using std::exp;
using std::log;
using std::sqrt;
using std::pow;
using std::sin;
using std::cos;
using std::acos;
using std::asin;
using std::atan2;
using std::floor;
using std::fabs;

template <class R>
struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a) : x(a), y(a), z(a) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(const V3& o) const { return V3(x * o.x, y * o.y, z * o.z); }
    V3 operator/(const V3& o) const { return V3(x / o.x, y / o.y, z / o.z); }
    V3 operator*(R s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(R s) const { return V3(x / s, y / s, z / s); }
    V3 operator-() const { return V3(-x, -y, -z); }
};
template <class R> inline V3<R> operator*(R s, const V3<R>& v) { return v * s; }
template <class R> inline R dot(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R length(const V3<R>& a) { return sqrt(dot(a, a)); }
template <class R> inline V3<R> normalize(const V3<R>& a) { return a / length(a); }
template <class R> inline V3<R> vexp(const V3<R>& a) { return V3<R>(exp(a.x), exp(a.y), exp(a.z)); }
template <class R> inline V3<R> vmax(const V3<R>& a, const V3<R>& b) {
    return V3<R>(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z));
}
template <class R> inline R max3(const V3<R>& v) { return std::max(v.x, std::max(v.y, v.z)); }

template <class R> inline R clampr(R x, R lo, R hi) { return std::min(std::max(x, lo), hi); }
template <class R> inline R smoothstep(R e0, R e1, R x) {
    R t = clampr<R>((x - e0) / (e1 - e0), R(0), R(1));
    return t * t * (R(3) - R(2) * t);
}
template <class R> inline R mixr(R a, R b, R t) { return a + (b - a) * t; }
template <class R> inline V3<R> mixv(const V3<R>& a, const V3<R>& b, R t) { return a + (b - a) * t; }
template <class R> inline R fractr(R x) { return x - floor(x); }

// Texel storage is a flat array of `ch` channels per texel, x fastest, then
// y, then z (the memory order of wgpu texture uploads).
template <class R, class S>
struct Tex {
    const S* data = nullptr;
    int w = 0, h = 1, d = 1, ch = 4;
    R at(int x, int y, int z, int c) const {
        return R(data[((size_t(z) * h + y) * w + x) * ch + c]);
    }
};

inline int wrap_index(int i, int n) { int m = i % n; return m < 0 ? m + n : m; }
inline int clamp_index(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

// Bilinear sample of a 2D texture; `repeat` selects Repeat, else ClampToEdge.
template <class R, class S>
inline void sample2d(const Tex<R, S>& t, R u, R v, bool repeat, R out[4]) {
    R cx = u * R(t.w) - R(0.5), cy = v * R(t.h) - R(0.5);
    R fx0 = floor(cx), fy0 = floor(cy);
    R fx = cx - fx0, fy = cy - fy0;
    int ix = int(fx0), iy = int(fy0);
    int x0, x1, y0, y1;
    if (repeat) {
        x0 = wrap_index(ix, t.w); x1 = wrap_index(ix + 1, t.w);
        y0 = wrap_index(iy, t.h); y1 = wrap_index(iy + 1, t.h);
    } else {
        x0 = clamp_index(ix, t.w); x1 = clamp_index(ix + 1, t.w);
        y0 = clamp_index(iy, t.h); y1 = clamp_index(iy + 1, t.h);
    }
    for (int c = 0; c < t.ch && c < 4; ++c) {
        R a = t.at(x0, y0, 0, c), b = t.at(x1, y0, 0, c);
        R e = t.at(x0, y1, 0, c), f = t.at(x1, y1, 0, c);
        R top = a + (b - a) * fx, bot = e + (f - e) * fx;
        out[c] = top + (bot - top) * fy;
    }
    for (int c = t.ch; c < 4; ++c) out[c] = (c == 3) ? R(1) : R(0);
}

// Trilinear sample of a 3D texture.
template <class R, class S>
inline void sample3d(const Tex<R, S>& t, R u, R v, R w, bool repeat, R out[4]) {
    R cx = u * R(t.w) - R(0.5), cy = v * R(t.h) - R(0.5), cz = w * R(t.d) - R(0.5);
    R fx0 = floor(cx), fy0 = floor(cy), fz0 = floor(cz);
    R fx = cx - fx0, fy = cy - fy0, fz = cz - fz0;
    int ix = int(fx0), iy = int(fy0), iz = int(fz0);
    int x0, x1, y0, y1, z0, z1;
    if (repeat) {
        x0 = wrap_index(ix, t.w); x1 = wrap_index(ix + 1, t.w);
        y0 = wrap_index(iy, t.h); y1 = wrap_index(iy + 1, t.h);
        z0 = wrap_index(iz, t.d); z1 = wrap_index(iz + 1, t.d);
    } else {
        x0 = clamp_index(ix, t.w); x1 = clamp_index(ix + 1, t.w);
        y0 = clamp_index(iy, t.h); y1 = clamp_index(iy + 1, t.h);
        z0 = clamp_index(iz, t.d); z1 = clamp_index(iz + 1, t.d);
    }
    for (int c = 0; c < t.ch && c < 4; ++c) {
        R c000 = t.at(x0, y0, z0, c), c100 = t.at(x1, y0, z0, c);
        R c010 = t.at(x0, y1, z0, c), c110 = t.at(x1, y1, z0, c);
        R c001 = t.at(x0, y0, z1, c), c101 = t.at(x1, y0, z1, c);
        R c011 = t.at(x0, y1, z1, c), c111 = t.at(x1, y1, z1, c);
        R a0 = c000 + (c100 - c000) * fx, a1 = c010 + (c110 - c010) * fx;
        R b0 = c001 + (c101 - c001) * fx, b1 = c011 + (c111 - c011) * fx;
        R p0 = a0 + (a1 - a0) * fy, p1 = b0 + (b1 - b0) * fy;
        out[c] = p0 + (p1 - p0) * fz;
    }
    for (int c = t.ch; c < 4; ++c) out[c] = (c == 3) ? R(1) : R(0);
}

}  // namespace oracle
