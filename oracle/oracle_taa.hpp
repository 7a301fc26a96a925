// FP64 restatement of the cloud TAA resolve (shaders/clouds/cloud_taa.wgsl)
// and of the RT sampling it relies on (the cloud RT sampler of
// crates/ps-clouds/src/lib.rs:256-265: linear filter, ClampToEdge). Test
// infrastructure only: the product's TAA kernel never calls this.
#pragma once

#include <algorithm>
#include <cmath>

#include "../include/pedalsky_hip.h"

namespace oracle {

// Bilinear tap of an RGBA RT (w x h, row-major, 4 doubles per texel) at uv,
// texel centres at (i + 0.5) / w, ClampToEdge.
inline void rt_bilinear(const double* rt, int w, int h, double u, double v, double out[4]) {
    const double sx = u * w - 0.5, sy = v * h - 0.5;
    const double fx0 = std::floor(sx), fy0 = std::floor(sy);
    const double fx = sx - fx0, fy = sy - fy0;
    auto cl = [](double i, int n) { return (int)std::min(std::max(i, 0.0), double(n - 1)); };
    const int x0 = cl(fx0, w), x1 = cl(fx0 + 1.0, w), y0 = cl(fy0, h), y1 = cl(fy0 + 1.0, h);
    for (int k = 0; k < 4; ++k) {
        const double a = rt[(size_t(y0) * w + x0) * 4 + k], b = rt[(size_t(y0) * w + x1) * 4 + k];
        const double c = rt[(size_t(y1) * w + x0) * 4 + k], d = rt[(size_t(y1) * w + x1) * 4 + k];
        out[k] = (a * (1.0 - fx) + b * fx) * (1.0 - fy) + (c * (1.0 - fx) + d * fx) * fy;
    }
}

// cloud_taa.wgsl:81-109: the view direction through RT uv (the current
// frame's inv_view_proj at NDC z = 1, minus the camera position,
// normalised), projected with w = 0 by prev_view_proj. Returns the
// inside-[0, 1]^2 flag; (0, 0) and false when the previous clip w <= 1e-4.
inline bool taa_reproject(const ps_frame_uniforms& f, double u, double v, double& pu, double& pv) {
    const double n[4] = {u * 2.0 - 1.0, 1.0 - v * 2.0, 1.0, 1.0};
    double h[4];
    for (int r = 0; r < 4; ++r) {
        h[r] = 0.0;
        for (int c = 0; c < 4; ++c) h[r] += double(f.inv_view_proj[c * 4 + r]) * n[c];
    }
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = h[k] / h[3] - double(f.camera_position_world[k]);
    const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) d[k] /= len;
    double p[4];
    for (int r = 0; r < 4; ++r) {
        p[r] = 0.0;
        for (int c = 0; c < 3; ++c) p[r] += double(f.prev_view_proj[c * 4 + r]) * d[c];
    }
    if (p[3] <= 1e-4) {
        pu = pv = 0.0;
        return false;
    }
    pu = p[0] / p[3] * 0.5 + 0.5;
    pv = 1.0 - (p[1] / p[3] * 0.5 + 0.5);
    return pu >= 0.0 && pv >= 0.0 && pu <= 1.0 && pv <= 1.0;
}

// cloud_taa.wgsl:142-197 for RT texel (x, y): history at the reprojected uv
// clamped into the current frame's 3x3 neighbourhood box (cloud_taa.wgsl:
// 115-137; the taps sit on texel centres, ClampToEdge), blended with the
// current sample, luminance >= 0 and transmittance in [0, 1]. With no valid
// history or a blend weight of 1 the output is the current sample.
inline void taa_resolve_texel(const ps_frame_uniforms& f, int w, int h, int x, int y, const double* cur_l,
                              const double* cur_t, const double* hist_l, const double* hist_t, double blend,
                              double history_valid, double out_l[4], double out_t[4]) {
    const size_t o = (size_t(y) * w + x) * 4;
    const double bw = std::min(std::max(blend, 0.0), 1.0);
    if (history_valid < 0.5 || bw >= 1.0) {
        for (int k = 0; k < 4; ++k) {
            out_l[k] = cur_l[o + k];
            out_t[k] = cur_t[o + k];
        }
        return;
    }
    const double u = (x + 0.5) / w, v = (y + 0.5) / h;
    double pu, pv;
    const bool inside = taa_reproject(f, u, v, pu, pv);
    double hl[4], ht[4];
    rt_bilinear(hist_l, w, h, pu, pv, hl);
    rt_bilinear(hist_t, w, h, pu, pv, ht);
    double lo_l[4], hi_l[4], lo_t[4], hi_t[4];
    for (int k = 0; k < 4; ++k) {
        lo_l[k] = lo_t[k] = 1e10;
        hi_l[k] = hi_t[k] = -1e10;
    }
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int xx = std::min(std::max(x + dx, 0), w - 1), yy = std::min(std::max(y + dy, 0), h - 1);
            const size_t q = (size_t(yy) * w + xx) * 4;
            for (int k = 0; k < 4; ++k) {
                lo_l[k] = std::min(lo_l[k], cur_l[q + k]);
                hi_l[k] = std::max(hi_l[k], cur_l[q + k]);
                lo_t[k] = std::min(lo_t[k], cur_t[q + k]);
                hi_t[k] = std::max(hi_t[k], cur_t[q + k]);
            }
        }
    const double eb = inside ? bw : 1.0;
    for (int k = 0; k < 4; ++k) {
        const double a = std::min(std::max(hl[k], lo_l[k]), hi_l[k]);
        const double b = std::min(std::max(ht[k], lo_t[k]), hi_t[k]);
        out_l[k] = std::max(a * (1.0 - eb) + cur_l[o + k] * eb, 0.0);
        out_t[k] = std::min(std::max(b * (1.0 - eb) + cur_t[o + k] * eb, 0.0), 1.0);
    }
}

}  // namespace oracle
