// Differentiation of the atmosphere LUT chain for gfx950: gradients of a
// loss on the sky + aerial-perspective image with respect to the Rayleigh,
// Mie scattering, Mie absorption and ozone coefficients (per channel) and
// the Rayleigh and Mie scale heights, for many parameter sets at once.
//
// Design
//  * 14 parameters against ~6e4 LUT texels per set: the chain is
//    differentiated in tangent (forward) mode. Every LUT texel carries, per
//    channel c, the derivatives of its channel-c value with respect to the
//    six parameters that reach channel c: beta_R[c], beta_Ms[c],
//    beta_Ma[c], beta_O[c], H_R, H_M (channels never mix inside the chain:
//    every extinction, scattering and transmittance term is per channel).
//    Tangent LUTs are FP32 [set][channel][6][texel]; the values are the
//    regular FP32 LUTs baked by lut_kernels.hip for the same sets.
//  * The tangent kernels restate the value kernels' FP64 arithmetic
//    (transmittance.comp.wgsl:18-31, multiscatter.comp.wgsl:58-156,
//    skyview.comp.wgsl:46-174, aerialperspective.comp.wgsl:41-127) with the
//    chain rule applied to each term, one channel per launch row (grid.z).
//  * The image pass is the adjoint side: per pixel it evaluates the image
//    from the value LUTs, forms the L2 residual against the target, and
//    contracts the cotangent with the tangent taps; per-block sums of the 18
//    channel-parameter products and the loss go to each set's accumulators.
//    The contraction equals the reverse-mode gradient; no tape, no atomics
//    into LUT space.
#include "ps_device.h"
#include "ps_launch.h"

namespace {

constexpr int NT = 6;  // tangents per channel: beta_R, beta_Ms, beta_Ma, beta_O, H_R, H_M

struct TD { double v, d[NT]; };  // a channel value and its six tangents

__device__ __forceinline__ TD td(double v) {
    TD r;
    r.v = v;
#pragma unroll
    for (int j = 0; j < NT; ++j) r.d[j] = 0.0;
    return r;
}

// geometry helpers (atmosphere.wgsl:95-132), FP64 as in lut_kernels.hip
struct V { double x, y, z; };
__device__ __forceinline__ V v3(double x, double y, double z) { V r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ V vadd(V a, V b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V vscl(V a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double vlen(V a) { return sqrt(vdot(a, a)); }

__device__ __forceinline__ bool sphere_hit(V ro, V rd, double r, double& t0, double& t1) {
    double b = vdot(ro, rd);
    double c = vdot(ro, ro) - r * r;
    double disc = b * b - c;
    if (disc < 0.0) return false;
    double sq = sqrt(disc);
    t0 = -b - sq; t1 = -b + sq;
    return true;
}
__device__ __forceinline__ double dist_top(const AtmD& a, V p, V d) {
    double t0, t1;
    if (!sphere_hit(p, d, a.top, t0, t1)) return 0.0;
    return fmax(t1, 0.0);
}
__device__ __forceinline__ double dist_boundary(const AtmD& a, V p, V d) {
    double tt0, tt1, tp0, tp1;
    bool ht = sphere_hit(p, d, a.top, tt0, tt1);
    bool hp = sphere_hit(p, d, a.R, tp0, tp1);
    if (!ht) return 0.0;
    double tm = tt1;
    if (hp && tp0 > 0.0) tm = fmin(tm, tp0);
    return fmax(tm, 0.0);
}

// Channel-c medium at altitude h with tangents (atmosphere.wgsl:23-61).
struct Medium { TD st, ss, sr, sm; };

__device__ __forceinline__ Medium medium(const AtmD& a, int c, double h) {
    const double dr = exp(-h / a.hr), dm = exp(-h / a.hm);
    const double dz = fmax(0.0, 1.0 - fabs(h - a.oz_c) / fmax(a.oz_t * 0.5, 1.0));
    const double ddr = dr * h / (a.hr * a.hr), ddm = dm * h / (a.hm * a.hm);
    const double br = a.ray_s[c], bs = a.mie_s[c], ba = a.mie_a[c], bo = a.ozone[c];
    Medium m;
    m.sr = td(br * dr); m.sr.d[0] = dr; m.sr.d[4] = br * ddr;
    m.sm = td(bs * dm); m.sm.d[1] = dm; m.sm.d[5] = bs * ddm;
    m.ss = td(br * dr + bs * dm);
    m.ss.d[0] = dr; m.ss.d[1] = dm; m.ss.d[4] = br * ddr; m.ss.d[5] = bs * ddm;
    m.st = td(br * dr + (bs + ba) * dm + bo * dz);
    m.st.d[0] = dr; m.st.d[1] = dm; m.st.d[2] = dm; m.st.d[3] = dz; m.st.d[4] = br * ddr; m.st.d[5] = (bs + ba) * ddm;
    return m;
}

// Bilinear ClampToEdge tap of channel c of a value LUT and its tangent LUT
// (texel-centred, as lut_kernels.hip's lut_d).
__device__ __forceinline__ TD tap2(const float4* __restrict__ val, const float* __restrict__ tan, int w, int h, int c,
                                   double u, double v) {
    double cx = u * w - 0.5, cy = v * h - 0.5;
    double fx0 = floor(cx), fy0 = floor(cy);
    double fx = cx - fx0, fy = cy - fy0;
    int ix = (int)fx0, iy = (int)fy0;
    int x0 = clampi(ix, w), x1 = clampi(ix + 1, w), y0 = clampi(iy, h), y1 = clampi(iy + 1, h);
    const int i00 = y0 * w + x0, i10 = y0 * w + x1, i01 = y1 * w + x0, i11 = y1 * w + x1;
    const double w00 = (1.0 - fx) * (1.0 - fy), w10 = fx * (1.0 - fy), w01 = (1.0 - fx) * fy, w11 = fx * fy;
    auto ch = [&](float4 q) { return c == 0 ? (double)q.x : (c == 1 ? (double)q.y : (double)q.z); };
    // the value follows the lerp form of lut_d (same rounding as the forward)
    const double A = ch(val[i00]), B = ch(val[i10]), Cc = ch(val[i01]), D = ch(val[i11]);
    const double tx = A + (B - A) * fx, bx = Cc + (D - Cc) * fx;
    TD r = td(tx + (bx - tx) * fy);
    const size_t n = (size_t)w * h;
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const float* t = tan + ((size_t)c * NT + j) * n;
        r.d[j] = w00 * t[i00] + w10 * t[i10] + w01 * t[i01] + w11 * t[i11];
    }
    return r;
}

// Transmittance-LUT tap toward `dir` from p with the Earth shadow
// (atmosphere_lut_sampling.wgsl:6-22).
__device__ __forceinline__ TD trans_tap(const LutParams& P, const float4* T, const float* TT, int c, V p, V dir) {
    const AtmD& a = P.a;
    double t0, t1;
    if (sphere_hit(p, dir, a.R, t0, t1) && t0 > 0.0) return td(0.0);
    double r = vlen(p);
    double H = sqrt(fmax(a.top * a.top - a.R * a.R, 0.0));
    double rho = sqrt(fmax(r * r - a.R * a.R, 0.0));
    double d = fmax(dist_top(a, p, dir), 0.0);
    double d_min = a.top - r, d_max = rho + H;
    double u = clampd((d - d_min) / fmax(d_max - d_min, 1.0), 0.0, 1.0);
    double v = rho / fmax(H, 1.0);
    return tap2(T, TT, P.tw, P.th, c, u, v);
}

// (1 - trs) / safe-extinction step integral with tangents:
// s = ins * (1 - e^{-st dt}) / max(st, 1e-7)
// (one FP64 reciprocal per call: the tangents multiply by it)
__device__ __forceinline__ TD step_integral(const TD& ins, const TD& st, double dt, const TD& trs) {
    const double safe = fmax(st.v, 1e-7);
    const double inv = 1.0 / safe;
    const double om = 1.0 - trs.v;
    TD s = td(ins.v * om / safe);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const double dom = -trs.d[j];
        const double dsafe = st.v > 1e-7 ? st.d[j] : 0.0;
        s.d[j] = ((ins.d[j] * om + ins.v * dom) - s.v * dsafe) * inv;
    }
    return s;
}

__device__ __forceinline__ TD exp_neg_scaled(const TD& st, double dt) {
    TD e = td(exp(-st.v * dt));
#pragma unroll
    for (int j = 0; j < NT; ++j) e.d[j] = -e.v * dt * st.d[j];
    return e;
}

__device__ __forceinline__ void store_tan(float* out, size_t n, int c, int idx, const TD& v) {
#pragma unroll
    for (int j = 0; j < NT; ++j) out[((size_t)c * NT + j) * n + idx] = (float)v.d[j];
}

// ---- transmittance tangents (transmittance.comp.wgsl:18-31) -------------------
__global__ __launch_bounds__(64) void k_transmittance_tan(const LutParams* __restrict__ Ps, float* __restrict__ TT) {
    const LutParams& P = Ps[blockIdx.y];
    const int c = blockIdx.z;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)P.tw * P.th;
    if (i >= (int)n) return;
    float* out = TT + (size_t)blockIdx.y * 3 * NT * n;
    const AtmD& a = P.a;
    int x = i % P.tw, y = i / P.tw;
    double u = (x + 0.5) / P.tw, v = (y + 0.5) / P.th;
    double H = sqrt(fmax(a.top * a.top - a.R * a.R, 0.0));
    double rho = v * H;
    double r = sqrt(rho * rho + a.R * a.R);
    double d_min = a.top - r, d_max = rho + H;
    double d = d_min + u * (d_max - d_min);
    double mu = 1.0;
    if (d != 0.0) mu = clampd((H * H - rho * rho - d * d) / fmax(2.0 * r * d, 1e-6), -1.0, 1.0);
    V p = v3(0, r, 0), dir = v3(sqrt(fmax(1.0 - mu * mu, 0.0)), mu, 0);
    const int N = 40;
    double t_max = dist_boundary(a, p, dir);
    TD od = td(0.0);
    if (t_max > 0.0) {
        double dt = t_max / N;
        double r0 = vlen(p);
        double cv = vdot(vscl(p, 1.0 / fmax(r0, 1.0)), dir);
        double h0 = r0 - a.R;
        TD prev = medium(a, c, h0).st;
        for (int k = 1; k <= N; ++k) {
            double t = k * dt;
            V pi = vadd(p, vscl(dir, t));
            double num = 2.0 * t * r0 * cv + t * t;
            double h = h0 + num / (fmax(r0, 1.0) + vlen(pi));
            TD cur = medium(a, c, h).st;
            od.v += (prev.v + cur.v) * (0.5 * dt);
#pragma unroll
            for (int j = 0; j < NT; ++j) od.d[j] += (prev.d[j] + cur.d[j]) * (0.5 * dt);
            prev = cur;
        }
    }
    const double T = exp(-od.v);
    TD t = td(T);
#pragma unroll
    for (int j = 0; j < NT; ++j) t.d[j] = -T * od.d[j];
    store_tan(out, n, c, i, t);
}

// ---- multi-scattering tangents (multiscatter.comp.wgsl:58-156): one wave per
// texel and channel, one lane per Fibonacci direction. ---------------------------
__global__ __launch_bounds__(64) void k_multiscatter_tan(const LutParams* __restrict__ Ps, const float4* __restrict__ T,
                                                         const float* __restrict__ TT, float* __restrict__ MST) {
    const LutParams& P = Ps[blockIdx.y];
    const int c = blockIdx.z;
    const int texel = blockIdx.x;
    const size_t nt = (size_t)P.tw * P.th, nm = (size_t)P.mw * P.mh;
    if (texel >= (int)nm) return;
    const size_t s = blockIdx.y;
    const float4* Ts = T + s * nt;
    const float* TTs = TT + s * 3 * NT * nt;
    float* out = MST + s * 3 * NT * nm;
    if (P.ms_disabled) {
        if (threadIdx.x == 0) store_tan(out, nm, c, texel, td(0.0));
        return;
    }
    const AtmD& a = P.a;
    int x = texel % P.mw, y = texel / P.mw;
    double u = (x + 0.5) / P.mw, v = (y + 0.5) / P.mh;
    double sun_cos = clampd(u * 2.0 - 1.0, -1.0, 1.0);
    double sun_sin = sqrt(fmax(1.0 - sun_cos * sun_cos, 0.0));
    double r = a.R + v * (a.top - a.R);
    V p0 = v3(0, r, 0), sun = v3(sun_sin, sun_cos, 0);
    double h0 = r - a.R;
    const int N_DIRS = 64, N_STEPS = 20;
    int di = threadIdx.x;
    double z = 1.0 - 2.0 * (di + 0.5) / N_DIRS;
    double rr = sqrt(fmax(1.0 - z * z, 0.0));
    double fr = di * 0.6180339887498949;
    double phi = 2.0 * PS_PI * (fr - floor(fr));
    V vd = v3(rr * cos(phi), rr * sin(phi), z);
    TD l = td(0.0), f = td(0.0);
    double t_max = dist_boundary(a, p0, vd);
    if (t_max > 0.0) {
        double dt = t_max / N_STEPS;
        double cv = vdot(vscl(p0, 1.0 / fmax(r, 1.0)), vd);
        TD tr = td(1.0);
        for (int st_i = 0; st_i < N_STEPS; ++st_i) {
            double t = (st_i + 0.5) * dt;
            V p = vadd(p0, vscl(vd, t));
            double num = 2.0 * t * r * cv + t * t;
            double r_delta = num / (fmax(r, 1.0) + vlen(p));
            double h = h0 + r_delta;
            const Medium m = medium(a, c, h);
            const TD trs = exp_neg_scaled(m.st, dt);
            double r_p = r + r_delta;
            double cos_s = vdot(p, sun) / fmax(r_p, 1.0);
            double ratio = a.R / r_p;
            TD vis = td(0.0);
            if (cos_s > -sqrt(fmax(1.0 - ratio * ratio, 0.0))) vis = trans_tap(P, Ts, TTs, c, p, sun);
            TD ins = td(m.ss.v * vis.v);
#pragma unroll
            for (int j = 0; j < NT; ++j) ins.d[j] = m.ss.d[j] * vis.v + m.ss.v * vis.d[j];
            const TD si = step_integral(ins, m.st, dt, trs);
            const TD fi = step_integral(m.ss, m.st, dt, trs);
            l.v += tr.v * si.v;
            f.v += tr.v * fi.v;
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                l.d[j] += tr.d[j] * si.v + tr.v * si.d[j];
                f.d[j] += tr.d[j] * fi.v + tr.v * fi.d[j];
            }
#pragma unroll
            for (int j = 0; j < NT; ++j) tr.d[j] = tr.d[j] * trs.v + tr.v * trs.d[j];
            tr.v *= trs.v;
        }
    }
    // wave64 sums of l, f and their tangents
    double acc[2 + 2 * NT];
    acc[0] = l.v; acc[1] = f.v;
#pragma unroll
    for (int j = 0; j < NT; ++j) { acc[2 + j] = l.d[j]; acc[2 + NT + j] = f.d[j]; }
#pragma unroll
    for (int k = 0; k < 2 + 2 * NT; ++k) {
        double sacc = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) sacc += __shfl_xor(sacc, off, 64);
        acc[k] = sacc;
    }
    if (threadIdx.x == 0) {
        const double norm = (4.0 * PS_PI) / N_DIRS * (1.0 / (4.0 * PS_PI));
        const double L2 = acc[0] * norm, Fraw = acc[1] * norm;
        const double F = clampd(Fraw, 0.0, 0.95);
        const bool lin = Fraw > 0.0 && Fraw < 0.95;
        TD psi = td(L2 / (1.0 - F));
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const double dL2 = acc[2 + j] * norm, dF = lin ? acc[2 + NT + j] * norm : 0.0;
            psi.d[j] = dL2 / (1.0 - F) + L2 * dF / ((1.0 - F) * (1.0 - F));
        }
        store_tan(out, nm, c, texel, psi);
    }
}

// In-scatter march of one channel with tangents (skyview.comp.wgsl:112-147,
// aerialperspective.comp.wgsl:259-293).
template <bool AP>
__device__ __forceinline__ void inscatter_tan(const LutParams& P, const float4* T, const float* TT, const float4* MS,
                                              const float* MST, int c, V p, double r, V vd, double dt, int n_steps,
                                              double pr, double pm, TD& lum, TD& tr) {
    const AtmD& a = P.a;
    V sun = v3(P.f.sun[0], P.f.sun[1], P.f.sun[2]);
    double cv = vdot(vscl(p, 1.0 / fmax(r, 1.0)), vd);
    double h0 = r - a.R;
    lum = td(0.0);
    tr = td(1.0);
    const double E = AP ? P.f.E[c] : 1.0;
    for (int s = 0; s < n_steps; ++s) {
        double t = (s + 0.5) * dt;
        V pi = vadd(p, vscl(vd, t));
        double num = 2.0 * t * r * cv + t * t;
        double r_delta = num / (fmax(r, 1.0) + vlen(pi));
        double h = h0 + r_delta;
        if (AP && h < 0.0) break;
        const Medium m = medium(a, c, h);
        const TD trs = exp_neg_scaled(m.st, dt);
        const TD vis = trans_tap(P, T, TT, c, pi, sun);
        double r_p = r + r_delta;
        double hn = clampd(h / fmax(a.top - a.R, 1.0), 0.0, 1.0);
        double sc = clampd(vdot(vscl(pi, 1.0 / fmax(r_p, 1.0)), sun), -1.0, 1.0);
        const TD ms = tap2(MS, MST, P.mw, P.mh, c, sc * 0.5 + 0.5, hn);
        // ins = ((sr pr + sm pm) vis + (sr + sm) ms) E
        const double kv = m.sr.v * pr + m.sm.v * pm, sv = m.sr.v + m.sm.v;
        TD ins = td((kv * vis.v + sv * ms.v) * E);
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const double dk = m.sr.d[j] * pr + m.sm.d[j] * pm, ds = m.sr.d[j] + m.sm.d[j];
            ins.d[j] = (dk * vis.v + kv * vis.d[j] + ds * ms.v + sv * ms.d[j]) * E;
        }
        const TD si = step_integral(ins, m.st, dt, trs);
        lum.v += tr.v * si.v;
#pragma unroll
        for (int j = 0; j < NT; ++j) lum.d[j] += tr.d[j] * si.v + tr.v * si.d[j];
#pragma unroll
        for (int j = 0; j < NT; ++j) tr.d[j] = tr.d[j] * trs.v + tr.v * trs.d[j];
        tr.v *= trs.v;
    }
}

// ---- sky-view tangents (skyview.comp.wgsl:46-174) ---------------------------------
__global__ __launch_bounds__(64) void k_skyview_tan(const LutParams* __restrict__ Ps, const float4* __restrict__ T,
                                                    const float* __restrict__ TT, const float4* __restrict__ MS,
                                                    const float* __restrict__ MST, float* __restrict__ SVT) {
    const LutParams& P = Ps[blockIdx.y];
    const int c = blockIdx.z;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nt = (size_t)P.tw * P.th, nm = (size_t)P.mw * P.mh, ns = (size_t)P.sw * P.sh;
    if (i >= (int)ns) return;
    const size_t sI = blockIdx.y;
    const float4* Ts = T + sI * nt;
    const float* TTs = TT + sI * 3 * NT * nt;
    const float4* MSs = MS + sI * nm;
    const float* MSTs = MST + sI * 3 * NT * nm;
    float* out = SVT + sI * 3 * NT * ns;
    const AtmD& a = P.a;
    int x = i % P.sw, y = i / P.sw;
    double u = (x + 0.5) / P.sw, v = (y + 0.5) / P.sh;
    V p = vadd(v3(P.f.cam[0], P.f.cam[1], P.f.cam[2]), v3(0, a.R, 0));
    double r = vlen(p);
    if (r < a.R + 1.0) {
        p = vscl(p, (a.R + 1.0) / r);
        r = a.R + 1.0;
    }
    double sin_h = clampd(a.R / r, 0.0, 1.0);
    double zha = PS_PI - asin(sin_h);
    double vza;
    if (v < 0.5) {
        double cc = v * 2.0;
        vza = cc * cc * zha;
    } else {
        double cc = (v - 0.5) * 2.0;
        vza = zha + cc * cc * (PS_PI - zha);
    }
    V sun = v3(P.f.sun[0], P.f.sun[1], P.f.sun[2]);
    double az = atan2(sun.x, sun.z) + u * 2.0 * PS_PI;
    double cvz = cos(vza), svz = sin(vza);
    V vd = v3(svz * sin(az), cvz, svz * cos(az));
    double ct = vdot(vd, sun);
    double pr = (3.0 / (16.0 * PS_PI)) * (1.0 + ct * ct);
    double g = a.mie_g, g2 = g * g;
    double den = fmax(1.0 + g2 - 2.0 * g * ct, 1e-4);
    double pm = (1.0 - g2) / (4.0 * PS_PI * den * sqrt(den));
    double t_max = dist_boundary(a, p, vd);
    if (t_max <= 0.0) {
        store_tan(out, ns, c, i, td(0.0));
        return;
    }
    TD lum, tr;
    inscatter_tan<false>(P, Ts, TTs, MSs, MSTs, c, p, r, vd, t_max / 32.0, 32, pr, pm, lum, tr);
    double tp0, tp1;
    if (sphere_hit(p, vd, a.R, tp0, tp1) && tp0 > 0.0) {
        V pg = vadd(p, vscl(vd, tp0));
        V ng = vscl(pg, 1.0 / vlen(pg));
        const TD sv = trans_tap(P, Ts, TTs, c, pg, sun);
        double ndl = fmax(vdot(ng, sun), 0.0);
        const double k = a.albedo[c] * (1.0 / 3.14159265) * ndl;
        lum.v += tr.v * k * sv.v;
#pragma unroll
        for (int j = 0; j < NT; ++j) lum.d[j] += tr.d[j] * k * sv.v + tr.v * k * sv.d[j];
    }
    store_tan(out, ns, c, i, lum);
}

// ---- aerial-perspective tangents (aerialperspective.comp.wgsl:41-127): the
// in-scatter (tangent rows 0..5) and the transmittance (rows 6..11) of the
// channel; the luminance-weighted alpha follows from the transmittances. ------
__global__ __launch_bounds__(64) void k_aerial_tan(const LutParams* __restrict__ Ps, const float4* __restrict__ T,
                                                   const float* __restrict__ TT, const float4* __restrict__ MS,
                                                   const float* __restrict__ MST, float* __restrict__ APT) {
    const LutParams& P = Ps[blockIdx.y];
    const int c = blockIdx.z;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nt = (size_t)P.tw * P.th, nm = (size_t)P.mw * P.mh, na = (size_t)P.aw * P.ah * P.ad;
    if (i >= (int)na) return;
    const size_t sI = blockIdx.y;
    const float4* Ts = T + sI * nt;
    const float* TTs = TT + sI * 3 * NT * nt;
    const float4* MSs = MS + sI * nm;
    const float* MSTs = MST + sI * 3 * NT * nm;
    float* out = APT + sI * 3 * 2 * NT * na;
    const AtmD& a = P.a;
    int x = i % P.aw, y = (i / P.aw) % P.ah, z = i / (P.aw * P.ah);
    double ndx = (x + 0.5) / P.aw * 2.0 - 1.0, ndy = (y + 0.5) / P.ah * 2.0 - 1.0;
    const double* m = P.f.inv_vp;
    double hv[4] = {ndx, -ndy, 1.0, 1.0}, nh[4];
    for (int rI = 0; rI < 4; ++rI) nh[rI] = m[rI] * hv[0] + m[4 + rI] * hv[1] + m[8 + rI] * hv[2] + m[12 + rI] * hv[3];
    V cam = v3(P.f.cam[0], P.f.cam[1], P.f.cam[2]);
    V np = v3(nh[0] / nh[3], nh[1] / nh[3], nh[2] / nh[3]);
    V vd = v3(np.x - cam.x, np.y - cam.y, np.z - cam.z);
    vd = vscl(vd, 1.0 / vlen(vd));
    double zn = (z + 0.5) / P.ad;
    double t_target = 50.0 * pow(100000.0 / 50.0, zn);
    V p0 = vadd(cam, v3(0, a.R, 0));
    double r0 = vlen(p0);
    V sun = v3(P.f.sun[0], P.f.sun[1], P.f.sun[2]);
    double ct = vdot(vd, sun);
    double pr = (3.0 / (16.0 * PS_PI)) * (1.0 + ct * ct);
    double g = a.mie_g, g2 = g * g;
    double den = fmax(1.0 + g2 - 2.0 * g * ct, 1e-4);
    double pm = (1.0 - g2) / (4.0 * PS_PI * den * sqrt(den));
    TD lum, tr;
    inscatter_tan<true>(P, Ts, TTs, MSs, MSTs, c, p0, r0, vd, t_target / 16.0, 16, pr, pm, lum, tr);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        out[((size_t)c * 2 * NT + j) * na + i] = (float)lum.d[j];
        out[((size_t)c * 2 * NT + NT + j) * na + i] = (float)tr.d[j];
    }
}

// Coefficients of each set from a device array (n x 14: beta_R[3],
// beta_Ms[3], beta_Ma[3], beta_O[3], H_R, H_M) into its LUT parameter block.
__global__ void k_set_coeffs(LutParams* __restrict__ Ps, const float* __restrict__ k, int n) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float* q = k + (size_t)s * 14;
    AtmD& a = Ps[s].a;
    for (int c = 0; c < 3; ++c) {
        a.ray_s[c] = q[c];
        a.mie_s[c] = q[3 + c];
        a.mie_a[c] = q[6 + c];
        a.ozone[c] = q[9 + c];
    }
    a.hr = q[12];
    a.hm = q[13];
}

__global__ void k_zero(float* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// ---- sky + AP image, its L2 loss against a target, and the gradient ----------------
// Per pixel: the sky pass (sky.wgsl:159-220: sky-view LUT x E, sun disk;
// no cloud layers, so no overcast blend) where the depth buffer is far, and
// L = L_ground * ap.a + ap.rgb at the surface distance where it holds a
// surface (the AP application of pbr.wgsl:653-655 to a constant-radiance
// surface). Loss sum over pixels and RGB of (L - target)^2.
struct SkyApArgs {
    const LutParams* lp;              // per set (frame shared)
    const float4 *T, *SV, *AP;        // value LUTs, set-major
    const float *TT, *SVT, *APT;      // tangent LUTs, set-major
    const float* depth;               // W*H reverse-Z NDC or null
    const float* target;              // W*H RGBA
    const float* weight;              // W*H per-pixel loss weights, or null (1)
    float* out;                       // per set 16 floats: 14 gradients, loss, unused
    float* image;                     // optional W*H RGBA image of set 0 (or null)
    int W, H;
    int want_grad;                    // 0: image only (no target, no tangents)
    float ground;                     // surface radiance
    // frame geometry (FP64 view rays, FP32 LUT coordinates as the sky pass)
    double inv_vp[16], cam[3], sun_d[3], cos_disk, sun_ang;
    float sun[3], E[3], cam_zha, sun_az, cam_alt, R, top, big_h, top_minus_R;
};

__device__ __forceinline__ float clampf_(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// transmittance-LUT coordinates from altitude (the cloud pass's form)
__device__ __forceinline__ bool trans_uv_h(const SkyApArgs& A, float h, float mu, float& u, float& v) {
    float r = A.R + h;
    float rho = sqrtf(fmaxf(h * (2.f * A.R + h), 0.f));
    if (mu <= -rho / r) return false;
    float dtop = A.top_minus_R - h;
    float disc = dtop * (A.top + r) + (r * mu) * (r * mu);
    float sq = sqrtf(fmaxf(disc, 0.f));
    float d = mu >= 0.f ? (dtop * (A.top + r)) / fmaxf(r * mu + sq, 1e-30f) : -r * mu + sq;
    d = fmaxf(d, 0.f);
    float d_min = dtop, d_max = rho + A.big_h;
    u = clampf_((d - d_min) / fmaxf(d_max - d_min, 1.f), 0.f, 1.f);
    v = rho / fmaxf(A.big_h, 1.f);
    return true;
}

// bilinear ClampToEdge corners/weights (float coordinates, as lut2d)
struct Tap4 { int i[4]; float w[4]; float fx, fy; };
__device__ __forceinline__ Tap4 tap4(int w, int h, float u, float v) {
    float cx = u * (float)w - 0.5f, cy = v * (float)h - 0.5f;
    float fx0 = floorf(cx), fy0 = floorf(cy);
    Tap4 t;
    t.fx = cx - fx0; t.fy = cy - fy0;
    int ix = (int)fx0, iy = (int)fy0;
    int x0 = clampi(ix, w), x1 = clampi(ix + 1, w), y0 = clampi(iy, h), y1 = clampi(iy + 1, h);
    t.i[0] = y0 * w + x0; t.i[1] = y0 * w + x1; t.i[2] = y1 * w + x0; t.i[3] = y1 * w + x1;
    t.w[0] = (1.f - t.fx) * (1.f - t.fy); t.w[1] = t.fx * (1.f - t.fy); t.w[2] = (1.f - t.fx) * t.fy;
    t.w[3] = t.fx * t.fy;
    return t;
}

constexpr int kPixPerThread = 4;

__global__ __launch_bounds__(256) void k_sky_ap_grad(SkyApArgs A) {
    const int set = blockIdx.y;
    const LutParams& P = A.lp[set];
    const size_t nt = (size_t)P.tw * P.th, ns = (size_t)P.sw * P.sh, na = (size_t)P.aw * P.ah * P.ad;
    const float4* T = A.T + set * nt;
    const float4* SV = A.SV + set * ns;
    const float4* AP = A.AP + set * na;
    const float* TT = A.TT + set * 3 * NT * nt;
    const float* SVT = A.SVT + set * 3 * NT * ns;
    const float* APT = A.APT + set * 3 * 2 * NT * na;
    const float lw[3] = {0.2126f, 0.7152f, 0.0722f};
    float G[3][NT];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < NT; ++j) G[c][j] = 0.f;
    float loss = 0.f;
    const int npix = A.W * A.H;
    const int base = (blockIdx.x * blockDim.x) * kPixPerThread + threadIdx.x;
    for (int k = 0; k < kPixPerThread; ++k) {
        const int pix = base + k * blockDim.x;
        if (pix >= npix) break;
        const int xc = pix % A.W, yc = pix / A.W;
        const double fx = xc + 0.5, fy = yc + 0.5;
        const double ndx = fx / A.W * 2.0 - 1.0, ndy = 1.0 - fy / A.H * 2.0;
        const double* m = A.inv_vp;
        const float4 tg = A.want_grad ? ((const float4*)A.target)[pix] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float wgt = A.weight ? A.weight[pix] : 1.f;
        float dn = A.depth ? A.depth[pix] : 0.f;
        if (dn > 0.f) {
            // surface distance (cloud_march.wgsl:764-775 unprojection, FP64)
            double w0 = m[0] * ndx + m[4] * ndy + m[8] * dn + m[12];
            double w1 = m[1] * ndx + m[5] * ndy + m[9] * dn + m[13];
            double w2 = m[2] * ndx + m[6] * ndy + m[10] * dn + m[14];
            double w3 = m[3] * ndx + m[7] * ndy + m[11] * dn + m[15];
            double qx = w0 / w3 - A.cam[0], qy = w1 / w3 - A.cam[1], qz = w2 / w3 - A.cam[2];
            const float d = (float)sqrt(qx * qx + qy * qy + qz * qz);
            const float ds = fmaxf(d, 50.f);
            const float apz = clampf_(__logf(ds / 50.f) * (float)(1.0 / 7.6009024595420822), 0.f, 1.f);
            const float u = (float)fx / A.W, v = (float)fy / A.H;
            const float cx = u * P.aw - 0.5f, cy = v * P.ah - 0.5f, cz = apz * P.ad - 0.5f;
            const float fx0 = floorf(cx), fy0 = floorf(cy), fz0 = floorf(cz);
            const float wx = cx - fx0, wy = cy - fy0, wz = cz - fz0;
            const int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
            const int xs[2] = {clampi(ix, P.aw), clampi(ix + 1, P.aw)}, ys[2] = {clampi(iy, P.ah), clampi(iy + 1, P.ah)},
                      zs[2] = {clampi(iz, P.ad), clampi(iz + 1, P.ad)};
            const int sxy = P.aw * P.ah;
            int id8[8];
            float4 cv8[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                id8[q] = zs[q >> 2] * sxy + ys[(q >> 1) & 1] * P.aw + xs[q & 1];
                cv8[q] = AP[id8[q]];
            }
            // value in the forward's lerp form
            float ap[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const float v000 = ((const float*)&cv8[0])[kk], v100 = ((const float*)&cv8[1])[kk];
                const float v010 = ((const float*)&cv8[2])[kk], v110 = ((const float*)&cv8[3])[kk];
                const float v001 = ((const float*)&cv8[4])[kk], v101 = ((const float*)&cv8[5])[kk];
                const float v011 = ((const float*)&cv8[6])[kk], v111 = ((const float*)&cv8[7])[kk];
                const float a0 = v000 + (v100 - v000) * wx, a1 = v010 + (v110 - v010) * wx;
                const float b0 = v001 + (v101 - v001) * wx, b1 = v011 + (v111 - v011) * wx;
                const float p0 = a0 + (a1 - a0) * wy, p1 = b0 + (b1 - b0) * wy;
                ap[kk] = p0 + (p1 - p0) * wz;
            }
            const float L[3] = {A.ground * ap[3] + ap[0], A.ground * ap[3] + ap[1], A.ground * ap[3] + ap[2]};
            if (A.image && set == 0) ((float4*)A.image)[pix] = make_float4(L[0], L[1], L[2], 1.f);
            if (!A.want_grad) continue;
            const float r[3] = {L[0] - tg.x, L[1] - tg.y, L[2] - tg.z};
            loss += wgt * (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            // dL_c = d ap_c + L_ground * sum_k lw_k dT_k: channel c's
            // parameters reach the in-scatter of channel c and the shared alpha
            const float gsum = 2.f * wgt * (r[0] + r[1] + r[2]) * A.ground;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float wq = (q & 1 ? wx : 1.f - wx) * ((q >> 1) & 1 ? wy : 1.f - wy) * (q >> 2 ? wz : 1.f - wz);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float gl = wq * 2.f * wgt * r[c], gt = wq * gsum * lw[c];
#pragma unroll
                    for (int j = 0; j < NT; ++j)
                        G[c][j] += gl * APT[((size_t)c * 2 * NT + j) * na + id8[q]] +
                                   gt * APT[((size_t)c * 2 * NT + NT + j) * na + id8[q]];
                }
            }
        } else {
            // sky pass without cloud layers: sky-view LUT x E plus the sun disk
            double dxd = m[0] * ndx + m[4] * ndy + m[8] + m[12], dyd = m[1] * ndx + m[5] * ndy + m[9] + m[13];
            double dzd = m[2] * ndx + m[6] * ndy + m[10] + m[14], dwd = m[3] * ndx + m[7] * ndy + m[11] + m[15];
            dxd = dxd / dwd - A.cam[0]; dyd = dyd / dwd - A.cam[1]; dzd = dzd / dwd - A.cam[2];
            const double il = 1.0 / sqrt(dxd * dxd + dyd * dyd + dzd * dzd);
            dxd *= il; dyd *= il; dzd *= il;
            const float dx = (float)dxd, dy = (float)dyd, dz = (float)dzd;
            const double cos_vs = dxd * A.sun_d[0] + dyd * A.sun_d[1] + dzd * A.sun_d[2];
            float rxz = sqrtf(dx * dx + dz * dz);
            float vza = atan2f(rxz, dy);
            float u, v;
            if (vza < A.cam_zha) v = sqrtf(clampf_(vza / fmaxf(A.cam_zha, 1e-6f), 0.f, 1.f)) * 0.5f;
            else v = 0.5f + sqrtf(clampf_((vza - A.cam_zha) / fmaxf((float)PS_PI - A.cam_zha, 1e-6f), 0.f, 1.f)) * 0.5f;
            float du = (atan2f(dx, dz) - A.sun_az) * (float)(1.0 / (2.0 * PS_PI));
            u = du - floorf(du);
            const float4 s = lut2d(SV, P.sw, P.sh, u, v);
            float L[3] = {s.x * A.E[0], s.y * A.E[1], s.z * A.E[2]};
            bool disk = false;
            float k2[3] = {0.f, 0.f, 0.f}, tu = 0.f, tv = 0.f;
            if (cos_vs > A.cos_disk) {
                double theta = acos(clampd(cos_vs, -1.0, 1.0));
                double tn = theta / fmax(A.sun_ang, 1e-6);
                double limb = 1.0 - 0.6 * (1.0 - sqrt(fmax(1.0 - tn * tn, 0.0)));
                double omega = PS_PI * A.sun_ang * A.sun_ang;
                double sc = limb / fmax(omega, 1e-6);
                float mu_c = (float)((A.cam[0] * A.sun_d[0] + (A.cam[1] + A.R) * A.sun_d[1] + A.cam[2] * A.sun_d[2]) /
                                     (A.R + (double)A.cam_alt));
                if (trans_uv_h(A, A.cam_alt, mu_c, tu, tv)) {
                    disk = true;
                    const float4 tt = lut2d(T, P.tw, P.th, tu, tv);
                    const float vis[3] = {tt.x, tt.y, tt.z};
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        k2[c] = (float)(A.E[c] * sc);
                        L[c] += k2[c] * vis[c];
                    }
                }
            }
            bool pos[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                pos[c] = L[c] > 0.f;  // max(L, 0) passes the derivative only above 0
                L[c] = fmaxf(L[c], 0.f);
            }
            if (A.image && set == 0) ((float4*)A.image)[pix] = make_float4(L[0], L[1], L[2], 1.f);
            if (!A.want_grad) continue;
            float gr[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float rc = L[c] - ((const float*)&tg)[c];
                loss += wgt * rc * rc;
                gr[c] = pos[c] ? 2.f * wgt * rc : 0.f;
            }
            const Tap4 tp = tap4(P.sw, P.sh, u, v);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float g = gr[c] * A.E[c];
#pragma unroll
                for (int j = 0; j < NT; ++j) {
                    const float* t = SVT + ((size_t)c * NT + j) * ns;
                    G[c][j] += g * (tp.w[0] * t[tp.i[0]] + tp.w[1] * t[tp.i[1]] + tp.w[2] * t[tp.i[2]] +
                                    tp.w[3] * t[tp.i[3]]);
                }
            }
            if (disk) {
                const Tap4 tq = tap4(P.tw, P.th, tu, tv);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float g = gr[c] * k2[c];
#pragma unroll
                    for (int j = 0; j < NT; ++j) {
                        const float* t = TT + ((size_t)c * NT + j) * nt;
                        G[c][j] += g * (tq.w[0] * t[tq.i[0]] + tq.w[1] * t[tq.i[1]] + tq.w[2] * t[tq.i[2]] +
                                        tq.w[3] * t[tq.i[3]]);
                    }
                }
            }
        }
    }
    if (!A.want_grad) return;
    // block reduction: wave sums, then one atomic per value per wave
    float vals[3 * NT + 1];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < NT; ++j) vals[c * NT + j] = G[c][j];
    vals[3 * NT] = loss;
#pragma unroll
    for (int q = 0; q < 3 * NT + 1; ++q) {
        float s = vals[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        vals[q] = s;
    }
    if ((threadIdx.x & 63) == 0) {
        float* o = A.out + (size_t)set * 16;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            atomicAdd(o + 0 + c, vals[c * NT + 0]);   // beta_R[c]
            atomicAdd(o + 3 + c, vals[c * NT + 1]);   // beta_Ms[c]
            atomicAdd(o + 6 + c, vals[c * NT + 2]);   // beta_Ma[c]
            atomicAdd(o + 9 + c, vals[c * NT + 3]);   // beta_O[c]
        }
        atomicAdd(o + 12, vals[4] + vals[NT + 4] + vals[2 * NT + 4]);  // H_R
        atomicAdd(o + 13, vals[5] + vals[NT + 5] + vals[2 * NT + 5]);  // H_M
        atomicAdd(o + 14, vals[3 * NT]);
    }
}

}  // namespace

hipError_t ps_launch_set_coeffs(LutParams* d_params, const float* coeffs, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_set_coeffs, dim3((n + 63) / 64), dim3(64), 0, s, d_params, coeffs, n);
    return hipGetLastError();
}

hipError_t ps_launch_lut_tangents(const LutParams* d_params, const LutParams& P, int n, const float4* T,
                                  const float4* MS, float* TT, float* MST, float* SVT, float* APT, hipStream_t s) {
    const unsigned ns = (unsigned)n;
    hipLaunchKernelGGL(k_transmittance_tan, dim3((P.tw * P.th + 63) / 64, ns, 3), dim3(64), 0, s, d_params, TT);
    hipLaunchKernelGGL(k_multiscatter_tan, dim3(P.mw * P.mh, ns, 3), dim3(64), 0, s, d_params, T, TT, MST);
    hipLaunchKernelGGL(k_skyview_tan, dim3((P.sw * P.sh + 63) / 64, ns, 3), dim3(64), 0, s, d_params, T, TT, MS, MST,
                       SVT);
    hipLaunchKernelGGL(k_aerial_tan, dim3((P.aw * P.ah * P.ad + 63) / 64, ns, 3), dim3(64), 0, s, d_params, T, TT, MS,
                       MST, APT);
    return hipGetLastError();
}

hipError_t ps_launch_sky_ap_grad(const SkyApLaunch& L, int n, hipStream_t s) {
    SkyApArgs A;
    A.lp = L.lp; A.T = L.T; A.SV = L.SV; A.AP = L.AP; A.TT = L.TT; A.SVT = L.SVT; A.APT = L.APT;
    A.depth = L.depth; A.target = L.target; A.weight = L.weight; A.out = L.out; A.image = L.image;
    A.W = L.W; A.H = L.H; A.ground = L.ground; A.want_grad = L.want_grad;
    for (int i = 0; i < 16; ++i) A.inv_vp[i] = L.inv_vp[i];
    for (int i = 0; i < 3; ++i) {
        A.cam[i] = L.cam[i]; A.sun_d[i] = L.sun_d[i]; A.sun[i] = L.sun[i]; A.E[i] = L.E[i];
    }
    A.cos_disk = L.cos_disk; A.sun_ang = L.sun_ang;
    A.cam_zha = L.cam_zha; A.sun_az = L.sun_az; A.cam_alt = L.cam_alt;
    A.R = L.R; A.top = L.top; A.big_h = L.big_h; A.top_minus_R = L.top_minus_R;
    if (L.want_grad) hipLaunchKernelGGL(k_zero, dim3((n * 16 + 255) / 256), dim3(256), 0, s, L.out, n * 16);
    const int npix = L.W * L.H;
    const int blocks = (npix + 256 * kPixPerThread - 1) / (256 * kPixPerThread);
    hipLaunchKernelGGL(k_sky_ap_grad, dim3(blocks, n), dim3(256), 0, s, A);
    return hipGetLastError();
}
