// Atmosphere LUT chain for gfx950: transmittance -> multi-scattering ->
// {sky-view, aerial perspective} (Hillaire 2020), the four compute passes of
// crates/ps-atmosphere/src/lib.rs:494-584.
//
// These kernels run in FP64. The whole chain is ~2.7e6 march steps per frame
// (vs ~1e8 for the cloud march) and MI355X executes FP64 FMA at the FP32
// non-packed rate, so double precision costs a few microseconds and removes
// the catastrophic cancellations of r^2 - R^2 at planet scale from the
// tables the march reads. Results are stored as FP32 RGBA.
#include "ps_device.h"
#include "ps_launch.h"

namespace {

struct D3 { double x, y, z; };
__device__ __forceinline__ D3 d3(double x, double y, double z) { D3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ D3 add(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 sub(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 mul(D3 a, D3 b) { return d3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ D3 scl(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double len(D3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ D3 vexp_neg(D3 a) { return d3(exp(-a.x), exp(-a.y), exp(-a.z)); }
__device__ __forceinline__ D3 ld3(const double* p) { return d3(p[0], p[1], p[2]); }

// atmosphere.wgsl:23-46
__device__ __forceinline__ void density(const AtmD& a, double h, double& dr, double& dm, double& dozn) {
    dr = exp(-h / a.hr);
    dm = exp(-h / a.hm);
    double half_t = a.oz_t * 0.5;
    double dz = fabs(h - a.oz_c);
    dozn = fmax(0.0, 1.0 - dz / fmax(half_t, 1.0));
}
__device__ __forceinline__ D3 extinction(const AtmD& a, double dr, double dm, double dozn) {
    return d3(a.ray_s[0] * dr + (a.mie_s[0] + a.mie_a[0]) * dm + a.ozone[0] * dozn,
              a.ray_s[1] * dr + (a.mie_s[1] + a.mie_a[1]) * dm + a.ozone[1] * dozn,
              a.ray_s[2] * dr + (a.mie_s[2] + a.mie_a[2]) * dm + a.ozone[2] * dozn);
}

// atmosphere.wgsl:67-75
__device__ __forceinline__ double phase_rayleigh(double c) { return (3.0 / (16.0 * PS_PI)) * (1.0 + c * c); }
__device__ __forceinline__ double phase_mie(double c, double g) {
    double g2 = g * g;
    double den = fmax(1.0 + g2 - 2.0 * g * c, 1e-4);
    return (1.0 - g2) / (4.0 * PS_PI * den * sqrt(den));
}

// atmosphere.wgsl:95-104
__device__ __forceinline__ bool sphere_hit(D3 ro, D3 rd, double r, double& t0, double& t1) {
    double b = dot(ro, rd);
    double c = dot(ro, ro) - r * r;
    double disc = b * b - c;
    if (disc < 0.0) return false;
    double sq = sqrt(disc);
    t0 = -b - sq; t1 = -b + sq;
    return true;
}
// atmosphere.wgsl:108-132
__device__ __forceinline__ double dist_top(const AtmD& a, D3 p, D3 d) {
    double t0, t1;
    if (!sphere_hit(p, d, a.top, t0, t1)) return 0.0;
    return fmax(t1, 0.0);
}
__device__ __forceinline__ double dist_boundary(const AtmD& a, D3 p, D3 d) {
    double tt0, tt1, tp0, tp1;
    bool ht = sphere_hit(p, d, a.top, tt0, tt1);
    bool hp = sphere_hit(p, d, a.R, tp0, tp1);
    if (!ht) return 0.0;
    double tm = tt1;
    if (hp && tp0 > 0.0) tm = fmin(tm, tp0);
    return fmax(tm, 0.0);
}

// Bilinear ClampToEdge tap of an FP32 LUT, accumulated in double.
__device__ __forceinline__ D3 lut_d(const float4* __restrict__ t, int w, int h, double u, double v) {
    double cx = u * w - 0.5, cy = v * h - 0.5;
    double fx0 = floor(cx), fy0 = floor(cy);
    double fx = cx - fx0, fy = cy - fy0;
    int ix = (int)fx0, iy = (int)fy0;
    int x0 = clampi(ix, w), x1 = clampi(ix + 1, w), y0 = clampi(iy, h), y1 = clampi(iy + 1, h);
    float4 A = t[y0 * w + x0], B = t[y0 * w + x1], C = t[y1 * w + x0], Dd = t[y1 * w + x1];
    D3 r;
    double tx, bx;
    tx = A.x + (B.x - (double)A.x) * fx; bx = C.x + (Dd.x - (double)C.x) * fx; r.x = tx + (bx - tx) * fy;
    tx = A.y + (B.y - (double)A.y) * fx; bx = C.y + (Dd.y - (double)C.y) * fx; r.y = tx + (bx - tx) * fy;
    tx = A.z + (B.z - (double)A.z) * fx; bx = C.z + (Dd.z - (double)C.z) * fx; r.z = tx + (bx - tx) * fy;
    return r;
}

// atmosphere.wgsl:146-159 + atmosphere_lut_sampling.wgsl:6-22
__device__ __forceinline__ D3 sample_trans(const LutParams& P, const float4* __restrict__ T, D3 p, D3 dir) {
    const AtmD& a = P.a;
    double t0, t1;
    if (sphere_hit(p, dir, a.R, t0, t1) && t0 > 0.0) return d3(0, 0, 0);
    double r = len(p);
    double H = sqrt(fmax(a.top * a.top - a.R * a.R, 0.0));
    double rho = sqrt(fmax(r * r - a.R * a.R, 0.0));
    double d = fmax(dist_top(a, p, dir), 0.0);
    double d_min = a.top - r, d_max = rho + H;
    double u = clampd((d - d_min) / fmax(d_max - d_min, 1.0), 0.0, 1.0);
    double v = rho / fmax(H, 1.0);
    return lut_d(T, P.tw, P.th, u, v);
}

__device__ __forceinline__ void store4(float4* out, int idx, D3 v, double w) {
    out[idx] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)w);
}

// ---------------------------------------------------------------------------
// Lane-parallel integration. A LUT bake is a few hundred waves of long
// serial FP64 loops (exp, sqrt, divides): one wave per SIMD at most, so the
// loop latency is exposed. Each texel's march is therefore split into SEG
// contiguous step ranges on SEG neighbouring lanes: optical depths add, and
// an in-scatter integral combines segment results (L_k, T_k), each computed
// from T = 1 at its segment start, as L = sum_k (prod_{j<k} T_j) L_k with an
// exclusive product scan over the segment lanes.
// ---------------------------------------------------------------------------
template <int SEG>
__device__ __forceinline__ double seg_sum(double v) {
#pragma unroll
    for (int o = 1; o < SEG; o <<= 1) v += __shfl_xor(v, o, SEG);
    return v;
}
template <int SEG>
__device__ __forceinline__ D3 seg_sum3(D3 v) { return d3(seg_sum<SEG>(v.x), seg_sum<SEG>(v.y), seg_sum<SEG>(v.z)); }

// Inclusive product scan over the SEG lanes of a group (lane `seg` gets
// prod_{j <= seg} v_j).
template <int SEG>
__device__ __forceinline__ D3 seg_scan_mul3(D3 v, int seg) {
#pragma unroll
    for (int d = 1; d < SEG; d <<= 1) {
        const D3 u = d3(__shfl_up(v.x, d, SEG), __shfl_up(v.y, d, SEG), __shfl_up(v.z, d, SEG));
        if (seg >= d) v = mul(v, u);
    }
    return v;
}

// Combine segment results: returns sum_k (prod_{j<k} T_j) L_k and the total
// transmittance prod_k T_k on every lane of the group.
template <int SEG>
__device__ __forceinline__ void seg_combine(D3 lum, D3 tr, int seg, D3& lum_all, D3& tr_all) {
    const D3 inc = seg_scan_mul3<SEG>(tr, seg);
    D3 exc = d3(__shfl_up(inc.x, 1, SEG), __shfl_up(inc.y, 1, SEG), __shfl_up(inc.z, 1, SEG));
    if (seg == 0) exc = d3(1, 1, 1);
    lum_all = seg_sum3<SEG>(mul(exc, lum));
    tr_all = d3(__shfl(inc.x, SEG - 1, SEG), __shfl(inc.y, SEG - 1, SEG), __shfl(inc.z, SEG - 1, SEG));
}

// ---- transmittance.comp.wgsl:18-31 with atmosphere.wgsl:162-214 ---------------
// 40 trapezoid steps on kTSeg lanes (steps k0 + 1 .. k0 + 40 / kTSeg each,
// the lane re-evaluating the extinction at its first node).
constexpr int kTSeg = 8;
__device__ __forceinline__ void transmittance_body(const LutParams& P, float4* __restrict__ out) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gid / kTSeg, seg = gid % kTSeg;
    const bool valid = i < P.tw * P.th;
    const int ic = valid ? i : 0;
    const AtmD& a = P.a;
    int x = ic % P.tw, y = ic / P.tw;
    double u = (x + 0.5) / P.tw, v = (y + 0.5) / P.th;
    double H = sqrt(fmax(a.top * a.top - a.R * a.R, 0.0));
    double rho = v * H;
    double r = sqrt(rho * rho + a.R * a.R);
    double d_min = a.top - r, d_max = rho + H;
    double d = d_min + u * (d_max - d_min);
    double mu = 1.0;
    if (d != 0.0) mu = clampd((H * H - rho * rho - d * d) / fmax(2.0 * r * d, 1e-6), -1.0, 1.0);
    D3 p = d3(0, r, 0), dir = d3(sqrt(fmax(1.0 - mu * mu, 0.0)), mu, 0);
    constexpr int N = 40, NS = N / kTSeg;
    static_assert(N % kTSeg == 0, "whole steps per segment lane");
    double t_max = dist_boundary(a, p, dir);
    D3 od = d3(0, 0, 0);
    if (valid && t_max > 0.0) {
        double dt = t_max / N;
        double r0 = len(p);
        double cv = dot(scl(p, 1.0 / fmax(r0, 1.0)), dir);
        double h0 = r0 - a.R;
        const int k0 = seg * NS;
        double h = h0;
        if (k0 > 0) {
            const double t = k0 * dt;
            const D3 pi = add(p, scl(dir, t));
            h = h0 + (2.0 * t * r0 * cv + t * t) / (fmax(r0, 1.0) + len(pi));
        }
        double dr, dm, dz;
        density(a, h, dr, dm, dz);
        D3 prev = extinction(a, dr, dm, dz);
        for (int k = k0 + 1; k <= k0 + NS; ++k) {
            double t = k * dt;
            D3 pi = add(p, scl(dir, t));
            double num = 2.0 * t * r0 * cv + t * t;
            double hk = h0 + num / (fmax(r0, 1.0) + len(pi));
            density(a, hk, dr, dm, dz);
            D3 cur = extinction(a, dr, dm, dz);
            od = add(od, scl(add(prev, cur), 0.5 * dt));
            prev = cur;
        }
    }
    od = seg_sum3<kTSeg>(od);
    if (valid && seg == 0) store4(out, i, vexp_neg(od), 1.0);
}

// ---- multiscatter.comp.wgsl:58-156: one 256-thread block per texel; wave w
// marches steps [5w, 5w + 5) of the 64 Fibonacci directions (one per lane);
// the four segments are combined through LDS, the directions by a wave
// reduction. ------------------------------------------------------------------------
constexpr int kMsSeg = 4;
__device__ __forceinline__ void multiscatter_body(const LutParams& P, const float4* __restrict__ T,
                                                  float4* __restrict__ out) {
    __shared__ double part[kMsSeg][9][64];
    int texel = blockIdx.x;
    if (texel >= P.mw * P.mh) return;
    const AtmD& a = P.a;
    const int di = threadIdx.x & 63, seg = threadIdx.x >> 6;  // blockDim.x == 64 * kMsSeg
    if (P.ms_disabled) {
        if (threadIdx.x == 0) out[texel] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    int x = texel % P.mw, y = texel / P.mw;
    double u = (x + 0.5) / P.mw, v = (y + 0.5) / P.mh;
    double sun_cos = clampd(u * 2.0 - 1.0, -1.0, 1.0);
    double sun_sin = sqrt(fmax(1.0 - sun_cos * sun_cos, 0.0));
    double r = a.R + v * (a.top - a.R);
    D3 p0 = d3(0, r, 0), sun = d3(sun_sin, sun_cos, 0);
    double h0 = r - a.R;
    constexpr int N_DIRS = 64, N_STEPS = 20, NS = N_STEPS / kMsSeg;
    static_assert(N_STEPS % kMsSeg == 0, "whole steps per segment");
    double z = 1.0 - 2.0 * (di + 0.5) / N_DIRS;
    double rr = sqrt(fmax(1.0 - z * z, 0.0));
    double fr = di * 0.6180339887498949;
    double phi = 2.0 * PS_PI * (fr - floor(fr));
    D3 vd = d3(rr * cos(phi), rr * sin(phi), z);
    D3 l_dir = d3(0, 0, 0), f_dir = d3(0, 0, 0), tr = d3(1, 1, 1);
    double t_max = dist_boundary(a, p0, vd);
    if (t_max > 0.0) {
        double dt = t_max / N_STEPS;
        double cv = dot(scl(p0, 1.0 / fmax(r, 1.0)), vd);
        for (int s = seg * NS; s < seg * NS + NS; ++s) {
            double t = (s + 0.5) * dt;
            D3 p = add(p0, scl(vd, t));
            double num = 2.0 * t * r * cv + t * t;
            double r_delta = num / (fmax(r, 1.0) + len(p));
            double h = h0 + r_delta;
            double dr, dm, dz;
            density(a, h, dr, dm, dz);
            D3 st = extinction(a, dr, dm, dz);
            D3 ss = d3(a.ray_s[0] * dr + a.mie_s[0] * dm, a.ray_s[1] * dr + a.mie_s[1] * dm,
                       a.ray_s[2] * dr + a.mie_s[2] * dm);
            D3 trs = vexp_neg(scl(st, dt));
            double r_p = r + r_delta;
            double cos_s = dot(p, sun) / fmax(r_p, 1.0);
            double ratio = a.R / r_p;
            D3 vis = d3(0, 0, 0);
            if (cos_s > -sqrt(fmax(1.0 - ratio * ratio, 0.0))) vis = sample_trans(P, T, p, sun);
            D3 safe = d3(fmax(st.x, 1e-7), fmax(st.y, 1e-7), fmax(st.z, 1e-7));
            D3 one_m = d3(1.0 - trs.x, 1.0 - trs.y, 1.0 - trs.z);
            D3 ins = mul(ss, vis);
            l_dir = add(l_dir, mul(tr, d3(ins.x * one_m.x / safe.x, ins.y * one_m.y / safe.y, ins.z * one_m.z / safe.z)));
            f_dir = add(f_dir, mul(tr, d3(ss.x * one_m.x / safe.x, ss.y * one_m.y / safe.y, ss.z * one_m.z / safe.z)));
            tr = mul(tr, trs);
        }
    }
    const double vals[9] = {l_dir.x, l_dir.y, l_dir.z, f_dir.x, f_dir.y, f_dir.z, tr.x, tr.y, tr.z};
#pragma unroll
    for (int k = 0; k < 9; ++k) part[seg][k][di] = vals[k];
    __syncthreads();
    if (seg != 0) return;
    // segments in march order: sums weighted by the transmittance in front
    D3 L = d3(0, 0, 0), F = d3(0, 0, 0), Tp = d3(1, 1, 1);
    for (int k = 0; k < kMsSeg; ++k) {
        L = add(L, mul(Tp, d3(part[k][0][di], part[k][1][di], part[k][2][di])));
        F = add(F, mul(Tp, d3(part[k][3][di], part[k][4][di], part[k][5][di])));
        Tp = mul(Tp, d3(part[k][6][di], part[k][7][di], part[k][8][di]));
    }
    // wave64 reduction of the six sums over the directions
    double v6[6] = {L.x, L.y, L.z, F.x, F.y, F.z};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double s = v6[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        v6[k] = s;
    }
    if (threadIdx.x == 0) {
        double norm = (4.0 * PS_PI) / N_DIRS * (1.0 / (4.0 * PS_PI));
        D3 L2 = d3(v6[0] * norm, v6[1] * norm, v6[2] * norm);
        D3 Fc = d3(clampd(v6[3] * norm, 0.0, 0.95), clampd(v6[4] * norm, 0.0, 0.95), clampd(v6[5] * norm, 0.0, 0.95));
        store4(out, texel, d3(L2.x / (1.0 - Fc.x), L2.y / (1.0 - Fc.y), L2.z / (1.0 - Fc.z)), 1.0);
    }
}

// One in-scatter march segment shared by the sky-view and aerial-perspective
// bakes (skyview.comp.wgsl:112-147, aerialperspective.comp.wgsl:259-293):
// steps [s0, s1) from transmittance 1. AP: stops at the first step below the
// ground (`broke`), as the reference's loop does.
template <bool AP>
__device__ __forceinline__ void inscatter_segment(const LutParams& P, const float4* __restrict__ T,
                                                  const float4* __restrict__ MS, D3 p, double r, D3 vd,
                                                  double dt, int s0, int s1, double phase_r, double phase_m,
                                                  D3& lum, D3& tr, bool& broke) {
    const AtmD& a = P.a;
    D3 sun = ld3(P.f.sun);
    double cv = dot(scl(p, 1.0 / fmax(r, 1.0)), vd);
    double h0 = r - a.R;
    lum = d3(0, 0, 0);
    tr = d3(1, 1, 1);
    broke = false;
    for (int s = s0; s < s1; ++s) {
        double t = (s + 0.5) * dt;
        D3 pi = add(p, scl(vd, t));
        double num = 2.0 * t * r * cv + t * t;
        double r_delta = num / (fmax(r, 1.0) + len(pi));
        double h = h0 + r_delta;
        if (AP && h < 0.0) {
            broke = true;
            break;
        }
        double dr, dm, dz;
        density(a, h, dr, dm, dz);
        D3 st = extinction(a, dr, dm, dz);
        D3 trs = vexp_neg(scl(st, dt));
        D3 vis = sample_trans(P, T, pi, sun);
        D3 sr = d3(a.ray_s[0] * dr, a.ray_s[1] * dr, a.ray_s[2] * dr);
        D3 sm = d3(a.mie_s[0] * dm, a.mie_s[1] * dm, a.mie_s[2] * dm);
        D3 sun_in = mul(add(scl(sr, phase_r), scl(sm, phase_m)), vis);
        double r_p = r + r_delta;
        double hn = clampd(h / fmax(a.top - a.R, 1.0), 0.0, 1.0);
        double sc = clampd(dot(scl(pi, 1.0 / fmax(r_p, 1.0)), sun), -1.0, 1.0);
        D3 ms = lut_d(MS, P.mw, P.mh, sc * 0.5 + 0.5, hn);
        D3 ins = add(sun_in, mul(add(sr, sm), ms));
        if (AP) ins = mul(ins, ld3(P.f.E));
        D3 safe = d3(fmax(st.x, 1e-7), fmax(st.y, 1e-7), fmax(st.z, 1e-7));
        D3 sint = d3(ins.x * (1.0 - trs.x) / safe.x, ins.y * (1.0 - trs.y) / safe.y, ins.z * (1.0 - trs.z) / safe.z);
        lum = add(lum, mul(tr, sint));
        tr = mul(tr, trs);
    }
}

// ---- skyview.comp.wgsl:46-174: 32 steps on kSvSeg lanes per texel ----------------
constexpr int kSvSeg = 8;
__device__ __forceinline__ void skyview_body(const LutParams& P, const float4* __restrict__ T,
                                             const float4* __restrict__ MS, float4* __restrict__ out, int block) {
    const int gid = block * (int)blockDim.x + (int)threadIdx.x;
    const int i = gid / kSvSeg, seg = gid % kSvSeg;
    const bool valid = i < P.sw * P.sh;
    const int ic = valid ? i : 0;
    const AtmD& a = P.a;
    int x = ic % P.sw, y = ic / P.sw;
    double u = (x + 0.5) / P.sw, v = (y + 0.5) / P.sh;
    D3 p = add(ld3(P.f.cam), d3(0, a.R, 0));
    double r = len(p);
    if (r < a.R + 1.0) {
        p = scl(p, (a.R + 1.0) / r);
        r = a.R + 1.0;
    }
    double sin_h = clampd(a.R / r, 0.0, 1.0);
    double zha = PS_PI - asin(sin_h);
    double vza;
    if (v < 0.5) {
        double c = v * 2.0;
        vza = c * c * zha;
    } else {
        double c = (v - 0.5) * 2.0;
        vza = zha + c * c * (PS_PI - zha);
    }
    D3 sun = ld3(P.f.sun);
    double az = atan2(sun.x, sun.z) + u * 2.0 * PS_PI;
    double cvz = cos(vza), svz = sin(vza);
    D3 vd = d3(svz * sin(az), cvz, svz * cos(az));
    double ct = dot(vd, sun);
    double pr = phase_rayleigh(ct), pm = phase_mie(ct, a.mie_g);
    double t_max = dist_boundary(a, p, vd);
    constexpr int N = 32, NS = N / kSvSeg;
    static_assert(N % kSvSeg == 0, "whole steps per segment lane");
    D3 lum_k = d3(0, 0, 0), tr_k = d3(1, 1, 1);
    bool broke;
    if (t_max > 0.0)
        inscatter_segment<false>(P, T, MS, p, r, vd, t_max / N, seg * NS, seg * NS + NS, pr, pm, lum_k, tr_k, broke);
    D3 lum, tr;
    seg_combine<kSvSeg>(lum_k, tr_k, seg, lum, tr);
    if (!valid || seg != 0) return;
    if (t_max <= 0.0) {
        out[i] = make_float4(0.f, 0.f, 0.f, 1.f);
        return;
    }
    double tp0, tp1;
    if (sphere_hit(p, vd, a.R, tp0, tp1) && tp0 > 0.0) {
        D3 pg = add(p, scl(vd, tp0));
        D3 ng = scl(pg, 1.0 / len(pg));
        D3 sv = sample_trans(P, T, pg, sun);
        double ndl = fmax(dot(ng, sun), 0.0);
        D3 bounce = scl(mul(ld3(a.albedo), sv), (1.0 / 3.14159265) * ndl);
        lum = add(lum, mul(tr, bounce));
    }
    store4(out, i, lum, 1.0);
}

// ---- aerialperspective.comp.wgsl:41-127: 16 steps on kApSeg lanes per froxel ------
constexpr int kApSeg = 4;
__device__ __forceinline__ void aerial_body(const LutParams& P, const float4* __restrict__ T,
                                            const float4* __restrict__ MS, float4* __restrict__ out, int block) {
    const int gid = block * (int)blockDim.x + (int)threadIdx.x;
    const int i = gid / kApSeg, seg = gid % kApSeg;
    const bool valid = i < P.aw * P.ah * P.ad;
    const int ic = valid ? i : 0;
    const AtmD& a = P.a;
    int x = ic % P.aw, y = (ic / P.aw) % P.ah, z = ic / (P.aw * P.ah);
    double ndx = (x + 0.5) / P.aw * 2.0 - 1.0, ndy = (y + 0.5) / P.ah * 2.0 - 1.0;
    const double* m = P.f.inv_vp;
    double hv[4] = {ndx, -ndy, 1.0, 1.0}, nh[4];
    for (int rI = 0; rI < 4; ++rI) nh[rI] = m[rI] * hv[0] + m[4 + rI] * hv[1] + m[8 + rI] * hv[2] + m[12 + rI] * hv[3];
    D3 cam = ld3(P.f.cam);
    D3 np = d3(nh[0] / nh[3], nh[1] / nh[3], nh[2] / nh[3]);
    D3 vd = sub(np, cam);
    vd = scl(vd, 1.0 / len(vd));
    double zn = (z + 0.5) / P.ad;
    double t_target = 50.0 * pow(100000.0 / 50.0, zn);
    D3 p0 = add(cam, d3(0, a.R, 0));
    double r0 = len(p0);
    D3 sun = ld3(P.f.sun);
    double ct = dot(vd, sun);
    double pr = phase_rayleigh(ct), pm = phase_mie(ct, a.mie_g);
    constexpr int N = 16, NS = N / kApSeg;
    static_assert(N % kApSeg == 0, "whole steps per segment lane");
    D3 lum_k, tr_k;
    bool broke;
    inscatter_segment<true>(P, T, MS, p0, r0, vd, t_target / N, seg * NS, seg * NS + NS, pr, pm, lum_k, tr_k, broke);
    // the march ends at the first segment that reached the ground: later
    // segments contribute nothing
    int first = broke ? seg : kApSeg;
#pragma unroll
    for (int o = 1; o < kApSeg; o <<= 1) first = min(first, __shfl_xor(first, o, kApSeg));
    if (seg > first) {
        lum_k = d3(0, 0, 0);
        tr_k = d3(1, 1, 1);
    }
    D3 lum, tr;
    seg_combine<kApSeg>(lum_k, tr_k, seg, lum, tr);
    if (!valid || seg != 0) return;
    double t_lum = tr.x * 0.2126 + tr.y * 0.7152 + tr.z * 0.0722;
    store4(out, i, lum, t_lum);
}

// Batched entry points: blockIdx.y selects the scene; its parameter block
// and its slice of each LUT array.
__global__ __launch_bounds__(64) void k_transmittance(const LutParams* __restrict__ Ps, float4* __restrict__ T) {
    const LutParams& P = Ps[blockIdx.y];
    transmittance_body(P, T + (size_t)blockIdx.y * P.tw * P.th);
}

__global__ __launch_bounds__(64 * kMsSeg) void k_multiscatter(const LutParams* __restrict__ Ps, const float4* __restrict__ T,
                                                     float4* __restrict__ MS) {
    const LutParams& P = Ps[blockIdx.y];
    multiscatter_body(P, T + (size_t)blockIdx.y * P.tw * P.th, MS + (size_t)blockIdx.y * P.mw * P.mh);
}

__global__ __launch_bounds__(64) void k_skyview(const LutParams* __restrict__ Ps, const float4* __restrict__ T,
                                                const float4* __restrict__ MS, float4* __restrict__ SV) {
    const LutParams& P = Ps[blockIdx.y];
    const size_t s = blockIdx.y;
    skyview_body(P, T + s * P.tw * P.th, MS + s * P.mw * P.mh, SV + s * P.sw * P.sh, blockIdx.x);
}

__global__ __launch_bounds__(64) void k_aerial(const LutParams* __restrict__ Ps, const float4* __restrict__ T,
                                               const float4* __restrict__ MS, float4* __restrict__ AP) {
    const LutParams& P = Ps[blockIdx.y];
    const size_t s = blockIdx.y;
    aerial_body(P, T + s * P.tw * P.th, MS + s * P.mw * P.mh, AP + s * P.aw * P.ah * P.ad, blockIdx.x);
}

// Sky-view and aerial-perspective LUTs in one launch (both read only the
// transmittance and multi-scattering LUTs): blocks [0, nsv) build sky-view
// texels, the rest AP froxels, so the two short, latency-bound passes
// overlap instead of running back to back.
__global__ __launch_bounds__(64) void k_skyview_aerial(const LutParams* __restrict__ Ps, const float4* __restrict__ T,
                                                       const float4* __restrict__ MS, float4* __restrict__ SV,
                                                       float4* __restrict__ AP, int nsv) {
    const LutParams& P = Ps[blockIdx.y];
    const size_t s = blockIdx.y;
    const int b = (int)blockIdx.x;
    if (b < nsv) skyview_body(P, T + s * P.tw * P.th, MS + s * P.mw * P.mh, SV + s * P.sw * P.sh, b);
    else aerial_body(P, T + s * P.tw * P.th, MS + s * P.mw * P.mh, AP + s * P.aw * P.ah * P.ad, b - nsv);
}

}  // namespace

hipError_t ps_launch_luts(const LutParams* d_params, const LutParams& P, int n, uint32_t flags, float4* T,
                          float4* MS, float4* SV, float4* AP, hipStream_t s, hipEvent_t* ev) {
    // ev (optional): 5 events recorded before T, MS, SV, AP and after AP.
    const unsigned ns = (unsigned)n;
    if (ev) hipEventRecord(ev[0], s);
    if (flags & 1u)
        hipLaunchKernelGGL(k_transmittance, dim3((P.tw * P.th * kTSeg + 63) / 64, ns), dim3(64), 0, s, d_params, T);
    if (ev) hipEventRecord(ev[1], s);
    if (flags & 2u)
        hipLaunchKernelGGL(k_multiscatter, dim3(P.mw * P.mh, ns), dim3(64 * kMsSeg), 0, s, d_params, T, MS);
    if (ev) hipEventRecord(ev[2], s);
    const int nsv = (P.sw * P.sh * kSvSeg + 63) / 64, nap = (P.aw * P.ah * P.ad * kApSeg + 63) / 64;
    if ((flags & 12u) == 12u) {
        // one launch for both (the sky-view time then covers both)
        hipLaunchKernelGGL(k_skyview_aerial, dim3(nsv + nap, ns), dim3(64), 0, s, d_params, T, MS, SV, AP, nsv);
        if (ev) hipEventRecord(ev[3], s);
    } else {
        if (flags & 4u) hipLaunchKernelGGL(k_skyview, dim3(nsv, ns), dim3(64), 0, s, d_params, T, MS, SV);
        if (ev) hipEventRecord(ev[3], s);
        if (flags & 8u) hipLaunchKernelGGL(k_aerial, dim3(nap, ns), dim3(64), 0, s, d_params, T, MS, AP);
    }
    if (ev) hipEventRecord(ev[4], s);
    return hipGetLastError();
}
