// Adjoint of the fused sky + cloud render (render_kernel.hip) for gfx950:
// gradients of a loss L(composite, cloud luminance, cloud transmittance)
// with respect to each layer's density_scale, extinction scale and
// coverage, and to the 128^3 density-perturbation volume.
//
// The forward pass being differentiated is cloud_march.wgsl:731-1054 (march,
// light march, Beer-Powder, multi-octave phase, energy-conserving step, AP
// composite) plus the dual-source composite of cloud_composite.wgsl:47-57.
//
// Design
//  * The training forward (k_render<RF_TRAIN>) stores, per pixel, the
//    transmittance at the first step of every checkpoint chunk (ck_len
//    primary steps, 16 by default) and the forward totals (pre-AP
//    luminance, final transmittance, the depth weights, the number of steps
//    executed).
//  * One lane per pixel, one wave per 8x8 tile, as in the forward. The
//    adjoint walks a pixel's chunks from the last executed one back to the
//    first. For each chunk it recomputes the chunk forward from its
//    checkpoint (density, light march, shading), which gives every term of
//    the gradient that does not depend on later samples, and writes a
//    6-float record per dense sample; it then sweeps the chunk's records in
//    reverse for the transmittance adjoint.
//  * Transmittance adjoint: with T_i the transmittance in front of dense
//    sample i, e_i its step transmittance and v_i = T_i (.) s_i (.) g_i the
//    sample's luminance times its luminance cotangent, the quantity
//    A_i = Tdot_i (.) T_i obeys A_i = A_{i+1} + v_i from A_N = Tdot_N (.) T_N,
//    and dL/drho_i through e_i is -step * sum_c sigma_t,c A_{i+1,c}. The
//    recurrence never divides by a transmittance, so deep, nearly opaque
//    samples cost no precision.
//  * Parameter gradients of the layer being marched are lane registers,
//    added to per-wave LDS sums after every chunk and to the global
//    accumulators once per wave. Perturbation-volume gradients (8 trilinear
//    corners per sample) are accumulated in a per-wave LDS window of texels
//    and added to an FP64 accumulator (delta_scatter_window).
#include "ps_device.h"
#include "ps_launch.h"

#include "ps_march.h"

namespace {

// Per chunk slot and lane, two 16-byte halves: the shading's results (the
// sample's luminance times its cotangent, dL/drho through the shading),
// written when the queue is shaded, and the sample (clamped shape, d density
// / d coverage factor), written when it is queued.
constexpr int kRecFields = 8;
enum { REC_V0 = 0, REC_V1, REC_V2, REC_GRHO, REC_C, REC_Q };

// (1 - (1 + x) e^-x) / x^2, the x-derivative factor of (1 - e^-x) / x:
// d/dx[(1 - e^-x)/x] = -psi(x). Series below x = 1/2 (the direct form
// cancels there), the closed form above.
__device__ __forceinline__ float psi_fn(float x) {
    if (x < 0.5f) {
        // sum_{m >= 2} (-1)^m (m - 1) / m! x^(m-2), through m = 10
        float p = 9.f / 3628800.f;
        p = p * x - 8.f / 362880.f;
        p = p * x + 7.f / 40320.f;
        p = p * x - 6.f / 5040.f;
        p = p * x + 5.f / 720.f;
        p = p * x - 4.f / 120.f;
        p = p * x + 3.f / 24.f;
        p = p * x - 2.f / 6.f;
        p = p * x + 0.5f;
        return p;
    }
    const float e = __expf(-x);
    return (1.f - (1.f + x) * e) / (x * x);
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// A value the compiler cannot see through, so that per-pixel FP64 work
// recomputed inside the chunk loop is not hoisted out of it (and kept live
// in registers across the whole loop).
__device__ __forceinline__ int opaque_int(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Perturbation gradients of many taps of a wave accumulate in per-wave LDS
// windows of 4 x 4 texel columns (x, z) by 32 height rows: the first window
// is placed around the first tap of a batch, the second around the first
// tap the first one misses. A tap whose eight corners all fall inside a
// window adds there with LDS float atomics; any other adds its eight weights
// straight to the FP64 accumulator. The windows are added to the
// accumulator (one FP64 atomic per non-zero texel) once per batch of taps: a
// chunk's primary samples, or one light-sample index of a queue flush. A tap
// in a window costs about one global atomic (the taps of a batch share
// texels), a direct tap eight: each global atomic lane is a separate memory
// transaction. A window is stored row-major (16 columns per row) and
// records which 8-row blocks its taps touched, so that a flush reads only
// those blocks (128 contiguous slots each).
constexpr int kWinCols = 4, kWinRows = 32, kWinSize = kWinCols * kWinCols * kWinRows, kWins = 2,
              kWinSlots = kWins * kWinSize, kWinBlockRows = 8;

struct DeltaWindow {
    uint32_t x0[kWins], z0[kWins], y0[kWins];  // per window: lowest texel column (x, z) and row
    uint32_t blocks[kWins];                    // per window: 8-row blocks touched (bit q: rows 8q..8q+7)
    int n;                                     // windows placed (wave-uniform)
};

__device__ __forceinline__ DeltaWindow no_window() {
    DeltaWindow w;
#pragma unroll
    for (int k = 0; k < kWins; ++k) w.x0[k] = w.z0[k] = w.y0[k] = w.blocks[k] = 0u;
    w.n = 0;
    return w;
}

// Window k around the tap of lane `leader`: one column before it in x and z,
// half the rows below it (clamped to the volume).
__device__ __forceinline__ void place_window(DeltaWindow& w, int k, uint32_t xi, uint32_t zi, uint32_t r0,
                                             int leader) {
    w.x0[k] = (__builtin_amdgcn_readlane(xi, leader) + 127u) & 127u;
    w.z0[k] = (__builtin_amdgcn_readlane(zi, leader) + 127u) & 127u;
    const uint32_t lr = __builtin_amdgcn_readlane(r0, leader);
    w.y0[k] = lr >= kWinRows / 2 ? min(lr - kWinRows / 2, 128u - kWinRows) : 0u;
}

__device__ __forceinline__ bool in_window(uint32_t x0, uint32_t z0, uint32_t y0, uint32_t xi, uint32_t zi,
                                          uint32_t r0, uint32_t r1) {
    const uint32_t lx = (xi - x0) & 127u, lz = (zi - z0) & 127u;
    return lx < kWinCols - 1 && lz < kWinCols - 1 && r0 >= y0 && r1 - y0 < (uint32_t)kWinRows;
}

// Add the windows' touched blocks to the accumulator and clear them. Called
// with all 64 lanes active.
__device__ __forceinline__ void delta_window_flush(double* __restrict__ g, float* win, DeltaWindow& w) {
    if (w.n == 0) return;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int lane = threadIdx.x;
#pragma unroll
    for (int k = 0; k < kWins; ++k) {
        if (k >= w.n) break;
        float* wk = win + k * kWinSize;
        uint32_t bm = w.blocks[k];
        while (bm != 0u) {
            const int q = __builtin_ctz(bm);
            bm &= bm - 1u;
#pragma unroll
            for (int j = 0; j < kWinBlockRows * kWinCols * kWinCols / 64; ++j) {
                const int slot = q * (kWinBlockRows * kWinCols * kWinCols) + j * 64 + lane;
                const float sum = wk[slot];
                if (sum != 0.f) {
                    const uint32_t row = (uint32_t)slot >> 4, col = (uint32_t)slot & 15u;
                    const uint32_t x = (w.x0[k] + col % kWinCols) & 127u, z = (w.z0[k] + col / kWinCols) & 127u;
                    atomic_add_g(g, (z << 14) | ((w.y0[k] + row) << 7) | x, (double)sum);
                    wk[slot] = 0.f;
                }
            }
        }
        w.blocks[k] = 0u;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    w.n = 0;
}

// Scatter v times the trilinear weights of each lane's tap d (lanes with
// has). Called with all 64 lanes active.
__device__ __forceinline__ void delta_scatter_window(double* __restrict__ g, float* win, DeltaWindow& w, bool has,
                                                     const DeltaTap& d, float v) {
    const uint32_t xi = d.x0, zi = d.z0 >> 14, r0 = d.y0 >> 7, r1 = d.y1 >> 7;
    const unsigned long long act = __ballot(has);
    if (act == 0ull) return;
    if (w.n == 0) {
        place_window(w, 0, xi, zi, r0, __builtin_ctzll(act));
        w.n = 1;
    }
    const bool in0 = has && in_window(w.x0[0], w.z0[0], w.y0[0], xi, zi, r0, r1);
    const unsigned long long miss = __ballot(has && !in0);
    if (miss != 0ull && w.n == 1) {
        place_window(w, 1, xi, zi, r0, __builtin_ctzll(miss));
        w.n = 2;
    }
    const bool in1 = has && !in0 && w.n == 2 && in_window(w.x0[1], w.z0[1], w.y0[1], xi, zi, r0, r1);
    const uint32_t x0 = in0 ? w.x0[0] : w.x0[1], z0 = in0 ? w.z0[0] : w.z0[1], y0 = in0 ? w.y0[0] : w.y0[1];
    const uint32_t a = r0 - y0, b = r1 - y0;  // window rows of the tap (meaningful when in0 || in1)
    // the 8-row blocks the window taps touch
#pragma unroll
    for (int q = 0; q < kWinRows / kWinBlockRows; ++q) {
        const bool tq = (a >> 3) == (uint32_t)q || (b >> 3) == (uint32_t)q;
        if (__ballot(in0 && tq) != 0ull) w.blocks[0] |= 1u << q;
        if (__ballot(in1 && tq) != 0ull) w.blocks[1] |= 1u << q;
    }
    if (!has) return;
    const float wx1 = d.fx, wx0 = 1.f - d.fx, wy1 = d.fy, wy0 = 1.f - d.fy, wz1 = d.fz, wz0 = 1.f - d.fz;
    if (in0 || in1) {
        const uint32_t lx = (xi - x0) & 127u, lz = (zi - z0) & 127u;
        // slot = window * kWinSize + row * 16 + (z * 4 + x)
        const uint32_t c = (in0 ? 0u : (uint32_t)kWinSize) + lz * kWinCols + lx;
        const uint32_t ra = c + a * (kWinCols * kWinCols), rb = c + b * (kWinCols * kWinCols);
        atomicAdd(win + ra, v * wz0 * wy0 * wx0);
        atomicAdd(win + ra + 1, v * wz0 * wy0 * wx1);
        atomicAdd(win + rb, v * wz0 * wy1 * wx0);
        atomicAdd(win + rb + 1, v * wz0 * wy1 * wx1);
        atomicAdd(win + ra + kWinCols, v * wz1 * wy0 * wx0);
        atomicAdd(win + ra + kWinCols + 1, v * wz1 * wy0 * wx1);
        atomicAdd(win + rb + kWinCols, v * wz1 * wy1 * wx0);
        atomicAdd(win + rb + kWinCols + 1, v * wz1 * wy1 * wx1);
    } else {
        const uint32_t off[8] = {d.z0 + d.y0 + d.x0, d.z0 + d.y0 + d.x1, d.z0 + d.y1 + d.x0, d.z0 + d.y1 + d.x1,
                                 d.z1 + d.y0 + d.x0, d.z1 + d.y0 + d.x1, d.z1 + d.y1 + d.x0, d.z1 + d.y1 + d.x1};
        const float wt[8] = {v * wz0 * wy0 * wx0, v * wz0 * wy0 * wx1, v * wz0 * wy1 * wx0, v * wz0 * wy1 * wx1,
                             v * wz1 * wy0 * wx0, v * wz1 * wy0 * wx1, v * wz1 * wy1 * wx0, v * wz1 * wy1 * wx1};
#pragma unroll
        for (int c = 0; c < 8; ++c) atomic_add_g(g, off[c], (double)wt[c]);
    }
}

// Scatter g (= dL/d od * density_scale) over each lane's light samples in
// `lin` (lanes with has). Called with all 64 lanes active; light sample i of
// every lane is scattered together.
__device__ __forceinline__ void light_delta_scatter_wave(const RenderParams& P, float* win, const LayerF& L,
                                                         bool has, float qx0, float qy0, float qz0, float px,
                                                         float py, float pz, float alt, uint32_t lin, float g) {
    const LightWalk w = light_walk(P, L, px, py, pz, alt);
    const int n = P.cone ? 6 : P.light_steps;
    unsigned long long any = __ballot(has && lin != 0u);
    if (any == 0ull) return;
    for (int i = 0; i < n; ++i) {
        const bool hi = has && ((lin >> i) & 1u);
        if (__ballot(hi) == 0ull) continue;
        DeltaTap d{};
        float wt = 0.f;
        if (hi) {
            float qx, qyy, qz, a;
            light_sample(P, w, i, qx0, qy0, qz0, px, py, pz, alt, qx, qyy, qz, a, wt);
            const float h = (a - L.base) * L.inv_thick;
            float dcx, dcy, dcz;
            delta_coords(P, qx - L.wind_off_x, qz - L.wind_off_z, h, dcx, dcy, dcz);
            d = delta_tap(dcx, dcy, dcz);
        }
        DeltaWindow dw = no_window();
        delta_scatter_window(P.g_delta64, win, dw, hi, d, g * wt);
        delta_window_flush(P.g_delta64, win, dw);
    }
}

// ---------------------------------------------------------------------------
// Deferred dense-sample adjoint shading. The light march, the shading and
// their adjoints need only the sample, the transmittance in front of it and
// the owning pixel's output seeds (not the later samples), so dense samples
// are queued per wave as in the forward and shaded 64 at a time by the whole
// wave; each result goes to the owner's record slot. Only the transmittance
// adjoint needs the reverse order; the owner runs it after the chunk's queue
// has drained.
// ---------------------------------------------------------------------------
constexpr int kBQCap = 128;  // >= 64 + 63: a full flush plus one step's arrivals
// BQ_QX..BQ_QZ: the sample's position in its segment's lookup frame (SegGeom)
enum { BQ_QX = 0, BQ_QY, BQ_QZ, BQ_ALT, BQ_WR, BQ_DENS, BQ_TR, BQ_TG, BQ_TB, BQ_T, BQ_STEP, BQ_META, BQ_FIELDS };
enum { SD_G0 = 0, SD_G1, SD_G2, SD_GTW, SD_GTWN, SD_OMC, SD_DX, SD_DY, SD_DZ, SD_FIELDS };

struct BackQueue {
    float f[BQ_FIELDS][kBQCap];
    float seed[SD_FIELDS][64];        // per owner lane: luminance cotangent, depth-weight cotangents, 1 - cos,
                                      // view ray
    float wacc[3 * PS_MAX_LAYERS];    // per-layer gradient sums of the wave
    float win[kWinSlots];             // perturbation-gradient windows of a batch (zero between uses)
};

// Shade queue entry j with its adjoint (cloud_march.wgsl:886-993): writes
// the owner's record (v, dL/drho through the shading) and adds the
// light-march and extinction contributions to the wave's layer sums.
// Light-march scatter work a shaded sample leaves for the wave (the sample's
// position and layer are read back from its queue entry).
struct LightScatter {
    float g;
    uint32_t lin;
};

template <bool DELTA>
__device__ __forceinline__ LightScatter shade_adj(const RenderParams& P, BackQueue& Q, int j, float* rec_tile,
                                                  const float* lrec_tile) {
    const float qx = Q.f[BQ_QX][j], qyl = Q.f[BQ_QY][j], qz = Q.f[BQ_QZ][j];
    const float alt = Q.f[BQ_ALT][j], wr = Q.f[BQ_WR][j], density = Q.f[BQ_DENS][j];
    const float T[3] = {Q.f[BQ_TR][j], Q.f[BQ_TG][j], Q.f[BQ_TB][j]};
    const float t = Q.f[BQ_T][j], step = Q.f[BQ_STEP][j];
    const uint32_t meta = __float_as_uint(Q.f[BQ_META][j]);
    const int owner = (int)(meta & 63u), li = (int)((meta >> 6) & 7u), slot = (int)((meta >> 12) & 31u);
    const uint32_t ctype = (meta >> 9) & 7u;
    const int kd = (int)((meta >> 17) & 255u);
    const NdfC nc = ndf_coef(ctype);
    const LayerC& LC = P.layers[li];
    const LayerF L = LC.f;
    const float G[3] = {Q.seed[SD_G0][owner], Q.seed[SD_G1][owner], Q.seed[SD_G2][owner]};
    const float g_tw = Q.seed[SD_GTW][owner], g_twn = Q.seed[SD_GTWN][owner];
    const float omcf = Q.seed[SD_OMC][owner];
    const float cosv = 1.f - omcf;
    // world position from the camera, as the forward's shading computes it
    const float px = P.cam_f[0] + Q.seed[SD_DX][owner] * t, py = P.cam_f[1] + Q.seed[SD_DY][owner] * t,
                pz = P.cam_f[2] + Q.seed[SD_DZ][owner] * t;
    const float lw[3] = {0.2126f, 0.7152f, 0.0722f};
    const float qy = py + P.R;
    float gds = 0.f, gcov = 0.f, gk = 0.f;
    // dense sample: light march and shading with their adjoints
    // the training forward's light march and partials (samples past its
    // record cap march again)
    float od, dod_ds, dod_cov;
    uint32_t lin;
    if (kd < P.lrec_cap) {
        const float4 r = *(const float4*)(lrec_tile + ((size_t)kd * 64 + owner) * LR_FIELDS);
        od = r.x; dod_ds = r.y; dod_cov = r.z;
        lin = __float_as_uint(r.w);
    } else {
        od = light_adj<DELTA>(P, L, nc, qx, qyl, qz, px, py, pz, alt, wr, ctype, dod_ds, dod_cov, lin);
    }
    const float hn = clampf((alt - L.base) * L.inv_thick, 0.f, 1.f);
    float st[3] = {0.f, 0.f, 0.f};
    {
        const float hh = fmaxf(alt, 1.f);
        const float mu = clampf((px * P.sun[0] + qy * P.sun[1] + pz * P.sun[2]) / (P.R + hh), -1.f, 1.f);
        float tu, tv;
        if (trans_uv_from_h(P, hh, mu, tu, tv)) {
            const float4 tt = lut2d(P.T, P.tw, P.th, tu, tv);
            st[0] = tt.x; st[1] = tt.y; st[2] = tt.z;
        }
    }
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
    float S1[3] = {0.f, 0.f, 0.f}, D1[3] = {0.f, 0.f, 0.f}, e0[3];
    for (int c = 0; c < 3; ++c) e0[c] = __expf(-L.sigma_t[c] * od);
    for (int n = 0; n < P.octaves; ++n) {
        const float ph = P.a_pow[n] * phase_octave_f(omcf, cosv, P.c_pow[n], g_hg, omg0, g_d, alpha,
                                                     w_d, ice);
        const float bn = P.b_pow[n];
        for (int c = 0; c < 3; ++c) {
            const float en = n == 0 ? e0[c] : __expf(-L.sigma_t[c] * od * bn);
            S1[c] += ph * en;
            D1[c] += ph * bn * en;
        }
    }
    float fbm = 1.f;
    if (P.forward_bias > 0.f) fbm = 1.f + P.forward_bias * smoothstepf(0.f, 1.f, fmaxf(cosv, 0.f));
    float amb[3];
    {
        const float rho = sqrtf(fmaxf(alt * (2.f * P.R + alt), 0.f));
        const float zha = (float)PS_PI - atan2f(P.R, rho);
        float u, v;
        skyview_uv_f(P, px, qy, pz, zha, u, v);
        const float4 sv = lut2d(P.SV, P.sw, P.sh, u, v);
        const float g = P.ambient * (0.3f + 0.7f * hn);
        amb[0] = sv.x * P.E[0] * g; amb[1] = sv.y * P.E[1] * g; amb[2] = sv.z * P.E[2] * g;
    }
    if (P.light_radius > 0.f) {
        const float lx = px - P.light_x, lz = pz - P.light_z;
        const float fo = clampf(1.f - sqrtf(lx * lx + lz * lz) / P.light_radius, 0.f, 1.f);
        const float wgt = fo * fo * (3.f - 2.f * fo);
        for (int c = 0; c < 3; ++c) amb[c] += P.light_rgb[c] * wgt;
    }
    const float pw = P.powder;
    float g_rho = 0.f, g_od = 0.f, v[3], e_step[3];
    for (int c = 0; c < 3; ++c) {
        const float sig = L.sigma_t[c];
        const float en = e0[c] + (2.f * e0[c] * (1.f - e0[c] * e0[c]) - e0[c]) * pw;
        const float den = -e0[c] * ((1.f - pw) + 2.f * pw * (1.f - 3.f * e0[c] * e0[c]));
        const float es = P.E[c] * st[c];
        const float sun = S1[c] * es * fbm, dsun = -D1[c] * es * fbm;
        const float Qc = sun * en + amb[c];
        const float dQ = dsun * en + sun * den;  // d Q / d (sigma_t od)
        const float in_c = P.sigma_s[c] * density * Qc;
        const float rs = density * sig;
        const bool lin_t = rs > 1e-6f;
        const float ls = fmaxf(rs, 1e-6f);
        const float xx = ls * step;
        const float ome = one_minus_exp_neg(xx);
        const float es_t = __expf(-xx);
        e_step[c] = es_t;
        const float s_int = in_c * ome / ls;
        const float g_sl = G[c] + lw[c] * (g_tw * t + g_twn);
        v[c] = T[c] * s_int * g_sl;
        const float g_s = T[c] * g_sl;
        const float K = P.sigma_s[c] * density * ome / ls;
        const float g_Q = g_s * K;
        const float dsdr = lin_t ? P.sigma_s[c] * Qc * step * es_t : P.sigma_s[c] * Qc * ome / ls;
        g_rho += g_s * dsdr;
        const float dsdsig = lin_t ? -P.sigma_s[c] * Qc * density * step * (density * step * psi_fn(xx)) : 0.f;
        const float g_sig = g_Q * od * dQ + g_s * dsdsig;
        g_od += g_Q * sig * dQ;
        gk += g_sig * LC.sigma_t0[c];
    }
    gds += g_od * dod_ds;
    gcov += g_od * dod_cov;
    *(float4*)(rec_tile + ((size_t)slot * 64 + owner) * kRecFields) = make_float4(v[0], v[1], v[2], g_rho);
    atomicAdd(&Q.wacc[li], gds);
    atomicAdd(&Q.wacc[PS_MAX_LAYERS + li], gk);
    atomicAdd(&Q.wacc[2 * PS_MAX_LAYERS + li], gcov);
    LightScatter ls;
    ls.lin = lin;
    ls.g = g_od * L.density_scale;
    return ls;
}

// Shade the first n (<= 64) queued samples with the whole wave, then move
// the entries [64, qn) to the front. Called with all 64 lanes active.
template <bool DELTA>
__device__ __forceinline__ void flush_adj(const RenderParams& P, BackQueue& Q, int lane, int n, int qn,
                                          float* rec_tile, const float* lrec_tile, bool want_delta) {
    wave_sync_lds();
    LightScatter ls{};
    if (lane < n) ls = shade_adj<DELTA>(P, Q, lane, rec_tile, lrec_tile);
    if (want_delta) {
        // the sample's lookup position, altitude and world position from its entry
        const int j = lane < n ? lane : 0;
        const uint32_t meta = __float_as_uint(Q.f[BQ_META][j]);
        const int owner = (int)(meta & 63u), li = (int)((meta >> 6) & 7u);
        const float t = Q.f[BQ_T][j];
        const float px = P.cam_f[0] + Q.seed[SD_DX][owner] * t, py = P.cam_f[1] + Q.seed[SD_DY][owner] * t,
                    pz = P.cam_f[2] + Q.seed[SD_DZ][owner] * t;
        light_delta_scatter_wave(P, Q.win, P.layers[li].f, lane < n, Q.f[BQ_QX][j], Q.f[BQ_QY][j],
                                 Q.f[BQ_QZ][j], px, py, pz, Q.f[BQ_ALT][j], ls.lin, ls.g);
    }
    wave_sync_lds();
    const int rest = qn - n;
    if (rest > 0) {
        float v[BQ_FIELDS];
        if (lane < rest) {
#pragma unroll
            for (int k = 0; k < BQ_FIELDS; ++k) v[k] = Q.f[k][64 + lane];
        }
        wave_sync_lds();
        if (lane < rest) {
#pragma unroll
            for (int k = 0; k < BQ_FIELDS; ++k) Q.f[k][lane] = v[k];
        }
        wave_sync_lds();
    }
}

template <bool DELTA>
__device__ __forceinline__ void backward_tile(const RenderParams& P, int tile_x, int tile_y, BackQueue& Q) {
    const int lane = threadIdx.x;
    const int x = tile_x * 8 + (lane & 7);
    const int y = tile_y * 8 + (lane >> 3);
    const bool valid = x < P.W && y < P.H;
    const int xc = min(x, P.W - 1), yc = min(y, P.H - 1);
    const size_t tile = (size_t)tile_y * P.n_tiles_x + tile_x;
    if (lane < 3 * PS_MAX_LAYERS) Q.wacc[lane] = 0.f;
    for (int k = lane; k < kWinSlots; k += 64) Q.win[k] = 0.f;

    const PixelRay ray = pixel_ray(P, xc, yc);
    const float dx = ray.dx, dy = ray.dy, dz = ray.dz;
    const float3 sky = sky_pass(P, ray);
    const float jitter = pixel_jitter(P, xc, yc);
    const double depth_t_max = pixel_depth_tmax(P, ray, xc, yc);
    int nseg = 0;
    const uint32_t order = valid ? layer_order(P, ray, depth_t_max, nseg) : 0u;
    (void)nseg;
    const float lw[3] = {0.2126f, 0.7152f, 0.0722f};

    // ---- forward totals and output cotangents ---------------------------------------
    const float* ps = P.px_state + tile * (PX_FIELDS * 64) + lane;
    const float lum_pre[3] = {ps[PX_LR * 64], ps[PX_LG * 64], ps[PX_LB * 64]};
    const float Tf[3] = {ps[PX_TR * 64], ps[PX_TG * 64], ps[PX_TB * 64]};
    const float tw = ps[PX_TW * 64], twn = ps[PX_TWN * 64];
    const int n_exec = valid ? (int)ps[PX_NEXEC * 64] : 0;
    const int n_dense = valid ? (int)ps[PX_NDENSE * 64] : 0;
    const size_t o = ((size_t)yc * P.W + xc) * 4;
    float g_lo[3] = {0.f, 0.f, 0.f}, g_T[3] = {0.f, 0.f, 0.f};
    const float skyv[3] = {sky.x, sky.y, sky.z};
    for (int c = 0; c < 3; ++c) {
        if (P.g_comp) {
            const float g = P.g_comp[o + c];
            g_lo[c] += g;
            g_T[c] += g * skyv[c];
        }
        if (P.g_lum) g_lo[c] += P.g_lum[o + c];
        if (P.g_trn) g_T[c] += P.g_trn[o + c];
    }

    // ---- adjoint of the AP composite (cloud_march.wgsl:1016-1048) --------------------
    float G[3] = {g_lo[0], g_lo[1], g_lo[2]};
    float g_tw = 0.f, g_twn = 0.f;
    const float opacity = 1.f - (Tf[0] * lw[0] + Tf[1] * lw[1] + Tf[2] * lw[2]);
    if (valid && opacity > 1e-4f) {
        const float tc = tw / fmaxf(twn, 1e-6f);
        const float dd = fmaxf(tc, 50.f);
        const float apz = clampf(__logf(dd / 50.f) * (float)(1.0 / 7.6009024595420822), 0.f, 1.f);
        const float u = (float)ray.fx * P.inv_W, v = (float)ray.fy * P.inv_H;
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
        float ap[4], dap[4];
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
            dap[k] = (p1 - p0) * (float)P.ad;  // d/d apz (0 where both slices clamp to one)
        }
        float g_ap_a = 0.f, g_op = 0.f, g_apz_rgb = 0.f;
        for (int c = 0; c < 3; ++c) {
            g_ap_a += g_lo[c] * lum_pre[c];
            g_op += g_lo[c] * ap[c];
            g_apz_rgb += g_lo[c] * dap[c];
            G[c] = g_lo[c] * ap[3];
        }
        for (int c = 0; c < 3; ++c) g_T[c] -= lw[c] * g_op;
        const float g_apz = g_ap_a * dap[3] + opacity * g_apz_rgb;
        const bool inside = tc > 50.f && apz > 0.f && apz < 1.f;
        const float g_tc = inside ? g_apz * (float)(1.0 / 7.6009024595420822) / tc : 0.f;
        const float twn_ = fmaxf(twn, 1e-6f);
        g_tw = g_tc / twn_;
        g_twn = twn > 1e-6f ? -g_tc * tc / twn_ : 0.f;
    }
    float A[3] = {g_T[0] * Tf[0], g_T[1] * Tf[1], g_T[2] * Tf[2]};

    Q.seed[SD_G0][lane] = G[0]; Q.seed[SD_G1][lane] = G[1]; Q.seed[SD_G2][lane] = G[2];
    Q.seed[SD_GTW][lane] = g_tw; Q.seed[SD_GTWN][lane] = g_twn; Q.seed[SD_OMC][lane] = ray.omcf;
    Q.seed[SD_DX][lane] = dx; Q.seed[SD_DY][lane] = dy; Q.seed[SD_DZ][lane] = dz;
    wave_sync_lds();

    // ---- chunks, last to first ----------------------------------------------------------
    const float wscale = P.inv_weather_scale, nscale = P.inv_base_scale * 128.f;
    const float gmax = DELTA ? delta_gmax(P) : 0.f;
    const int steps = P.steps, cps = P.ck_per_seg, ckl = P.ck_len;
    int chunk = -1;
    if (n_exec > 0) {
        const int g_last = n_exec - 1;
        chunk = (g_last / steps) * cps + (g_last % steps) / ckl;
    }
    int cur_seg = -1, li = 0;
    SegGeom g{};
    LayerF L;
    // perturbation-volume gradients need no volume: at delta = 0 they are
    // the gradients with respect to an additive perturbation
    const bool want_delta = P.g_delta != nullptr;
    float* rec_tile = P.rec + tile * (size_t)(ckl * kRecFields * 64);
    float* rec = rec_tile + lane * kRecFields;
    const float* lrec_tile = P.lrec + tile * (size_t)P.lrec_cap * (LR_FIELDS * 64);
    int qn = 0;  // queued samples (wave-uniform)
    int kd_next = n_dense;  // dense samples before the chunk after this one

    while (__ballot(chunk >= 0) != 0ull) {
        const bool act = chunk >= 0;
        int s0 = 0, s1 = 0;
        float T[3] = {1.f, 1.f, 1.f};
        uint32_t dmask = 0u;
        int kd = 0;                   // the pixel's dense samples before this one (forward order)
        int kd0 = 0;
        if (act) {
            const int seg = chunk / cps, sub = chunk % cps;
            if (seg != cur_seg) {
                cur_seg = seg;
                li = (int)((order >> (3 * seg)) & 7u);
                const LayerC& LC = P.layers[li];
                L = LC.f;
                // the FP64 ray again (keeping it live across the chunk loop costs registers)
                const int xo = opaque_int(xc), yo = opaque_int(yc);
                const PixelRay ray2 = pixel_ray(P, xo, yo);
                seg_geom(P, LC, ray2, pixel_depth_tmax(P, ray2, xo, yo), pixel_jitter(P, xo, yo), g);
            }
            s0 = sub * ckl;
            s1 = min(min(steps, s0 + ckl), n_exec - seg * steps);
            const float4 ck = *(const float4*)(P.ck + (tile * P.ck_max + chunk) * PS_CK_STRIDE + lane * 4);
            T[0] = ck.x; T[1] = ck.y; T[2] = ck.z;
            kd = (int)ck.w;
            kd0 = kd;
        }
        // The chunk's dense samples in forward order, from the training
        // forward's records when they hold every one of them, else by
        // marching the chunk again from its checkpoint; they are queued for
        // the adjoint shading either way.
        const bool use_rec = act && kd_next <= P.lrec_cap;
        const int n_rec = kd_next - kd0;
        for (int k = 0; k < ckl; ++k) {
            const int s = s0 + k;
            const bool more = act && (use_rec ? k < n_rec : s < s1);
            if (__ballot(more) == 0ull) break;
            bool dense = false;
            int slot = k;                 // the sample's step within the chunk
            float t = 0.f, px = 0.f, py = 0.f, pz = 0.f, alt = 0.f, wr = 0.f, density = 0.f;
            uint32_t ctype = 0u;
            if (more && use_rec) {
                // the sample half of the record: step word, clamped shape, coverage partial, weather
                const float4 r = *(const float4*)(lrec_tile + ((size_t)(kd0 + k) * 64 + lane) * LR_FIELDS + 4);
                const uint32_t sr = __float_as_uint(r.x);
                const int sk = (int)(sr & 0xffffu);
                ctype = sr >> 24;
                slot = sk - s0;
                const float cval = r.y;
                wr = r.w;
                const float tl = (float)sk * g.step;
                t = g.t0 + tl;
                seg_sample(P, g, dx, dy, dz, tl, px, py, pz, alt);
                density = cval * L.density_scale;
                dense = true;
                *(float2*)(rec + (size_t)slot * (kRecFields * 64) + REC_C) = make_float2(cval, r.z);
                dmask |= 1u << slot;
            } else if (more) {
                const float tl = (float)s * g.step;
                t = g.t0 + tl;
                seg_sample(P, g, dx, dy, dz, tl, px, py, pz, alt);
                const float h = (alt - L.base) * L.inv_thick;
                const float wx = px - L.wind_off_x, wz = pz - L.wind_off_z;
                float wfx, wfy;
                const uint32_t woff = weather_tap(P, wx * wscale + 0.5f, wz * wscale + 0.5f, wfx, wfy);
                const float ax = wx - L.skew_x * h, ay = py, az = wz - L.skew_z * h;
                const float cx = ax * nscale - 0.5f, cy = ay * nscale - 0.5f, cz = az * nscale - 0.5f;
                const uint32_t boff = bbound_index(cx, cy, cz) * 2u;
                const uint2 wq = ld_g<uint2>(P.weather, woff);
                const uint32_t bb = ld_g<uint16_t>(P.bbound, boff);
                float dcx = 0.f, dcy = 0.f, dcz = 0.f;
                if constexpr (DELTA) delta_coords(P, wx, wz, h, dcx, dcy, dcz);
                wr = weather_blend(wq, wfx, wfy);
                ctype = L.type;
                NdfC nc = layer_ndf(L);
                if (P.use_type_grid) {
                    float gu = (wx + 16000.f) / 32000.f, gv = (wz + 16000.f) / 32000.f;
                    const uint32_t gx = (uint32_t)(int)floorf(gu * 128.f) & 127u,
                                   gy = (uint32_t)(int)floorf(gv * 128.f) & 127u;
                    const uint32_t g = ld_g<uint8_t>(P.type_grid, gy * 128u + gx);
                    if (g != 255u) { ctype = g; nc = ndf_coef(g); }
                }
                // sample_density with its partials (cloud_march.wgsl:432-530)
                float cval = 0.f, q = 0.f;
                if (h >= 0.f && h <= 1.f + 1e-5f) {
                    const float wc = wr * L.coverage;
                    const float cov = clampf(wc, 0.f, 1.f);
                    if (cov >= 0.01f) {
                        const float ndfv = ndf_eval(nc, ctype, h, L.anvil);
                        float thr = DELTA ? (1.f - cov - gmax) - 1e-5f : (1.f - cov) * (1.f - 1e-5f);
                        float2 db = make_float2(0.f, 0.f);
                        if (DELTA && ndfv > thr) {
                            db = delta_bound(P, dcx, dcy, dcz);
                            thr = (1.f - cov - db.x) - 1e-5f;
                        }
                        float cloud = -1.f, dsh = 1.f;
                        if (ndfv > thr && brick_may_exceed(bb, L.esb * 0.5f, ndfv, thr)) {
                            const float ds_ = P.inv_detail_scale;
                            const float2 cu = curl_fetch(P.curl, ax * ds_ * 128.f - 0.5f, az * ds_ * 128.f - 0.5f);
                            const float2 b = base_fetch(P.base_rl, cx, cy, cz, P.base_rs, P.base_ls);
                            const float lf = clampf(b.y + L.esb * 0.5f, 0.f, 1.f);
                            const float base_cloud = (b.x + (1.f - lf)) * frcp(fmaxf(2.f - lf, 1e-5f));
                            cloud = base_cloud * ndfv - (1.f - cov);
                            if (cloud > 0.f) {
                                const float ddx = ax + cu.x * P.curl_off_scale, ddz = az + cu.y * P.curl_off_scale;
                                const float hf = detail_fetch(P.detail_hf, ddx * ds_ * 32.f - 0.5f,
                                                              ay * ds_ * 32.f - 0.5f, ddz * ds_ * 32.f - 0.5f);
                                const float dmod = hf + ((1.f - hf) - hf) * clampf(h * 10.f, 0.f, 1.f);
                                const float e = dmod * L.ld;
                                dsh = frcp(fmaxf(1.f - e, 1e-5f));
                                cloud = (cloud - e) * dsh;
                            }
                            if constexpr (DELTA) cloud += delta_value(P, db, dcx, dcy, dcz);
                        }
                        cval = clampf(cloud, 0.f, 1.f);
                        density = cval * L.density_scale;
                        if (cloud > 0.f && cloud < 1.f && wc > 0.f && wc < 1.f) q = L.density_scale * dsh * wr;
                    }
                }
                dense = density > 1e-3f;
                if (dense) {
                    *(float2*)(rec + (size_t)k * (kRecFields * 64) + REC_C) = make_float2(cval, q);
                    dmask |= 1u << k;
                }
            }
            const unsigned long long bal = __ballot(dense);
            if (bal != 0ull) {
                if (dense) {
                    const int pos = qn + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                    Q.f[BQ_QX][pos] = px; Q.f[BQ_QY][pos] = py; Q.f[BQ_QZ][pos] = pz;
                    Q.f[BQ_ALT][pos] = alt; Q.f[BQ_WR][pos] = wr; Q.f[BQ_DENS][pos] = density;
                    Q.f[BQ_TR][pos] = T[0]; Q.f[BQ_TG][pos] = T[1]; Q.f[BQ_TB][pos] = T[2];
                    Q.f[BQ_T][pos] = t; Q.f[BQ_STEP][pos] = g.step;
                    Q.f[BQ_META][pos] = __uint_as_float((uint32_t)lane | ((uint32_t)li << 6) | (ctype << 9) |
                                                        ((uint32_t)slot << 12) | ((uint32_t)min(kd, 255) << 17));
                    ++kd;
                    // step transmittance (the forward's update)
                    T[0] *= __expf(-fmaxf(density * L.sigma_t[0], 1e-6f) * g.step);
                    T[1] *= __expf(-fmaxf(density * L.sigma_t[1], 1e-6f) * g.step);
                    T[2] *= __expf(-fmaxf(density * L.sigma_t[2], 1e-6f) * g.step);
                }
                qn += __popcll(bal);
            }
            if (qn >= 64) {
                flush_adj<DELTA>(P, Q, lane, 64, qn, rec_tile, lrec_tile, want_delta);
                qn -= 64;
            }
        }
        if (qn > 0) {
            flush_adj<DELTA>(P, Q, lane, qn, qn, rec_tile, lrec_tile, want_delta);
            qn = 0;
        }
        // the records written by other lanes are visible after the fence
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        {
            // reverse sweep over the chunk's dense samples (transmittance
            // adjoint), slot by slot for the whole wave so the perturbation
            // scatter can combine lanes
            const LayerC& LC = P.layers[li];
            float gds = 0.f, gk = 0.f, gcov = 0.f;
            DeltaWindow win = no_window();
            for (int slot = ckl - 1; slot >= 0; --slot) {
                const bool mine = act && ((dmask >> slot) & 1u);
                if (__ballot(mine) == 0ull) continue;
                float g_rho = 0.f, cval = 0.f;
                DeltaTap dtap{};
                bool scat = false;
                if (mine) {
                    const float* rp = rec + (size_t)slot * (kRecFields * 64);
                    const float4 ra = *(const float4*)rp;
                    const float2 rb = *(const float2*)(rp + REC_C);
                    const float v0 = ra.x, v1 = ra.y, v2 = ra.z;
                    cval = rb.x;
                    const float g_rho_s = ra.w, q = rb.y;
                    const float rho = cval * L.density_scale;
                    g_rho = g_rho_s;
                    for (int c = 0; c < 3; ++c) {
                        const float sig = L.sigma_t[c];
                        if (rho * sig > 1e-6f) {
                            g_rho -= g.step * sig * A[c];
                            gk -= rho * g.step * A[c] * LC.sigma_t0[c];
                        }
                    }
                    gds += g_rho * cval;
                    gcov += g_rho * q;
                    A[0] += v0; A[1] += v1; A[2] += v2;
                    if (want_delta && cval < 1.f) {
                        const int s = s0 + slot;
                        float px, py, pz, alt;
                        seg_sample(P, g, dx, dy, dz, (float)s * g.step, px, py, pz, alt);
                        const float h = (alt - L.base) * L.inv_thick;
                        float dcx, dcy, dcz;
                        delta_coords(P, px - L.wind_off_x, pz - L.wind_off_z, h, dcx, dcy, dcz);
                        dtap = delta_tap(dcx, dcy, dcz);
                        scat = true;
                    }
                }
                if (want_delta) delta_scatter_window(P.g_delta64, Q.win, win, scat, dtap, g_rho * L.density_scale);
            }
            if (want_delta) delta_window_flush(P.g_delta64, Q.win, win);
            if (act) {
                atomicAdd(&Q.wacc[li], gds);
                atomicAdd(&Q.wacc[PS_MAX_LAYERS + li], gk);
                atomicAdd(&Q.wacc[2 * PS_MAX_LAYERS + li], gcov);
                --chunk;
                kd_next = kd0;
            }
        }
    }
    wave_sync_lds();
    if (lane < 3 * PS_MAX_LAYERS) {
        const float v = Q.wacc[lane];
        if (v != 0.f) atomic_add_g(P.g_layers, (uint32_t)lane, v);
    }
}

// Three waves per SIMD: the register budget is 168 VGPRs (the kernel fits
// without spilling).
template <bool DELTA>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_render_backward(
    const RenderParams* __restrict__ Pp) {
    __shared__ BackQueue Qs;
    int tx = blockIdx.x, ty = blockIdx.y;
    xcd_grouped_tile(tx, ty);
    backward_tile<DELTA>(*Pp, tx, ty, Qs);
}

// Bounds (max, min) of the perturbation volume over each 4^3 brick and its
// +1 halo (Repeat in x and z, ClampToEdge in height): every trilinear tap
// whose lower corner lies in the brick reads only texels inside this range.
// Each block of 256 bricks also writes its own bounds after the volume
// entry (kDeltaBricks + 1 + block) for k_delta_global.
constexpr int kDeltaBricks = 32 * 32 * 32;
constexpr int kDeltaBoundBlocks = kDeltaBricks / 256;

__global__ __launch_bounds__(256) void k_delta_bounds(const float* __restrict__ d, float2* __restrict__ bnd) {
    __shared__ float smx[256], smn[256];
    const int t = threadIdx.x;
    const int b = blockIdx.x * 256 + t;
    const int bx = b & 31, by = (b >> 5) & 31, bz = b >> 10;
    float mx = -3.0e38f, mn = 3.0e38f;
    for (int k = 0; k < 5; ++k) {
        const int z = (bz * 4 + k) & 127;
        for (int j = 0; j < 5; ++j) {
            const int yy = min(by * 4 + j, 127);
            for (int i = 0; i < 5; ++i) {
                const int xx = (bx * 4 + i) & 127;
                const float v = d[(z * 128 + yy) * 128 + xx];
                mx = fmaxf(mx, v);
                mn = fminf(mn, v);
            }
        }
    }
    bnd[b] = make_float2(mx, mn);
    smx[t] = mx;
    smn[t] = mn;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
            smx[t] = fmaxf(smx[t], smx[t + o]);
            smn[t] = fminf(smn[t], smn[t + o]);
        }
        __syncthreads();
    }
    if (t == 0) bnd[kDeltaBricks + 1 + blockIdx.x] = make_float2(smx[0], smn[0]);
}

// Volume bounds from the per-block bounds, written after the brick entries.
__global__ __launch_bounds__(64) void k_delta_global(float2* __restrict__ bnd) {
    const int t = threadIdx.x;
    const float2 a = bnd[kDeltaBricks + 1 + t], c = bnd[kDeltaBricks + 1 + 64 + t];
    float mx = fmaxf(a.x, c.x), mn = fminf(a.y, c.y);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        mx = fmaxf(mx, __shfl_xor(mx, o, 64));
        mn = fminf(mn, __shfl_xor(mn, o, 64));
    }
    if (t == 0) bnd[kDeltaBricks] = make_float2(mx, mn);
}

__global__ void k_zero_f64(double* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// dst += src, and src back to zero for the next backward
__global__ void k_flush_f64_to_f32(float* __restrict__ dst, double* __restrict__ src, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        dst[i] += (float)src[i];
        src[i] = 0.0;
    }
}

}  // namespace

hipError_t ps_launch_render_backward(const RenderParams* d_params, int W, int H, bool delta, float* g_delta,
                                     double* g_delta64, hipStream_t s) {
    const int nv = 128 * 128 * 128;
    dim3 grid((W + 7) / 8, (H + 7) / 8), block(64);
    if (delta) hipLaunchKernelGGL(k_render_backward<true>, grid, block, 0, s, d_params);
    else hipLaunchKernelGGL(k_render_backward<false>, grid, block, 0, s, d_params);
    if (g_delta) hipLaunchKernelGGL(k_flush_f64_to_f32, dim3(nv / 256), dim3(256), 0, s, g_delta, g_delta64, nv);
    return hipGetLastError();
}

hipError_t ps_launch_zero_f64(double* p, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_f64, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t ps_launch_delta_bounds(const float* delta, float2* bnd, hipStream_t s) {
    static_assert(kDeltaBoundBlocks == 128, "k_delta_global reduces 128 block bounds with one wave");
    hipLaunchKernelGGL(k_delta_bounds, dim3(kDeltaBoundBlocks), dim3(256), 0, s, delta, bnd);
    hipLaunchKernelGGL(k_delta_global, dim3(1), dim3(64), 0, s, bnd);
    return hipGetLastError();
}
