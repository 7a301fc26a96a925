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

#include "ps_march.h"


namespace {

// render_tile variants: event counting, the density-perturbation volume, and
// the training forward (transmittance checkpoints + per-pixel totals for the
// backward).
// RF_RT: the march writes the cloud RT (luminance and transmittance) of the
// half-resolution path or of a frame that goes through the TAA resolve; no
// sky pass (the full-resolution composite adds it).
enum { RF_COUNT = 1, RF_DELTA = 2, RF_TRAIN = 4, RF_RT = 8 };

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
// A record holds the sample's lookup-frame position (the segment frame of
// SegGeom), altitude, weather coverage, density, the transmittance in front
// of it, its ray distance and step, and the owner lane, layer and cloud
// type. World-frame quantities (local up, lightning distance) come from the
// owner's view ray, read from the owner lane when the queue is flushed.
// The training forward's queue also carries the sample's record fields
// (step word, clamped shape, coverage partial), written with the light
// march's record when the sample is shaded.
enum { QF_QX = 0, QF_QY, QF_QZ, QF_ALT, QF_WR, QF_DENS, QF_TR, QF_TG, QF_TB, QF_T, QF_STEP, QF_META, QF_SREC, QF_CVAL,
       QF_QC };
template <bool TRAIN>
struct WaveQueue {
    static constexpr int kFields = TRAIN ? 15 : 12;
    float f[kFields][kQCap];
    float acc[5][64];             // per-pixel luminance rgb, sum(w t), sum(w)
};

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Shade one queued dense sample (cloud_march.wgsl:886-993) and add its
// luminance and depth-weight contributions to the owning pixel. (odx, ody,
// odz) is the owner's view ray and omc its 1 - cos(view, sun). TRAIN: the
// light march also takes its parameter partials, and the results go to the
// owner's light-march record k (the backward's input) when k < lrec_cap.
template <bool DELTA, bool TRAIN>
__device__ __forceinline__ void shade_record(const RenderParams& P, WaveQueue<TRAIN>& Q, int j, Cnt& cnt, float odx,
                                             float ody, float odz, float omc, float* lrec_tile) {
    const float qx = Q.f[QF_QX][j], qy = Q.f[QF_QY][j], qz = Q.f[QF_QZ][j];
    const float alt = Q.f[QF_ALT][j], wr = Q.f[QF_WR][j], density = Q.f[QF_DENS][j];
    const float Tr = Q.f[QF_TR][j], Tg = Q.f[QF_TG][j], Tb = Q.f[QF_TB][j];
    const float t = Q.f[QF_T][j], step = Q.f[QF_STEP][j];
    const uint32_t meta = __float_as_uint(Q.f[QF_META][j]);
    const int owner = (int)(meta & 63u), li = (int)((meta >> 6) & 7u);
    const uint32_t ctype = (meta >> 9) & 7u;
    const LayerC& LC = P.layers[li];
    const LayerF L = LC.f;
    // world position from the camera (the local up, sun angle and lightning
    // distance need only its direction from the planet centre)
    const float px = P.cam_f[0] + odx * t, py = P.cam_f[1] + ody * t, pz = P.cam_f[2] + odz * t;
    cnt.dense++;
    float od;
    if constexpr (TRAIN) {
        float dod_ds, dod_cov;
        uint32_t lin;
        od = light_adj<DELTA>(P, L, ndf_coef(ctype), qx, qy, qz, px, py, pz, alt, wr, ctype, dod_ds, dod_cov, lin);
        const int k = (int)((meta >> 12) & 255u);
        if (k < P.lrec_cap) {
            float4* r = (float4*)(lrec_tile + ((size_t)k * 64 + owner) * LR_FIELDS);
            r[0] = make_float4(od, dod_ds, dod_cov, __uint_as_float(lin));
            r[1] = make_float4(Q.f[QF_SREC][j], Q.f[QF_CVAL][j], Q.f[QF_QC][j], wr);
        }
    } else {
        od = march_light<DELTA>(P, L, ndf_coef(ctype), qx, qy, qz, px, py, pz, alt, wr, ctype, cnt);
    }
    const float sig0 = L.sigma_t[0], sig1 = L.sigma_t[1], sig2 = L.sigma_t[2];
    // Beer-Powder (cloud_march.wgsl:888-892); exp(-2x) = exp(-x)^2
    const float e0r = __expf(-sig0 * od), e0g = __expf(-sig1 * od), e0b = __expf(-sig2 * od);
    const float pw = P.powder;
    const float enr = e0r + (2.f * e0r * (1.f - e0r * e0r) - e0r) * pw;
    const float eng = e0g + (2.f * e0g * (1.f - e0g * e0g) - e0g) * pw;
    const float enb = e0b + (2.f * e0b * (1.f - e0b * e0b) - e0b) * pw;
    const float hn = clampf((alt - L.base) * L.inv_thick, 0.f, 1.f);
    const float wy = py + P.R;
    // sun transmittance (cloud_march.wgsl:328-354)
    float str = 0.f, stg = 0.f, stb = 0.f;
    {
        const float hh = fmaxf(alt, 1.f);
        const float mu = clampf((px * P.sun[0] + wy * P.sun[1] + pz * P.sun[2]) / (P.R + hh), -1.f, 1.f);
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
        skyview_uv_f(P, px, wy, pz, zha, u, v);
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
// the entries [64, qn) to the front. Called with all 64 lanes active; (dx,
// dy, dz) and omc are this lane's own view ray and 1 - cos(view, sun).
template <bool DELTA, bool TRAIN>
__device__ __forceinline__ void flush_queue(const RenderParams& P, WaveQueue<TRAIN>& Q, int lane, int n, int qn, Cnt& cnt,
                                            float dx, float dy, float dz, float omc, float* lrec_tile) {
    wave_sync_lds();
    const int owner = lane < n ? (int)(__float_as_uint(Q.f[QF_META][lane]) & 63u) : lane;
    const float odx = __shfl(dx, owner), ody = __shfl(dy, owner), odz = __shfl(dz, owner);
    const float oomc = __shfl(omc, owner);
    if (lane < n) shade_record<DELTA, TRAIN>(P, Q, lane, cnt, odx, ody, odz, oomc, lrec_tile);
    wave_sync_lds();
    const int rest = qn - n;
    if (rest > 0) {
        constexpr int NF = WaveQueue<TRAIN>::kFields;
        float v[NF];
        if (lane < rest) {
#pragma unroll
            for (int k = 0; k < NF; ++k) v[k] = Q.f[k][64 + lane];
        }
        wave_sync_lds();
        if (lane < rest) {
#pragma unroll
            for (int k = 0; k < NF; ++k) Q.f[k][lane] = v[k];
        }
        wave_sync_lds();
    }
}

// Keep a value the compiler cannot see through, so that per-pixel FP64 work
// recomputed inside the march is not hoisted out of it (and kept live in
// registers across the whole march).
__device__ __forceinline__ int opaque_int(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// One wave tile of TW x (64 / TW) pixels: everything from the view ray to
// the written output pixels. Called by one full wave.
//
// The march runs in wave-uniform layer rounds. Each lane walks its own
// layer segments in entry order; a round takes the next layer of the first
// lane that still has one, and every lane whose next segment is that layer
// marches it, all from step 0 together. The layer index is then uniform, so
// the layer's constants are scalar loads, and the step counter is uniform.
// For the usual layer stacks every lane enters the layers in the same order
// and each layer takes one round.
template <int FLAGS, int TW>
__device__ __forceinline__ void render_tile(const RenderParams& P, int tile_x, int tile_y,
                                            WaveQueue<(FLAGS & RF_TRAIN) != 0>& Q) {
    static_assert(TW == 4 || TW == 8 || TW == 16, "wave tiles are 4x16, 8x8 or 16x4 pixels");
    constexpr bool COUNT = (FLAGS & RF_COUNT) != 0, DELTA = (FLAGS & RF_DELTA) != 0, TRAIN = (FLAGS & RF_TRAIN) != 0;
    constexpr bool RT = (FLAGS & RF_RT) != 0;
    static_assert(!TRAIN || TW == 8, "the training forward uses 8x8 tiles");
    const int lane = threadIdx.x;
    const int x = tile_x * TW + (lane % TW);
    const int y = tile_y * (64 / TW) + lane / TW;
    // Lanes outside the image stay in the wave (they help shade queued
    // samples) but march nothing and write nothing.
    const bool valid = x < P.W && y < P.H;
    const int xc = min(x, P.W - 1), yc = min(y, P.H - 1);
    Cnt cnt;
    const unsigned long long t_start = COUNT ? __builtin_amdgcn_s_memrealtime() : 0ull;
#pragma unroll
    for (int k = 0; k < 5; ++k) Q.acc[k][lane] = 0.f;

    // ---- per-pixel setup: view ray and the entry order of the layer segments ----------
    float dx, dy, dz, omcf;
    int nseg = 0;
    uint32_t order = 0u;
    {
        const PixelRay ray = pixel_ray(P, xc, yc);
        dx = ray.dx; dy = ray.dy; dz = ray.dz;
        omcf = ray.omcf;
        if (valid) order = layer_order(P, ray, pixel_depth_tmax(P, ray, xc, yc), nseg);
    }

    // ---- primary march in wave-uniform layer rounds ----------------------------------------
    const float wscale = P.inv_weather_scale, nscale = P.inv_base_scale * 128.f;
    const float gmax = DELTA ? delta_gmax(P) : 0.f;
    float tr = 1.f, tg = 1.f, tbb = 1.f;
    int next = 0;                 // the lane's next segment (rank in its entry order)
    bool alive = nseg > 0;        // not yet opaque
    int qn = 0;                   // queued samples (wave-uniform)
    int kd = 0;                   // training forward: the pixel's dense samples so far
    float* lrec_tile = TRAIN ? P.lrec + (size_t)(tile_y * P.n_tiles_x + tile_x) * P.lrec_cap * (LR_FIELDS * 64)
                             : nullptr;
    for (;;) {
        const bool want = alive && next < nseg;
        const unsigned long long wb = __ballot(want);
        if (wb == 0ull) break;
        const int my_li = (int)((order >> (3 * next)) & 7u);
        const int li = __builtin_amdgcn_readlane(my_li, (int)__builtin_ctzll(wb));
        const bool in_round = want && my_li == li;
        bool on = in_round;
        const LayerC& LC = P.layers[li];
        const LayerF& L = LC.f;
        SegGeom g{};
        if (on) {
            const int xo = opaque_int(xc), yo = opaque_int(yc);
            const PixelRay ray = pixel_ray(P, xo, yo);
            on = seg_geom(P, LC, ray, pixel_depth_tmax(P, ray, xo, yo), pixel_jitter(P, xo, yo), g);
        }
        // training checkpoints of this segment: chunk next * ck_per_seg + s / ck_len
        float* ckp = nullptr;
        if constexpr (TRAIN)
            ckp = P.ck + ((size_t)(tile_y * P.n_tiles_x + tile_x) * P.ck_max + next * P.ck_per_seg) * PS_CK_STRIDE +
                  lane * 4;
        NdfC ncl = layer_ndf(L);
        for (int s = 0; s < P.steps; ++s) {
            if (__ballot(on) == 0ull) break;
            if constexpr (TRAIN) {
                if (on && s % P.ck_len == 0)
                    *(float4*)(ckp + (size_t)(s / P.ck_len) * PS_CK_STRIDE) = make_float4(tr, tg, tbb, (float)kd);
            }
            bool dense = false;
            float t = 0.f, qx = 0.f, qy = 0.f, qz = 0.f, alt = 0.f, wr = 0.f, density = 0.f;
            float cval = 0.f, qc = 0.f;   // training forward: the clamped shape and its coverage partial
            uint32_t ctype = 0u;
            if (on) {
                cnt.primary++;
                const float tl = (float)s * g.step;
                t = g.t0 + tl;
                seg_sample(P, g, dx, dy, dz, tl, qx, qy, qz, alt);
                const float h = (alt - L.base) * L.inv_thick;
                // the weather quad and the brick bound depend on the position
                // only: both loads are issued before either is used
                const float wx = qx - L.wind_off_x, wz = qz - L.wind_off_z;
                float wfx, wfy;
                const uint32_t woff = weather_tap(P, wx * wscale + 0.5f, wz * wscale + 0.5f, wfx, wfy);
                const float ax = wx - L.skew_x * h, ay = qy, az = wz - L.skew_z * h;
                const float cx = ax * nscale - 0.5f, cy = ay * nscale - 0.5f, cz = az * nscale - 0.5f;
                const uint32_t boff = bbound_index(cx, cy, cz) * 2u;
                const uint2 wq = ld_g<uint2>(P.weather, woff);
                const uint32_t bb = ld_g<uint16_t>(P.bbound, boff);
                float dcx = 0.f, dcy = 0.f, dcz = 0.f;
                if constexpr (DELTA) delta_coords(P, wx, wz, h, dcx, dcy, dcz);
                wr = weather_blend(wq, wfx, wfy);
                ctype = L.type;
                NdfC nc = ncl;
                if (P.use_type_grid) {
                    float gu = (wx + 16000.f) / 32000.f, gv = (wz + 16000.f) / 32000.f;
                    const uint32_t gx = (uint32_t)(int)floorf(gu * 128.f) & 127u, gy = (uint32_t)(int)floorf(gv * 128.f) & 127u;
                    const uint32_t gt = ld_g<uint8_t>(P.type_grid, gy * 128u + gx);
                    if (gt != 255u) { ctype = gt; nc = ndf_coef(gt); }
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
                        // with the perturbation volume, skip only when shape + max(delta) <= 0:
                        // the volume's maximum first, then the brick's (see shape())
                        float thr = DELTA ? (1.f - cov - gmax) - 1e-5f : (1.f - cov) * (1.f - 1e-5f);
                        float2 db = make_float2(0.f, 0.f);
                        if (DELTA && ndfv > thr) {
                            db = delta_bound(P, dcx, dcy, dcz);
                            thr = (1.f - cov - db.x) - 1e-5f;
                        }
                        float cloud = -1.f, dsh = 1.f;
                        if (ndfv > thr && brick_may_exceed(bb, L.esb * 0.5f, ndfv, thr)) {
                            const float ds = P.inv_detail_scale;
                            const float2 cu = curl_fetch(P.curl, ax * ds * 128.f - 0.5f, az * ds * 128.f - 0.5f);
                            const float2 b = base_fetch(P.base_rl, cx, cy, cz, P.base_rs, P.base_ls);
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
                                dsh = frcp(fmaxf(1.f - e, 1e-5f));
                                cloud = (cloud - e) * dsh;
                            }
                            if constexpr (DELTA) cloud += delta_value(P, db, dcx, dcy, dcz);
                        }
                        density = clampf(cloud, 0.f, 1.f) * L.density_scale;
                        if constexpr (TRAIN) {
                            // the record the backward reads instead of recomputing
                            // the step (its d density / d coverage factor as there)
                            cval = clampf(cloud, 0.f, 1.f);
                            const float wc = wr * L.coverage;
                            if (cloud > 0.f && cloud < 1.f && wc > 0.f && wc < 1.f) qc = L.density_scale * dsh * wr;
                        }
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
                    Q.f[QF_QX][pos] = qx; Q.f[QF_QY][pos] = qy; Q.f[QF_QZ][pos] = qz;
                    Q.f[QF_ALT][pos] = alt; Q.f[QF_WR][pos] = wr; Q.f[QF_DENS][pos] = density;
                    Q.f[QF_TR][pos] = tr; Q.f[QF_TG][pos] = tg; Q.f[QF_TB][pos] = tbb;
                    Q.f[QF_T][pos] = t; Q.f[QF_STEP][pos] = g.step;
                    Q.f[QF_META][pos] = __uint_as_float((uint32_t)lane | ((uint32_t)li << 6) | (ctype << 9) |
                                                        ((uint32_t)min(kd, 255) << 12));
                    if constexpr (TRAIN) {
                        Q.f[QF_SREC][pos] = __uint_as_float((uint32_t)s | ((uint32_t)next << 16) | (ctype << 24));
                        Q.f[QF_CVAL][pos] = cval;
                        Q.f[QF_QC][pos] = qc;
                        ++kd;
                    }
                    tr *= __expf(-fmaxf(density * L.sigma_t[0], 1e-6f) * g.step);
                    tg *= __expf(-fmaxf(density * L.sigma_t[1], 1e-6f) * g.step);
                    tbb *= __expf(-fmaxf(density * L.sigma_t[2], 1e-6f) * g.step);
                    if (fmaxf(tr, fmaxf(tg, tbb)) < 0.01f) {
                        on = false;
                        alive = false;
                    }
                }
                qn += __popcll(bal);
            }
            if (qn >= 64) {
                flush_queue<DELTA, TRAIN>(P, Q, lane, 64, qn, cnt, dx, dy, dz, omcf, lrec_tile);
                qn -= 64;
            }
        }
        if (in_round) ++next;
    }
    if (qn > 0) flush_queue<DELTA, TRAIN>(P, Q, lane, qn, qn, cnt, dx, dy, dz, omcf, lrec_tile);
    wave_sync_lds();
    cnt.layer_marches = (uint32_t)nseg;

    // ---- aerial perspective on the cloud luminance (cloud_march.wgsl:1016-1048) ----
    const float lr = Q.acc[0][lane], lg = Q.acc[1][lane], lb = Q.acc[2][lane];
    const float tw = Q.acc[3][lane], twn = Q.acc[4][lane];
    if constexpr (TRAIN) {
        float* ps = P.px_state + (size_t)(tile_y * P.n_tiles_x + tile_x) * (PX_FIELDS * 64) + lane;
        ps[PX_LR * 64] = lr; ps[PX_LG * 64] = lg; ps[PX_LB * 64] = lb;
        ps[PX_TR * 64] = tr; ps[PX_TG * 64] = tg; ps[PX_TB * 64] = tbb;
        ps[PX_TW * 64] = tw; ps[PX_TWN * 64] = twn;
        ps[PX_NEXEC * 64] = (float)cnt.primary;
        ps[PX_NDENSE * 64] = (float)kd;
    }
    const float t_lum = tr * 0.2126f + tg * 0.7152f + tbb * 0.0722f;
    const float opacity = 1.f - t_lum;
    float olr = lr, olg = lg, olb = lb;
    if (valid && opacity > 1e-4f) {
        cnt.ap++;
        const float tc = tw / fmaxf(twn, 1e-6f);
        const float ds = fmaxf(tc, 50.f);
        const float apz = clampf(__logf(ds / 50.f) * (float)(1.0 / 7.6009024595420822), 0.f, 1.f);
        const float u = (float)((xc + 0.5) * P.rt_sx) * P.inv_W, v = (float)((yc + 0.5) * P.rt_sy) * P.inv_H;
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

    // ---- outputs: composite dst = lum + sky * T (the sky pass runs here, after
    // the march, so that nothing of it is live during the march) ---------------
    if (valid) {
        float3 sky = make_float3(0.f, 0.f, 0.f);
        if constexpr (!RT) {
            const int xo = opaque_int(xc), yo = opaque_int(yc);
            sky = sky_pass(P, pixel_ray(P, xo, yo));
        }
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
            if (P.tile_times) {
                const unsigned long long t_end = __builtin_amdgcn_s_memrealtime();
                unsigned long long* tt = P.tile_times + 2 * ((size_t)tile_y * P.n_tiles_x + tile_x);
                tt[0] = t_start;
                tt[1] = t_end;
            }
        }
    }
}

// Waves per SIMD: more waves hide more of the gather latency. This file is
// compiled without machine-level loop-invariant code motion, which would
// otherwise hoist constant materialisations out of the march loops and keep
// them in registers for the whole tile (120 VGPRs instead of 96 for the
// plain render). The plain render (and the RT march, and the batched tile
// kernel) then fits five waves per SIMD (a 96-VGPR budget) with no scratch.
// The training forward with the perturbation volume (the differentiation
// workload's) also runs five, spilling ~30 VGPRs, 1.5 % faster than four
// without spills; the other perturbation-volume, training and counting
// variants keep four (128 VGPRs), where they need no scratch.
template <int FLAGS>
constexpr int render_waves() {
    return (FLAGS & ~RF_RT) == 0 || FLAGS == (RF_TRAIN | RF_DELTA) ? 5 : 4;
}

template <int FLAGS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(render_waves<FLAGS>(), 8))) void k_render(
    const RenderParams* __restrict__ Pp) {
    __shared__ WaveQueue<(FLAGS & RF_TRAIN) != 0> Qs;
    int tx = blockIdx.x, ty = blockIdx.y;
    xcd_grouped_tile(tx, ty);
    render_tile<FLAGS, 8>(*Pp, tx, ty, Qs);
}

// Tiles of many scenes in one launch. Each entry of `tiles` is
// scene << 20 | tile_y << 10 | tile_x. Persistent: every one-wave block
// pulls the next entry from the counter until the list is exhausted (the
// exit test is wave-uniform and the counter only grows, so every wave
// leaves). Otherwise block b renders entry b. Five waves per SIMD, as the
// plain k_render.
template <int TW, bool PERSISTENT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(render_waves<0>(), 8))) void k_render_tiles(const RenderParams* __restrict__ Ps,
                                                     const uint32_t* __restrict__ tiles, int n_tiles,
                                                     int grab, unsigned* __restrict__ next) {
    __shared__ WaveQueue<false> Qs;
    if (PERSISTENT) {
        // entries [t, end) are this worker's; one counter increment takes
        // `grab` of them (every worker contends for the one counter, so a
        // million one-tile increments serialise there)
        unsigned t = 0, end = 0;
        for (;;) {
            if (t >= end) {
                // readfirstlane keeps the tile index, the entry and hence the
                // scene's parameter block in scalar registers
                unsigned t0 = 0;
                if (threadIdx.x == 0) t0 = atomicAdd(next, (unsigned)grab);
                t = __builtin_amdgcn_readfirstlane(t0);
                if (t >= (unsigned)n_tiles) break;
                end = min(t + (unsigned)grab, (unsigned)n_tiles);
            }
            const uint32_t e = __builtin_amdgcn_readfirstlane(tiles[t]);
            render_tile<0, TW>(Ps[e >> 20], (int)(e & 1023u), (int)((e >> 10) & 1023u), Qs);
            ++t;
        }
    } else {
        const uint32_t e = __builtin_amdgcn_readfirstlane(tiles[blockIdx.x]);
        render_tile<0, TW>(Ps[e >> 20], (int)(e & 1023u), (int)((e >> 10) & 1023u), Qs);
    }
}

__global__ void k_zero_counter(unsigned* p) { *p = 0u; }

// ---- cloud_composite_halfres.wgsl:47-126: Catmull-Rom upsample of the cloud RT -----
// The reference samples the RT with 9 bilinear taps at offsets that make the
// hardware filter reproduce the 16-tap Catmull-Rom kernel (B = 0, C = 0.5);
// here the 16 taps are read directly (ClampToEdge), which is the same
// kernel without the filter-weight quantisation.
__device__ __forceinline__ void catmull_rom_w(float f, float w[4]) {
    w[0] = f * (-0.5f + f * (1.f - 0.5f * f));
    w[1] = 1.f + f * f * (-2.5f + 1.5f * f);
    w[2] = f * (0.5f + f * (2.f - 1.5f * f));
    w[3] = f * f * (-0.5f + 0.5f * f);
}

__device__ __forceinline__ void catmull_rom2(const float4* __restrict__ a, const float4* __restrict__ b, int w, int h,
                                             float u, float v, float4& ra, float4& rb) {
    const float sx = u * (float)w, sy = v * (float)h;
    const float bx = floorf(sx - 0.5f), by = floorf(sy - 0.5f);
    float wx[4], wy[4];
    catmull_rom_w(sx - (bx + 0.5f), wx);
    catmull_rom_w(sy - (by + 0.5f), wy);
    const int ix = (int)bx, iy = (int)by;
    ra = make_float4(0.f, 0.f, 0.f, 0.f);
    rb = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int yy = clampi(iy - 1 + j, h);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int xx = clampi(ix - 1 + i, w);
            const float k = wx[i] * wy[j];
            const float4 p = a[yy * w + xx], q = b[yy * w + xx];
            ra.x += p.x * k; ra.y += p.y * k; ra.z += p.z * k;
            rb.x += q.x * k; rb.y += q.y * k; rb.z += q.z * k;
        }
    }
}

// Full-resolution composite over a cloud RT: the sky pass plus the RT's
// luminance and transmittance, dst = lum + sky * T. UPSAMPLE (the
// half-resolution path, cloud_composite_halfres.wgsl): Catmull-Rom of the
// RT, luminance clamped >= 0 and transmittance to [0, 1] as the reference
// does. Otherwise the RT has the frame's size and its texel is read as is
// (cloud_composite.wgsl). P holds the full-resolution parameters and the RT.
template <bool UPSAMPLE>
__global__ __launch_bounds__(256) void k_composite_rt(const RenderParams* __restrict__ Pp) {
    const RenderParams& P = *Pp;
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= P.W || y >= P.H) return;
    const PixelRay ray = pixel_ray(P, x, y);
    const float3 sky = sky_pass(P, ray);
    float4 lum, trn;
    if (UPSAMPLE) {
        catmull_rom2(P.rt_lum, P.rt_trn, P.rt_W, P.rt_H, ((float)x + 0.5f) * P.inv_W, ((float)y + 0.5f) * P.inv_H,
                     lum, trn);
        lum.x = fmaxf(lum.x, 0.f); lum.y = fmaxf(lum.y, 0.f); lum.z = fmaxf(lum.z, 0.f);
        trn.x = clampf(trn.x, 0.f, 1.f); trn.y = clampf(trn.y, 0.f, 1.f); trn.z = clampf(trn.z, 0.f, 1.f);
    } else {
        lum = P.rt_lum[(size_t)y * P.rt_W + x];
        trn = P.rt_trn[(size_t)y * P.rt_W + x];
    }
    const size_t o = (size_t)y * P.W + x;
    const float4 comp = make_float4(lum.x + sky.x * trn.x, lum.y + sky.y * trn.y, lum.z + sky.z * trn.z, 1.f);
    const float4 lum4 = make_float4(lum.x, lum.y, lum.z, 1.f);
    const float4 tr4 = make_float4(trn.x, trn.y, trn.z, 1.f);
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

// ---- cloud_taa.wgsl: temporal resolve of one cloud RT ----------------------------
// One thread per RT texel in 16 x 16 blocks. The block's current luminance
// and transmittance with a one-texel halo (edge texels repeated: the
// reference's ClampToEdge taps) are staged in LDS for the 3x3 box. The
// reprojection (cloud_taa.wgsl:81-109) runs in FP64: the direction is the
// difference of the unprojected near-plane point and the camera position,
// a cancellation that loses ~1e-4 of the direction in FP32 for a camera a
// few hundred metres from the origin.
__device__ __forceinline__ float4 bilinear_rt(const float4* __restrict__ t, int w, int h, double u, double v) {
    const double sx = u * w - 0.5, sy = v * h - 0.5;
    const double x0d = floor(sx), y0d = floor(sy);
    const float fx = (float)(sx - x0d), fy = (float)(sy - y0d);
    const double wm = (double)(w - 1), hm = (double)(h - 1);
    const int x0 = (int)fmin(fmax(x0d, 0.0), wm), x1 = (int)fmin(fmax(x0d + 1.0, 0.0), wm);
    const int y0 = (int)fmin(fmax(y0d, 0.0), hm), y1 = (int)fmin(fmax(y0d + 1.0, 0.0), hm);
    const float4 a = t[(size_t)y0 * w + x0], b = t[(size_t)y0 * w + x1];
    const float4 c = t[(size_t)y1 * w + x0], d = t[(size_t)y1 * w + x1];
    const float gx = 1.f - fx, gy = 1.f - fy;
    return make_float4((a.x * gx + b.x * fx) * gy + (c.x * gx + d.x * fx) * fy,
                       (a.y * gx + b.y * fx) * gy + (c.y * gx + d.y * fx) * fy,
                       (a.z * gx + b.z * fx) * gy + (c.z * gx + d.z * fx) * fy,
                       (a.w * gx + b.w * fx) * gy + (c.w * gx + d.w * fx) * fy);
}

__device__ __forceinline__ float4 min4(float4 a, float4 b) {
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), fminf(a.w, b.w));
}
__device__ __forceinline__ float4 max4(float4 a, float4 b) {
    return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w));
}
__device__ __forceinline__ float4 clamp4(float4 v, float4 lo, float4 hi) { return min4(max4(v, lo), hi); }

__global__ __launch_bounds__(256) void k_taa_resolve(const TaaParams p) {
    __shared__ float4 sl[18][18], st[18][18];
    const int bx = blockIdx.x * 16, by = blockIdx.y * 16;
    for (int i = threadIdx.x; i < 18 * 18; i += 256) {
        const int ly = i / 18, lx = i - ly * 18;
        const size_t g = (size_t)clampi(by + ly - 1, p.h) * p.w + clampi(bx + lx - 1, p.w);
        sl[ly][lx] = p.cur_lum[g];
        st[ly][lx] = p.cur_trn[g];
    }
    __syncthreads();
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int x = bx + tx, y = by + ty;
    if (x >= p.w || y >= p.h) return;
    const float4 cl = sl[ty + 1][tx + 1], ct = st[ty + 1][tx + 1];
    float4 ol = cl, ot = ct;
    const float bw = clampf(p.blend, 0.f, 1.f);
    if (p.history_valid && bw < 1.f) {
        const double u = ((double)x + 0.5) / p.w, v = ((double)y + 0.5) / p.h;
        const double n0 = u * 2.0 - 1.0, n1 = 1.0 - v * 2.0;
        double hq[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) hq[r] = p.inv_vp[r] * n0 + p.inv_vp[4 + r] * n1 + p.inv_vp[8 + r] + p.inv_vp[12 + r];
        const double iw = 1.0 / hq[3];
        double d0 = hq[0] * iw - p.cam[0], d1 = hq[1] * iw - p.cam[1], d2 = hq[2] * iw - p.cam[2];
        const double il = 1.0 / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        d0 *= il; d1 *= il; d2 *= il;
        const double cx = p.prev_vp[0] * d0 + p.prev_vp[4] * d1 + p.prev_vp[8] * d2;
        const double cy = p.prev_vp[1] * d0 + p.prev_vp[5] * d1 + p.prev_vp[9] * d2;
        const double cw = p.prev_vp[3] * d0 + p.prev_vp[7] * d1 + p.prev_vp[11] * d2;
        double pu = 0.0, pv = 0.0;
        bool inside = false;
        if (cw > 1e-4) {
            pu = cx / cw * 0.5 + 0.5;
            pv = 1.0 - (cy / cw * 0.5 + 0.5);
            inside = pu >= 0.0 && pv >= 0.0 && pu <= 1.0 && pv <= 1.0;
        }
        const float eb = inside ? bw : 1.f;
        const float4 hl = bilinear_rt(p.hist_lum, p.w, p.h, pu, pv), ht = bilinear_rt(p.hist_trn, p.w, p.h, pu, pv);
        float4 lo_l = cl, hi_l = cl, lo_t = ct, hi_t = ct;
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                lo_l = min4(lo_l, sl[ty + dy][tx + dx]); hi_l = max4(hi_l, sl[ty + dy][tx + dx]);
                lo_t = min4(lo_t, st[ty + dy][tx + dx]); hi_t = max4(hi_t, st[ty + dy][tx + dx]);
            }
        const float4 a = clamp4(hl, lo_l, hi_l), b = clamp4(ht, lo_t, hi_t);
        const float ga = 1.f - eb;
        ol = make_float4(fmaxf(a.x * ga + cl.x * eb, 0.f), fmaxf(a.y * ga + cl.y * eb, 0.f),
                         fmaxf(a.z * ga + cl.z * eb, 0.f), fmaxf(a.w * ga + cl.w * eb, 0.f));
        ot = make_float4(clampf(b.x * ga + ct.x * eb, 0.f, 1.f), clampf(b.y * ga + ct.y * eb, 0.f, 1.f),
                         clampf(b.z * ga + ct.z * eb, 0.f, 1.f), clampf(b.w * ga + ct.w * eb, 0.f, 1.f));
    }
    const size_t o = (size_t)y * p.w + x;
    p.out_lum[o] = ol;
    p.out_trn[o] = ot;
}

}  // namespace

hipError_t ps_launch_render(const RenderParams* d_params, int W, int H, bool count, bool delta, bool train,
                            hipStream_t s, bool rt) {
    dim3 grid((W + 7) / 8, (H + 7) / 8), block(64);
    const int f = (count ? RF_COUNT : 0) | (delta ? RF_DELTA : 0) | (train ? RF_TRAIN : 0) | (rt ? RF_RT : 0);
    switch (f) {
        case RF_RT: hipLaunchKernelGGL(k_render<RF_RT>, grid, block, 0, s, d_params); break;
        case RF_RT | RF_DELTA: hipLaunchKernelGGL(k_render<RF_RT | RF_DELTA>, grid, block, 0, s, d_params); break;
        case 0: hipLaunchKernelGGL(k_render<0>, grid, block, 0, s, d_params); break;
        case RF_DELTA: hipLaunchKernelGGL(k_render<RF_DELTA>, grid, block, 0, s, d_params); break;
        case RF_COUNT: hipLaunchKernelGGL(k_render<RF_COUNT>, grid, block, 0, s, d_params); break;
        case RF_DELTA | RF_COUNT: hipLaunchKernelGGL(k_render<RF_DELTA | RF_COUNT>, grid, block, 0, s, d_params); break;
        case RF_TRAIN: hipLaunchKernelGGL(k_render<RF_TRAIN>, grid, block, 0, s, d_params); break;
        case RF_TRAIN | RF_DELTA: hipLaunchKernelGGL(k_render<RF_TRAIN | RF_DELTA>, grid, block, 0, s, d_params); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t ps_launch_render_tiles(const RenderParams* d_params, const uint32_t* tiles, int n_tiles, int tile_w,
                                  int persistent, int blocks, int grab, unsigned* next, hipStream_t s) {
    if (n_tiles <= 0) return hipSuccess;
    if (persistent) {
        hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, s, next);
        const dim3 g(blocks), b(64);
        if (tile_w == 4) hipLaunchKernelGGL((k_render_tiles<4, true>), g, b, 0, s, d_params, tiles, n_tiles, grab, next);
        else if (tile_w == 16) hipLaunchKernelGGL((k_render_tiles<16, true>), g, b, 0, s, d_params, tiles, n_tiles, grab, next);
        else hipLaunchKernelGGL((k_render_tiles<8, true>), g, b, 0, s, d_params, tiles, n_tiles, grab, next);
    } else {
        const dim3 g(n_tiles), b(64);
        if (tile_w == 4) hipLaunchKernelGGL((k_render_tiles<4, false>), g, b, 0, s, d_params, tiles, n_tiles, grab, next);
        else if (tile_w == 16) hipLaunchKernelGGL((k_render_tiles<16, false>), g, b, 0, s, d_params, tiles, n_tiles, grab, next);
        else hipLaunchKernelGGL((k_render_tiles<8, false>), g, b, 0, s, d_params, tiles, n_tiles, grab, next);
    }
    return hipGetLastError();
}

hipError_t ps_launch_composite_rt(const RenderParams* d_params, int W, int H, bool upsample, hipStream_t s) {
    const dim3 g((W + 15) / 16, (H + 15) / 16), b(256);
    if (upsample) hipLaunchKernelGGL(k_composite_rt<true>, g, b, 0, s, d_params);
    else hipLaunchKernelGGL(k_composite_rt<false>, g, b, 0, s, d_params);
    return hipGetLastError();
}

hipError_t ps_launch_taa_resolve(const TaaParams& p, hipStream_t s) {
    hipLaunchKernelGGL(k_taa_resolve, dim3((p.w + 15) / 16, (p.h + 15) / 16), dim3(256), 0, s, p);
    return hipGetLastError();
}
