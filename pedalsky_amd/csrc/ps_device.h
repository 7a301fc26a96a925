// Shared definitions for the gfx950 kernels of the sky + cloud transport
// path: kernel parameter blocks (built on the host, read by every lane
// through the scalar cache) and the device helpers common to the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#define PS_MAX_LAYERS 8
#define PS_MAX_OCTAVES 8
#define PS_PI 3.14159265358979

// ---------------------------------------------------------------------------
// Atmosphere parameters in double precision (the LUT kernels run in FP64;
// there are ~1e6 LUT steps per frame against ~1e8 march steps).
// shaders/common/uniforms.wgsl:39-58.
// ---------------------------------------------------------------------------
struct AtmD {
    double R, top, hr, hm;
    double ray_s[3], mie_s[3], mie_a[3], ozone[3], albedo[3];
    double mie_g, oz_c, oz_t;
};

// Frame data the view-dependent LUTs (sky-view, aerial perspective) need.
struct LutFrameD {
    double inv_vp[16];
    double cam[3];
    double sun[3];
    double E[3];
};

struct LutParams {
    AtmD a;
    LutFrameD f;
    int tw, th, mw, mh, sw, sh, aw, ah, ad;
    int ms_disabled;
};

// ---------------------------------------------------------------------------
// Render parameters: one block per scene. Everything that is constant over a
// frame is folded here by the host (in double, then rounded once).
// ---------------------------------------------------------------------------
// The per-layer values the density evaluation reads (copied into registers
// when a lane enters a layer).
struct LayerF {
    float base, top, inv_thick, coverage;
    float density_scale, esb, ld, anvil;
    float sigma_t[3];
    float wind_off_x, wind_off_z;     // drift offset (m)
    float skew_x, skew_z;             // skew_dir * skew_thickness (m per unit h)
    uint32_t type;
    // NDF of types 0..6 as s(clamp(h * ndf_ia)) * (ndf_a + ndf_b * s(clamp((h - ndf_e0) * ndf_ie)))
    // with s(t) = t^2 (3 - 2t); type 7 (cumulonimbus) has its own profile.
    float ndf_ia, ndf_e0, ndf_ie, ndf_a, ndf_b;
};

struct LayerC {
    LayerF f;
    float sigma_t0[3];                // sigma_t before the per-layer extinction scale (adjoint)
    uint32_t is_cb;                   // layer type 7: phase depends on height
    float droplet;                    // raw layer droplet diameter (um)
    // phase constants for non-Cb layers (cloud_march.wgsl:256-272), folded in
    // FP64; omg_hg = 1 - g_hg without cancellation
    float g_hg, omg_hg, g_d, alpha, w_d, ice;
    // setup-side (FP64)
    double r_bottom, r_top;
};

struct RenderParams {
    int W, H;
    float inv_W, inv_H;
    double inv_vp[16];
    double cam[3];
    double sun_d[3];
    double cos_disk, sun_ang;
    float sun[3];
    float E[3];
    float toa;
    // atmosphere (FP32 copies)
    float R, top, big_h, top_minus_R;
    // sky pass
    float cam_zha;                    // zenith-horizon angle at the camera radius
    float sun_az;
    float overcast_lum;               // overcast_diffuse_luminance (grey)
    float cam_f[3];
    float cam_alt;                    // max(|cam_atm|, R + 1) - R
    // clouds
    int n_layers, steps, light_steps, octaves;
    int cone, temporal_jitter, use_type_grid, has_depth;
    uint32_t frame_index;
    float powder, forward_bias, ambient;
    float curl_off_scale;             // curl_strength * detail_scale_m
    float inv_detail_scale, inv_base_scale, inv_weather_scale;
    float sigma_s[3];
    float ms_a, ms_b, ms_c, droplet_bias;
    float b_pow[PS_MAX_OCTAVES];      // multi_scatter_b^n
    float a_pow[PS_MAX_OCTAVES];      // multi_scatter_a^n
    float c_pow[PS_MAX_OCTAVES];      // multi_scatter_c^n
    float light_rgb[3], light_radius, light_x, light_z;
    double r0;                        // camera radius |cam - centre|
    double h0;                        // r0 - R
    // Common periods (m) of every position-periodic lookup of the march:
    // x and z (base and detail volumes, curl map, weather map, type grid,
    // perturbation volume) and y (base and detail volumes); 0 when the
    // scales share no period below 4096 km. A layer segment's sample
    // positions are taken relative to its entry point minus a whole number
    // of periods, so they stay small (and exact to a few cm) even for rays
    // that reach a layer shell on the far side of the planet.
    double wrap_xz, wrap_y;
    LayerC layers[PS_MAX_LAYERS];
    // resources
    const float4* T; int tw, th;
    const float4* SV; int sw, sh;
    const float4* AP; int aw, ah, ad;
    // Folded noise, each texel stored with the neighbours a filtered tap also
    // reads (one 16-byte load per x corner instead of one load per corner):
    // base_rl: 128^3 texels in 4x4x4 bricks, per texel 4 x (u16 R | u16 LF << 16)
    // for (x, y, z), (x, y+1, z), (x, y, z+1), (x, y+1, z+1), LF = 0.625G + 0.25B + 0.125A;
    const uint32_t* base_rl;
    float base_rs, base_ls;           // fixed-point scales of base_rl's R and LF (base_fold_scales)
    // Empty-space bound of the base shape per brick of PS_BB x PS_BB x PS_BB
    // texels (PS_BN^3 bricks): u8 max R (rounded up) | u8 min LF (rounded
    // down) << 8, both in 1/255, over the brick and its +1 halo.
    const uint16_t* bbound;
    const float* detail_hf;           // 32^3, per texel 4 floats 0.625R + 0.25G + 0.125B at the same 4 texels
    const float2* curl;               // 128^2, per texel 2 float2 (2*rg - 1) at (x, y), (x, y+1)
    const uint2* weather;             // ww*wh texels, each the binary16 2x2 quad (x, y), (x+1, y) | (x, y+1), (x+1, y+1)
    int ww, wh;
    int weather_pow2;                 // ww and wh are powers of two
    int ww_log2;                      // log2(ww) when weather_pow2
    float ww_f, wh_f;
    const uint8_t* type_grid;         // 128^2
    const uint8_t* overcast;          // 128^2
    const uint8_t* blue;              // 64^2
    const float* depth;               // full_W*full_H reverse-Z NDC or null
    void* out_comp;
    void* out_lum;
    void* out_trans;
    int out_f16;
    int _pad;
    unsigned long long* counters;     // 10 x u64 when counting
    unsigned long long* tile_times;   // counting: per wave tile (row-major) its start and end on the
                                      // 100 MHz real-time clock, or null
    // ---- half-resolution cloud march (crates/ps-clouds/src/lib.rs:748-760,
    // pipeline.rs:277-319) ----
    // The march covers W x H pixels (the cloud RT); its view rays pass
    // through the viewport coordinates (x + 0.5) * rt_sx, (y + 0.5) * rt_sy
    // of the full_W x full_H viewport, and the depth buffer (full size) is
    // read at (int)(x * rt_sx), (int)(y * rt_sy). Full resolution: W =
    // full_W, rt_sx = rt_sy = 1. The composite over an RT (half resolution,
    // or any frame that goes through the TAA resolve) reads rt_lum, rt_trn
    // (rt_W x rt_H float4).
    int full_W, full_H;
    double rt_sx, rt_sy;
    const float4* rt_lum;
    const float4* rt_trn;
    int rt_W, rt_H;
    // ---- differentiation extensions (no reference counterpart; the
    // neutral values, no volume and unit extinction scales, are the
    // reference) ----
    // Additive density perturbation: 128^3 floats, x fastest, then the
    // layer-normalised height, then z; x and z span the weather tile
    // (Repeat), height is ClampToEdge.
    const float* delta;
    // (max, min) of delta over each 4^3 brick and its +1 halo (32^3
    // float2), then over the whole volume (1), then 128 per-block partials
    const float2* delta_bnd;
    float delta_su;                   // 128 / weather_scale: delta texels per metre
    // ---- training forward (checkpoints) and backward ----
    // Per wave tile (tiles row-major, 64 lanes): transmittance checkpoints
    // [tile][chunk][lane][rgb + dense count] at the first step of every ck_len-step chunk
    // of every layer segment, and the per-pixel forward totals
    // [tile][PX_FIELDS][lane].
    float* ck;
    float* px_state;
    // Records of the training forward, per wave tile and per pixel's k-th
    // dense sample (k < lrec_cap, in march order): [tile][k][lane][LR_FIELDS]
    // (two 16-byte halves: the light march, then the sample).
    // The backward reads the sample (step, shape, coverage partial, weather
    // value) instead of marching the chunk again, and its light march
    // instead of marching toward the sun again; a chunk with samples beyond
    // the cap is marched again.
    float* lrec;
    int lrec_cap;
    int ck_len;                       // checkpoint interval (primary steps), 1..32
    int ck_per_seg;                   // ceil(steps / ck_len)
    int ck_max;                       // chunk slots per pixel (n_layers * ck_per_seg)
    int n_tiles_x;
    // cotangents of the three outputs (W*H RGBA FP32) or null
    const float* g_comp;
    const float* g_lum;
    const float* g_trn;
    float* g_layers;                  // 3 x 8 accumulated: density_scale, extinction_scale, coverage
    float* g_delta;                   // 128^3 accumulated, or null
    double* g_delta64;                // FP64 accumulator the kernel scatters into (added to g_delta after)
    float* rec;                       // backward records [tile][slot][lane][8]
};

// The cloud TAA resolve of one RT (shaders/clouds/cloud_taa.wgsl), passed by
// value: the current march, the history slot it reads and the slot it
// writes, all w x h float4; the current frame's inv_view_proj and camera
// position and the previous frame's view_proj (column-major, FP64).
struct TaaParams {
    const float4* cur_lum;
    const float4* cur_trn;
    const float4* hist_lum;
    const float4* hist_trn;
    float4* out_lum;
    float4* out_trn;
    int w, h;
    float blend;                      // current-sample weight, clamped to [0, 1]
    int history_valid;
    double inv_vp[16];
    double prev_vp[16];
    double cam[3];
};

// Per-pixel forward totals stored by the training forward (PX_NDENSE: the
// number of dense samples).
enum { PX_LR = 0, PX_LG, PX_LB, PX_TR, PX_TG, PX_TB, PX_TW, PX_TWN, PX_NEXEC, PX_NDENSE, PX_FIELDS };
// A transmittance checkpoint: rgb and the number of dense samples the pixel
// had before the chunk (as a float), one 16-byte vector per lane.
#define PS_CK_STRIDE 256
// Stored per dense sample: the light march's optical depth, its partials in
// density_scale and coverage, and the bit mask of light samples inside the
// density clamp's linear range; the sample's step (bits 0-15), segment rank
// (16-23) and cloud type (24-31) as integer bits, its clamped shape, its
// d density / d coverage factor and its weather value.
enum { LR_OD = 0, LR_DDS, LR_DCOV, LR_LIN, LR_SREC, LR_CVAL, LR_QC, LR_WR, LR_FIELDS };
#ifndef PS_LREC_CAP
#define PS_LREC_CAP 128
#endif
// Edge of an empty-space brick of the base volume, in texels, and the number
// of bricks per axis.
#ifndef PS_BB_LOG2
#define PS_BB_LOG2 4
#endif
#define PS_BB (1 << PS_BB_LOG2)
#define PS_BN (128 >> PS_BB_LOG2)

// ---------------------------------------------------------------------------
// Small vector helpers.
// ---------------------------------------------------------------------------
#define PS_GLOBAL __attribute__((address_space(1)))

// Load a T at a 32-bit byte offset from a global-memory base. Resource
// pointers are read from the parameter block, so the compiler sees them as
// generic; this names the global address space (global_load with a scalar
// base and a 32-bit vector offset instead of flat loads with 64-bit
// address arithmetic).
template <class T>
__device__ __forceinline__ T ld_g(const void* base, uint32_t byte_off) {
    return *(const PS_GLOBAL T*)((const PS_GLOBAL char*)base + byte_off);
}
// Relaxed device-scope atomic add through the global address space: the
// global_atomic form, which (unlike the generic flat form) does not hold the
// wave's next LDS access until the atomic has completed.
template <class T>
__device__ __forceinline__ void atomic_add_g(T* base, uint32_t idx, T v) {
    __hip_atomic_fetch_add((PS_GLOBAL T*)base + idx, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
typedef float ps_f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t ps_u32x2 __attribute__((ext_vector_type(2)));
template <>
__device__ __forceinline__ float2 ld_g<float2>(const void* base, uint32_t byte_off) {
    const ps_f32x2 v = *(const PS_GLOBAL ps_f32x2*)((const PS_GLOBAL char*)base + byte_off);
    return make_float2(v.x, v.y);
}
template <>
__device__ __forceinline__ uint2 ld_g<uint2>(const void* base, uint32_t byte_off) {
    const ps_u32x2 v = *(const PS_GLOBAL ps_u32x2*)((const PS_GLOBAL char*)base + byte_off);
    return make_uint2(v.x, v.y);
}
typedef uint32_t ps_u32x4 __attribute__((ext_vector_type(4)));
typedef float ps_f32x4 __attribute__((ext_vector_type(4)));
template <>
__device__ __forceinline__ uint4 ld_g<uint4>(const void* base, uint32_t byte_off) {
    const ps_u32x4 v = *(const PS_GLOBAL ps_u32x4*)((const PS_GLOBAL char*)base + byte_off);
    return make_uint4(v.x, v.y, v.z, v.w);
}
template <>
__device__ __forceinline__ float4 ld_g<float4>(const void* base, uint32_t byte_off) {
    const ps_f32x4 v = *(const PS_GLOBAL ps_f32x4*)((const PS_GLOBAL char*)base + byte_off);
    return make_float4(v.x, v.y, v.z, v.w);
}

// NDF coefficients of cloud types 0..6 (rising edge 1/a; falling part
// e0, 1/(e1 - e0), a, b) and a zero row for every other type; one table for
// the host (per-layer copies) and the device (per-sample type grid).
#define PS_NDF_ROWS                                                                                        \
    {{1.f / 0.07f, 1.f, 1.f / (0.2f - 1.f), 0.f, 1.f},  {1.f / 0.10f, 0.6f, 1.f / (1.f - 0.6f), 1.f, -1.f},  \
     {1.f / 0.15f, 1.f, 1.f / (0.4f - 1.f), 0.f, 1.f},  {1.f / 0.20f, 1.f, 1.f / (0.3f - 1.f), 0.f, 1.f},    \
     {1.f / 0.30f, 0.7f, 1.f / (1.f - 0.7f), 1.f, -1.f}, {1.f / 0.40f, 0.6f, 1.f / (1.f - 0.6f), 1.f, -1.f}, \
     {1.f / 0.50f, 0.8f, 1.f / (1.f - 0.8f), 1.f, -1.f}, {0.f, 0.f, 0.f, 0.f, 0.f},                         \
     {0.f, 0.f, 0.f, 0.f, 0.f}}

// Hardware reciprocal and square root (1 ulp); used in the FP32 march where
// the correctly rounded sequences would dominate the step cost.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float s3(float t) { return t * t * (3.f - 2.f * t); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ double clampd(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }
__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
    float t = clampf((x - e0) / (e1 - e0), 0.f, 1.f);
    return t * t * (3.f - 2.f * t);
}
__device__ __forceinline__ double smoothstepd(double e0, double e1, double x) {
    double t = clampd((x - e0) / (e1 - e0), 0.0, 1.0);
    return t * t * (3.0 - 2.0 * t);
}
__device__ __forceinline__ int wrapi(int i, int n) { int m = i % n; return m < 0 ? m + n : m; }
// XCD-grouped tile order for one-wave-per-tile launches over a 2-D grid of
// tiles. Workgroup b runs on XCD b mod 8 (round-robin dispatch), so in
// raster order each XCD would take every 8th tile column and its L2 would
// hold the whole frame's noise working set. Here the m-th workgroup of XCD x
// takes tile ((m / 16) * 8 + x) * 16 + m % 16 of the raster order: each XCD
// renders runs of 16 adjacent tiles, and the runs still interleave across
// XCDs, so the load stays balanced. Workgroups past the last full group of
// 8 x 16 keep their own tile (the map is a bijection on the grid). 1.1 %
// faster than raster order at 1280x720 with identical output; a run of 4 or
// 32 tiles is no better.
__device__ __forceinline__ void xcd_grouped_tile(int& tx, int& ty) {
    constexpr int kRun = 16, kXcds = 8;
    const int gx = (int)gridDim.x, n = gx * (int)gridDim.y, b = ty * gx + tx;
    if (b < (n / (kXcds * kRun)) * (kXcds * kRun)) {
        const int x = b % kXcds, m = b / kXcds;
        const int t = ((m / kRun) * kXcds + x) * kRun + m % kRun;
        tx = t % gx;
        ty = t / gx;
    }
}

__device__ __forceinline__ int clampi(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

// Bilinear ClampToEdge sample of an RGBA float LUT in float.
__device__ __forceinline__ float4 lut2d(const float4* __restrict__ t, int w, int h, float u, float v) {
    float cx = u * (float)w - 0.5f, cy = v * (float)h - 0.5f;
    float fx0 = floorf(cx), fy0 = floorf(cy);
    float fx = cx - fx0, fy = cy - fy0;
    int ix = (int)fx0, iy = (int)fy0;
    int x0 = clampi(ix, w), x1 = clampi(ix + 1, w), y0 = clampi(iy, h), y1 = clampi(iy + 1, h);
    float4 a = t[y0 * w + x0], b = t[y0 * w + x1], c = t[y1 * w + x0], d = t[y1 * w + x1];
    float4 r;
    float tx, bx;
    tx = a.x + (b.x - a.x) * fx; bx = c.x + (d.x - c.x) * fx; r.x = tx + (bx - tx) * fy;
    tx = a.y + (b.y - a.y) * fx; bx = c.y + (d.y - c.y) * fx; r.y = tx + (bx - tx) * fy;
    tx = a.z + (b.z - a.z) * fx; bx = c.z + (d.z - c.z) * fx; r.z = tx + (bx - tx) * fy;
    tx = a.w + (b.w - a.w) * fx; bx = c.w + (d.w - c.w) * fx; r.w = tx + (bx - tx) * fy;
    return r;
}

// Texel index in the 4x4x4-bricked 128^3 base volume: bricks are laid out
// x fastest (32 per axis), 64 texels per brick x fastest; the index is the
// sum of one term per axis.
__device__ __forceinline__ int brick_x(int x) { return ((x >> 2) << 6) | (x & 3); }
__device__ __forceinline__ int brick_y(int y) { return ((y >> 2) << 11) | ((y & 3) << 2); }
__device__ __forceinline__ int brick_z(int z) { return ((z >> 2) << 16) | ((z & 3) << 4); }

// Accurate 1 - exp(-x) for x >= 0 (the energy-conserving step factor).
// For small x the direct form loses all digits in FP32; a short series is
// used there, the exponential elsewhere.
__device__ __forceinline__ float one_minus_exp_neg(float x) {
    if (x < 0.0625f) {
        // x - x^2/2 + x^3/6 - x^4/24 + x^5/120: error < x^6/720 ~ 8e-11 at x = 1/16
        float p = 1.f / 120.f;
        p = p * (-x) + 1.f / 24.f;
        p = p * (-x) + 1.f / 6.f;
        p = p * (-x) + 0.5f;
        p = p * (-x) + 1.f;
        return x * p;
    }
    return 1.f - __expf(-x);
}
