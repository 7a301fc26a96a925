// Host-side launchers of the gfx950 kernels (one per translation unit).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct LutParams;
struct RenderParams;

// LUT chain of n scenes: d_params holds n LutParams (device); `dims` is a
// host copy of one of them (every scene of a launch has the same LUT sizes);
// scene s writes T + s*tw*th, MS + s*mw*mh, SV + s*sw*sh, AP + s*aw*ah*ad.
// flags: bit0 transmittance, bit1 multi-scatter, bit2 sky-view, bit3 AP.
// ev: optional 5 events recorded around the four passes.
hipError_t ps_launch_luts(const LutParams* d_params, const LutParams& dims, int n, uint32_t flags, float4* T,
                          float4* MS, float4* SV, float4* AP, hipStream_t s, hipEvent_t* ev);

// Single-scene render: `delta` enables the density-perturbation volume,
// `train` also stores the checkpoints and totals the backward reads.
// `rt`: the march writes the cloud RT (W x H = the RT size; no sky pass,
// luminance and transmittance to out_lum / out_trans).
hipError_t ps_launch_render(const RenderParams* d_params, int W, int H, bool count, bool delta, bool train,
                            hipStream_t s, bool rt = false);
// Full-resolution composite over the cloud RT (W x H viewport): Catmull-Rom
// upsample of a half-resolution RT, or the texel of a full-size one.
hipError_t ps_launch_composite_rt(const RenderParams* d_params, int W, int H, bool upsample, hipStream_t s);
// Cloud TAA resolve of one RT.
hipError_t ps_launch_taa_resolve(const TaaParams& p, hipStream_t s);
// Backward of a training render (the same parameter block, with cotangents
// and gradient outputs set).
// With g_delta non-null the perturbation-volume gradient is accumulated in
// FP64 (g_delta64: zero on entry, zero again on return) and then added into
// g_delta.
hipError_t ps_launch_render_backward(const RenderParams* d_params, int W, int H, bool delta, float* g_delta,
                                     double* g_delta64, hipStream_t s);
// p[0, n) = 0 (the FP64 gradient accumulator, once after allocation).
hipError_t ps_launch_zero_f64(double* p, int n, hipStream_t s);
// Per-brick maxima of the 128^3 density-perturbation volume (32^3 bricks of
// 4^3 texels plus a +1 halo: Repeat in x and z, ClampToEdge in height).
hipError_t ps_launch_delta_bounds(const float* delta, float2* bnd, hipStream_t s);

// Many scenes in one launch: `tiles` lists (scene << 20 | tile_y << 10 |
// tile_x) for every wave tile of every scene; a wave tile is tile_w x
// (64 / tile_w) pixels (tile_w in {4, 8, 16}). persistent != 0: `blocks`
// one-wave workers pull tiles from the counter `next` (zeroed by this call,
// with a kernel, so the sequence can be captured in a graph); else one block
// per tile.
// Persistent workers take `grab` consecutive list entries per counter increment.
hipError_t ps_launch_render_tiles(const RenderParams* d_params, const uint32_t* tiles, int n_tiles, int tile_w,
                                  int persistent, int blocks, int grab, unsigned* next, hipStream_t s);

// Noise: bake the stored volumes (format 0 = UNORM8, 1 = FP16), then fold
// them into the render layouts.
hipError_t ps_launch_noise_bake(uint32_t seed, int format, void* base, void* detail, void* curl, hipStream_t s);
// The factors that turn the folded base volume's fixed-point R and LF
// channels back into values, for the noise storage format (0 UNORM8, 1
// binary16).
inline void base_fold_scales(int format, float& r_scale, float& lf_scale) {
    r_scale = format == 0 ? 1.f / 65535.f : 1.f / 32768.f;
    lf_scale = format == 0 ? 1.f / 2040.f : 1.f / 65535.f;
}

hipError_t ps_launch_noise_fold(int format, const void* base, const void* detail, const void* curl,
                                uint32_t* base_rl, float* detail_hf, float2* curl2, uint16_t* bbound,
                                hipStream_t s);
// Presentation (post_kernel.hip). Images are W*H FP32 RGBA.
// Ground shading (surface pixels of the depth buffer): `ground` radiance, or
// the ground pass itself with `pbr`; `ground` NULL and `pbr` 0 keep the
// composite. The AP, the cloud composite and (with `ldr`) the tone map follow.
struct PresentArgs {
    const float4* composite;
    const float4* luminance;
    const float4* transmittance;
    const float4* ground;
    float4* hdr;
    uint32_t* ldr;                    // RGBA8 sRGB, R in the low byte
    float exposure;
    int mode;                         // bit 0: clamp (else ACES filmic); bit 1: Rgba8Unorm target (else sRGB)
    // ground pass (pbr.wgsl fs_main) inputs
    int pbr;
    uint32_t material;
    float albedo_scale[3];            // world ground_albedo / 0.18
    float wetness, puddle_start, puddle_coverage, precip, temperature, snow_depth;
    float sim_seconds;
};
hipError_t ps_launch_present(const RenderParams* d_params, int W, int H, const PresentArgs& a, hipStream_t s);
// Post-process chain on the HDR image `hdr` (W x H, updated in place):
// godrays (when `godrays`: radial blur into `rays`, rw x rh, toward the sun
// at `sun_uv`, then added) and bloom (when `bloom`: bright pass, pyramid in
// `levels` of sizes lw[i] x lh[i], i < 3), then the final add of the bloom,
// the optional HDR copy to `out_hdr` and the tone map to `ldr`.
struct PostArgs {
    float4* hdr;
    int W, H;
    int godrays;
    float4* rays;
    int rw, rh;
    float sun_u, sun_v;
    int samples;
    float decay, rays_intensity, bright_threshold;
    int bloom;
    float4* levels[3];
    int lw[3], lh[3];
    float bloom_threshold, bloom_knee, bloom_intensity;
    float4* out_hdr;
    uint32_t* ldr;
    float exposure;
    int mode;
};
hipError_t ps_launch_post(const PostArgs& a, hipStream_t s);
// Auto-exposure statistics (sum of log2 luminance, pixel count) of a W x H
// image into out[0..1].
hipError_t ps_launch_exposure_stats(const float4* hdr, int W, int H, float* out, hipStream_t s);
// Weather: binary16 channel 0 of a w*h*channels map -> per-texel 2x2 quads
// (Repeat wrap) of the coverage plane, still binary16.
hipError_t ps_launch_weather_fold(const uint16_t* src, int w, int h, int channels, uint2* dst, hipStream_t s);
// Weather-map synthesis: the 128x128 RGBA binary16 map (R from the coverage
// grid when `grid` is non-NULL, else `gate`) and, when `type_src` is non-NULL,
// the 128x128 cloud-type grid resampled from it (same gw x gh source).
hipError_t ps_launch_weather_synth(const float* grid, int gw, int gh, float extent, float gate, float alpha,
                                   const uint8_t* type_src, uint16_t* out_rgba, uint8_t* out_types, hipStream_t s);
// 64x64 blue-noise rank map (R8 bytes) from the Gaussian splat weights by
// squared texel distance, w[r2] = exp(-r2 / (2 * 1.9^2)) as the host computes
// them (r2 = 0..72).
constexpr int kBlueWeights = 73;
struct BlueNoiseWeights {
    float w[kBlueWeights];
};
hipError_t ps_launch_blue_noise(const BlueNoiseWeights& w, uint8_t* out, hipStream_t s);

// ---- atmosphere differentiation (atmo_adjoint.hip) ----------------------------
// Overwrite the 14 differentiated coefficients of n LUT parameter blocks from
// a device array (n x 14: beta_R[3], beta_Ms[3], beta_Ma[3], beta_O[3], H_R,
// H_M).
hipError_t ps_launch_set_coeffs(LutParams* d_params, const float* coeffs, int n, hipStream_t s);
// Tangent LUTs of n sets (value LUTs T and MS of the same sets must be baked):
// per set [channel][6 tangents][texel] for T, MS, SV and [channel][12][froxel]
// for AP (in-scatter tangents, then transmittance tangents).
hipError_t ps_launch_lut_tangents(const LutParams* d_params, const LutParams& dims, int n, const float4* T,
                                  const float4* MS, float* TT, float* MST, float* SVT, float* APT, hipStream_t s);

// One frame's sky + AP image over n parameter sets: per set, 14 gradients and
// the L2 loss against `target` into out[set * 16 + 0..14].
struct SkyApLaunch {
    const LutParams* lp;
    const float4 *T, *SV, *AP;
    const float *TT, *SVT, *APT;
    const float* depth;
    const float* target;
    const float* weight;              // per-pixel loss weights or null
    float* out;
    float* image;                     // optional image of set 0
    int W, H;
    int want_grad;                    // 0: image only (target, tangents and out unused)
    float ground;
    double inv_vp[16], cam[3], sun_d[3], cos_disk, sun_ang;
    float sun[3], E[3], cam_zha, sun_az, cam_alt, R, top, big_h, top_minus_R;
};
hipError_t ps_launch_sky_ap_grad(const SkyApLaunch& L, int n, hipStream_t s);
