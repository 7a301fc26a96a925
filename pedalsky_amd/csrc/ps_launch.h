// Host-side launchers of the gfx950 kernels (one per translation unit).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct LutParams;
struct RenderParams;

// flags: bit0 transmittance, bit1 multi-scatter, bit2 sky-view, bit3 AP.
// ev: optional 5 events recorded around the four passes.
hipError_t ps_launch_luts(const LutParams& P, uint32_t flags, float4* T, float4* MS, float4* SV, float4* AP,
                          hipStream_t s, hipEvent_t* ev);

hipError_t ps_launch_render(const RenderParams* d_params, int W, int H, bool count, hipStream_t s);

// Noise: bake the stored volumes (format 0 = UNORM8, 1 = FP16), then fold
// them into the render layouts.
hipError_t ps_launch_noise_bake(uint32_t seed, int format, void* base, void* detail, void* curl, hipStream_t s);
hipError_t ps_launch_noise_fold(int format, const void* base, const void* detail, const void* curl,
                                uint32_t* base_rl, float* detail_hf, float2* curl2, uint8_t* rmax, hipStream_t s);
// Weather: binary16 channel 0 of a w*h*channels map -> per-texel 2x2 quads
// (Repeat wrap) of the coverage plane, still binary16.
hipError_t ps_launch_weather_fold(const uint16_t* src, int w, int h, int channels, uint2* dst, hipStream_t s);
