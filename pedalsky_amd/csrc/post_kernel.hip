// Presentation of a rendered frame for gfx950: aerial perspective applied to
// an opaque surface's radiance, the dual-source cloud composite over it, and
// the tone map to the sRGB swapchain the reference reads its goldens from.
//
//  * Surface pixels (depth buffer holds a surface): L = lit * ap.a + ap.rgb,
//    the AP froxel tap at the pixel's viewport uv and the surface distance
//    mapped as atmosphere_lut_sampling.wgsl:33-38 (shaders/ground/pbr.wgsl:
//    653-666); then hdr = cloud_luminance + L * cloud_transmittance
//    (cloud_composite.wgsl:47-57). Sky pixels keep the render's composite.
//  * Tone map (shaders/postprocess/tonemap.wgsl:37-58): exposure
//    1 / (1.2 * 2^EV100), ACES filmic (Narkowicz fit) or clamp, then the sRGB
//    encode of the *Srgb swapchain, rounded to 8 bits.
// One thread per pixel: the pass is bandwidth-bound (reads 16-64 B, writes
// 4-20 B per pixel).
#include "ps_device.h"
#include "ps_launch.h"

#include "ps_march.h"

namespace {

// Trilinear ClampToEdge tap of the AP LUT (RGBA32F, aw x ah x ad) at uvw.
__device__ __forceinline__ float4 ap_tap(const RenderParams& P, float u, float v, float w) {
    const float cx = u * P.aw - 0.5f, cy = v * P.ah - 0.5f, cz = w * P.ad - 0.5f;
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
    auto tri = [&](float v000, float v100, float v010, float v110, float v001, float v101, float v011,
                   float v111) {
        const float a0 = v000 + (v100 - v000) * wx, a1 = v010 + (v110 - v010) * wx;
        const float b0 = v001 + (v101 - v001) * wx, b1 = v011 + (v111 - v011) * wx;
        const float p0 = a0 + (a1 - a0) * wy, p1 = b0 + (b1 - b0) * wy;
        return p0 + (p1 - p0) * wz;
    };
    return make_float4(tri(c000.x, c100.x, c010.x, c110.x, c001.x, c101.x, c011.x, c111.x),
                       tri(c000.y, c100.y, c010.y, c110.y, c001.y, c101.y, c011.y, c111.y),
                       tri(c000.z, c100.z, c010.z, c110.z, c001.z, c101.z, c011.z, c111.z),
                       tri(c000.w, c100.w, c010.w, c110.w, c001.w, c101.w, c011.w, c111.w));
}

// ACES filmic, Narkowicz's fit (tonemap.wgsl:37-44), saturated.
__device__ __forceinline__ float aces(float x) {
    return fminf(fmaxf((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.f), 1.f);
}

// Linear -> sRGB 8-bit (the encode an Rgba8UnormSrgb target applies on
// store), rounded to nearest.
__device__ __forceinline__ uint32_t srgb8(float c) {
    c = fminf(fmaxf(c, 0.f), 1.f);
    const float e = c <= 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.f / 2.4f) - 0.055f;
    return (uint32_t)fminf(floorf(e * 255.f + 0.5f), 255.f);
}

__global__ __launch_bounds__(256) void k_present(const RenderParams* __restrict__ Pp, PresentArgs A) {
    const RenderParams& P = *Pp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.W * P.H) return;
    const int x = i % P.W, y = i / P.W;
    float4 hdr = A.composite[i];
    if (A.ground && P.has_depth && P.depth[i] > 0.f) {
        const PixelRay ray = pixel_ray(P, x, y);
        const float d = (float)pixel_depth_tmax(P, ray, x, y);
        const float ds = fmaxf(d, 50.f);
        const float apz = clampf(__logf(ds / 50.f) * (float)(1.0 / 7.6009024595420822), 0.f, 1.f);
        const float4 ap = ap_tap(P, ((float)x + 0.5f) * P.inv_W, ((float)y + 0.5f) * P.inv_H, apz);
        const float4 g = A.ground[i], lum = A.luminance[i], tr = A.transmittance[i];
        hdr.x = lum.x + (g.x * ap.w + ap.x) * tr.x;
        hdr.y = lum.y + (g.y * ap.w + ap.y) * tr.y;
        hdr.z = lum.z + (g.z * ap.w + ap.z) * tr.z;
        hdr.w = 1.f;
    }
    if (A.hdr) A.hdr[i] = hdr;
    if (A.ldr) {
        float r = hdr.x * A.exposure, gg = hdr.y * A.exposure, b = hdr.z * A.exposure;
        if (A.mode == 0) {
            r = aces(r); gg = aces(gg); b = aces(b);
        }
        A.ldr[i] = srgb8(r) | (srgb8(gg) << 8) | (srgb8(b) << 16) | (255u << 24);
    }
}

}  // namespace

hipError_t ps_launch_present(const RenderParams* d_params, int W, int H, const PresentArgs& a, hipStream_t s) {
    const int n = W * H;
    hipLaunchKernelGGL(k_present, dim3((n + 255) / 256), dim3(256), 0, s, d_params, a);
    return hipGetLastError();
}
