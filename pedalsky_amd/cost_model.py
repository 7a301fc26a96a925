"""Frozen per-event cost table of the sky + cloud path (the roofline model).

Algorithmic work of a frame = event counts (``ps_count_work``) x the costs
below. The costs are hand counts of the reference WGSL as written (its
unfolded channels: the base volume's four channels, the detail volume's
three, the phase constants per dense sample), not of this build's folded
implementation. Convention (the one the MI355X FP32 peak uses): an FMA is 2
FLOP, so a texture lerp a + (b - a) w is 3 FLOP and a trilinear tap is 7
lerps = 21 FLOP per channel; a transcendental, square root or divide is 1
FLOP; compares, min/max and selects are 1 FLOP. Bytes count the texels a tap
reads in the benchmark formats (FP16 noise, RG FP16 weather map, FP32 LUTs),
the algorithmic gather volume before any cache. Hardware counters never feed
these numbers.

Per event (cloud_march.wgsl / cloud_density_common.wgsl line ranges):
  primary_steps  : position (6), altitude_from_entry (14), weather uv +
                   1-channel bilinear (21), cloud-type grid (6), height and
                   coverage gates (7) (:869-883, :432-447)           54 FLOP, 17 B
  covered_steps  : skew/advection (10), base coordinates (9), RGBA base
                   trilinear (84), LF-FBM (8), remap (6), NDF (18),
                   coverage remap (8) (:448-494)                   145 FLOP, 64 B
  eroded_steps   : curl bilinear RG (18) + coordinates (3), curl offset (6),
                   detail coordinates (6), RGB detail trilinear (63), HF FBM
                   (5), height mix (5), strength (4), erosion remap (6)
                   (:502-527)                                      116 FLOP, 80 B
  light_steps    : light position (8), altitude (14), gates (4)
                   (:570-577)                                       26 FLOP,  0 B
  light_covered  : sample_density_light after the gates (advection, base
                   trilinear, LF-FBM, remaps, NDF) + od accumulation
                   (density_common :384-414)                       145 FLOP, 64 B
  dense_samples  : Beer-Powder (30), phase height (4), phase constants (18),
                   sun transmittance with Earth shadow and LUT tap (70),
                   height (4), sky ambient uv + tap (61), in-scatter (12),
                   energy-conserving step, depth weight, transmittance (51)
                   (:886-993 minus the octave loop)                250 FLOP, 128 B
  octaves        : HG (11) + Draine (16) + mix (3) + sun term and octave
                   factors (20) (:919-927, :151-168, :278-296)      50 FLOP,  0 B
  pixels         : view ray (45), sky pass (95), jitter + depth unprojection
                   (43), layer intervals, sort and chromatic factors (85),
                   AP depth + RGBA trilinear (95), composite (12)  375 FLOP, 236 B
  ap_taps        : counted inside `pixels`                           0 FLOP,  0 B
LUT chain per frame (FP64, per step of each bake): transmittance 40 steps x
(256*64) x 60, multiscatter 20 x 64 x (32*32) x 120, sky-view 32 x
(192*108) x 110, AP 16 x (32*32*32) x 110.
"""
from __future__ import annotations

FLOPS = {
    "pixels": 375.0,
    "layer_marches": 0.0,
    "primary_steps": 54.0,
    "covered_steps": 145.0,
    "eroded_steps": 116.0,
    "dense_samples": 250.0,
    "light_steps": 26.0,
    "light_covered": 145.0,
    "octaves": 50.0,
    "ap_taps": 0.0,
}

BYTES = {
    "pixels": 236.0,
    "layer_marches": 0.0,
    "primary_steps": 17.0,
    "covered_steps": 64.0,
    "eroded_steps": 80.0,
    "dense_samples": 128.0,
    "light_steps": 0.0,
    "light_covered": 64.0,
    "octaves": 0.0,
    "ap_taps": 0.0,
}

# MI355X published peaks. The FP32 vector figure counts packed FP32 (two
# values per lane per VALU op) and an FMA as 2 FLOP; the kernels here are
# built without packed FP32, so 256 CUs x 4 SIMD x 16 lanes x 2 FLOP x 2.4
# GHz = 78.6 TFLOP/s is the rate they can reach.
PEAK_FP32_VALU_TFLOPS = 157.3
PEAK_FP32_VALU_UNPACKED_TFLOPS = 78.6
PEAK_HBM_GBPS = 8000.0

# Inputs one frame reads at least once from HBM: the base volume (128^3 texels,
# R and LF as 2 x u16), its brick bounds (8^3 x u16), detail (32^3 f32), curl
# (128^2 x 2 f32), the 512^2 weather quads (4 x binary16) and the four FP32
# LUTs of the benchmark dimensions; the per-pixel depth (f32) is added when a
# depth buffer is read. The folded copies that store each texel with its filter
# neighbours hold the same values four (base, detail) and two (curl) times;
# those repeats are a layout choice, not algorithmic bytes.
INPUT_BYTES = (128 ** 3 * 4 + 8 ** 3 * 2 + 32 ** 3 * 4 + 128 ** 2 * 8 + 512 ** 2 * 8
               + (256 * 64 + 32 * 32 + 192 * 108 + 32 ** 3) * 16)


def render_hbm_bytes(width: int, height: int, out_bytes_per_px: int = 16, depth: bool = True) -> float:
    """Algorithmic HBM bytes of one render: inputs once, depth in (when
    there is a depth buffer), outputs out."""
    return float(INPUT_BYTES + width * height * ((4 if depth else 0) + out_bytes_per_px))


def lut_flops(dims) -> float:
    return (40 * dims.transmittance_w * dims.transmittance_h * 60
            + 20 * 64 * dims.multiscatter_w * dims.multiscatter_h * 120
            + 32 * dims.skyview_w * dims.skyview_h * 110
            + 16 * dims.ap_w * dims.ap_h * dims.ap_d * 110)


def render_flops(counters) -> float:
    return sum(getattr(counters, k) * v for k, v in FLOPS.items())


def render_bytes(counters) -> float:
    return sum(getattr(counters, k) * v for k, v in BYTES.items())
