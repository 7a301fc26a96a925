"""Frozen per-event cost table of the sky + cloud path (the roofline model).

Algorithmic work of a frame = event counts (``ps_count_work``) x the costs
below. FLOPs count arithmetic VALU operations of the path as written in the
reference WGSL (one per add/mul/fma/min/max/compare-select, one per
transcendental or sqrt/div); bytes count the texels a trilinear/bilinear tap
reads (the algorithmic gather volume, before any cache). Hardware counters
never feed these numbers.

Per event (cloud_march.wgsl / cloud_density_common.wgsl line ranges):
  primary_steps  : march position, stable altitude, weather bilinear, gates
                   (:869-883)                                   48 FLOP, 16 B
  covered_steps  : base trilinear (RGBA folded to 2 channels), LF-FBM, remap,
                   NDF, coverage remap (:458-494)               78 FLOP, 64 B
  eroded_steps   : curl bilinear, detail trilinear, erosion remap (:502-527)
                                                                62 FLOP, 64 B
  light_steps    : light-march position + altitude + gates (:570-577)
                                                                22 FLOP,  0 B
  light_covered  : base trilinear + shape in the light march (density_common
                   :384-414)                                    74 FLOP, 64 B
  dense_samples  : Beer-Powder, sun transmittance tap, sky ambient tap,
                   lightning, energy-conserving step, luminance-weighted depth
                   (:886-993 minus the octave loop)            230 FLOP, 128 B
  octaves        : one multi-scatter octave (:919-927; phase hoisted per
                   pixel)                                        8 FLOP,  0 B
  pixels         : view ray, sky pass (sky.wgsl:159-220), layer intervals +
                   sort, per-layer phase setup, AP trilinear, composite
                                                               420 FLOP, 213 B
  ap_taps        : counted inside `pixels`                       0 FLOP,  0 B
LUT chain per frame (FP64, per step of each bake): transmittance 40 steps x
(256*64) x 60, multiscatter 20 x 64 x (32*32) x 120, sky-view 32 x
(192*108) x 110, AP 16 x (32*32*32) x 110.
"""
from __future__ import annotations

FLOPS = {
    "pixels": 420.0,
    "layer_marches": 0.0,
    "primary_steps": 48.0,
    "covered_steps": 78.0,
    "eroded_steps": 62.0,
    "dense_samples": 230.0,
    "light_steps": 22.0,
    "light_covered": 74.0,
    "octaves": 8.0,
    "ap_taps": 0.0,
}

BYTES = {
    "pixels": 213.0,
    "layer_marches": 0.0,
    "primary_steps": 16.0,
    "covered_steps": 64.0,
    "eroded_steps": 64.0,
    "dense_samples": 128.0,
    "light_steps": 0.0,
    "light_covered": 64.0,
    "octaves": 0.0,
    "ap_taps": 0.0,
}

# MI355X published peaks. The FP32 vector figure counts packed FP32 (two
# lanes per VALU op); the kernels here are built without packed FP32, so
# 256 CUs x 4 SIMD x 16 lanes x 2 FLOP x 2.4 GHz = 78.6 TFLOP/s is the rate
# they can reach.
PEAK_FP32_VALU_TFLOPS = 157.3
PEAK_FP32_VALU_UNPACKED_TFLOPS = 78.6
PEAK_HBM_GBPS = 8000.0

# Inputs one frame reads at least once from HBM: base volume (128^3 x 2 x u16),
# its brick maxima (32^3 u8), detail (32^3 f32), curl (128^2 x 2 f32), the
# 512^2 weather quads (4 x binary16) and the four FP32 LUTs of the benchmark
# dimensions, plus the per-pixel depth (f32).
INPUT_BYTES = (128 ** 3 * 4 + 32 ** 3 + 32 ** 3 * 4 + 128 ** 2 * 8 + 512 ** 2 * 8
               + (256 * 64 + 32 * 32 + 192 * 108 + 32 ** 3) * 16)


def render_hbm_bytes(width: int, height: int, out_bytes_per_px: int = 16) -> float:
    """Algorithmic HBM bytes of one render: inputs once, depth in, output out."""
    return float(INPUT_BYTES + width * height * (4 + out_bytes_per_px))


def lut_flops(dims) -> float:
    return (40 * dims.transmittance_w * dims.transmittance_h * 60
            + 20 * 64 * dims.multiscatter_w * dims.multiscatter_h * 120
            + 32 * dims.skyview_w * dims.skyview_h * 110
            + 16 * dims.ap_w * dims.ap_h * dims.ap_d * 110)


def render_flops(counters) -> float:
    return sum(getattr(counters, k) * v for k, v in FLOPS.items())


def render_bytes(counters) -> float:
    return sum(getattr(counters, k) * v for k, v in BYTES.items())
