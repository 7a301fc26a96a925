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

# MI355X published peaks
PEAK_FP32_VALU_TFLOPS = 157.3
PEAK_HBM_GBPS = 8000.0


def lut_flops(dims) -> float:
    return (40 * dims.transmittance_w * dims.transmittance_h * 60
            + 20 * 64 * dims.multiscatter_w * dims.multiscatter_h * 120
            + 32 * dims.skyview_w * dims.skyview_h * 110
            + 16 * dims.ap_w * dims.ap_h * dims.ap_d * 110)


def render_flops(counters) -> float:
    return sum(getattr(counters, k) * v for k, v in FLOPS.items())


def render_bytes(counters) -> float:
    return sum(getattr(counters, k) * v for k, v in BYTES.items())
