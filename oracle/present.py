"""Oracle: presentation of a frame (surface aerial perspective, composite,
tone map), restated. Test infrastructure only (tests/, smoke() and bench.py's
CPU baseline may use it; the product never does).

  * shaders/ground/pbr.wgsl:653-666: L = lit * ap.a + ap.rgb, the AP froxel
    at the pixel's viewport uv and the surface distance;
  * cloud_composite.wgsl:47-57: hdr = cloud_luminance + L * transmittance;
  * shaders/postprocess/tonemap.wgsl:37-58: exposure 1 / (1.2 * 2^EV100),
    ACES filmic (Narkowicz) or clamp; then the sRGB encode of the *Srgb
    swapchain (IEC 61966-2-1) rounded to 8 bits.

FP64. Pinned by closed forms: ACES at 0 and 1 and its saturation, the
exposure at EV100 0, and the sRGB encode at the segment joint and at 0.5 /
1.0 (tests/test_present_oracle.py); the AP sampling it is fed with is the
oracle's own, pinned where it is defined.
"""
from __future__ import annotations

import numpy as np


def exposure(ev100: float) -> float:
    return 1.0 / (1.2 * 2.0 ** ev100)


def aces(x):
    x = np.asarray(x, np.float64)
    return np.clip((x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14), 0.0, 1.0)


def srgb_encode(c):
    c = np.clip(np.asarray(c, np.float64), 0.0, 1.0)
    return np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(c, 1.0 / 2.4) - 0.055)


def srgb8(c):
    return np.minimum(np.floor(srgb_encode(c) * 255.0 + 0.5), 255.0).astype(np.uint8)


def present(comp, lum, trn, depth, ground_lit, ap_rgb, ap_a, ev100, tonemap=0):
    """(hdr (h, w, 3), ldr (h, w, 3) u8). ap_rgb / ap_a: the AP froxel at each
    surface pixel (only read where depth > 0)."""
    hdr = np.array(comp[..., :3], np.float64)
    if ground_lit is not None:
        g = np.asarray(depth) > 0
        base = np.asarray(ground_lit, np.float64)[..., :3] * np.asarray(ap_a, np.float64)[..., None] + ap_rgb
        hdr[g] = np.asarray(lum, np.float64)[..., :3][g] + base[g] * np.asarray(trn, np.float64)[..., :3][g]
    s = hdr * exposure(ev100)
    m = aces(s) if tonemap == 0 else np.clip(s, 0.0, 1.0)
    return hdr, srgb8(m)
