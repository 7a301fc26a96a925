"""Oracle: the reference's blue-noise rank map, restated in numpy.

Test infrastructure only (tests/, smoke() and bench.py's CPU baseline may use
it; the product never does). generate_blue_noise(64)
(crates/ps-clouds/src/noise.rs:646-704): 4096 times, the untaken texel of
lowest energy (the first in scan order among equals) takes the next rank and
adds a Gaussian splat of Chebyshev radius 6 (sigma 1.9) to the energy field
on the 64x64 torus; ranks quantise to round(r / 4095 * 255).

float32 throughout, in the reference's order: the splat weight
exp(-r2 * (1 / (2 sigma^2))) is the correctly rounded float32 exponential of
the float32 product (tests/test_blue_noise.py pins these 73 weights against
the C library's expf, which Rust's f32::exp calls), and every texel gains at
most one weight per placement, so the energies are the reference's bit for
bit.

Pinned by tests/test_oracle_pins.py's blue-noise checks (the first nine
placements 7 texels apart along row 0, the quantised-ramp histogram, the
suppressed low frequencies).
"""
from __future__ import annotations

import math

import numpy as np

SIZE = 64
RADIUS = 6
SIGMA = np.float32(1.9)


def splat_weights() -> np.ndarray:
    """w[r2] for r2 = 0..72 (noise.rs:655-675)."""
    inv = np.float32(1.0) / (np.float32(2.0) * SIGMA * SIGMA)
    out = np.zeros(73, np.float32)
    for r2 in range(73):
        x = np.float32(-np.float32(r2)) * inv
        out[r2] = np.float32(math.exp(float(x)))
    return out


def generate(size: int = SIZE) -> np.ndarray:
    """(size * size,) uint8 ranks in row-major [y, x] order."""
    n = size * size
    w = splat_weights()
    dy, dx = np.mgrid[-RADIUS:RADIUS + 1, -RADIUS:RADIUS + 1]
    wk = w[(dx * dx + dy * dy).ravel()]
    dx, dy = dx.ravel(), dy.ravel()
    energy = np.zeros(n, np.float32)
    masked = np.zeros(n, np.float32)       # energy, +inf where taken
    ranks = np.zeros(n, np.int64)
    for placement in range(n):
        best = int(np.argmin(masked))      # first index of the minimum
        ranks[best] = placement
        masked[best] = np.inf
        cx, cy = best % size, best // size
        idx = ((cy + dy) % size) * size + (cx + dx) % size
        energy[idx] = energy[idx] + wk      # distinct texels: one add each
        live = np.isfinite(masked[idx])
        masked[idx[live]] = energy[idx[live]]
    r = ranks.astype(np.float32) / np.float32(n - 1) * np.float32(255.0)
    fl = np.floor(r)
    rounded = np.where(r - fl >= np.float32(0.5), fl + np.float32(1.0), fl)  # half away from zero
    return np.clip(rounded, 0, 255).astype(np.uint8)
