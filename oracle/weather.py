"""Oracle: the reference's weather-map and cloud-type-grid synthesis, restated.

Test infrastructure only (tests/, smoke() and bench.py's CPU baseline may use
it; the product never does). numpy float32 arithmetic in the reference's
order (numpy never fuses a * b + c, and neither does Rust), so results are
bit-identical to crates/ps-synthesis/src/weather_map.rs:46-170,
coverage_grid.rs:31-61 and cloud_type_grid.rs:41-92.

Pinned by the reference's own tests (weather_map.rs:202-260: size and
extent, R = 1 with layers and 0 without, pixel_to_world centring), by
hash2 recomputed with Python integers, by the value noise equalling the hash
at lattice points, and by the bilinear sampler returning the grid values at
texel centres and a linear ramp exactly.
"""
from __future__ import annotations

import numpy as np

SIZE = 128
EXTENT_M = np.float32(32000.0)
SENTINEL = 255
f32 = np.float32


def pixel_to_world(x, y):
    """weather_map.rs:134-139."""
    half = EXTENT_M * f32(0.5)
    fx = (np.asarray(x, np.float32) + f32(0.5)) / f32(SIZE)
    fy = (np.asarray(y, np.float32) + f32(0.5)) / f32(SIZE)
    return fx * EXTENT_M - half, fy * EXTENT_M - half


def hash2(x, y):
    """weather_map.rs:160-167 (u32 wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        h = np.asarray(x, np.int64).astype(np.uint32) * np.uint32(2654435761)
        h ^= np.asarray(y, np.int64).astype(np.uint32) * np.uint32(2246822519)
        h ^= h >> np.uint32(16)
        h = h * np.uint32(2654435761)
        h ^= h >> np.uint32(13)
    # u32::MAX as f32 rounds to 2^32
    return np.clip(h.astype(np.float32) / f32(4294967296.0), f32(0.0), f32(1.0))


def value_noise_2d(x, y):
    """weather_map.rs:143-158."""
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    xi, yi = np.floor(x), np.floor(y)
    xf, yf = x - xi, y - yi
    ix, iy = xi.astype(np.int32), yi.astype(np.int32)
    h00, h10 = hash2(ix, iy), hash2(ix + 1, iy)
    h01, h11 = hash2(ix, iy + 1), hash2(ix + 1, iy + 1)
    sx = xf * xf * (f32(3.0) - f32(2.0) * xf)
    sy = yf * yf * (f32(3.0) - f32(2.0) * yf)
    bottom = h00 + (h10 - h00) * sx
    top = h01 + (h11 - h01) * sx
    return bottom + (top - bottom) * sy


def grid_sample_world(data, w, h, extent_m, x, z):
    """LoadedCoverageGrid::sample_world + sample (coverage_grid.rs:31-61)."""
    data = np.asarray(data, np.float32).reshape(-1)
    extent = f32(extent_m)
    half = extent * f32(0.5)
    u = (np.asarray(x, np.float32) + half) / extent
    v = (np.asarray(z, np.float32) + half) / extent
    fx = np.clip(u, f32(0), f32(1)) * (f32(w) - f32(1.0))
    fy = np.clip(v, f32(0), f32(1)) * (f32(h) - f32(1.0))
    x0 = np.floor(fx).astype(np.uint32)
    y0 = np.floor(fy).astype(np.uint32)
    x1 = np.minimum(x0 + 1, w - 1)
    y1 = np.minimum(y0 + 1, h - 1)
    tx = fx - x0.astype(np.float32)
    ty = fy - y0.astype(np.float32)
    v00, v10 = data[y0 * w + x0], data[y0 * w + x1]
    v01, v11 = data[y1 * w + x0], data[y1 * w + x1]
    bottom = v00 + (v10 - v00) * tx
    top = v01 + (v11 - v01) * tx
    return np.clip(bottom + (top - bottom) * ty, f32(0), f32(1))


def synthesise_weather(has_layers=True, precip_mm_per_h=0.0, ground_wetness=0.0, grid=None, grid_w=0, grid_h=0,
                       grid_extent_m=32000.0):
    """WeatherMap::synthesise (weather_map.rs:46-101): (128, 128, 4) float16."""
    y, x = np.meshgrid(np.arange(SIZE), np.arange(SIZE), indexing="ij")
    if grid is not None:
        gx, gy = pixel_to_world(x, y)
        cov = grid_sample_world(grid, grid_w, grid_h, grid_extent_m, gx, gy)
    else:
        cov = np.full((SIZE, SIZE), f32(1.0 if has_layers else 0.0), np.float32)
    bias = value_noise_2d(x.astype(np.float32) * f32(0.05), y.astype(np.float32) * f32(0.05)) * f32(2.0) - f32(1.0)
    precip = np.clip(f32(precip_mm_per_h) / f32(50.0), f32(0), f32(1))
    wet = np.clip(f32(ground_wetness), f32(0), f32(1))
    alpha = max(precip, wet)
    out = np.zeros((SIZE, SIZE, 4), np.float16)
    out[..., 0] = cov.astype(np.float16)
    out[..., 2] = bias.astype(np.float16)
    out[..., 3] = np.float16(alpha)
    return out


def synthesise_type_grid(src=None, grid_w=0, grid_h=0):
    """CloudTypeGrid::synthesise (cloud_type_grid.rs:41-92): (128, 128) u8."""
    if src is None:
        return np.full((SIZE, SIZE), SENTINEL, np.uint8)
    src = np.asarray(src, np.uint8).reshape(-1)
    y, x = np.meshgrid(np.arange(SIZE), np.arange(SIZE), indexing="ij")
    sx = np.minimum((x.astype(np.float32) / f32(SIZE) * f32(grid_w)).astype(np.uint32), grid_w - 1)
    sy = np.minimum((y.astype(np.float32) / f32(SIZE) * f32(grid_h)).astype(np.uint32), grid_h - 1)
    v = src[sy * grid_w + sx]
    return np.where(v <= 7, v, SENTINEL).astype(np.uint8)
