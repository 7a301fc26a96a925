"""The weather-map / cloud-type-grid oracle (oracle/weather.py) against the
reference's own tests (crates/ps-synthesis/src/weather_map.rs:202-260) and
closed forms; the entry points are exported. CPU only."""
import numpy as np
import pytest

import oracle.weather as W


def _hash2_int(x: int, y: int) -> float:
    """hash2 with Python integers (weather_map.rs:160-167)."""
    m = 0xFFFFFFFF
    h = ((x & m) * 2654435761) & m
    h ^= ((y & m) * 2246822519) & m
    h ^= h >> 16
    h = (h * 2654435761) & m
    h ^= h >> 13
    return float(np.float32(h)) / 4294967296.0


def test_reference_dimensions_and_gate():
    # weather_map.rs:202-209, 211-240, 242-250
    wm = W.synthesise_weather(has_layers=True)
    assert wm.shape == (128, 128, 4) and W.SIZE == 128 and W.EXTENT_M == 32000.0
    flat = wm.reshape(-1)
    for i in (0, 4 * 64 * 128 + 4 * 64, flat.size - 4):
        assert abs(float(flat[i]) - 1.0) < 1e-3
    flat0 = W.synthesise_weather(has_layers=False).reshape(-1)
    for i in (0, 4 * 64 * 128 + 4 * 64, flat0.size - 4):
        assert float(flat0[i]) == 0.0
    assert np.all(wm[..., 1] == 0)


def test_pixel_to_world_centres_at_origin():
    # weather_map.rs:252-259
    cx, cy = W.pixel_to_world(63, 63)
    assert abs(cx) < 32000.0 / 128 and abs(cy) < 32000.0 / 128
    x0, _ = W.pixel_to_world(0, 0)
    assert x0 == np.float32(-16000.0 + 125.0)


def test_hash_matches_integer_arithmetic():
    rng = np.random.default_rng(3)
    xs = rng.integers(-100000, 100000, 300)
    ys = rng.integers(-100000, 100000, 300)
    got = W.hash2(xs, ys)
    for x, y, g in zip(xs, ys, got):
        assert float(g) == np.float32(_hash2_int(int(x), int(y)))


def test_value_noise_lattice_and_smoothstep():
    for x, y in ((0, 0), (3, 5), (-2, 7), (6, 6)):
        assert float(W.value_noise_2d(np.float32(x), np.float32(y))) == float(W.hash2(x, y))
    # at xf = 0.25 the smoothstep weight is 0.15625 (not the linear 0.25)
    h00, h10 = _hash2_int(4, 9), _hash2_int(5, 9)
    v = float(W.value_noise_2d(np.float32(4.25), np.float32(9.0)))
    assert abs(v - (h00 + (h10 - h00) * 0.15625)) < 1e-6
    b = W.synthesise_weather()[..., 2].astype(np.float32)
    assert b.min() >= -1.0 and b.max() <= 1.0 and b.std() > 0.1


def test_grid_sampler_closed_forms():
    w, h, ext = 9, 5, 20000.0
    xs = np.linspace(-12000.0, 12000.0, 97, dtype=np.float32)
    zs = np.float32(1234.0)
    # a constant grid is returned exactly; a ramp in x is reproduced (edge-clamped)
    assert np.all(W.grid_sample_world(np.full(w * h, 0.375, np.float32), w, h, ext, xs, zs) == np.float32(0.375))
    ramp = np.tile(np.arange(w, dtype=np.float32) / np.float32(w - 1), h)
    got = W.grid_sample_world(ramp, w, h, ext, xs, zs)
    want = np.clip((xs.astype(np.float64) + ext / 2) / ext, 0.0, 1.0)
    assert np.max(np.abs(got - want)) < 1e-6
    # the corner texels sit at +-extent/2
    g = np.arange(w * h, dtype=np.float32) / np.float32(w * h)
    assert W.grid_sample_world(g, w, h, ext, np.float32(-ext / 2), np.float32(-ext / 2)) == g[0]
    assert W.grid_sample_world(g, w, h, ext, np.float32(ext / 2), np.float32(ext / 2)) == g[-1]


def test_alpha_and_type_grid():
    assert np.all(W.synthesise_weather(precip_mm_per_h=10.0, ground_wetness=0.1)[..., 3] == np.float16(0.2))
    assert np.all(W.synthesise_weather(precip_mm_per_h=80.0)[..., 3] == np.float16(1.0))
    assert np.all(W.synthesise_weather(ground_wetness=0.6)[..., 3] == np.float16(0.6))
    src = (np.arange(128 * 128) % 11).astype(np.uint8)
    t = W.synthesise_type_grid(src, 128, 128).reshape(-1)
    assert np.all(t == np.where(src <= 7, src, 255))
    assert np.all(W.synthesise_type_grid(None) == 255)
    # nearest resample: target x reads source floor(x * w / 128)
    src2 = np.arange(5 * 3, dtype=np.uint8) % 8
    t2 = W.synthesise_type_grid(src2, 5, 3)
    assert t2[0, 0] == src2[0] and t2[127, 127] == src2[2 * 5 + 4] and t2[64, 26] == src2[1 * 5 + 1]


def test_entry_points_exported():
    from pedalsky_amd import abi
    lib = abi.load()
    for name in ("ps_synthesise_weather", "ps_read_weather"):
        assert hasattr(lib, name)


def _cumulus(cov):
    from pedalsky_amd import scene
    return scene.synth_layer(0, 1500.0, 2300.0, cov, density_scale=1.0)


def test_overcast_mask_reference_cases():
    """TopDownMask::synthesise, the reference's tests (density_mask.rs:
    187-246): no layers give zero, cumulus gives a non-zero mask growing with
    coverage, and a 4x4 grid with full coverage on its left half and none on
    the right gives a mask that is higher on the left and near zero on the
    right."""
    from pedalsky_amd import scene
    assert not scene.overcast_mask([]).any()
    assert scene.overcast_mask([_cumulus(0.5)])[0] > 0
    assert scene.overcast_mask([_cumulus(0.9)])[0] > scene.overcast_mask([_cumulus(0.2)])[0]
    grid = np.array([[1.0 if (i % 4) < 2 else 0.0 for i in range(4 * r, 4 * r + 4)] for r in range(4)], np.float32)
    m = scene.overcast_mask([_cumulus(0.5)], grid, 32000.0).reshape(128, 128)
    left, right = int(m[64, 32]), int(m[64, 96])
    assert left > right and right < 8


def test_overcast_mask_grid_is_the_uniform_mask_at_constant_coverage():
    """The grid's coverage replaces each layer's own (synthesis-remapped)
    coverage: a constant grid holding a layer's coverage gives, at every
    texel, that layer's uniform mask; outside the grid's extent the edge
    values hold (ClampToEdge)."""
    from pedalsky_amd import scene
    for c in (0.0, 0.35, 1.0):
        for layer in (_cumulus(c), scene.synth_layer(5, 8000.0, 9000.0, c, density_scale=0.3)):
            want = scene.overcast_mask([layer])[0]
            got = scene.overcast_mask([layer], np.full((7, 5), layer.coverage, np.float32), 20000.0)
            assert (got == want).all(), (c, want, np.unique(got))
    # a grid over 8 km whose columns step from 0 to 1: beyond its extent the
    # mask takes the edge columns' values
    layers = [_cumulus(1.0)]
    g = np.tile(np.linspace(0.0, 1.0, 8, dtype=np.float32), (8, 1))
    m = scene.overcast_mask(layers, g, 8000.0).reshape(128, 128)
    full = scene.overcast_mask(layers, np.ones((2, 2), np.float32), 8000.0)[0]
    assert (m[:, :40] == 0).all() and (m[:, 88:] == full).all()
    assert (np.diff(m[64].astype(int)) >= 0).all()
