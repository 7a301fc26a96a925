"""GPU parity of the optional branches of the sky + cloud path against the FP64
oracle, through the C ABI, one case per feature the default workloads leave
off. Each case also checks that the feature changes the image, so a branch
that silently does nothing cannot pass.

Bar (same as the plain-path tests): relative radiance error 99.9th
percentile <= 1e-3, median <= 1e-5; cloud transmittance absolute error
99.9th percentile <= 1e-3.
"""
import sys

import numpy as np
import pytest

import oracle
from pedalsky_amd import abi, scene
from pedalsky_amd.renderer import Renderer

pytestmark = pytest.mark.gpu

W, H = 256, 144


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def renderer():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU (no CPU fallback exists)"
    r = Renderer(0)
    yield r
    r.close()


def _base(**kw):
    args = dict(width=W, height=H, steps=48)
    args.update(kw)
    return scene.workload_small(**args)


def _type_grid():
    rng = scene.pcg_uniforms(0x7E, 128 * 128)
    g = np.floor(rng * 10.0).astype(np.int64)          # 0..9
    g = np.where(g >= 8, 255, g).astype(np.uint8)       # 8, 9 -> sentinel (layer type)
    return g


def _variant(name):
    """(workload, description) for one feature case."""
    if name == "cone_light_sampling":
        w = _base()
        w.clouds.cone_light_sampling = 1
        return w, "Nubis cone light march (cloud_march.wgsl:599-666)"
    if name == "cloud_type_grid":
        w = _base()
        w.type_grid = _type_grid()
        return w, "per-texel cloud types (cloud_density_common.wgsl:233-251)"
    if name == "cumulonimbus_ice":
        layers = [scene.synth_layer(7, 1500.0, 9000.0, 0.6, anvil_bias=1.5)]
        w = _base(layers=layers, sun_zenith=30.0, pitch=10.0)
        w.winds = scene.layer_winds(layers)
        w.overcast = scene.overcast_mask(layers)
        return w, "Cb NDF with anvil, height-blended droplets to the ice halo (cloud_march.wgsl:256-296)"
    if name == "three_layers_low_sun":
        layers = [scene.synth_layer(1, 800.0, 1200.0, 0.7), scene.synth_layer(3, 3000.0, 4000.0, 0.6),
                  scene.synth_layer(6, 9000.0, 10000.0, 0.5)]
        w = _base(layers=layers, sun_zenith=85.0, pitch=5.0, yaw=160.0)
        w.winds = scene.layer_winds(layers)
        w.clouds.cloud_layer_count = 3
        w.overcast = scene.overcast_mask(layers)
        return w, "St + Ac + Cs stack, sun 5 deg above the horizon"
    if name == "lightning":
        w = _base()
        w.frame.lightning_illuminance[:] = [4.0e4, 4.0e4, 5.0e4, 6000.0]
        w.frame.lightning_origin_world[:] = [500.0, 1800.0, 5000.0, 0.0]
        return w, "lightning emission term (cloud_march.wgsl:968-977)"
    if name == "temporal_jitter":
        w = _base()
        w.clouds.temporal_jitter = 1
        w.frame.frame_index = 5
        return w, "blue-noise XOR by frame index (cloud_march.wgsl:745-756)"
    if name == "drift_skew_diurnal":
        layers = [scene.synth_layer(0, 1500.0, 2300.0, 0.5, density_scale=1.0, shape_bias=0.3, detail_bias=-0.2)]
        w = _base(layers=layers)
        w.clouds.wind_drift_strength = 1.0
        w.clouds.wind_skew_strength = 0.8
        w.clouds.diurnal_strength = 0.7
        w.frame.simulated_seconds = 900.0
        return w, "wind drift offset, skew and diurnal shape/detail bias (cloud_density_common.wgsl:154-219)"
    if name == "depth_clip_mid_layer":
        w = _base()
        # opaque surfaces 2.5-7 km away over the middle rows (the march
        # interval ends inside the layer, cloud_march.wgsl:758-775); the
        # ground elsewhere, as in the plain workloads
        t = scene.ground_distance(w.frame)
        band = slice(H // 4, 3 * H // 4)
        t[band, :] = np.minimum(t[band, :], 2500.0 + 4500.0 * (np.arange(W) / (W - 1))[None, :])
        w.depth = scene.distance_depth(w.frame, t)
        return w, "scene depth clipping the march mid-layer (cloud_march.wgsl:758-806)"
    if name == "sun_disk_in_view":
        # sun 20 deg above the horizon, 5 deg above the view centre; a 1.5 deg
        # disk radius so the disk and its limb cover several pixels
        w = _base(sun_zenith=70.0, pitch=15.0, yaw=180.0)
        w.frame.sun_direction[3] = np.radians(1.5)
        return w, "sun disk with limb darkening in the sky pass (sky.wgsl:186-215)"
    if name == "phase_and_scattering_knobs":
        layers = [scene.synth_layer(0, 1500.0, 2300.0, 0.5, density_scale=1.0, droplet_diameter_um=8.0)]
        w = _base(layers=layers)
        w.clouds.multi_scatter_octaves = 2
        w.clouds.powder_strength = 0.0
        w.clouds.forward_bias = 0.6
        w.clouds.ambient_strength = 0.4
        w.clouds.droplet_diameter_bias = 1.3
        return w, "chromatic Mie (d < 20 um), 2 octaves, no powder, forward bias"
    raise KeyError(name)


VARIANTS = ["cone_light_sampling", "cloud_type_grid", "cumulonimbus_ice", "three_layers_low_sun", "lightning",
            "temporal_jitter", "drift_skew_diurnal", "depth_clip_mid_layer", "sun_disk_in_view",
            "phase_and_scattering_knobs"]


def _render_both(r: Renderer, w):
    r.load_workload(w)
    out = r.render_numpy(which=("composite", "cloud_transmittance"))
    fmt = w.noise_format
    base = oracle.decode(r.read_noise(abi.PS_NOISE_BASE, fmt), fmt)
    det = oracle.decode(r.read_noise(abi.PS_NOISE_DETAIL, fmt), fmt)
    curl = oracle.decode(r.read_noise(abi.PS_NOISE_CURL, fmt), fmt)
    fr = oracle.Frame(w, base, det, curl)
    comp, _, trn, _ = fr.render_region(0, 0, w.width, w.height)
    return out, comp, trn


@pytest.mark.parametrize("name", VARIANTS)
def test_feature_matches_oracle(renderer, name):
    w, what = _variant(name)
    _log(f"case: {name}: {what}")
    out, comp, trn = _render_both(renderer, w)
    e = oracle.relative_error(out["composite"], comp)
    p999, med = float(np.percentile(e, 99.9)), float(np.median(e))
    td = np.abs(out["cloud_transmittance"][..., :3] - trn[..., :3]).max(-1)
    tp = float(np.percentile(td, 99.9))
    _log(f"  composite p99.9 {p999:.3e} median {med:.3e} max {e.max():.3e}; "
         f"transmittance abs p99.9 {tp:.3e} max {td.max():.3e}")
    assert p999 <= 1e-3
    assert med <= 1e-5
    assert tp <= 1e-3
    # the feature is live: the oracle image differs from the plain small workload's
    ref_w = _base()
    rb = oracle.Frame(ref_w, *(oracle.decode(renderer.read_noise(k, w.noise_format), w.noise_format)
                               for k in (abi.PS_NOISE_BASE, abi.PS_NOISE_DETAIL, abi.PS_NOISE_CURL)))
    ref_comp, _, _, _ = rb.render_region(0, 0, ref_w.width, ref_w.height)
    changed = float(np.abs(comp[..., :3] - ref_comp[..., :3]).max() / max(np.abs(ref_comp[..., :3]).max(), 1e-30))
    _log(f"  change vs plain workload: {changed:.3e}")
    assert changed > 1e-4
