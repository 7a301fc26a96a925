"""CPU tests of the C ABI boundary: struct layouts, exported symbols, error
behaviour without a device, and the host-side synthesis helpers against the
numbers the reference's own unit tests pin."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

from pedalsky_amd import abi, scene

HEADER = Path(__file__).resolve().parent.parent / "include" / "pedalsky_hip.h"


def test_struct_layouts_match_reference():
    # frame_uniforms.rs:195-201 (448), weather.rs:544-555 (144, 48), params.rs:221-231 (112)
    abi.check_layouts()
    assert C.sizeof(abi.FrameUniforms) == 448
    assert C.sizeof(abi.AtmosphereParams) == 144
    assert C.sizeof(abi.CloudParams) == 112
    assert C.sizeof(abi.CloudLayer) == 48
    for cls in (abi.FrameUniforms, abi.AtmosphereParams, abi.CloudParams, abi.CloudLayer):
        assert C.sizeof(cls) % 16 == 0
    # batch descriptors (static_assert'ed in batch.cpp)
    assert C.sizeof(abi.SceneDesc) == 96 and abi.SceneDesc.out.offset == 64
    assert C.sizeof(abi.Tuning) == 20
    # batch step schedule: count, four arrays of 8 int32, the measured time
    assert C.sizeof(abi.StepSchedule) == 136 and abi.StepSchedule.batch_ms.offset == 132
    # differentiation outputs: two device pointers
    assert C.sizeof(abi.CloudGrads) == 16 and abi.CloudGrads.density_delta.offset == 8


def test_library_exports_every_header_symbol(lib):
    text = HEADER.read_text()
    declared = set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(ps_\w+)\s*\(", text, flags=re.M))
    assert declared, "no declarations parsed"
    assert declared == set(abi.EXPORTED), declared ^ set(abi.EXPORTED)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.ps_abi_version() == 1


def test_errors_without_device(lib, gpu_available):
    if gpu_available:
        pytest.skip("device present: covered by the GPU tests")
    h = C.c_void_p()
    st = lib.ps_ctx_create(0, C.byref(h))
    assert st != abi.PS_OK
    assert lib.ps_last_error()  # message present
    assert lib.ps_ctx_create(0, None) == abi.PS_ERR_INVALID_ARGUMENT
    # the step knobs refuse a missing context or batch before touching a device
    assert lib.ps_set_march_steps(None, 64, 6) == abi.PS_ERR_INVALID_ARGUMENT
    t = abi.StepSchedule()
    assert lib.ps_batch_get_step_schedule(None, C.byref(t)) == abi.PS_ERR_INVALID_ARGUMENT
    assert lib.ps_batch_set_step_schedule(None, C.byref(t)) == abi.PS_ERR_INVALID_ARGUMENT
    assert lib.ps_batch_autotune_steps(None, C.c_float(10.0), 1, C.byref(t)) == abi.PS_ERR_INVALID_ARGUMENT


def test_default_params_match_reference(lib):
    a = abi.AtmosphereParams()
    lib.ps_default_atmosphere(a)
    # weather.rs:60-84
    assert a.planet_radius_m == 6_360_000.0 and a.atmosphere_top_m == 6_460_000.0
    assert a.rayleigh_scale_height_m == 8000.0 and a.mie_scale_height_m == 1200.0
    assert np.allclose(a.rayleigh_scattering[:3], [5.802e-6, 13.558e-6, 33.1e-6])
    assert np.allclose(a.mie_absorption[:3], [4.4e-6] * 3)
    assert abs(a.mie_g - 0.8) < 1e-7
    assert a.ozone_center_m == 25000.0 and a.ozone_thickness_m == 30000.0
    p = abi.CloudParams()
    lib.ps_default_cloud_params(p)
    # params.rs:142-213
    assert np.allclose(p.sigma_s[:], [0.03, 0.04, 0.06])
    assert p.cloud_steps == 192 and p.light_steps == 6 and p.multi_scatter_octaves == 4
    assert abs(p.detail_strength - 0.05) < 1e-7 and abs(p.curl_strength - 0.1) < 1e-7
    assert p.base_scale_m == 4500.0 and p.detail_scale_m == 800.0 and p.weather_scale_m == 32000.0
    assert p.wind_drift_strength == 1.0 and p.wind_skew_strength == 0.5 and p.diurnal_strength == 1.0


def test_coverage_remap_reference_pins(lib):
    # ps-synthesis/src/cloud_layers.rs:278-327
    f = lib.ps_remap_coverage
    assert abs(f(0.0)) < 1e-6
    assert abs(f(1.0) - 0.60) < 1e-6
    assert abs(f(0.10) - 0.40) < 1e-5
    assert abs(f(0.50) - (0.40 + (0.50 - 0.10) / 0.9 * 0.2)) < 1e-5
    prev = 0.0
    for i in range(101):
        v = f(i / 100.0)
        assert v >= prev - 1e-6
        assert v - prev <= 0.06
        prev = v


@pytest.mark.parametrize("t,ds,d", [(0, 1.0, 20.0), (1, 1.0, 16.0), (2, 1.0, 16.0), (3, 0.85, 14.0),
                                    (4, 0.7, 30.0), (5, 0.55, 50.0), (6, 0.4, 50.0), (7, 1.4, 20.0)])
def test_per_type_defaults(lib, t, ds, d):
    # cloud_layers.rs:190-252
    L = scene.synth_layer(t, 100.0, 200.0, 0.5)
    # exact f32 equality, as the Rust test's assert_eq on f32
    assert L.density_scale == np.float32(ds)
    assert L.droplet_diameter_um == np.float32(d)
    assert L.anvil_bias == (1.0 if t == 7 else 0.0)
    assert L.cloud_type == t


def test_explicit_overrides_and_overlap(lib):
    L = scene.synth_layer(5, 8000.0, 9000.0, 0.5, density_scale=2.5, droplet_diameter_um=35.0)
    assert L.density_scale == 2.5 and L.droplet_diameter_um == 35.0
    ok = (abi.CloudLayer * 2)(scene.synth_layer(1, 200, 600, 0.5), scene.synth_layer(0, 1500, 2300, 0.5))
    assert lib.ps_check_layers_overlap(ok, 2, None, None) == 0
    bad = (abi.CloudLayer * 2)(scene.synth_layer(0, 1000, 2000, 0.5), scene.synth_layer(1, 1500, 2500, 0.5))
    i, j = C.c_int32(), C.c_int32()
    assert lib.ps_check_layers_overlap(bad, 2, C.byref(i), C.byref(j)) == 1
    assert (i.value, j.value) == (0, 1)


def test_blue_noise_histogram_flat():
    # crates/ps-clouds/src/noise.rs:710-726
    px = scene.blue_noise()
    assert px.shape == (4096,)
    hist = np.bincount(px >> 4, minlength=16)
    avg = px.size / 16.0
    assert np.all(np.abs(hist - avg) / avg < 0.5)
    # every rank 0..4095 maps to a distinct quantised bucket spread: full range used
    assert px.min() == 0 and px.max() == 255


def test_layer_wind_profile(lib):
    # wind_field.rs:167-178: 1/7 power law, veer to 30 deg at 12 km; the tap
    # at a layer's mid altitude interpolates two of the 16 levels.
    out = (C.c_float * 2)()
    lib.ps_layer_wind(5.0, 240.0, 1500.0, 2300.0, out)
    u, v = out[0], out[1]
    # level centres y=2 (1875 m) and y=3 (2625 m); mid 1900 m -> weight 0.0333
    def level(alt):
        s = 5.0 * (max(alt, 10.0) / 10.0) ** 0.143
        veer = min(max(alt / 12000.0, 0.0), 1.0) ** 0.7
        a = math.radians(240.0 + 180.0) + veer * math.radians(30.0)
        return s * math.sin(a), -s * math.cos(a)
    u2, v2 = level(1875.0)
    u3, v3 = level(2625.0)
    fy = 1900.0 / 12000.0 * 16 - 0.5 - 2
    assert u == pytest.approx(u2 + (u3 - u2) * fy, rel=2e-3)
    assert v == pytest.approx(v2 + (v3 - v2) * fy, rel=2e-3)
    speed = math.hypot(u, v)
    assert 10.0 < speed < 11.0


def test_overcast_mask_value(lib):
    # density_mask.rs:134-165 with the CPU NDF (ndf.rs:20-58)
    layers = [scene.synth_layer(0, 1500.0, 2300.0, 0.5, density_scale=1.0)]
    v = scene.overcast_mask(layers)[0]
    n = 8
    hs = (np.arange(n) + 0.5) / n

    def ss(e0, e1, x):
        t = np.clip((x - e0) / (e1 - e0), 0, 1)
        return t * t * (3 - 2 * t)
    mean = np.mean(ss(0, 0.07, hs) * ss(1, 0.2, hs))
    col = min(max(mean * 0.8 * layers[0].coverage * 1.0 / 10.0, 0), 1)
    assert v == int(col * 255.0)


def test_workload_synthesis():
    w = scene.workload_small()
    assert (w.width, w.height) == (320, 180)
    assert len(w.layers) == 1 and w.layers[0].coverage == pytest.approx(0.48889, abs=1e-5)
    cov = w.weather.view(np.float16)[..., 0].astype(np.float64)
    assert cov.shape == (512, 512)
    # rank-uniform coverage: mean 0.5, spans [0, 1]
    assert abs(cov.mean() - 0.5) < 1e-3 and cov.min() < 1e-3 and cov.max() > 0.999
    # smoothed: neighbouring texels close
    assert np.abs(np.diff(cov, axis=1)).mean() < 0.02
    assert w.depth.shape == (180, 320) and w.depth.dtype == np.float32
    f = scene.workload_full_frame(0)
    assert (f.width, f.height) == (1280, 720) and len(f.layers) == 2
    p, y = scene.full_frame_camera(3)
    assert -10.0 <= p <= 30.0 and 0.0 <= y < 360.0


def test_camera_matrices_inverse():
    # frame_uniforms.rs:203-224: inv_view_proj * view_proj ~ I within 1e-3
    f = scene.make_frame(1280, 720, pitch_deg=12.0, yaw_deg=33.0)
    vp = np.array(f.view_proj[:], dtype=np.float64).reshape(4, 4)
    inv = np.array(f.inv_view_proj[:], dtype=np.float64).reshape(4, 4)
    prod = (inv.T @ vp.T).T
    assert np.allclose(prod, np.eye(4), atol=1e-3)
    # reverse-Z: a point 10 m ahead maps to depth near/10
    fwd = scene.fly_camera_forward(math.radians(12.0), math.radians(33.0))
    p = np.array(f.camera_position_world[:3]) + 10.0 * fwd
    clip = np.append(p, 1.0) @ vp
    assert clip[2] / clip[3] == pytest.approx(0.1 / 10.0, rel=1e-4)


def test_pcg32_reference_stream():
    # PCG32 demo output for seed 42, stream 54 (pcg-c reference pcg32-demo)
    r = scene.Pcg32(42, 54)
    got = [r.next_u32() for _ in range(6)]
    assert got == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
