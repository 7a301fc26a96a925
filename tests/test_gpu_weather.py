"""GPU weather-map and cloud-type-grid synthesis (ps_synthesise_weather)
against the oracle restatement of crates/ps-synthesis/src/weather_map.rs,
coverage_grid.rs and cloud_type_grid.rs: every texel bit-identical; and the
synthesised inputs drive the march exactly as the same texels uploaded
through ps_upload_weather."""
import sys

import numpy as np
import pytest

import oracle.weather as W
from pedalsky_amd import abi, scene
from pedalsky_amd.renderer import Renderer

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def r():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU (no CPU fallback exists)"
    rr = Renderer(0)
    yield rr
    rr.close()


def _cases():
    rng = np.random.default_rng(0x3E)
    g1 = rng.random((150, 200), dtype=np.float32)
    t1 = rng.integers(0, 10, (150, 200), dtype=np.uint8)
    g2 = np.linspace(0.0, 1.0, 64 * 64, dtype=np.float32).reshape(64, 64)
    return [
        ("layers, no grid, precip 10 mm/h, wetness 0.3", dict(has_layers=True, precip_mm_per_h=10.0,
                                                              ground_wetness=0.3)),
        ("no layers", dict(has_layers=False)),
        ("200x150 coverage + type grid, 40 km", dict(coverage_grid=g1, type_grid=t1, grid_extent_m=40000.0)),
        ("64x64 ramp grid, 20 km, precip 80 mm/h", dict(coverage_grid=g2, grid_extent_m=20000.0,
                                                       precip_mm_per_h=80.0)),
        ("type grid only (200x150)", dict(type_grid=t1)),
    ]


def _oracle(kw):
    g, t = kw.get("coverage_grid"), kw.get("type_grid")
    shape = g.shape if g is not None else (t.shape if t is not None else (0, 0))
    wm = W.synthesise_weather(kw.get("has_layers", True), kw.get("precip_mm_per_h", 0.0),
                              kw.get("ground_wetness", 0.0), None if g is None else g.reshape(-1),
                              shape[1], shape[0], kw.get("grid_extent_m", 32000.0))
    tg = W.synthesise_type_grid(None if t is None else t.reshape(-1), shape[1], shape[0])
    return wm, tg


@pytest.mark.parametrize("name,kw", _cases())
def test_synthesis_bit_identical(r, name, kw):
    _log(f"case: weather synthesis, {name}")
    r.synthesise_weather(**kw)
    rgba, types = r.read_weather()
    wm, tg = _oracle(kw)
    bad = int(np.count_nonzero(rgba != wm.view(np.uint16)))
    _log(f"  texels differing from the oracle: {bad}")
    assert bad == 0
    assert np.array_equal(types, tg)


def test_synthesised_inputs_drive_the_march(r):
    import torch
    w = scene.workload_small()
    r.load_workload(w)
    r.bake_luts()
    kw = _cases()[2][1]
    _log("case: render 320x180 with synthesised vs uploaded weather (200x150 grids)")
    r.synthesise_weather(**kw, overcast=w.overcast)
    bufs = r.alloc_outputs(w.width, w.height, abi.PS_PIXEL_F32, ("composite",))
    r.render(r.image_out(bufs))
    r.synchronize()
    a = bufs["composite"].cpu().numpy()
    wm, tg = _oracle(kw)
    r.upload_weather(wm.view(np.uint16), tg, w.overcast)
    r.render(r.image_out(bufs))
    r.synchronize()
    torch.cuda.synchronize()
    b = bufs["composite"].cpu().numpy()
    _log(f"  composites identical: {np.array_equal(a, b)}; cloudy fraction {float(np.mean(a[..., 3] > 0)):.3f}")
    assert np.array_equal(a, b)


def test_errors(r):
    import ctypes as C
    lib = r.lib
    bad = abi.WeatherSynth(1, 0.0, 0.0, None, (C.c_uint8 * 4)(0, 1, 2, 3), 0, 2, 32000.0)
    assert lib.ps_synthesise_weather(r.ctx, C.byref(bad), None) == abi.PS_ERR_INVALID_ARGUMENT
    g = (C.c_float * 4)(0.0, 0.5, 1.0, 0.5)
    bad = abi.WeatherSynth(1, 0.0, 0.0, g, None, 2, 2, 0.0)
    assert lib.ps_synthesise_weather(r.ctx, C.byref(bad), None) == abi.PS_ERR_INVALID_ARGUMENT
    assert lib.ps_synthesise_weather(r.ctx, None, None) == abi.PS_ERR_INVALID_ARGUMENT
