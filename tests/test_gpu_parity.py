"""GPU parity of the gfx950 path against the FP64 oracle, through the C ABI.

Bars (relative radiance error = max over RGB of |gpu - oracle| /
max(|oracle|, 1e-4 * max frame radiance)):
  * LUTs: every texel within 2e-5 (FP64 compute, FP32 storage);
  * rendered frames: relative radiance error 99.9th percentile <= 1e-3, median
    <= 1e-5;
  * noise bakes: integer hash exact; stored texels within one storage step.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle
from pedalsky_amd import abi, scene
from pedalsky_amd.renderer import Renderer

pytestmark = pytest.mark.gpu


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


def _oracle_frame(r: Renderer, w):
    fmt = w.noise_format
    base = oracle.decode(r.read_noise(abi.PS_NOISE_BASE, fmt), fmt)
    det = oracle.decode(r.read_noise(abi.PS_NOISE_DETAIL, fmt), fmt)
    curl = oracle.decode(r.read_noise(abi.PS_NOISE_CURL, fmt), fmt)
    return oracle.Frame(w, base, det, curl)


def _rel_stats(got, ref):
    e = oracle.relative_error(got, ref)
    return float(np.percentile(e, 99.9)), float(np.median(e)), float(e.max())


def test_luts_match_oracle(renderer):
    w = scene.workload_small()
    _log("case: LUT chain, small workload")
    renderer.load_workload(w)
    renderer.bake_luts()
    ref = oracle.bake_luts(w.frame, w.atmosphere, w.lut_dims)
    for which, key in ((abi.PS_LUT_TRANSMITTANCE, "T"), (abi.PS_LUT_MULTISCATTER, "MS"),
                       (abi.PS_LUT_SKYVIEW, "SV"), (abi.PS_LUT_AERIAL_PERSPECTIVE, "AP")):
        got = renderer.read_lut(which).astype(np.float64)
        exp = ref[key]
        assert got.shape == exp.shape, key
        floor = 1e-6 * np.abs(exp[..., :3]).max()
        err = np.abs(got[..., :3] - exp[..., :3]) / np.maximum(np.abs(exp[..., :3]), floor)
        _log(f"  {key}: max rel err {err.max():.3e}")
        assert err.max() < 2e-5, key
        if key == "AP":
            assert np.abs(got[..., 3] - exp[..., 3]).max() < 1e-6


def test_multiscatter_disabled_clears_lut(renderer):
    w = scene.workload_small()
    renderer.load_workload(w)
    renderer.bake_luts(abi.PS_LUT_ALL | abi.PS_LUT_MULTISCATTER_DISABLED)
    ms = renderer.read_lut(abi.PS_LUT_MULTISCATTER)
    assert np.all(ms == 0.0)
    ref = oracle.bake_luts(w.frame, w.atmosphere, w.lut_dims, ms_disabled=True)
    sv = renderer.read_lut(abi.PS_LUT_SKYVIEW).astype(np.float64)
    assert np.abs(sv[..., :3] - ref["SV"][..., :3]).max() / np.abs(ref["SV"][..., :3]).max() < 2e-5


@pytest.mark.parametrize("fmt", [abi.PS_NOISE_UNORM8, abi.PS_NOISE_FP16])
def test_noise_bake_matches_oracle(renderer, fmt):
    _log(f"case: noise bake format {fmt}")
    renderer.generate_noise(0, fmt)
    for which, ref_fn in ((abi.PS_NOISE_DETAIL, lambda: oracle.noise_detail(0)),
                          (abi.PS_NOISE_CURL, lambda: oracle.noise_curl(0)),
                          (abi.PS_NOISE_BASE, lambda: oracle.noise_base(0, 0, 16))):
        got = renderer.read_noise(which, fmt)
        exp = oracle.quantize(ref_fn(), fmt)
        if which == abi.PS_NOISE_BASE:
            got = got[:16]
        if fmt == abi.PS_NOISE_UNORM8:
            d = np.abs(got.astype(np.int32) - exp.astype(np.int32))
            frac = float((d > 0).mean())
            _log(f"  volume {which}: max step diff {d.max()}, differing texels {frac:.2e}")
            assert d.max() <= 1
        else:
            # the GPU bake runs the WGSL's FP32 arithmetic: allow one binary16
            # step at the value, or 1e-5 absolute where the value is near 0
            # (the curl map's central difference scales FP32 noise by 64)
            g = got.view(np.float16).astype(np.float64)
            e = exp.view(np.float16).astype(np.float64)
            step = np.maximum(np.spacing(np.abs(e).astype(np.float16)).astype(np.float64), 1e-5)
            d = np.abs(g - e) / step
            frac = float((d > 0).mean())
            _log(f"  volume {which}: max diff {d.max():.3f} steps, differing texels {frac:.2e}")
            assert d.max() <= 1.0 + 1e-9
        assert frac < 2e-2


@pytest.mark.parametrize("ground", [True, False], ids=["ground", "all-far"])
def test_render_small_matches_oracle(renderer, ground):
    w = scene.workload_small(ground=ground)
    _log(f"case: render 320x180 one cumulus layer, {'planet ground depth' if ground else 'no depth buffer'}")
    renderer.load_workload(w)
    out = renderer.render_numpy()
    fr = _oracle_frame(renderer, w)
    comp, lum, trn, _ = fr.render_region(0, 0, w.width, w.height)
    p999, med, mx = _rel_stats(out["composite"], comp)
    _log(f"  composite: p99.9 {p999:.3e} median {med:.3e} max {mx:.3e}")
    assert p999 <= 1e-3
    assert med <= 1e-5
    tp = float(np.percentile(np.abs(out["cloud_transmittance"][..., :3] - trn[..., :3]).max(-1), 99.9))
    _log(f"  transmittance abs p99.9 {tp:.3e}")
    assert tp <= 1e-3
    # clouds are actually present in this frame
    assert (trn[..., 1] < 0.9).mean() > 0.02


# Sweep frames of the full-frame workload: with the planet-ground depth
# buffer, and with none (depth all far: rays below the horizon pass through
# the planet and march the layer shells on its far side, ~1.3e4 km away).
FULL_FRAME_CASES = [(0, True), (7, True), (12, True), (0, False), (5, False), (10, False), (15, False)]


@pytest.mark.parametrize("index, ground", FULL_FRAME_CASES,
                         ids=[f"frame{i}-{'ground' if g else 'all-far'}" for i, g in FULL_FRAME_CASES])
def test_render_full_frame_matches_oracle(renderer, index, ground):
    w = scene.workload_full_frame(index, ground=ground)
    _log(f"case: render 1280x720 cumulus + cirrus, sweep frame {index}, "
         f"{'planet ground depth' if ground else 'no depth buffer'}")
    renderer.load_workload(w)
    out = renderer.render_numpy(which=("composite", "cloud_transmittance"))
    fr = _oracle_frame(renderer, w)
    comp, lum, trn, _ = fr.render_region(0, 0, w.width, w.height)
    p999, med, mx = _rel_stats(out["composite"], comp)
    tp = float(np.percentile(np.abs(out["cloud_transmittance"][..., :3] - trn[..., :3]).max(-1), 99.9))
    _log(f"  composite: p99.9 {p999:.3e} median {med:.3e} max {mx:.3e}; transmittance abs p99.9 {tp:.3e}; "
         f"cloudy {(trn[..., 1] < 0.9).mean():.3f}")
    assert p999 <= 1e-3
    assert med <= 1e-5
    assert tp <= 1e-3


def test_f16_output_rounds_f32(renderer):
    w = scene.workload_small()
    renderer.load_workload(w)
    a = renderer.render_numpy(which=("composite",), fmt=abi.PS_PIXEL_F32)["composite"]
    b = renderer.render_numpy(which=("composite",), fmt=abi.PS_PIXEL_F16, bake=False)["composite"]
    assert np.array_equal(a.astype(np.float16).astype(np.float32), b)


def test_work_counters_match_oracle(renderer):
    w = scene.workload_small()
    renderer.load_workload(w)
    renderer.bake_luts()
    c = renderer.count_work()
    fr = _oracle_frame(renderer, w)
    _, _, _, cnt = fr.render_region(0, 0, w.width, w.height, counters=True)
    names = [f[0] for f in abi.WorkCounters._fields_]
    for i, n in enumerate(names):
        g, o = getattr(c, n), int(cnt[i])
        _log(f"  {n}: gpu {g} oracle {o}")
        assert abs(g - o) <= max(2, 2e-3 * o), n


def test_error_paths(renderer):
    lib = abi.load()
    h = C.c_void_p()
    abi.check(lib.ps_ctx_create(0, C.byref(h)))
    try:
        out = abi.ImageOut()
        assert lib.ps_render(h, None, C.byref(out)) == abi.PS_ERR_NOT_READY
        assert b"ps_set_atmosphere" in lib.ps_last_error()
        bad = (abi.CloudLayer * 2)(scene.synth_layer(0, 1000, 2000, 0.5), scene.synth_layer(1, 1500, 2500, 0.5))
        p = scene.cloud_params(n_layers=2)
        assert lib.ps_set_clouds(h, C.byref(p), bad, 2, None) == abi.PS_ERR_OVERLAPPING_LAYERS
        assert lib.ps_set_clouds(h, C.byref(p), bad, 9, None) == abi.PS_ERR_INVALID_ARGUMENT
        p.multi_scatter_octaves = 9
        ok = (abi.CloudLayer * 1)(scene.synth_layer(0, 1000, 2000, 0.5))
        assert lib.ps_set_clouds(h, C.byref(p), ok, 1, None) == abi.PS_ERR_UNSUPPORTED
        assert lib.ps_ctx_create(99, C.byref(C.c_void_p())) == abi.PS_ERR_INVALID_ARGUMENT
    finally:
        lib.ps_ctx_destroy(h)
