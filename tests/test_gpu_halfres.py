"""GPU parity of the half-resolution cloud path (the reference's
clouds.half_res_render: crates/ps-clouds/src/lib.rs:748-760,
crates/ps-clouds/src/pipeline.rs:277-319, cloud_composite_halfres.wgsl)
against the FP64 oracle, through the C ABI.

Bar (same as the plain-path tests): relative radiance error 99.9th
percentile <= 1e-3, median <= 1e-5; cloud transmittance absolute error
99.9th percentile <= 1e-3. An even and an odd frame size (the odd one has a
non-integer RT scale: floor(W / 2) RT columns over W pixels).
"""
import sys

import numpy as np
import pytest

import oracle
from pedalsky_amd import abi, scene
from pedalsky_amd.abi import PedalskyError
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


def _frame(r, w):
    fmt = w.noise_format
    return oracle.Frame(w, *(oracle.decode(r.read_noise(k, fmt), fmt)
                             for k in (abi.PS_NOISE_BASE, abi.PS_NOISE_DETAIL, abi.PS_NOISE_CURL)))


@pytest.mark.parametrize("size", [(256, 144), (257, 145)])
def test_halfres_matches_oracle(renderer, size):
    W, H = size
    w = scene.workload_small(width=W, height=H, steps=48)
    _log(f"case: half resolution {W}x{H} (RT {W // 2}x{H // 2})")
    renderer.load_workload(w)
    full = renderer.render_numpy(which=("composite",))
    renderer.set_half_res(True)
    try:
        out = renderer.render_numpy()
    finally:
        renderer.set_half_res(False)
    fr = _frame(renderer, w)
    comp, lum, trn = fr.render_halfres(0, 0, W, H)
    e = oracle.relative_error(out["composite"], comp)
    p999, med = float(np.percentile(e, 99.9)), float(np.median(e))
    td = np.abs(out["cloud_transmittance"][..., :3] - trn[..., :3]).max(-1)
    tp = float(np.percentile(td, 99.9))
    el = oracle.relative_error(out["cloud_luminance"], lum)
    _log(f"  composite p99.9 {p999:.3e} median {med:.3e} max {e.max():.3e}; "
         f"luminance p99.9 {np.percentile(el, 99.9):.3e}; transmittance abs p99.9 {tp:.3e} max {td.max():.3e}")
    assert p999 <= 1e-3
    assert med <= 1e-5
    assert tp <= 1e-3
    assert (trn[..., 1] < 0.9).mean() > 0.02, "the frame should hold clouds"
    # the path is live: the upsampled image is not the full-resolution one
    diff = float(np.abs(out["composite"][..., :3] - full["composite"][..., :3]).max()
                 / np.abs(full["composite"][..., :3]).max())
    _log(f"  change vs full-resolution render: {diff:.3e}")
    assert diff > 1e-3
    # switching back restores the full-resolution path bit for bit
    again = renderer.render_numpy(which=("composite",))
    assert np.array_equal(again["composite"], full["composite"])


def test_halfres_f16_output(renderer):
    W, H = 160, 90
    w = scene.workload_small(width=W, height=H, steps=32)
    renderer.load_workload(w)
    renderer.set_half_res(True)
    try:
        o32 = renderer.render_numpy(which=("composite",))
        o16 = renderer.render_numpy(which=("composite",), fmt=abi.PS_PIXEL_F16)
    finally:
        renderer.set_half_res(False)
    ref = o32["composite"].astype(np.float16).astype(np.float32)
    assert np.array_equal(o16["composite"], ref)


def test_halfres_rejects_full_resolution_only_calls(renderer):
    w = scene.workload_small(width=64, height=32, steps=16)
    renderer.load_workload(w)
    renderer.set_half_res(True)
    try:
        bufs = renderer.alloc_outputs(64, 32)
        out = renderer.image_out(bufs)
        renderer.bake_luts()
        with pytest.raises(PedalskyError) as ei:
            renderer.render_train(out)
        assert ei.value.status == abi.PS_ERR_UNSUPPORTED
        with pytest.raises(PedalskyError) as ei:
            renderer.count_work()
        assert ei.value.status == abi.PS_ERR_UNSUPPORTED
        from pedalsky_amd.batch import Batch
        with pytest.raises(PedalskyError) as ei:
            Batch(renderer, [w])
        assert ei.value.status == abi.PS_ERR_UNSUPPORTED
        renderer.render(out)     # the half-resolution render itself still works
        renderer.synchronize()
    finally:
        renderer.set_half_res(False)
