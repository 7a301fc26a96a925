"""The tile kernels dispatch their one-wave tiles in an XCD-grouped order
(runs of 16 tiles per XCD, tiles past the last full group of 128 in place).
Every pixel must still be written exactly once: outputs are pre-filled with
NaN and rendered at grid sizes just below, at and above a full group, and in
shapes whose tile rows do not divide the runs; the training forward (its
checkpoints are indexed by tile) is checked the same way, and the plain
render against the oracle on one of them.
"""
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


# (W, H): 127, 128, 129 and 257 tiles of 8x8 in one row; 16 x 8 = 128 and
# 17 x 9 = 153 tiles in two dimensions; a ragged 100 x 37 frame
SIZES = [(1016, 8), (1024, 8), (1032, 8), (2056, 8), (128, 64), (136, 72), (100, 37)]


def _render(r, train=False):
    import torch
    W, H = r.workload.width, r.workload.height
    bufs = {k: torch.full((H, W, 4), float("nan"), dtype=torch.float32, device="cuda:0")
            for k in ("composite", "cloud_luminance", "cloud_transmittance")}
    torch.cuda.synchronize()
    out = r.image_out(bufs, abi.PS_PIXEL_F32)
    r.bake_luts()
    (r.render_train if train else r.render)(out)
    r.synchronize()
    return {k: v.cpu().numpy() for k, v in bufs.items()}


@pytest.mark.parametrize("size", SIZES)
def test_every_pixel_is_written(renderer, size):
    W, H = size
    _log(f"case: tile coverage {W}x{H} ({((W + 7) // 8) * ((H + 7) // 8)} tiles)")
    w = scene.workload_small(width=W, height=H, steps=16)
    renderer.load_workload(w)
    a = _render(renderer)
    b = _render(renderer, train=True)
    for k in a:
        assert np.isfinite(a[k]).all(), f"render left pixels of {k} unwritten"
        assert np.array_equal(a[k], b[k]), f"training forward differs from the render in {k}"


def test_grouped_order_matches_oracle(renderer):
    W, H = 136, 72
    w = scene.workload_small(width=W, height=H, steps=32)
    renderer.load_workload(w)
    out = _render(renderer)
    fmt = w.noise_format
    fr = oracle.Frame(w, *(oracle.decode(renderer.read_noise(k, fmt), fmt)
                           for k in (abi.PS_NOISE_BASE, abi.PS_NOISE_DETAIL, abi.PS_NOISE_CURL)))
    comp, _, _, _ = fr.render_region(0, 0, W, H)
    e = oracle.relative_error(out["composite"], comp)
    _log(f"  composite p99.9 {np.percentile(e, 99.9):.3e} max {e.max():.3e}")
    assert np.percentile(e, 99.9) <= 1e-3
