"""GPU test of ps_render_noalloc (the render entry point that allocates
nothing) through the C ABI: on two calls in a row, the second with a smaller
image, its output must equal ps_render's bit for bit (the same kernel reads
the same parameter block, staged in the per-GPU scratch array instead of the
context's ring)."""
import ctypes as C
import sys

import numpy as np
import pytest

from pedalsky_amd import abi, scene
from pedalsky_amd.renderer import Renderer

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _render_both(r, w):
    """The workload's frame through ps_render and ps_render_noalloc."""
    import torch
    r.set_frame(w.frame)
    r.set_depth(w.depth)
    r.bake_luts()
    outs = []
    for fn in (r.lib.ps_render, r.lib.ps_render_noalloc):
        bufs = r.alloc_outputs(w.width, w.height, abi.PS_PIXEL_F32, ("composite", "cloud_luminance",
                                                                        "cloud_transmittance"))
        torch.cuda.synchronize()
        dp = None if r.depth_ptr is None else C.c_void_p(r.depth_ptr)
        abi.check(fn(r.ctx, dp, C.byref(r.image_out(bufs))))
        r.synchronize()
        outs.append({k: v.cpu().numpy() for k, v in bufs.items()})
    return outs


def test_render_noalloc_matches_render():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU (no CPU fallback exists)"
    with Renderer(0) as r:
        w = scene.workload_full_frame(3, width=640, height=360)
        r.load_workload(w)
        for size, (W, H) in enumerate(((640, 360), (200, 120))):
            pitch, yaw = scene.full_frame_camera(3 + size)
            w.frame = scene.make_frame(W, H, camera=(0.0, 500.0, 0.0), pitch_deg=pitch, yaw_deg=yaw,
                                       sun_zenith_deg=65.0)
            w.depth = scene.ground_depth(w.frame) if size == 0 else None
            _log(f"case: ps_render_noalloc call {size + 1}: {W}x{H}, pitch {pitch:.1f} yaw {yaw:.1f}, "
                 f"{'ground depth' if w.depth is not None else 'no depth buffer'}")
            a, b = _render_both(r, w)
            for k in a:
                same = np.array_equal(a[k], b[k])
                _log(f"  {k}: bit-identical {same}")
                assert same, k
            assert (a["cloud_transmittance"][..., 1] < 0.99).mean() > 0.01


def test_render_noalloc_rejects_what_needs_allocation():
    import torch
    assert torch.cuda.is_available()
    with Renderer(0) as r:
        w = scene.workload_small()
        r.load_workload(w)
        bufs = r.alloc_outputs(w.width, w.height, abi.PS_PIXEL_F32, ("composite",))
        out = r.image_out(bufs)
        # LUTs not baked yet
        assert r.lib.ps_render_noalloc(r.ctx, None, C.byref(out)) == abi.PS_ERR_NOT_READY
        r.bake_luts()
        r.set_half_res(True)
        assert r.lib.ps_render_noalloc(r.ctx, None, C.byref(out)) == abi.PS_ERR_UNSUPPORTED
        r.set_half_res(False)
        assert r.lib.ps_render_noalloc(r.ctx, None, None) == abi.PS_ERR_INVALID_ARGUMENT
        abi.check(r.lib.ps_render_noalloc(r.ctx, None, C.byref(out)))
        r.synchronize()
