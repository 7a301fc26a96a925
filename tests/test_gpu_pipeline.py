"""GPU tests of ps_frame's frame pipelining: when a frame rebakes every LUT,
ps_frame bakes it into a second LUT set on a side stream while the previous
frame still renders from the current set, then the sets swap. The output
must be exactly what baking into the one set and rendering does, frame after
frame, also when ps_frame calls mix with partial bakes, plain renders and
LUT reads, and the in-loop timing must still split LUT and render time.
"""
import sys

import numpy as np
import pytest

from pedalsky_amd import abi, scene
from pedalsky_amd.renderer import Renderer

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _frames(w, n):
    rng = np.random.default_rng(7)
    return [scene.make_frame(w.width, w.height, camera=(0.0, 500.0, 0.0), pitch_deg=float(rng.uniform(-10, 30)),
                             yaw_deg=float(rng.uniform(0, 360)), sun_zenith_deg=float(rng.uniform(20, 80)))
            for _ in range(n)]


@pytest.fixture(scope="module")
def rend():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU (no CPU fallback exists)"
    w = scene.workload_full_frame(0, width=640, height=360, ground=False)
    r = Renderer(0)
    r.load_workload(w)
    yield r, w
    r.close()


def _run(r, w, frames, plan):
    """plan per frame: 'frame' (ps_frame, every LUT), 'partial' (ps_frame with
    the sky-view and AP LUTs only), 'bake+render' (ps_bake_luts + ps_render)."""
    import torch
    outs = []
    for f, how in zip(frames, plan):
        r.set_frame(f)
        bufs = r.alloc_outputs(w.width, w.height, abi.PS_PIXEL_F32, ("composite", "cloud_transmittance"))
        out = r.image_out(bufs)
        if how == "frame":
            r.frame(out)
        elif how == "partial":
            r.frame(out, abi.PS_LUT_SKYVIEW | abi.PS_LUT_AERIAL_PERSPECTIVE)
        else:
            r.bake_luts()
            r.render(out)
        outs.append(bufs)
    r.synchronize()
    torch.cuda.synchronize()
    return [{k: v.cpu().numpy() for k, v in b.items()} for b in outs]


def test_pipelined_frames_match_single_set(rend):
    r, w = rend
    frames = _frames(w, 8)
    _log("case: 8 frames at 640x360 (all far), pipelined ps_frame vs bake + render into one set")
    ref = _run(r, w, frames, ["bake+render"] * 8)
    got = _run(r, w, frames, ["frame"] * 8)
    same = [all(np.array_equal(g[k], e[k]) for k in g) for g, e in zip(got, ref)]
    _log(f"  bit-identical frames: {sum(same)} / {len(same)}")
    assert all(same)


def test_pipelined_frames_mixed_with_other_calls(rend):
    r, w = rend
    frames = _frames(w, 8)
    plan = ["frame", "frame", "partial", "bake+render", "frame", "partial", "frame", "frame"]
    _log(f"case: mixed sequence {plan}")
    # the reference sequence bakes the same LUTs into the one set
    ref_plan = ["bake+render" if p == "frame" else p for p in plan]
    ref = _run(r, w, frames, ref_plan)
    got = _run(r, w, frames, plan)
    same = [all(np.array_equal(g[k], e[k]) for k in g) for g, e in zip(got, ref)]
    _log(f"  bit-identical frames: {sum(same)} / {len(same)}")
    assert all(same)
    # the LUTs read back are the last frame's
    lut_after = {k: r.read_lut(k) for k in (abi.PS_LUT_TRANSMITTANCE, abi.PS_LUT_SKYVIEW,
                                            abi.PS_LUT_AERIAL_PERSPECTIVE)}
    r.set_frame(frames[-1])
    r.bake_luts()
    for k, v in lut_after.items():
        assert np.array_equal(v, r.read_lut(k)), k


def test_pipelined_timing_splits_luts_and_render(rend):
    r, w = rend
    frames = _frames(w, 6)
    bufs = r.alloc_outputs(w.width, w.height, abi.PS_PIXEL_F32, ("composite",))
    out = r.image_out(bufs)
    r.timing_begin(6)
    for f in frames:
        r.set_frame(f)
        r.frame(out)
    t = r.timing_end()
    _log(f"  timing over 6 pipelined frames: T {t.transmittance_ms:.4f} MS {t.multiscatter_ms:.4f} "
         f"SV+AP {t.skyview_ms + t.aerial_perspective_ms:.4f} render {t.render_ms:.4f} frame {t.frame_ms:.4f} ms")
    assert t.frames == 6
    for v in (t.transmittance_ms, t.multiscatter_ms, t.skyview_ms, t.render_ms, t.frame_ms):
        assert v > 0.0
    lut_ms = t.transmittance_ms + t.multiscatter_ms + t.skyview_ms + t.aerial_perspective_ms
    assert lut_ms < 1.0 and t.render_ms > lut_ms
    # frame_ms: the first bake's start to the last render's end, per frame; with the bakes
    # hidden behind the previous renders it stays below render + LUT time
    assert t.frame_ms <= (t.render_ms + lut_ms) * 1.05


def test_pipelined_training_steps_match(rend):
    """ps_frame_train (pipelined LUT chain + training forward) then
    ps_backward, over several optimisation-like steps with new cameras:
    outputs bit-identical to ps_bake_luts + ps_render_train, gradients equal
    to the order of the atomic additions (FP32 per-layer sums: 1e-6)."""
    import ctypes as C
    import torch
    r, w = rend
    frames = _frames(w, 4)
    dev = "cuda:0"

    def steps(pipelined):
        res = []
        for f in frames:
            r.set_frame(f)
            bufs = r.alloc_outputs(w.width, w.height, abi.PS_PIXEL_F32, ("composite",))
            out = r.image_out(bufs)
            if pipelined:
                abi.check(r.lib.ps_frame_train(r.ctx, abi.PS_LUT_ALL, C.c_void_p(r.depth_ptr) if r.depth_ptr else None,
                                               C.byref(out)))
            else:
                r.bake_luts()
                r.render_train(out)
            g = torch.ones_like(bufs["composite"])
            lg = torch.zeros(24, dtype=torch.float32, device=dev)
            torch.cuda.synchronize()
            r.backward(g_composite=g, layer_grads=lg)
            r.synchronize()
            res.append((bufs["composite"].cpu().numpy(), lg.cpu().numpy()))
        return res

    ref, got = steps(False), steps(True)
    for k, ((c0, g0), (c1, g1)) in enumerate(zip(ref, got)):
        same = bool(np.array_equal(c0, c1))
        gd = float(np.abs(g1 - g0).max() / max(np.abs(g0).max(), 1e-30))
        _log(f"  step {k}: forward identical {same}, layer gradients max rel diff {gd:.2e}")
        assert same and gd <= 1e-6


def test_backward_after_later_pipelined_frames(rend):
    """A backward issued after the next pipelined frame reads the LUT set
    its training forward used, and the frame after that (which rebakes that
    set on the side stream) waits for it: the gradients equal those of a
    backward issued straight after its training forward, and the later
    frames equal bake + render."""
    import ctypes as C
    import torch
    r, w = rend
    frames = _frames(w, 3)
    dev = "cuda:0"
    _log("case: ps_frame_train, ps_frame (new sun), ps_backward of the first, ps_frame (new sun)")

    def grads_buffers():
        return torch.zeros(24, dtype=torch.float32, device=dev)

    r.set_frame(frames[0])
    b0 = r.alloc_outputs(w.width, w.height, abi.PS_PIXEL_F32, ("composite",))
    r.bake_luts()
    r.render_train(r.image_out(b0))
    g = torch.ones_like(b0["composite"])
    lg_ref = grads_buffers()
    torch.cuda.synchronize()
    r.backward(g_composite=g, layer_grads=lg_ref)
    ref_later = _run(r, w, frames[1:], ["bake+render"] * 2)

    r.set_frame(frames[0])
    b0p = r.alloc_outputs(w.width, w.height, abi.PS_PIXEL_F32, ("composite",))
    depth = C.c_void_p(r.depth_ptr) if r.depth_ptr else None
    abi.check(r.lib.ps_frame_train(r.ctx, abi.PS_LUT_ALL, depth, C.byref(r.image_out(b0p))))
    later = [r.alloc_outputs(w.width, w.height, abi.PS_PIXEL_F32, ("composite", "cloud_transmittance"))
             for _ in range(2)]
    r.set_frame(frames[1])
    r.frame(r.image_out(later[0]))
    lg = grads_buffers()
    torch.cuda.synchronize()
    r.backward(g_composite=g, layer_grads=lg)
    r.set_frame(frames[2])
    r.frame(r.image_out(later[1]))
    r.synchronize()
    torch.cuda.synchronize()
    gd = float((lg - lg_ref).abs().max() / lg_ref.abs().max().clamp_min(1e-30))
    _log(f"  layer gradients max rel diff {gd:.2e}")
    assert np.array_equal(b0p["composite"].cpu().numpy(), b0["composite"].cpu().numpy())
    assert gd <= 1e-6
    for k, (e, bufs) in enumerate(zip(ref_later, later)):
        assert all(np.array_equal(bufs[n].cpu().numpy(), e[n]) for n in e), k
