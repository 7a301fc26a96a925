"""CPU pins of the oracle's half-resolution path (the reference's
clouds.half_res_render: crates/ps-clouds/src/lib.rs:748-760,
crates/ps-clouds/src/pipeline.rs:277-319, cloud_composite_halfres.wgsl).

The RT texel (x, y) of a W x H frame is the full-resolution march of the same
camera at viewport floor(W/2) x floor(H/2) whose depth buffer holds the full
frame's depth at ((int)(x * W / rw), (int)(y * H / rh)); the composite is the
full-resolution sky plus that RT upsampled with the Catmull-Rom kernel
(B = 0, C = 0.5, ClampToEdge), luminance clamped >= 0 and transmittance to
[0, 1]. The test rebuilds the composite from those pieces (an independent
numpy Catmull-Rom) and compares it with the oracle's own half-resolution
entry point.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from pedalsky_amd import scene

from test_oracle_adjoint import _smooth_noise


def _catmull_rom_up(rt, W, H):
    """rt: (rh, rw, C) -> (H, W, C) at pixel centres, ClampToEdge."""
    rh, rw = rt.shape[:2]

    def axis(n_out, n_in):
        s = (np.arange(n_out) + 0.5) / n_out * n_in
        b = np.floor(s - 0.5)
        f = s - (b + 0.5)
        w = np.stack([f * (-0.5 + f * (1.0 - 0.5 * f)), 1.0 + f * f * (-2.5 + 1.5 * f),
                      f * (0.5 + f * (2.0 - 1.5 * f)), f * f * (-0.5 + 0.5 * f)], -1)
        idx = np.clip(b[:, None].astype(int) - 1 + np.arange(4)[None, :], 0, n_in - 1)
        m = np.zeros((n_out, n_in))
        for k in range(4):
            np.add.at(m, (np.arange(n_out), idx[:, k]), w[:, k])
        return m
    my, mx = axis(H, rh), axis(W, rw)
    return np.einsum("yi,ijc,xj->yxc", my, rt, mx)


@pytest.mark.parametrize("size", [(96, 54), (97, 55)])
def test_halfres_composite_is_the_upsampled_half_size_march(size):
    W, H = size
    w = scene.workload_small(width=W, height=H, steps=24)
    base, det, curl = _smooth_noise()
    rw, rh = max(1, int(np.float32(W) * np.float32(0.5))), max(1, int(np.float32(H) * np.float32(0.5)))
    # the RT: a march of the same camera at the RT's viewport, reading the
    # full frame's depth at the scaled coordinates
    small = dataclasses.replace(w)
    small.frame = type(w.frame).from_buffer_copy(w.frame)
    small.frame.viewport_size[0] = rw
    small.frame.viewport_size[1] = rh
    if w.depth is not None:
        ys = (np.arange(rh) * (H / rh)).astype(int)
        xs = (np.arange(rw) * (W / rw)).astype(int)
        small.depth = np.ascontiguousarray(w.depth[ys][:, xs])
    fr_small = oracle.Frame(small, base, det, curl)
    _, l_rt, t_rt, _ = fr_small.render_region(0, 0, rw, rh)
    # the full-resolution sky: the same frame without cloud layers
    clear = dataclasses.replace(w)
    clear.clouds = type(w.clouds).from_buffer_copy(w.clouds)
    clear.clouds.cloud_layer_count = 0
    sky, _, _, _ = oracle.Frame(clear, base, det, curl).render_region(0, 0, W, H)
    lum = np.maximum(_catmull_rom_up(l_rt[..., :3], W, H), 0.0)
    trn = np.clip(_catmull_rom_up(t_rt[..., :3], W, H), 0.0, 1.0)
    comp = lum + sky[..., :3] * trn
    fr = oracle.Frame(w, base, det, curl)
    c_h, l_h, t_h = fr.render_halfres(0, 0, W, H)
    assert (t_rt[..., 1] < 0.9).mean() > 0.02, "the RT should hold clouds"
    scale = np.abs(comp).max()
    assert np.abs(c_h[..., :3] - comp).max() <= 1e-9 * scale
    assert np.abs(l_h[..., :3] - lum).max() <= 1e-9 * np.abs(lum).max()
    assert np.abs(t_h[..., :3] - trn).max() <= 1e-9
    # a sub-region is the same computation
    c_s, _, _ = fr.render_halfres(13, 7, 40, 30)
    assert np.array_equal(c_s, c_h[7:37, 13:53])


def test_catmull_rom_reproduces_linear_fields():
    """The Catmull-Rom kernel interpolates exactly and reproduces linear
    functions away from the clamped border."""
    rw, rh, W, H = 20, 12, 40, 24
    yy, xx = np.meshgrid(np.arange(rh) + 0.5, np.arange(rw) + 0.5, indexing="ij")
    rt = (2.0 * xx + 3.0 * yy)[..., None]
    up = _catmull_rom_up(rt, W, H)[..., 0]
    Y, X = np.meshgrid((np.arange(H) + 0.5) / H * rh, (np.arange(W) + 0.5) / W * rw, indexing="ij")
    inner = (X > 2) & (X < rw - 2) & (Y > 2) & (Y < rh - 2)
    assert np.allclose(up[inner], (2.0 * X + 3.0 * Y)[inner], atol=1e-12)
