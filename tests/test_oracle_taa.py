"""CPU pins of the oracle's cloud TAA resolve (shaders/clouds/cloud_taa.wgsl)
and of its RT march / RT composite entry points.

The reprojection is checked against an independent pinhole model built from
the camera's yaw, pitch and field of view (no matrices), read back through a
history that is a linear ramp in texel coordinates (bilinear filtering
reproduces it exactly) and a current frame whose 3x3 neighbourhood box holds
the whole ramp. The remaining checks are closed forms: a constant current
frame clamps any history to itself, rejected reprojections (behind the
previous camera, off its screen) and invalid history return the current
sample, and a fixed camera turns the resolve into the exponential average
that converges as 0.875^n.
"""
import dataclasses
import math

import numpy as np
import pytest

import oracle
from pedalsky_amd import scene

from test_oracle_adjoint import _smooth_noise

RW, RH = 64, 36
K = 1.0e6


def _frame_pair(yaw0, yaw1, pitch0=0.0, pitch1=0.0, camera=(0.0, 1.0, 0.0)):
    """A workload whose frame has camera yaw1/pitch1 and prev_view_proj of
    a camera at yaw0/pitch0 (same position)."""
    w = scene.workload_small(width=RW, height=RH, steps=8)
    f1 = scene.make_frame(RW, RH, camera=camera, pitch_deg=pitch1, yaw_deg=yaw1)
    f0 = scene.make_frame(RW, RH, camera=camera, pitch_deg=pitch0, yaw_deg=yaw0)
    f1.prev_view_proj[:] = f0.view_proj[:]
    w = dataclasses.replace(w)
    w.frame = f1
    w.depth = None
    return w


@pytest.fixture(scope="module")
def noise():
    return _smooth_noise()


def _oracle(w, noise):
    return oracle.Frame(w, *noise)


def _basis(pitch_deg, yaw_deg):
    f = scene.fly_camera_forward(math.radians(pitch_deg), math.radians(yaw_deg))
    s = np.cross(f, [0.0, 1.0, 0.0])
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    return f, s, u


def _pinhole_prev_uv(yaw0, yaw1, pitch0, pitch1, fov_deg=60.0):
    """Previous-frame uv of every RT texel centre by the pinhole model."""
    ty = math.tan(math.radians(fov_deg) / 2)
    tx = ty * RW / RH
    f1, s1, u1 = _basis(pitch1, yaw1)
    f0, s0, u0 = _basis(pitch0, yaw0)
    ys, xs = np.meshgrid(np.arange(RH) + 0.5, np.arange(RW) + 0.5, indexing="ij")
    nx, ny = xs / RW * 2 - 1, 1 - ys / RH * 2
    d = f1[None, None] + (nx * tx)[..., None] * s1 + (ny * ty)[..., None] * u1
    z = d @ f0
    px, py = (d @ s0) / z / tx, (d @ u0) / z / ty
    return px * 0.5 + 0.5, 0.5 - py * 0.5


def _columns_box():
    """Current frame whose every 3x3 neighbourhood spans [0.5 K, 1.5 K]."""
    cols = np.where(np.arange(RW) % 2 == 0, 1.5 * K, 0.5 * K)
    c = np.broadcast_to(cols[None, :, None], (RH, RW, 4)).copy()
    return c


def _ramp():
    ys, xs = np.meshgrid(np.arange(RH), np.arange(RW), indexing="ij")
    r = K + xs + 100.0 * ys
    return np.repeat(r[..., None], 4, axis=-1).astype(np.float64)


@pytest.mark.parametrize("yaw0,yaw1,pitch0,pitch1", [(30.0, 31.5, 5.0, 5.0), (200.0, 199.0, 10.0, 11.0)])
def test_reprojection_matches_pinhole_model(noise, yaw0, yaw1, pitch0, pitch1):
    w = _frame_pair(yaw0, yaw1, pitch0, pitch1)
    fr = _oracle(w, noise)
    cur = _columns_box()
    hist = _ramp()
    b = 1.0 / 8.0
    out_l, _ = fr.taa_resolve(cur, np.full_like(cur, 0.5), hist, np.full_like(cur, 0.5), b, 1.0)
    h_read = (out_l - b * cur) / (1.0 - b)          # the history value the resolve read
    pu, pv = _pinhole_prev_uv(yaw0, yaw1, pitch0, pitch1)
    px, py = pu * RW - 0.5, pv * RH - 0.5            # continuous texel coordinates
    expect = K + px + 100.0 * py
    inner = (px >= 0) & (px <= RW - 1) & (py >= 0) & (py <= RH - 1)
    assert inner.mean() > 0.5
    err = np.abs(h_read[..., 0] - expect)[inner]
    assert err.max() < 2e-2, err.max()              # 1e-4 texel in x, 2e-4 in y
    # texels whose direction left the previous screen keep the current value
    outside = (pu < -1e-3) | (pu > 1 + 1e-3) | (pv < -1e-3) | (pv > 1 + 1e-3)
    assert outside.any()
    assert np.array_equal(out_l[outside], cur[outside])


def test_constant_current_frame_clamps_any_history(noise):
    w = _frame_pair(10.0, 11.0)
    fr = _oracle(w, noise)
    rng = np.random.default_rng(7)
    cur_l = np.full((RH, RW, 4), 0.3)
    cur_t = np.full((RH, RW, 4), 0.6)
    out_l, out_t = fr.taa_resolve(cur_l, cur_t, rng.uniform(-5, 5, cur_l.shape), rng.uniform(-5, 5, cur_l.shape),
                                  0.125, 1.0)
    assert np.allclose(out_l, 0.3, rtol=0, atol=1e-15)
    assert np.allclose(out_t, 0.6, rtol=0, atol=1e-15)


@pytest.mark.parametrize("yaw0", [190.0, 130.0])
def test_rejected_reprojection_returns_current(noise, yaw0):
    """Behind the previous camera (180 deg) or outside its frustum (120 deg; the horizontal field is 91 deg):
    the output is the current sample, clamped."""
    w = _frame_pair(yaw0, 10.0)
    fr = _oracle(w, noise)
    rng = np.random.default_rng(3)
    cur_l, cur_t = rng.uniform(-0.2, 2, (RH, RW, 4)), rng.uniform(-0.2, 1.2, (RH, RW, 4))
    out_l, out_t = fr.taa_resolve(cur_l, cur_t, rng.uniform(0, 1, cur_l.shape), rng.uniform(0, 1, cur_l.shape),
                                  0.125, 1.0)
    assert np.array_equal(out_l, np.maximum(cur_l, 0.0))
    assert np.array_equal(out_t, np.clip(cur_t, 0.0, 1.0))


def test_no_history_returns_current_unclamped(noise):
    w = _frame_pair(10.0, 10.5)
    fr = _oracle(w, noise)
    rng = np.random.default_rng(5)
    cur_l, cur_t = rng.uniform(-1, 2, (RH, RW, 4)), rng.uniform(-1, 2, (RH, RW, 4))
    h = rng.uniform(0, 1, cur_l.shape)
    for blend, valid in ((0.125, 0.0), (1.0, 1.0), (3.0, 1.0)):
        out_l, out_t = fr.taa_resolve(cur_l, cur_t, h, h, blend, valid)
        assert np.array_equal(out_l, cur_l) and np.array_equal(out_t, cur_t)


def test_fixed_camera_is_an_exponential_average(noise):
    w = _frame_pair(20.0, 20.0)
    fr = _oracle(w, noise)
    rng = np.random.default_rng(11)
    cur = _columns_box() * 1e-6                      # every box spans [0.5, 1.5]
    hist = 1.0 + rng.uniform(-0.4, 0.4, cur.shape)
    out, _ = fr.taa_resolve(cur, np.full_like(cur, 0.5), hist, np.full_like(cur, 0.5), 0.125, 1.0)
    assert np.abs(out - (0.875 * hist + 0.125 * cur)).max() < 1e-6
    # repeated frames: the distance to the current frame shrinks as 0.875^n
    h = hist
    for _ in range(6):
        h, _ = fr.taa_resolve(cur, np.full_like(cur, 0.5), h, np.full_like(cur, 0.5), 0.125, 1.0)
    assert np.abs((h - cur) - 0.875 ** 6 * (hist - cur)).max() < 1e-5


def test_rt_march_and_composite_reproduce_the_renders():
    W, H = 96, 54
    w = scene.workload_small(width=W, height=H, steps=24)
    fr = oracle.Frame(w, *_smooth_noise())
    comp, lum, trn, _ = fr.render_region(0, 0, W, H)
    l_rt, t_rt = fr.march_rt(W, H)
    c2, l2, t2 = fr.composite_rt(l_rt, t_rt, upsample=False)
    scale = np.abs(comp).max()
    assert np.abs(c2[..., :3] - comp[..., :3]).max() <= 1e-12 * scale
    assert np.abs(l2[..., :3] - lum[..., :3]).max() <= 1e-12 * np.abs(lum).max()
    assert np.abs(t2[..., :3] - trn[..., :3]).max() <= 1e-12
    ch, lh, th = fr.render_halfres(0, 0, W, H)
    l_h, t_h = fr.march_rt(W // 2, H // 2)
    c3, l3, t3 = fr.composite_rt(l_h, t_h, upsample=True)
    assert np.array_equal(c3, ch) and np.array_equal(l3, lh) and np.array_equal(t3, th)


def test_rt_march_with_extensions_reproduces_render_ext():
    W, H = 64, 36
    w = scene.workload_adjoint(width=W, height=H)
    fr = oracle.Frame(w, *_smooth_noise())
    tgt = scene.adjoint_target(w.layers)
    diff = oracle.Diff(2, density_scale=tgt.density_scale, extinction_scale=tgt.extinction_scale,
                       coverage=tgt.coverage, delta=tgt.delta)
    comp, lum, trn, _ = fr.render_ext(diff, 0, 0, W, H)
    c2, l2, t2 = fr.composite_rt(*fr.march_rt(W, H, diff), upsample=False)
    assert np.abs(c2[..., :3] - comp[..., :3]).max() <= 1e-12 * np.abs(comp).max()
    assert np.abs(t2[..., :3] - trn[..., :3]).max() <= 1e-12
    plain, _, _, _ = fr.render_region(0, 0, W, H)
    assert np.abs(plain[..., :3] - comp[..., :3]).max() > 1e-3 * np.abs(comp).max()
