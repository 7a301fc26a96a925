"""GPU tests of the march step counts: ps_set_march_steps (the reference's
CloudsSubsystem::set_march_steps, crates/ps-clouds/src/lib.rs:701-712, which
clamps the primary steps to [8, 256] and the light steps to [1, 16]) and the
batch step schedule (ps_batch_*_step_schedule, ps_batch_autotune_steps).

A render at any step counts must match the FP64 oracle marching with the same
counts (relative radiance error 99.9th percentile <= 1e-3), and setting the
counts through the knob must be bit-identical to passing them in the cloud
parameters. A batch under a step schedule renders each scene exactly as
ps_frame does with its class's counts (8x8 tiles: bit-identical).
"""
import copy
import ctypes as C
import sys

import numpy as np
import pytest

import oracle
from pedalsky_amd import abi, scene
from pedalsky_amd.batch import Batch
from pedalsky_amd.renderer import Renderer

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rend():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU (no CPU fallback exists)"
    w = scene.workload_small()
    r = Renderer(0)
    r.load_workload(w)
    yield r, w
    r.close()


def _with_steps(w, cloud, light):
    w2 = copy.copy(w)
    w2.clouds = abi.CloudParams.from_buffer_copy(w.clouds)
    w2.clouds.cloud_steps, w2.clouds.light_steps = cloud, light
    return w2


def _render(r, w):
    r.set_frame(w.frame)
    r.set_atmosphere(w.atmosphere)
    r.set_depth(w.depth)
    return r.render_numpy(which=("composite",))["composite"]


def _noise(r, fmt):
    return tuple(oracle.decode(r.read_noise(k, fmt), fmt)
                 for k in (abi.PS_NOISE_BASE, abi.PS_NOISE_DETAIL, abi.PS_NOISE_CURL))


def test_set_march_steps_clamps_as_the_reference(rend):
    r, w = rend
    for asked, clamped in (((300, 20), (256, 16)), ((4, 0), (8, 1)), ((40, 3), (40, 3))):
        _log(f"case: set_march_steps{asked} on {w.name}")
        r.set_clouds(w.clouds, w.layers, w.winds)
        r.set_march_steps(*asked)
        got = _render(r, w)
        w2 = _with_steps(w, *clamped)
        r.set_clouds(w2.clouds, w2.layers, w2.winds)
        want = _render(r, w2)
        _log(f"  same as cloud params with {clamped}: {bool(np.array_equal(got, want))}")
        assert np.array_equal(got, want)
    r.set_clouds(w.clouds, w.layers, w.winds)


def test_march_steps_match_oracle(rend):
    r, w = rend
    base, det, curl = _noise(r, w.noise_format)
    for cloud, light in ((24, 2), (200, 12)):
        _log(f"case: {w.name} at {cloud} primary / {light} light steps vs the oracle")
        r.set_clouds(w.clouds, w.layers, w.winds)
        r.set_march_steps(cloud, light)
        got = _render(r, w)
        comp, _, _, _ = oracle.Frame(_with_steps(w, cloud, light), base, det, curl).render_region(
            0, 0, w.width, w.height)
        e = oracle.relative_error(got, comp)
        p999 = float(np.percentile(e, 99.9))
        _log(f"  composite p99.9 {p999:.3e} median {float(np.median(e)):.3e}")
        assert p999 <= 1e-3
    r.set_clouds(w.clouds, w.layers, w.winds)
    lib = r.lib
    assert lib.ps_set_march_steps(None, 64, 6) == abi.PS_ERR_INVALID_ARGUMENT


@pytest.fixture(scope="module")
def batch_setup(rend):
    r, _ = rend
    ws = scene.workload_batch(n=6, seed=0xB7)
    for w in ws:
        w.overcast = ws[0].overcast
    r.load_workload(ws[0])
    return r, ws


def _single(r, ws, steps):
    import torch
    outs = []
    for w, (cloud, light) in zip(ws, steps):
        r.set_atmosphere(w.atmosphere)
        r.set_frame(w.frame)
        r.set_clouds(w.clouds, w.layers, w.winds)
        r.set_march_steps(cloud, light)
        r.set_depth(w.depth)
        bufs = r.alloc_outputs(w.width, w.height, abi.PS_PIXEL_F32, ("composite",))
        r.frame(r.image_out(bufs))
        r.synchronize()
        outs.append(bufs["composite"].cpu().numpy())
    torch.cuda.synchronize()
    return outs


def test_batch_step_schedule_is_per_scene_set_march_steps(batch_setup):
    import torch
    r, ws = batch_setup
    b = Batch(r, ws, tuning=abi.Tuning(8, 0, 8, 1))
    t = b.step_schedule()
    sizes = [(t.width[k], t.height[k]) for k in range(t.n_classes)]
    _log(f"case: 6-scene batch step schedule, classes {sizes}")
    assert sorted(set((w.width, w.height) for w in ws), key=lambda s: (-s[0] * s[1], -s[0])) == sizes
    assert all(t.cloud_steps[k] == 0 and t.light_steps[k] == 0 for k in range(t.n_classes))
    # a count for every class (one out of range: clamped as set_march_steps clamps)
    plan = [(64, 4), (300, 6), (32, 2), (48, 20)]
    for k in range(t.n_classes):
        t.cloud_steps[k], t.light_steps[k] = plan[k % len(plan)]
    b.set_step_schedule(t)
    t2 = b.step_schedule()
    steps = []
    for w in ws:
        k = sizes.index((w.width, w.height))
        steps.append((t2.cloud_steps[k], t2.light_steps[k]))
    assert all(8 <= c <= 256 and 1 <= l <= 16 for c, l in steps)
    b.capture()
    b.replay()
    r.synchronize()
    torch.cuda.synchronize()
    got = [bb["composite"].cpu().numpy() for bb in b.bufs]
    ref = _single(r, ws, steps)
    same = [bool(np.array_equal(g, e)) for g, e in zip(got, ref)]
    _log(f"  per-scene steps {steps}: graph replay bit-identical to ps_frame: {sum(same)} / {len(same)}")
    assert all(same)
    # the smallest scene against the oracle at its class's counts
    i = int(np.argmin([w.width * w.height for w in ws]))
    w = _with_steps(ws[i], *steps[i])
    base, det, curl = _noise(r, w.noise_format)
    comp, _, _, _ = oracle.Frame(w, base, det, curl).render_region(0, 0, w.width, w.height)
    e = oracle.relative_error(got[i], comp)
    p999 = float(np.percentile(e, 99.9))
    _log(f"  scene {i} at {steps[i]} vs the oracle: composite p99.9 {p999:.3e}")
    assert p999 <= 1e-3
    # 0 restores the scenes' own counts
    for k in range(t.n_classes):
        t.cloud_steps[k] = t.light_steps[k] = 0
    b.set_step_schedule(t)
    b.replay()
    r.synchronize()
    torch.cuda.synchronize()
    own = [bb["composite"].cpu().numpy() for bb in b.bufs]
    ref = _single(r, ws, [(w.clouds.cloud_steps, w.clouds.light_steps) for w in ws])
    assert all(np.array_equal(g, e) for g, e in zip(own, ref))
    # errors: a schedule that is not the batch's
    lib = r.lib
    bad = b.step_schedule()
    bad.width[0] += 1
    assert lib.ps_batch_set_step_schedule(b.h, C.byref(bad)) == abi.PS_ERR_INVALID_ARGUMENT
    bad = b.step_schedule()
    bad.n_classes += 1
    assert lib.ps_batch_set_step_schedule(b.h, C.byref(bad)) == abi.PS_ERR_INVALID_ARGUMENT
    bad = b.step_schedule()
    bad.cloud_steps[0] = -1
    assert lib.ps_batch_set_step_schedule(b.h, C.byref(bad)) == abi.PS_ERR_INVALID_ARGUMENT
    assert lib.ps_batch_autotune_steps(b.h, C.c_float(0.0), 1, None) == abi.PS_ERR_INVALID_ARGUMENT
    b.close()


def test_batch_autotune_steps_meets_budget(batch_setup):
    r, ws = batch_setup
    b = Batch(r, ws)
    b.capture()
    # a generous budget keeps every class at its scenes' own counts
    full = b.autotune_steps(1e6, reps=2)
    own = {}
    for w in ws:
        k = [(full.width[j], full.height[j]) for j in range(full.n_classes)].index((w.width, w.height))
        own[k] = max(own.get(k, (0, 0))[0], w.clouds.cloud_steps), max(own.get(k, (0, 0))[1], w.clouds.light_steps)
    _log(f"case: autotune_steps, generous budget: {full.as_dict()}")
    for k in range(full.n_classes):
        assert (full.cloud_steps[k], full.light_steps[k]) == (min(own[k][0], 256), min(own[k][1], 16))
    # half the full-quality time: the schedule gives up march steps, largest work first
    budget = 0.5 * full.batch_ms
    t = b.autotune_steps(budget, reps=2)
    _log(f"  budget {budget:.2f} ms: {t.as_dict()}")
    for k in range(t.n_classes):
        assert 8 <= t.cloud_steps[k] <= 256 and 1 <= t.light_steps[k] <= 16
        assert t.cloud_steps[k] <= full.cloud_steps[k] and t.light_steps[k] <= full.light_steps[k]
    at_min = all(t.cloud_steps[k] == 8 and t.light_steps[k] == 1 for k in range(t.n_classes))
    assert t.batch_ms <= budget or at_min
    assert b.step_schedule().as_dict()["classes"] == t.as_dict()["classes"]
    # the applied schedule runs at the reported time (replay of the captured graph)
    again = b.time(3)
    _log(f"  replay at the chosen schedule: {again:.2f} ms (tuner measured {t.batch_ms:.2f})")
    assert again <= 1.15 * t.batch_ms + 0.5
    b.close()
