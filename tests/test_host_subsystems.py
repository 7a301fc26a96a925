"""The C++ host layer over the C ABI (include/pedalsky_subsystems.hpp): the
reference's RenderSubsystem boundary (crates/ps-core/src/subsystem.rs) for
AtmosphereSubsystem and CloudsSubsystem plus the executor's pass order,
driven by tests/host/subsystem_frame.cpp (built by build()).

CPU: the executor's order (prepare in registration order, passes stable-
sorted by PassStage) and the error contract (a failing call throws with the
library's status and message). GPU: two frames through the executor (the
second with the same weather revision, so the static LUTs are not baked
again) write exactly what the Python ABI path's ps_frame writes, with and
without a clouds subsystem.
"""
import ctypes as C
import subprocess
import sys

import numpy as np
import pytest

from pedalsky_amd import abi, scene
from pedalsky_amd.build import HOST_TEST_BIN


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _binary():
    assert HOST_TEST_BIN.exists(), f"{HOST_TEST_BIN} is missing: run build() first"
    return str(HOST_TEST_BIN)


def test_host_layer_executor_order_and_errors():
    res = subprocess.run([_binary(), "--selftest"], capture_output=True, text=True, timeout=60)
    _log(res.stdout + res.stderr)
    assert res.returncode == 0 and "selftest ok" in res.stdout


def _write_scene(path, w):
    i32 = lambda v: np.int32(v).tobytes()  # noqa: E731
    with open(path, "wb") as f:
        f.write(b"PSSCENE1")
        f.write(bytes(w.frame))
        f.write(bytes(w.atmosphere))
        f.write(bytes(w.clouds))
        f.write(i32(len(w.layers)))
        for L in w.layers:
            f.write(bytes(L))
        f.write(np.ascontiguousarray(w.winds, np.float32).reshape(-1)[:2 * len(w.layers)].tobytes())
        f.write(i32(w.noise_seed) + i32(w.noise_format))
        f.write(bytes(w.lut_dims))
        wt = np.ascontiguousarray(w.weather, np.uint16)
        f.write(i32(wt.shape[1]) + i32(wt.shape[0]) + i32(wt.shape[2]) + wt.tobytes())
        f.write(i32(0 if w.type_grid is None else 1))
        grid = np.zeros(128 * 128, np.uint8) if w.type_grid is None else np.ascontiguousarray(w.type_grid, np.uint8)
        f.write(grid.reshape(-1).tobytes())
        f.write(np.ascontiguousarray(w.overcast, np.uint8).reshape(-1).tobytes())
        f.write(np.ascontiguousarray(w.blue, np.uint8).reshape(-1).tobytes())
        f.write(i32(w.width) + i32(w.height) + i32(0 if w.depth is None else 1))
        if w.depth is not None:
            f.write(np.ascontiguousarray(w.depth, np.float32).tobytes())


@pytest.mark.gpu
@pytest.mark.parametrize("sky_only", [False, True])
def test_host_layer_frame_matches_ps_frame(tmp_path, sky_only):
    import torch
    from pedalsky_amd.renderer import Renderer
    assert torch.cuda.is_available(), "GPU tests need a GPU (no CPU fallback exists)"
    w = scene.workload_small()
    _log(f"case: {w.name} through the C++ executor ({'atmosphere only' if sky_only else 'atmosphere + clouds'})")
    _write_scene(tmp_path / "scene.bin", w)
    cmd = [_binary(), "--scene", str(tmp_path / "scene.bin"), "--out", str(tmp_path / "out.bin")]
    if sky_only:
        cmd.append("--sky-only")
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    _log(res.stdout + res.stderr)
    assert res.returncode == 0 and "frame ok" in res.stdout
    got = np.fromfile(tmp_path / "out.bin", np.float32).reshape(w.height, w.width, 4)
    with Renderer(0) as r:
        r.load_workload(w)
        if sky_only:
            p = abi.CloudParams()
            r.lib.ps_default_cloud_params(C.byref(p))
            r.set_clouds(p, [])
        want = r.render_numpy(which=("composite",))["composite"]
    same = bool(np.array_equal(got, want))
    _log(f"  composite identical to ps_frame: {same} (max abs diff {float(np.abs(got - want).max()):.3e})")
    assert same
