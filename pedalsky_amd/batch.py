"""Many scenes per launch through the C ABI (ps_batch_*).

The context holds the inputs every scene shares (noise, weather map, blue
noise, LUT sizes); each scene brings its own uniforms, overcast mask, depth
and output images. ``Batch`` keeps every buffer a descriptor points at alive
for the lifetime of the batch.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import Optional, Sequence

import numpy as np

from . import abi
from .abi import check


class Batch:
    def __init__(self, renderer, workloads: Sequence, fmt: int = abi.PS_PIXEL_F32,
                 outputs=("composite",), tuning: Optional[abi.Tuning] = None):
        import torch
        self.r = renderer
        self.lib = renderer.lib
        self.ws = list(workloads)
        self.fmt = fmt
        dev = f"cuda:{renderer.device}"
        self.bufs = [renderer.alloc_outputs(w.width, w.height, fmt, outputs) for w in self.ws]
        self.depths = [None if w.depth is None else torch.from_numpy(np.ascontiguousarray(w.depth, np.float32)).to(dev)
                       for w in self.ws]
        torch.cuda.synchronize()
        self._keep = []
        self.descs = (abi.SceneDesc * len(self.ws))()
        self._fill()
        h = C.c_void_p()
        check(self.lib.ps_batch_create(renderer.ctx, self.descs, len(self.ws), C.byref(h)))
        self.h = h
        if tuning is not None:
            self.set_tuning(tuning)

    def _fill(self) -> None:
        self._keep = []
        for i, w in enumerate(self.ws):
            d = self.descs[i]
            layers = (abi.CloudLayer * max(len(w.layers), 1))(*w.layers)
            winds = np.ascontiguousarray(w.winds, dtype=np.float32).reshape(-1)
            oc = np.ascontiguousarray(w.overcast, dtype=np.uint8)
            self._keep += [layers, winds, oc]
            d.frame = C.addressof(w.frame)
            d.atmosphere = C.addressof(w.atmosphere)
            d.clouds = C.addressof(w.clouds)
            d.layers = C.addressof(layers)
            d.n_layers = len(w.layers)
            d.layer_wind_xz = winds.ctypes.data if len(winds) else None
            d.overcast_mask_128 = oc.ctypes.data
            d.depth = None if self.depths[i] is None else self.depths[i].data_ptr()
            d.out = self.r.image_out(self.bufs[i], self.fmt)

    @property
    def pixels(self) -> int:
        return sum(w.width * w.height for w in self.ws)

    def update(self, workloads: Sequence) -> None:
        self.ws = list(workloads)
        self._fill()
        check(self.lib.ps_batch_update(self.h, self.descs, len(self.ws)))

    def set_tuning(self, t: abi.Tuning) -> None:
        check(self.lib.ps_batch_set_tuning(self.h, C.byref(t)))

    def tuning(self) -> abi.Tuning:
        t = abi.Tuning()
        check(self.lib.ps_batch_get_tuning(self.h, C.byref(t)))
        return t

    def autotune(self, reps: int = 2) -> abi.Tuning:
        t = abi.Tuning()
        check(self.lib.ps_batch_autotune(self.h, reps, C.byref(t)))
        return t

    def step_schedule(self) -> abi.StepSchedule:
        t = abi.StepSchedule()
        check(self.lib.ps_batch_get_step_schedule(self.h, C.byref(t)))
        return t

    def set_step_schedule(self, t: abi.StepSchedule) -> None:
        check(self.lib.ps_batch_set_step_schedule(self.h, C.byref(t)))

    def autotune_steps(self, budget_ms: float, reps: int = 2) -> abi.StepSchedule:
        """Step counts per resolution class for a device-time budget per
        run (ps_batch_autotune_steps); the schedule stays applied."""
        t = abi.StepSchedule()
        check(self.lib.ps_batch_autotune_steps(self.h, float(budget_ms), reps, C.byref(t)))
        return t

    def run(self) -> None:
        check(self.lib.ps_batch_run(self.h))

    def capture(self) -> None:
        check(self.lib.ps_batch_capture(self.h))

    def replay(self) -> None:
        check(self.lib.ps_batch_replay(self.h))

    def time(self, reps: int) -> float:
        ms = C.c_float()
        check(self.lib.ps_batch_time(self.h, reps, C.byref(ms)))
        return float(ms.value)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.ps_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def time_sequential(renderer, workloads: Sequence, bufs, depths, reps: int = 1) -> float:
    """Wall time (ms) of the plain path over the scenes, one after another:
    per scene the uniform uploads, the LUT chain and the render (ps_frame),
    with one device synchronisation at the end of each pass over the list."""
    import torch
    lib = renderer.lib
    layer_arrays = [(abi.CloudLayer * max(len(w.layers), 1))(*w.layers) for w in workloads]
    winds = [np.ascontiguousarray(w.winds, dtype=np.float32).reshape(-1) for w in workloads]
    outs = [renderer.image_out(b) for b in bufs]
    wptr = [(C.c_float * len(x))(*x.tolist()) if len(x) else None for x in winds]
    t0 = 0.0
    for rep in range(reps + 1):
        if rep == 1:
            renderer.synchronize()
            t0 = time.perf_counter()
        for i, w in enumerate(workloads):
            check(lib.ps_set_atmosphere(renderer.ctx, C.byref(w.atmosphere)))
            check(lib.ps_set_frame(renderer.ctx, C.byref(w.frame)))
            check(lib.ps_set_clouds(renderer.ctx, C.byref(w.clouds), layer_arrays[i], len(w.layers), wptr[i]))
            dp = None if depths[i] is None else C.c_void_p(depths[i].data_ptr())
            check(lib.ps_frame(renderer.ctx, abi.PS_LUT_ALL, dp, C.byref(outs[i])))
    renderer.synchronize()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps
