"""Python handle on one library context (``ps_ctx``): every call goes through
the C ABI in ``include/pedalsky_hip.h``. Device buffers are torch tensors
(torch is plumbing for device memory here; all compute is in the HIP
library)."""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import abi
from .abi import check


def _torch():
    import torch
    return torch


class Renderer:
    def __init__(self, device: int = 0):
        self.lib = abi.load()
        self.device = device
        h = C.c_void_p()
        check(self.lib.ps_ctx_create(device, C.byref(h)))
        self.ctx = h
        self.workload = None
        self._depth_dev = None
        self._density_delta = None  # borrowed by the library (set_cloud_extensions)

    # lifecycle ---------------------------------------------------------------
    def close(self) -> None:
        if self.ctx:
            self.lib.ps_ctx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self) -> None:
        check(self.lib.ps_ctx_synchronize(self.ctx))

    def set_stream(self, stream_handle: Optional[int]) -> None:
        check(self.lib.ps_ctx_set_stream(self.ctx, stream_handle))

    # state --------------------------------------------------------------------
    def set_lut_dims(self, dims: abi.LutDims) -> None:
        check(self.lib.ps_ctx_set_lut_dims(self.ctx, C.byref(dims)))

    def lut_dims(self) -> abi.LutDims:
        d = abi.LutDims()
        check(self.lib.ps_ctx_get_lut_dims(self.ctx, C.byref(d)))
        return d

    def set_atmosphere(self, a: abi.AtmosphereParams) -> None:
        check(self.lib.ps_set_atmosphere(self.ctx, C.byref(a)))

    def set_frame(self, f: abi.FrameUniforms) -> None:
        check(self.lib.ps_set_frame(self.ctx, C.byref(f)))

    def set_clouds(self, params: abi.CloudParams, layers, winds=None) -> None:
        n = len(layers)
        arr = (abi.CloudLayer * max(n, 1))(*layers)
        w = None
        if winds is not None:
            wa = np.ascontiguousarray(winds, dtype=np.float32).reshape(-1)
            w = (C.c_float * len(wa))(*wa.tolist())
        check(self.lib.ps_set_clouds(self.ctx, C.byref(params), arr, n, w))

    def set_march_steps(self, cloud_steps: int, light_steps: int) -> None:
        """CloudsSubsystem::set_march_steps: clamped to [8, 256] / [1, 16]."""
        check(self.lib.ps_set_march_steps(self.ctx, int(cloud_steps), int(light_steps)))

    # inputs -------------------------------------------------------------------
    def generate_noise(self, seed: int = 0, fmt: int = abi.PS_NOISE_FP16) -> None:
        check(self.lib.ps_generate_noise(self.ctx, seed, fmt))

    def upload_noise(self, fmt: int, base: np.ndarray, detail: np.ndarray, curl: np.ndarray) -> None:
        base, detail, curl = (np.ascontiguousarray(x) for x in (base, detail, curl))
        check(self.lib.ps_upload_noise(self.ctx, fmt, base.ctypes.data, detail.ctypes.data, curl.ctypes.data))

    def read_noise(self, which: int, fmt: int) -> np.ndarray:
        dt = np.uint8 if fmt == abi.PS_NOISE_UNORM8 else np.uint16
        shape = {abi.PS_NOISE_BASE: (128, 128, 128, 4), abi.PS_NOISE_DETAIL: (32, 32, 32, 4),
                 abi.PS_NOISE_CURL: (128, 128, 2)}[which]
        out = np.zeros(shape, dtype=dt)
        check(self.lib.ps_read_noise(self.ctx, which, out.ctypes.data, out.nbytes))
        return out

    def upload_blue_noise(self, ranks: np.ndarray) -> None:
        r = np.ascontiguousarray(ranks, dtype=np.uint8)
        check(self.lib.ps_upload_blue_noise(self.ctx, r.ctypes.data))

    def synthesise_blue_noise(self) -> None:
        """Generate the 64x64 blue-noise ranks on the GPU into the context."""
        check(self.lib.ps_synthesise_blue_noise(self.ctx))

    def read_blue_noise(self) -> np.ndarray:
        """The context's (64, 64) blue-noise ranks."""
        out = np.zeros((64, 64), np.uint8)
        check(self.lib.ps_read_blue_noise(self.ctx, out.ctypes.data))
        return out

    def upload_weather(self, weather: np.ndarray, type_grid: Optional[np.ndarray] = None,
                       overcast: Optional[np.ndarray] = None) -> None:
        w = np.ascontiguousarray(weather, dtype=np.uint16)
        tg = None if type_grid is None else np.ascontiguousarray(type_grid, dtype=np.uint8)
        oc = None if overcast is None else np.ascontiguousarray(overcast, dtype=np.uint8)
        check(self.lib.ps_upload_weather(self.ctx, w.ctypes.data, w.shape[1], w.shape[0], w.shape[2],
                                         None if tg is None else tg.ctypes.data,
                                         None if oc is None else oc.ctypes.data))

    def synthesise_weather(self, has_layers: bool = True, precip_mm_per_h: float = 0.0,
                           ground_wetness: float = 0.0, coverage_grid: Optional[np.ndarray] = None,
                           type_grid: Optional[np.ndarray] = None, grid_extent_m: float = 32000.0,
                           overcast: Optional[np.ndarray] = None) -> None:
        """The reference's weather map and cloud-type grid, synthesised on the
        GPU (ps_synthesise_weather); grids are (grid_h, grid_w) arrays."""
        shape = coverage_grid.shape if coverage_grid is not None else (
            type_grid.shape if type_grid is not None else (0, 0))
        g = None if coverage_grid is None else np.ascontiguousarray(coverage_grid, dtype=np.float32)
        t = None if type_grid is None else np.ascontiguousarray(type_grid, dtype=np.uint8)
        if g is not None and t is not None and g.shape != t.shape:
            raise ValueError("coverage and type grids must have the same size")
        oc = None if overcast is None else np.ascontiguousarray(overcast, dtype=np.uint8)
        ws = abi.WeatherSynth(int(bool(has_layers)), float(precip_mm_per_h), float(ground_wetness),
                              None if g is None else g.ctypes.data_as(C.POINTER(C.c_float)),
                              None if t is None else t.ctypes.data_as(C.POINTER(C.c_uint8)),
                              int(shape[1]), int(shape[0]), float(grid_extent_m))
        check(self.lib.ps_synthesise_weather(self.ctx, C.byref(ws), None if oc is None else oc.ctypes.data))

    def read_weather(self):
        """(128, 128, 4) uint16 binary16 bits of the synthesised map and the
        (128, 128) cloud-type grid."""
        rgba = np.zeros((128, 128, 4), np.uint16)
        types = np.zeros((128, 128), np.uint8)
        check(self.lib.ps_read_weather(self.ctx, rgba.ctypes.data, types.ctypes.data))
        return rgba, types

    def set_depth(self, depth: Optional[np.ndarray]) -> None:
        # work queued on the library's stream may still read the current
        # buffer: wait for it before torch's allocator may reuse the block
        if getattr(self, "_depth_dev", None) is not None:
            self.synchronize()
        if depth is None:
            self._depth_dev = None
        else:
            torch = _torch()
            self._depth_dev = torch.from_numpy(np.ascontiguousarray(depth, dtype=np.float32)).to(
                f"cuda:{self.device}")
            torch.cuda.synchronize(self.device)

    @property
    def depth_ptr(self):
        return None if self._depth_dev is None else self._depth_dev.data_ptr()

    def load_workload(self, w) -> None:
        """Upload every input of a scene.Workload (noise baked on the GPU)."""
        self.set_lut_dims(w.lut_dims)
        self.set_atmosphere(w.atmosphere)
        self.set_frame(w.frame)
        self.set_clouds(w.clouds, w.layers, w.winds)
        self.generate_noise(w.noise_seed, w.noise_format)
        self.upload_blue_noise(w.blue)
        self.upload_weather(w.weather, w.type_grid, w.overcast)
        self.set_depth(w.depth)
        self.workload = w

    # passes -------------------------------------------------------------------
    def bake_luts(self, flags: int = abi.PS_LUT_ALL) -> None:
        check(self.lib.ps_bake_luts(self.ctx, flags))

    def read_lut(self, which: int) -> np.ndarray:
        d = self.lut_dims()
        shape = {abi.PS_LUT_TRANSMITTANCE: (d.transmittance_h, d.transmittance_w, 4),
                 abi.PS_LUT_MULTISCATTER: (d.multiscatter_h, d.multiscatter_w, 4),
                 abi.PS_LUT_SKYVIEW: (d.skyview_h, d.skyview_w, 4),
                 abi.PS_LUT_AERIAL_PERSPECTIVE: (d.ap_d, d.ap_h, d.ap_w, 4)}[which]
        out = np.zeros(shape, dtype=np.float32)
        check(self.lib.ps_read_lut(self.ctx, which, out.ctypes.data, out.size))
        return out

    def alloc_outputs(self, width: int, height: int, fmt: int = abi.PS_PIXEL_F32, which=("composite",)):
        torch = _torch()
        dt = torch.float32 if fmt == abi.PS_PIXEL_F32 else torch.float16
        bufs = {k: torch.empty((height, width, 4), dtype=dt, device=f"cuda:{self.device}") for k in which}
        torch.cuda.synchronize(self.device)
        return bufs

    def image_out(self, bufs, fmt: int = abi.PS_PIXEL_F32) -> abi.ImageOut:
        o = abi.ImageOut()
        o.composite = bufs["composite"].data_ptr() if "composite" in bufs else None
        o.cloud_luminance = bufs["cloud_luminance"].data_ptr() if "cloud_luminance" in bufs else None
        o.cloud_transmittance = bufs["cloud_transmittance"].data_ptr() if "cloud_transmittance" in bufs else None
        o.format = fmt
        return o

    def render(self, out: abi.ImageOut) -> None:
        check(self.lib.ps_render(self.ctx, self.depth_ptr, C.byref(out)))

    def frame(self, out: abi.ImageOut, lut_flags: int = abi.PS_LUT_ALL) -> None:
        check(self.lib.ps_frame(self.ctx, lut_flags, self.depth_ptr, C.byref(out)))

    def time_frames(self, frames: int, out: abi.ImageOut, lut_flags: int = abi.PS_LUT_ALL) -> abi.FrameTiming:
        t = abi.FrameTiming()
        check(self.lib.ps_time_frames(self.ctx, frames, lut_flags, self.depth_ptr, C.byref(out), C.byref(t)))
        return t

    def timing_begin(self, max_frames: int) -> None:
        check(self.lib.ps_timing_begin(self.ctx, max_frames))

    def timing_end(self) -> abi.FrameTiming:
        t = abi.FrameTiming()
        check(self.lib.ps_timing_end(self.ctx, C.byref(t)))
        return t

    def count_work(self) -> abi.WorkCounters:
        c = abi.WorkCounters()
        check(self.lib.ps_count_work(self.ctx, self.depth_ptr, C.byref(c)))
        return c

    # differentiation ----------------------------------------------------------
    def set_cloud_extensions(self, extinction_scale=None, density_delta=None) -> None:
        """Per-layer extinction scales (host sequence, None = all 1) and the
        128^3 density-perturbation volume (a contiguous float32 device
        tensor the library reads on every render, or None). The library
        borrows the tensor's memory until the next call: the renderer keeps
        a reference, so a temporary stays alive (and its memory is not handed
        to another tensor) while renders read it."""
        k = None
        n = 0
        if extinction_scale is not None:
            vals = [float(v) for v in extinction_scale]
            n = len(vals)
            k = (C.c_float * max(n, 1))(*vals)
        ptr = None
        if density_delta is not None:
            if density_delta.dtype != _torch().float32 or not density_delta.is_contiguous() or \
                    density_delta.numel() != 128 ** 3 or not density_delta.is_cuda:
                raise ValueError("density_delta must be a contiguous float32 CUDA tensor of 128^3 elements")
            ptr = density_delta.data_ptr()
        check(self.lib.ps_set_cloud_extensions(self.ctx, k, n, ptr))
        self._density_delta = density_delta

    def set_half_res(self, enabled: bool) -> None:
        """Half-resolution cloud march with a Catmull-Rom upsample in the
        full-resolution composite (the reference's clouds.half_res_render)."""
        check(self.lib.ps_set_half_res(self.ctx, 1 if enabled else 0))

    def set_taa(self, enabled: bool) -> None:
        """Cloud temporal anti-aliasing (the reference's clouds.temporal_taa;
        pass temporal_taa and not freeze_time)."""
        check(self.lib.ps_set_taa(self.ctx, 1 if enabled else 0))

    def invalidate_taa_history(self) -> None:
        check(self.lib.ps_invalidate_taa_history(self.ctx))

    def set_active_view(self, view: int) -> None:
        """The view (0 or 1) whose TAA history the next renders use."""
        check(self.lib.ps_set_active_view(self.ctx, view))

    def set_checkpoint_interval(self, steps: int) -> None:
        check(self.lib.ps_set_checkpoint_interval(self.ctx, steps))

    def render_train(self, out: abi.ImageOut) -> None:
        check(self.lib.ps_render_train(self.ctx, self.depth_ptr, C.byref(out)))

    def backward(self, g_composite=None, g_luminance=None, g_transmittance=None, layer_grads=None,
                 delta_grad=None) -> None:
        """Accumulate gradients of the last render_train into the device
        buffers `layer_grads` (24 floats) and `delta_grad` (128^3)."""
        def ptr(t):
            return None if t is None else t.data_ptr()
        g = abi.CloudGrads(ptr(layer_grads), ptr(delta_grad))
        check(self.lib.ps_backward(self.ctx, ptr(g_composite), ptr(g_luminance), ptr(g_transmittance),
                                   C.byref(g)))

    def render_numpy(self, which=("composite", "cloud_luminance", "cloud_transmittance"),
                     fmt: int = abi.PS_PIXEL_F32, bake: bool = True) -> dict:
        """Bake the LUTs, render, and return the outputs as numpy arrays."""
        W, H = int(self.workload.width), int(self.workload.height)
        bufs = self.alloc_outputs(W, H, fmt, which)
        out = self.image_out(bufs, fmt)
        if bake:
            self.bake_luts()
        self.render(out)
        self.synchronize()
        return {k: v.float().cpu().numpy() for k, v in bufs.items()}
