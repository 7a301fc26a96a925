"""Inverse calibration of the atmosphere coefficients from a sky + aerial-
perspective image, on the GPU.

The image, its L2 loss against a target and the loss gradient with respect
to the 14 coefficients (Rayleigh, Mie scattering, Mie absorption and ozone
per channel; Rayleigh and Mie scale heights) of many parameter sets come
from one ``ps_atmosphere_grad`` call (tangent LUT chain + image pass, see
include/pedalsky_hip.h). The optimiser works on six log-multipliers per set
(one per coefficient group) with torch's Adam; torch holds the device
buffers and the optimiser state, the library does the rendering and
differentiation.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import abi, scene

# coefficient groups of the 14-vector: rayleigh, mie scattering, mie absorption,
# ozone (3 channels each), rayleigh and mie scale heights
GROUPS = ((0, 1, 2), (3, 4, 5), (6, 7, 8), (9, 10, 11), (12,), (13,))
GROUP_NAMES = ("rayleigh_scattering", "mie_scattering", "mie_absorption", "ozone_absorption",
               "rayleigh_scale_height_m", "mie_scale_height_m")


def coeffs_of(atm: abi.AtmosphereParams) -> np.ndarray:
    return np.array(list(atm.rayleigh_scattering[:3]) + list(atm.mie_scattering[:3]) +
                    list(atm.mie_absorption[:3]) + list(atm.ozone_absorption[:3]) +
                    [atm.rayleigh_scale_height_m, atm.mie_scale_height_m], dtype=np.float64)


def expand_multipliers(mult: np.ndarray) -> np.ndarray:
    """(n, 6) group multipliers -> (n, 14) per-coefficient multipliers."""
    mult = np.asarray(mult, dtype=np.float64)
    out = np.empty(mult.shape[:-1] + (14,))
    for g, idx in enumerate(GROUPS):
        for i in idx:
            out[..., i] = mult[..., g]
    return out


def atmosphere_with(base: abi.AtmosphereParams, coeffs) -> abi.AtmosphereParams:
    a = abi.AtmosphereParams.from_buffer_copy(base)
    k = [float(v) for v in coeffs]
    a.rayleigh_scattering[:3] = k[0:3]
    a.mie_scattering[:3] = k[3:6]
    a.mie_absorption[:3] = k[6:9]
    a.ozone_absorption[:3] = k[9:12]
    a.rayleigh_scale_height_m = k[12]
    a.mie_scale_height_m = k[13]
    return a


@dataclasses.dataclass
class CalibrationSetup:
    workload: scene.Workload
    base: np.ndarray          # 14 Earth coefficients
    hidden: np.ndarray        # 6 hidden multipliers
    init: np.ndarray          # (n, 6) starting multipliers
    ground_radiance: float


def calibration_setup(n_sets: int = 256, width: int = 1280, height: int = 720, seed: int = 0xC5,
                      hidden_seed: int = 0xC5A, ground_radiance: float = 1000.0) -> CalibrationSetup:
    """The atmosphere calibration workload: the full-frame camera (sweep
    frame 0, sun zenith 65 deg, camera at 500 m) over the planet ground, no
    cloud layers; `n_sets` starting points with every group multiplier drawn
    U[0.5, 1.5] (PCG32 `seed`) and one hidden set drawn the same way (PCG32
    `hidden_seed`)."""
    w = scene.workload_full_frame(0, width=width, height=height)
    w.name = f"atmosphere calibration {width}x{height}, sky + AP over the planet ground"
    w.layers = []
    w.clouds = scene.cloud_params(n_layers=0)
    w.winds = np.zeros((0, 2), np.float32)
    w.overcast = np.zeros(128 * 128, np.uint8)
    init = 0.5 + scene.Pcg32(seed).uniform(6 * n_sets).reshape(n_sets, 6)
    hidden = 0.5 + scene.Pcg32(hidden_seed).uniform(6)
    return CalibrationSetup(w, coeffs_of(w.atmosphere), hidden, init, ground_radiance)


class SkyApProblem:
    """Batched sky + AP image gradients on one renderer context."""

    def __init__(self, renderer, workload, ground_radiance: float = 1000.0):
        import torch
        self.r = renderer
        self.w = workload
        self.ground = float(ground_radiance)
        r = renderer
        r.set_lut_dims(workload.lut_dims)
        r.set_atmosphere(workload.atmosphere)
        r.set_frame(workload.frame)
        r.set_depth(workload.depth)
        r.workload = workload
        self.W, self.H = workload.width, workload.height
        self.dev = f"cuda:{r.device}"
        self.torch = torch

    def image(self, atm=None):
        """The image for `atm` (default: the workload's atmosphere)."""
        torch = self.torch
        self.r.set_atmosphere(atm if atm is not None else self.w.atmosphere)
        out = torch.empty((self.H, self.W, 4), dtype=torch.float32, device=self.dev)
        torch.cuda.synchronize(self.r.device)
        abi.check(self.r.lib.ps_sky_ap_image(self.r.ctx, self.r.depth_ptr, self.ground, out.data_ptr()))
        self.r.synchronize()
        self.r.set_atmosphere(self.w.atmosphere)
        return out

    def grad(self, coeffs, target, out=None, image=None, sync=True, weight=None):
        """coeffs: (n, 14) float32 CUDA tensor; returns (n, 16) float32 CUDA:
        14 gradients, the loss, padding. `weight`: optional (H, W) float32
        CUDA per-pixel loss weights."""
        torch = self.torch
        n = coeffs.shape[0]
        if out is None:
            out = torch.empty((n, 16), dtype=torch.float32, device=self.dev)
        abi.check(self.r.lib.ps_atmosphere_grad(self.r.ctx, n, coeffs.data_ptr(), self.r.depth_ptr, self.ground,
                                                target.data_ptr(), None if weight is None else weight.data_ptr(),
                                                out.data_ptr(), None if image is None else image.data_ptr()))
        if sync:
            self.r.synchronize()
        return out


def relative_weights(target):
    """Per-pixel weights that make the L2 loss a relative one:
    1 / (mean RGB of the target)^2, the mean floored at 1e-3 of the frame
    mean. Without them the few brightest pixels (around the sun) dominate
    the loss and the parameters they are insensitive to drift."""
    m = target[..., :3].mean(dim=-1)
    return (1.0 / m.clamp_min(1e-3 * float(m.mean())) ** 2).contiguous()


def calibrate(problem: SkyApProblem, target, init_mult: np.ndarray, iters: int = 200, lr: float = 1e-2,
              weight=None):
    """Adam on the log group multipliers of every set toward `target`
    (optionally per-pixel weighted L2). Returns (final multipliers (n, 6),
    loss history (iters, n))."""
    torch = problem.torch
    dev = problem.dev
    base = torch.tensor(coeffs_of(problem.w.atmosphere), dtype=torch.float32, device=dev)
    gmap = torch.tensor([g for g, idx in enumerate(GROUPS) for _ in idx], device=dev)
    s = torch.tensor(np.log(init_mult), dtype=torch.float32, device=dev, requires_grad=True)
    opt = torch.optim.Adam([s], lr=lr)
    n = s.shape[0]
    out = torch.empty((n, 16), dtype=torch.float32, device=dev)
    # the library runs on torch's stream for the duration: no host sync per step
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    problem.r.set_stream(stream.cuda_stream)
    hist = []
    try:
        with torch.cuda.stream(stream):
            for _ in range(iters):
                with torch.no_grad():
                    mult = torch.exp(s)[:, gmap]            # (n, 14)
                    coeffs = (base * mult).contiguous()
                problem.grad(coeffs, target, out=out, sync=False, weight=weight)
                with torch.no_grad():
                    g14 = out[:, :14] * coeffs              # d loss / d log coeff
                    g6 = torch.zeros_like(s).index_add_(1, gmap, g14)
                    s.grad = g6
                    hist.append(out[:, 14].clone())
                opt.step()
        torch.cuda.synchronize()
    finally:
        problem.r.set_stream(None)
    return torch.exp(s).detach().cpu().numpy(), torch.stack(hist).cpu().numpy()
