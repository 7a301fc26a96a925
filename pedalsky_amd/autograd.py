"""torch.autograd binding of the differentiable cloud render.

``CloudRender.apply`` runs the training forward (``ps_render_train``) of a
:class:`~pedalsky_amd.renderer.Renderer` whose inputs are already loaded
(noise, weather, LUTs, frame) and returns the composite, cloud luminance and
cloud transmittance images; its backward calls ``ps_backward`` and returns
the gradients of the per-layer density_scale, extinction scale and coverage
and of the 128^3 density-perturbation volume. Torch is plumbing here (device
buffers and the autograd graph); forward and adjoint are the gfx950 kernels.
"""
from __future__ import annotations

import torch

from . import abi


class CloudRender(torch.autograd.Function):
    """Inputs: renderer, cloud params (abi.CloudParams), layer structs (list),
    winds, density_scale / extinction_scale / coverage (float32 tensors with
    one entry per layer), density_delta (contiguous float32 128^3 CUDA
    tensor or None). Outputs: three (H, W, 4) float32 CUDA tensors."""

    @staticmethod
    def forward(ctx, renderer, cloud_params, layers, winds, density_scale, extinction_scale, coverage,
                density_delta):
        n = len(layers)
        ds = density_scale.detach().float().cpu().tolist()
        cov = coverage.detach().float().cpu().tolist()
        kap = extinction_scale.detach().float().cpu().tolist()
        if not (len(ds) == len(cov) == len(kap) == n):
            raise ValueError("one density_scale / extinction_scale / coverage entry per layer")
        ls = []
        for L, d, c in zip(layers, ds, cov):
            L2 = type(L).from_buffer_copy(L)
            L2.density_scale = d
            L2.coverage = c
            ls.append(L2)
        renderer.set_clouds(cloud_params, ls, winds)
        delta = None if density_delta is None else density_delta.detach().contiguous()
        renderer.set_cloud_extensions(kap, delta)
        W, H = int(renderer.workload.width), int(renderer.workload.height)
        bufs = renderer.alloc_outputs(W, H, abi.PS_PIXEL_F32,
                                      ("composite", "cloud_luminance", "cloud_transmittance"))
        renderer.render_train(renderer.image_out(bufs, abi.PS_PIXEL_F32))
        ctx.renderer = renderer
        ctx.n = n
        ctx.devices = (density_scale.device, extinction_scale.device, coverage.device)
        ctx.dtypes = (density_scale.dtype, extinction_scale.dtype, coverage.dtype)
        ctx.has_delta = density_delta is not None
        ctx.delta_shape = None if density_delta is None else density_delta.shape
        renderer.synchronize()
        return bufs["composite"], bufs["cloud_luminance"], bufs["cloud_transmittance"]

    @staticmethod
    def backward(ctx, g_comp, g_lum, g_trn):
        r = ctx.renderer
        dev = f"cuda:{r.device}"

        def cot(g):
            return None if g is None else g.detach().float().contiguous()
        layer_grads = torch.zeros(24, dtype=torch.float32, device=dev)
        want_delta = ctx.has_delta and ctx.needs_input_grad[7]
        delta_grad = torch.zeros(128 ** 3, dtype=torch.float32, device=dev) if want_delta else None
        torch.cuda.synchronize(r.device)
        gc, gl, gt = cot(g_comp), cot(g_lum), cot(g_trn)
        r.backward(gc, gl, gt, layer_grads, delta_grad)
        r.synchronize()
        n = ctx.n
        (dd, dk, dc), (td, tk, tc) = ctx.devices, ctx.dtypes
        g_ds = layer_grads[:n].to(device=dd, dtype=td)
        g_k = layer_grads[8:8 + n].to(device=dk, dtype=tk)
        g_cov = layer_grads[16:16 + n].to(device=dc, dtype=tc)
        gd = None if delta_grad is None else delta_grad.view(ctx.delta_shape)
        return None, None, None, None, g_ds, g_k, g_cov, gd


def render(renderer, cloud_params, layers, winds, density_scale, extinction_scale, coverage, density_delta=None):
    """Differentiable render: (composite, cloud_luminance, cloud_transmittance)."""
    return CloudRender.apply(renderer, cloud_params, layers, winds, density_scale, extinction_scale, coverage,
                             density_delta)
