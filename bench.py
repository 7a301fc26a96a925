"""Benchmark of the sky + cloud path on one MI355X.

One step = one full frame at 1280x720: the four LUT bakes (transmittance,
multi-scattering, sky-view, 32^3 aerial perspective), then the fused sky
pass + two-layer cloud march (cumulus 1500-2300 m, cirrus 8000-9000 m;
128 primary / 6 light steps; FP16 noise) + AP + composite, writing the
composited radiance and the cloud transmittance. Consecutive steps walk the
seeded 16-camera sweep (pitch U[-10, 30] deg, yaw U[0, 360) deg). The depth
buffer is all far (no ground: rays below the horizon march the layer shells
on the far side of the planet); --depth ground times the same frames with
the planet-ground depth buffer instead. Inputs are resident in HBM before
the timed region.

Prints one JSON line (see the round contract): value = Mpix/s, roofline of
the dominant kernel (the render), and a CPU baseline (the oracle's float
instantiation of the reference shaders on the host cores, a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(w, base, detail, curl, budget_s: float, threads: int) -> dict:
    import numpy as np  # noqa: F401
    import oracle
    frames = 0
    t0 = time.perf_counter()
    while True:
        oracle.render_f32_frame(w, base, detail, curl, nthreads=threads)
        frames += 1
        el = time.perf_counter() - t0
        if el >= budget_s or frames >= 64:
            break
    mpix = frames * w.width * w.height / el / 1e6
    return {"value": round(mpix, 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
            "sample": f"{frames} whole {w.width}x{w.height} frame(s) of sweep camera 0 (LUT chain + sky + "
                      f"2-layer march + AP + composite), float instantiation of the CPU restatement, "
                      f"{el:.1f} s",
            "host": host_cores()}


def host_cores() -> dict:
    """The host's core counts: what the OS reports, what this process may run
    on (its affinity mask) and the OpenMP thread limit of the environment."""
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def baseline_threads() -> int:
    """Every core this process may use, capped by an OpenMP thread limit the
    environment sets (the GPU box fixes its per-GPU CPU share that way)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(n, 1)


def bench_batch(r, n_scenes: int) -> dict:
    """Uneven scenes (PCG32 seed 0xC3: resolutions 256x144 .. 1920x1080, 1-3
    layers of any type, scaled atmospheres) rendered by one autotuned,
    graph-captured batch launch sequence, and the same scenes through the
    plain per-scene path (uniform uploads + LUT chain + render per scene)."""
    from pedalsky_amd import cost_model, scene
    from pedalsky_amd.batch import Batch, time_sequential
    log(f"batch: generating {n_scenes} scenes")
    ws = scene.workload_batch(n=n_scenes, seed=0xC3)
    b = Batch(r, ws)
    mpix = b.pixels / 1e6
    log(f"batch: {mpix:.1f} Mpix; autotuning")
    tuning = b.autotune(reps=2)
    b.capture()
    b.replay()
    r.synchronize()
    ms = b.time(5)
    log(f"batch: {ms:.2f} ms per graph replay; timing the per-scene path")
    seq_ms = time_sequential(r, ws, b.bufs, b.depths, reps=2)
    flops = batch_flops(r, ws, b.depths)
    # the step-schedule autotuner (the reference's adaptive-quality loop over
    # the batch): the schedule that fits half the full-quality replay time
    log("batch: step schedule for half the replay time")
    sched = b.autotune_steps(0.5 * ms, reps=2)
    b.close()
    tflops = flops / (ms * 1e-3) / 1e12
    return {"scenes": n_scenes, "megapixels": round(mpix, 3), "tuning": tuning.as_dict(),
            "ms_per_batch": round(ms, 3), "value": round(mpix / (ms * 1e-3), 3), "unit": "Mpix/s",
            "sequential_ms": round(seq_ms, 3), "sequential_mpix_s": round(mpix / (seq_ms * 1e-3), 3),
            "speedup_vs_sequential": round(seq_ms / ms, 3), "graph": True,
            "roofline": {"bound": "valu", "achieved": round(tflops, 3), "peak": cost_model.PEAK_FP32_VALU_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(tflops / cost_model.PEAK_FP32_VALU_TFLOPS, 4),
                         "peak_unpacked": cost_model.PEAK_FP32_VALU_UNPACKED_TFLOPS,
                         "frac_unpacked": round(tflops / cost_model.PEAK_FP32_VALU_UNPACKED_TFLOPS, 4),
                         "flop_per_batch": flops,
                         "note": "algorithmic FLOP of every scene's LUT chain and render (event counts x the "
                                 "cost table) over the graph replay time"},
            "step_schedule": {"budget_ms": round(0.5 * ms, 3), **sched.as_dict(),
                              "mpix_s": round(mpix / (sched.batch_ms * 1e-3), 3),
                              "note": "ps_batch_autotune_steps: primary / light steps per resolution class within "
                                      "set_march_steps' clamps for half the full-quality replay time; the headline "
                                      "batched value above keeps every scene's own 128/6 steps"},
            "note": "every scene bakes its own LUT set; noise and weather map shared"}


def batch_flops(r, ws, depths) -> float:
    """Algorithmic FLOP of one batch: per scene its LUT chain and the event
    counts of its render (ps_count_work, outside any timed region) x the cost
    table."""
    import ctypes as C
    import numpy as np
    from pedalsky_amd import abi, cost_model
    lib = r.lib
    total = 0.0
    for w, d in zip(ws, depths):
        layers = (abi.CloudLayer * max(len(w.layers), 1))(*w.layers)
        winds = np.ascontiguousarray(w.winds, dtype=np.float32).reshape(-1)
        wp = (C.c_float * len(winds))(*winds.tolist()) if len(winds) else None
        abi.check(lib.ps_set_atmosphere(r.ctx, C.byref(w.atmosphere)))
        abi.check(lib.ps_set_frame(r.ctx, C.byref(w.frame)))
        abi.check(lib.ps_set_clouds(r.ctx, C.byref(w.clouds), layers, len(w.layers), wp))
        abi.check(lib.ps_bake_luts(r.ctx, abi.PS_LUT_ALL))
        cw = abi.WorkCounters()
        abi.check(lib.ps_count_work(r.ctx, None if d is None else C.c_void_p(d.data_ptr()), C.byref(cw)))
        total += cost_model.render_flops(cw) + cost_model.lut_flops(r.lut_dims())
    r.synchronize()
    return total


def bench_adjoint(r, W: int, H: int, steps: int, warmup: int) -> dict:
    """Forward + adjoint of the two-layer march: per step the LUT chain, the
    training forward (transmittance checkpoints every 16 steps), the L2
    cotangent against a target rendered with perturbed density_scale,
    extinction scale, coverage and perturbation volume, and the backward
    (per-layer gradients and the 128^3 perturbation-volume gradient).
    Timed beside the plain forward frame at the same size."""
    import ctypes as C
    import numpy as np
    import torch
    from pedalsky_amd import abi, scene
    w = scene.workload_adjoint(0, width=W, height=H)
    r.load_workload(w)
    tgt = scene.adjoint_target(w.layers)
    dev = "cuda:0"

    def set_params(ds, cov, kap, delta):
        ls = []
        for L, d, c in zip(w.layers, ds, cov):
            L2 = abi.CloudLayer.from_buffer_copy(L)
            L2.density_scale, L2.coverage = float(d), float(c)
            ls.append(L2)
        r.set_clouds(w.clouds, ls, w.winds)
        r.set_cloud_extensions([float(k) for k in kap], delta)

    d_tgt = torch.from_numpy(tgt.delta.reshape(-1)).to(dev)
    set_params(tgt.density_scale, tgt.coverage, tgt.extinction_scale, d_tgt)
    bufs = r.alloc_outputs(W, H, abi.PS_PIXEL_F32, ("composite",))
    out = r.image_out(bufs, abi.PS_PIXEL_F32)
    r.frame(out)
    r.synchronize()
    target = bufs["composite"].clone()
    d0 = torch.zeros(128 ** 3, dtype=torch.float32, device=dev)
    set_params([L.density_scale for L in w.layers], [L.coverage for L in w.layers], [1.0] * len(w.layers), d0)
    lg = torch.zeros(24, dtype=torch.float32, device=dev)
    dg = torch.zeros(128 ** 3, dtype=torch.float32, device=dev)
    g = torch.zeros_like(target)
    # dL/d composite of sum((composite - target)^2) = 2 composite - 2 target, one kernel per step
    neg2_target = target * -2.0
    lib = r.lib
    grads = abi.CloudGrads(lg.data_ptr(), dg.data_ptr())
    depth = C.c_void_p(r.depth_ptr)
    # one stream for the library and torch's cotangent kernels: no host
    # synchronisation inside a step
    r.synchronize()
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()  # default priority
    r.set_stream(stream.cuda_stream)

    def fwd_adj():
        abi.check(lib.ps_frame_train(r.ctx, abi.PS_LUT_ALL, depth, C.byref(out)))
        torch.add(neg2_target, bufs["composite"], alpha=2.0, out=g)  # the backward reads RGB only
        abi.check(lib.ps_backward(r.ctx, C.c_void_p(g.data_ptr()), None, None, C.byref(grads)))

    grads_layers = abi.CloudGrads(lg.data_ptr(), None)

    def fwd_adj_layers():
        abi.check(lib.ps_frame_train(r.ctx, abi.PS_LUT_ALL, depth, C.byref(out)))
        torch.add(neg2_target, bufs["composite"], alpha=2.0, out=g)  # the backward reads RGB only
        abi.check(lib.ps_backward(r.ctx, C.c_void_p(g.data_ptr()), None, None, C.byref(grads_layers)))

    def fwd():
        abi.check(lib.ps_frame(r.ctx, abi.PS_LUT_ALL, depth, C.byref(out)))

    res = {}
    for name, fn, ext in (("fwd", fwd, False), ("fwd_delta", fwd, True), ("fwd_adjoint", fwd_adj, True),
                          ("fwd_adjoint_layers", fwd_adj_layers, True)):
        if ext:
            r.set_cloud_extensions([1.0] * len(w.layers), d0)
        else:
            r.set_cloud_extensions(None, None)
        with torch.cuda.stream(stream):
            for _ in range(warmup):
                fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                fn()
            torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / steps * 1e3
    r.set_stream(None)
    r.set_cloud_extensions(None, None)
    mp = W * H / 1e6
    return {"workload": f"{W}x{H} cumulus + cirrus, 128/6 steps, perturbation volume + extinction scales, "
                        f"L2 loss vs a perturbed target, checkpoints every 16 steps",
            "steps": steps, "ms_fwd": round(res["fwd"], 4), "ms_fwd_delta": round(res["fwd_delta"], 4),
            "ms_fwd_adjoint": round(res["fwd_adjoint"], 4),
            "ms_fwd_adjoint_layer_grads_only": round(res["fwd_adjoint_layers"], 4),
            "value": round(mp / (res["fwd_adjoint"] * 1e-3), 3), "unit": "Mpix/s",
            "fwd_mpix_s": round(mp / (res["fwd"] * 1e-3), 3),
            "fwd_adjoint_over_fwd": round(res["fwd_adjoint"] / res["fwd"], 3),
            "note": "fwd = plain frame (LUT chain + render); fwd_delta = the same with the perturbation "
                    "volume and extinction scales; fwd_adjoint = LUT chain + training forward (checkpoints) + "
                    "L2 cotangent + ps_backward (layer and volume gradients), all on one stream"}


def bench_present(r, W: int, H: int, steps: int) -> dict:
    """Presentation of a rendered 1280x720 frame (ps_present): the ground pass
    on the flat startup ground, AP and cloud composite, godrays (64 samples,
    sun on screen), the 3-level bloom pyramid, ACES tone map to sRGB8 and the
    HDR copy. Timed on its own (the frame's render is not in the timed
    region), with HIP events around the launches."""
    import ctypes as C
    import torch
    from pedalsky_amd import abi, scene
    w = scene.workload_small(width=W, height=H, sun_zenith=75.0, pitch=10.0)
    w.depth = scene.flat_ground_depth(w.frame)
    r.load_workload(w)
    bufs = r.alloc_outputs(W, H, abi.PS_PIXEL_F32, ("composite", "cloud_luminance", "cloud_transmittance"))
    out = r.image_out(bufs)
    r.frame(out)
    r.synchronize()
    ldr = torch.zeros((H, W, 4), dtype=torch.uint8, device="cuda:0")
    hdr = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda:0")
    pp = abi.default_present_params()
    pp.ground_model = abi.PS_GROUND_PBR
    stream = torch.cuda.Stream()  # default priority
    r.set_stream(stream.cuda_stream)
    res = {"workload": f"{W}x{H} rendered frame (one cumulus layer, sun 15 deg up and on screen), ground pass on "
                       f"the flat ground, godrays 64 samples, bloom, ACES + sRGB8 and the HDR copy"}
    for name, rays, bloom in (("ms_ground_ap_tonemap", 0, 0), ("ms_full_chain", 1, 1)):
        pp.godrays, pp.bloom = rays, bloom

        def call():
            abi.check(r.lib.ps_present(r.ctx, r.depth_ptr, None, C.byref(out), C.byref(pp), ldr.data_ptr(),
                                       hdr.data_ptr()))
        with torch.cuda.stream(stream):
            for _ in range(5):
                call()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(steps):
                call()
            e1.record(stream)
            e1.synchronize()
        res[name] = round(e0.elapsed_time(e1) / steps, 4)
    r.set_stream(None)
    return res


def bench_calibration(r, n_sets: int, iters: int) -> dict:
    """Atmosphere inverse calibration: `n_sets` seeded starting points
    (group multipliers U[0.5, 1.5], PCG32 seed 0xC5) fitted by Adam (lr 1e-2
    on the log multipliers) to the 1280x720 sky + AP image of a hidden set
    (seed 0xC5A) with the relative L2 loss; every iteration bakes all sets'
    LUT chains with tangents and evaluates all sets' images and gradients in
    one call."""
    import numpy as np
    import torch
    from pedalsky_amd import calibration as cal
    su = cal.calibration_setup(n_sets=n_sets)
    prob = cal.SkyApProblem(r, su.workload, su.ground_radiance)
    hidden = cal.atmosphere_with(su.workload.atmosphere, su.base * cal.expand_multipliers(su.hidden))
    target = prob.image(hidden)
    w = cal.relative_weights(target)
    cal.calibrate(prob, target, su.init[:2], iters=2, weight=w)  # warm-up (allocations)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    final, hist = cal.calibrate(prob, target, su.init, iters=iters, weight=w)
    dt = time.perf_counter() - t0
    err = np.abs(final / su.hidden[None, :] - 1.0)
    drop = hist[-1] / hist[0]
    mpix = n_sets * su.workload.width * su.workload.height * iters / 1e6
    return {"workload": f"{n_sets} parameter sets x {iters} Adam iterations, 1280x720 sky + AP image, "
                        f"relative L2 vs a hidden set; per iteration: LUT chains + tangent LUTs + image gradients "
                        f"of all sets",
            "seconds": round(dt, 3), "ms_per_iteration": round(dt / iters * 1e3, 3),
            "value": round(mpix / dt, 3), "unit": "Mpix/s (image gradient evaluations)",
            "loss_ratio_median": float(np.median(drop)),
            "recovery_error_median": float(np.median(err)), "recovery_error_p90": float(np.percentile(err, 90)),
            "recovery_error_median_per_group": {k: float(v) for k, v in zip(cal.GROUP_NAMES, np.median(err, axis=0))}}


def measured_traffic():
    """HBM bytes per k_render launch from the committed counter summary
    (FETCH_SIZE x its copy-calibrated factor + WRITE_SIZE, rocprofv3 --pmc
    runs of this bench); None when the summary is absent."""
    p = Path(__file__).resolve().parent / "profiles" / "k_render_fullframe_1280x720_counters.md"
    try:
        m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", p.read_text(), re.M)
    except OSError:
        return None
    return int(m.group(1)) if m else None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched measurement")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward + adjoint measurement")
    ap.add_argument("--adjoint-steps", type=int, default=16)
    ap.add_argument("--no-calibration", action="store_true", help="skip the atmosphere calibration run")
    ap.add_argument("--no-half-res", action="store_true", help="skip the half-resolution cloud path measurement")
    ap.add_argument("--no-taa", action="store_true", help="skip the cloud TAA measurement")
    ap.add_argument("--no-present", action="store_true", help="skip the presentation-chain measurement")
    ap.add_argument("--calibration-sets", type=int, default=256)
    ap.add_argument("--calibration-iters", type=int, default=200)
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--depth", choices=("none", "ground"), default="none",
                    help="depth buffer of the headline frames: none (all far, the benchmark configuration) or "
                         "the planet ground")
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=720)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 256 MiB device-to-device copy (a known-size reference for the "
                         "FETCH_SIZE / WRITE_SIZE hardware counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this path is benchmarked on a single MI355X "
              "(no multi-GPU run is part of this build); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from pedalsky_amd import abi, cost_model, scene
    from pedalsky_amd.renderer import Renderer

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    W, H = args.width, args.height
    sweep = [scene.workload_full_frame(i, width=W, height=H, ground=True) for i in range(16)]
    w0 = sweep[0]
    r = Renderer(0)
    r.load_workload(w0)
    ground_depths = [torch.from_numpy(w.depth).to("cuda:0") for w in sweep]
    depths = [None] * 16 if args.depth == "none" else ground_depths
    bufs = r.alloc_outputs(W, H, abi.PS_PIXEL_F32, ("composite", "cloud_transmittance"))
    out = r.image_out(bufs, abi.PS_PIXEL_F32)
    lib = r.lib
    import ctypes as C

    def dptr(d):
        return None if d is None else C.c_void_p(d.data_ptr())

    def frame(k: int, ds=None) -> None:
        w = sweep[k % 16]
        d = (depths if ds is None else ds)[k % 16]
        abi.check(lib.ps_set_frame(r.ctx, C.byref(w.frame)))
        abi.check(lib.ps_frame(r.ctx, abi.PS_LUT_ALL, dptr(d), C.byref(out)))

    if args.calibrate_copy:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 268435456 bytes read, 268435456 bytes written")
        del src, dst

    log(f"warmup {args.warmup} frames")
    for k in range(args.warmup):
        frame(k)
    r.synchronize()
    torch.cuda.synchronize()

    log(f"timing {args.steps} frames")
    r.timing_begin(args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        frame(k)
    r.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    timing = r.timing_end()
    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    mpix = args.steps * W * H / elapsed / 1e6

    # algorithmic work over the sweep (counting pass, outside the timed region)
    log("counting algorithmic work over the sweep")
    flops, gbytes, cnt_sum = 0.0, 0.0, {}
    used = [k % 16 for k in range(args.steps)]
    per_cam = {}
    for i in sorted(set(used)):
        abi.check(lib.ps_set_frame(r.ctx, C.byref(sweep[i].frame)))
        abi.check(lib.ps_bake_luts(r.ctx, abi.PS_LUT_ALL))
        cw = abi.WorkCounters()
        abi.check(lib.ps_count_work(r.ctx, dptr(depths[i]), C.byref(cw)))
        per_cam[i] = cw
    for i in used:
        cw = per_cam[i]
        flops += cost_model.render_flops(cw)
        gbytes += cost_model.render_bytes(cw)
        for f, _ in abi.WorkCounters._fields_:
            cnt_sum[f] = cnt_sum.get(f, 0) + getattr(cw, f)
    flops_per_launch = flops / len(used)
    bytes_per_launch = gbytes / len(used)
    render_s = timing.render_ms * 1e-3
    hbm_bytes = cost_model.render_hbm_bytes(W, H, out_bytes_per_px=32, depth=args.depth != "none")
    achieved_tflops = flops_per_launch / render_s / 1e12
    achieved_gbps = bytes_per_launch / render_s / 1e9

    result = {
        "metric": "cloud+sky Mpix/s @1280x720 (fwd, fwd+adjoint, batched) at <=1e-3 rel. radiance",
        "value": round(mpix, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded noise bake, PCG32 weather map, 16-camera seeded sweep)",
        "config": {"workload": f"forward frame {W}x{H}: LUT chain + sky + cumulus/cirrus march "
                               f"128/6 steps + AP 32^3 + composite, FP16 noise, "
                               f"{'depth all far (no ground)' if args.depth == 'none' else 'planet-ground depth'}",
                   "width": W, "height": H, "layers": 2, "primary_steps": 128, "light_steps": 6,
                   "lut": "T 256x64, MS 32x32, SV 192x108, AP 32^3 (FP32)", "variant": "fwd",
                   "depth": "all far" if args.depth == "none" else "planet ground",
                   "outputs": "composite RGBA32F + cloud transmittance RGBA32F"},
        "kernel_ms": {"transmittance": round(timing.transmittance_ms, 5),
                      "multiscatter": round(timing.multiscatter_ms, 5),
                      "skyview_aerial": round(timing.skyview_ms + timing.aerial_perspective_ms, 5),
                      "render": round(timing.render_ms, 5), "frame_events": round(timing.frame_ms, 5),
                      "note": "sky-view and aerial-perspective LUTs are baked by one fused launch; ps_frame bakes "
                              "each frame's LUT chain into a second LUT set on a side stream while the previous "
                              "frame renders, so the LUT times are wall times shared with that render (frame_events "
                              "is the whole sequence's time per frame)"},
        "roofline": {"bound": "valu", "kernel": "k_render", "achieved": round(achieved_tflops, 3),
                     "peak": cost_model.PEAK_FP32_VALU_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved_tflops / cost_model.PEAK_FP32_VALU_TFLOPS, 4),
                     "traffic": measured_traffic(),
                     "peak_unpacked": cost_model.PEAK_FP32_VALU_UNPACKED_TFLOPS,
                     "frac_unpacked": round(achieved_tflops / cost_model.PEAK_FP32_VALU_UNPACKED_TFLOPS, 4),
                     "gather_GBps": round(achieved_gbps, 1),
                     "flop_per_launch": flops_per_launch, "gather_bytes_per_launch": bytes_per_launch,
                     "algorithmic_hbm_bytes_per_launch": hbm_bytes,
                     "hbm_GBps": round(hbm_bytes / render_s / 1e9, 1)},
        "work_per_frame": {k: v / len(used) for k, v in cnt_sum.items()},
    }
    if args.depth == "none":
        log("the same frames with the planet-ground depth buffer")
        for k in range(args.warmup):
            frame(k, ground_depths)
        r.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            frame(k, ground_depths)
        r.synchronize()
        gs = (time.perf_counter() - t0) / args.steps
        result["ground_depth"] = {"ms_per_frame": round(gs * 1e3, 4), "mpix_s": round(W * H / gs / 1e6, 3),
                                  "workload": "the same frames with the depth buffer of an opaque planet "
                                              "ground (clips the far-side layer shells)"}

    def time_mode(half: bool, taa: bool) -> float:
        r.set_half_res(half)
        r.set_taa(taa)
        r.invalidate_taa_history()
        for k in range(args.warmup):
            frame(k)
        r.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            frame(k)
        r.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        r.set_half_res(False)
        r.set_taa(False)
        return dt

    if not args.no_half_res:
        log("half-resolution cloud path")
        hs = time_mode(True, False)
        result["half_res"] = {"ms_per_frame": round(hs * 1e3, 4), "mpix_s": round(W * H / hs / 1e6, 3),
                              "speedup_vs_full": round(ms_per_step / (hs * 1e3), 3),
                              "rt": f"{max(1, W // 2)}x{max(1, H // 2)}",
                              "workload": "the same frames: cloud march into the RT, Catmull-Rom composite "
                                          "with the sky pass at full resolution"}
    if not args.no_taa:
        log("cloud TAA")
        ts = time_mode(False, True)
        ths = time_mode(True, True)
        result["taa"] = {"ms_per_frame": round(ts * 1e3, 4), "mpix_s": round(W * H / ts / 1e6, 3),
                         "overhead_vs_full": round(ts * 1e3 / ms_per_step, 3),
                         "half_res_ms_per_frame": round(ths * 1e3, 4),
                         "workload": "the same frames: march into a full-size (and a half-size) RT, TAA resolve "
                                     "against the history, composite"}
    if not args.no_present:
        log("presentation chain")
        result["present"] = bench_present(r, W, H, args.steps)
        r.load_workload(w0)
    if not args.no_adjoint:
        log("forward + adjoint")
        result["fwd_adjoint"] = bench_adjoint(r, W, H, args.adjoint_steps, 3)
        result["fwd_adjoint_640x360"] = bench_adjoint(r, 640, 360, args.adjoint_steps, 3)
        r.load_workload(w0)
    if not args.no_calibration:
        log("atmosphere calibration")
        result["atmosphere_calibration"] = bench_calibration(r, args.calibration_sets, args.calibration_iters)
        r.load_workload(w0)
    if not args.no_batch:
        result["batched"] = bench_batch(r, args.batch_scenes)
    if not args.no_cpu_baseline:
        log("CPU baseline")
        fmt = w0.noise_format
        import oracle
        base = oracle.decode(r.read_noise(abi.PS_NOISE_BASE, fmt), fmt)
        det = oracle.decode(r.read_noise(abi.PS_NOISE_DETAIL, fmt), fmt)
        curl = oracle.decode(r.read_noise(abi.PS_NOISE_CURL, fmt), fmt)
        wc = scene.workload_full_frame(0, width=W, height=H, ground=args.depth != "none")
        result["cpu_baseline"] = cpu_baseline(wc, base, det, curl, args.cpu_seconds, baseline_threads())
    r.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
