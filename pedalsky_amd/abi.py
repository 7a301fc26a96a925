"""ctypes mirror of ``include/pedalsky_hip.h`` and the loader of the gfx950 library.

The structs are byte-identical to the reference's GPU uploads
(crates/ps-core/src/frame_uniforms.rs:123-162, crates/ps-core/src/weather.rs:21-134,
crates/ps-clouds/src/params.rs:31-140); ``check_layouts`` pins sizes and
offsets against the numbers the reference's own layout tests assert.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libpedalsky_hip.so"

PS_OK = 0
PS_ERR_INVALID_ARGUMENT = 1
PS_ERR_DEVICE = 2
PS_ERR_OUT_OF_MEMORY = 3
PS_ERR_NOT_READY = 4
PS_ERR_OVERLAPPING_LAYERS = 5
PS_ERR_UNSUPPORTED = 6

PS_NOISE_UNORM8 = 0
PS_NOISE_FP16 = 1
PS_PIXEL_F32 = 0
PS_PIXEL_F16 = 1
PS_TONEMAP_ACES = 0
PS_TONEMAP_CLAMP = 1
PS_TARGET_SRGB8 = 0
PS_TARGET_UNORM8 = 1
PS_GROUND_RADIANCE = 0
PS_GROUND_PBR = 1
PS_MATERIAL_GRASS, PS_MATERIAL_BARE_SOIL, PS_MATERIAL_TARMAC, PS_MATERIAL_SAND, PS_MATERIAL_WATER_EDGE = range(5)

PS_LUT_TRANSMITTANCE = 1
PS_LUT_MULTISCATTER = 2
PS_LUT_SKYVIEW = 4
PS_LUT_AERIAL_PERSPECTIVE = 8
PS_LUT_ALL = 15
PS_LUT_MULTISCATTER_DISABLED = 16

PS_NOISE_BASE = 0
PS_NOISE_DETAIL = 1
PS_NOISE_CURL = 2

MAX_CLOUD_LAYERS = 8

F4 = C.c_float * 4
F16 = C.c_float * 16


class Utc(C.Structure):
    """ps_utc: a UTC instant at whole seconds."""

    _fields_ = [(n, C.c_int32) for n in ("year", "month", "day", "hour", "minute", "second")]


class SunPosition(C.Structure):
    """ps_sun_position (SunPositionPrecise, spa.rs:29-45, plus the Julian day)."""

    _fields_ = [(n, C.c_double) for n in (
        "altitude_deg", "azimuth_deg", "zenith_deg", "distance_au", "right_ascension_deg", "declination_deg",
        "apparent_sidereal_time_deg", "julian_day")]


class WorldSun(C.Structure):
    """ps_world_sun_state (the WorldState sun fields, world.rs:30-37)."""

    _fields_ = [("altitude_rad", C.c_float), ("azimuth_rad", C.c_float), ("distance_au", C.c_float),
                ("direction_world", C.c_float * 3), ("toa_illuminance_lux", C.c_float)]


class WeatherSynth(C.Structure):
    """ps_weather_synth (WeatherMap / CloudTypeGrid synthesis inputs)."""

    _fields_ = [("has_layers", C.c_int32), ("precip_mm_per_h", C.c_float), ("ground_wetness", C.c_float),
                ("coverage_grid", C.POINTER(C.c_float)), ("type_grid", C.POINTER(C.c_uint8)),
                ("grid_w", C.c_int32), ("grid_h", C.c_int32), ("grid_extent_m", C.c_float)]


class SurfaceParams(C.Structure):
    """SurfaceParams (64 B, crates/ps-core/src/weather.rs:159-201)."""

    _fields_ = [(n, C.c_float) for n in (
        "visibility_m", "temperature_c", "dewpoint_c", "pressure_hpa", "wind_dir_deg", "wind_speed_mps",
        "ground_wetness", "puddle_coverage", "snow_depth_m", "puddle_start", "precip_intensity_mm_per_h",
        "precip_kind", "material")] + [("_pad", C.c_float * 3)]


class PresentParams(C.Structure):
    """ps_present_params: tone-map uniforms (tonemap.wgsl:7-13), ground model,
    godrays and bloom tunables, surface parameters. Fields not given are 0
    (caller radiance, godrays and bloom off)."""

    _fields_ = [("ev100", C.c_float), ("tonemap", C.c_int32), ("ground_model", C.c_int32),
                ("godrays", C.c_int32), ("godrays_samples", C.c_uint32), ("godrays_decay", C.c_float),
                ("godrays_intensity", C.c_float), ("godrays_bright_threshold", C.c_float),
                ("bloom", C.c_int32), ("bloom_threshold_ev100", C.c_float), ("bloom_intensity", C.c_float),
                ("bloom_knee_ev", C.c_float), ("surface", SurfaceParams), ("target", C.c_int32)]


def default_present_params() -> PresentParams:
    p = PresentParams()
    load().ps_default_present_params(C.byref(p))
    return p


class FrameUniforms(C.Structure):
    """FrameUniformsGpu (448 B)."""

    _fields_ = [
        ("view", F16), ("proj", F16), ("view_proj", F16), ("inv_view_proj", F16),
        ("camera_position_world", F4), ("camera_velocity_world", F4),
        ("sun_direction", F4), ("sun_illuminance", F4),
        ("lightning_illuminance", F4), ("lightning_origin_world", F4),
        ("viewport_size", F4),
        ("time_seconds", C.c_float), ("simulated_seconds", C.c_float),
        ("frame_index", C.c_uint32), ("ev100", C.c_float),
        ("prev_view_proj", F16),
    ]


class AtmosphereParams(C.Structure):
    """AtmosphereParams / WorldUniformsGpu (144 B)."""

    _fields_ = [
        ("planet_radius_m", C.c_float), ("atmosphere_top_m", C.c_float),
        ("rayleigh_scale_height_m", C.c_float), ("mie_scale_height_m", C.c_float),
        ("rayleigh_scattering", F4), ("mie_scattering", F4), ("mie_absorption", F4),
        ("mie_g", C.c_float), ("_pad_after_mie_g", C.c_float * 3),
        ("ozone_absorption", F4),
        ("ozone_center_m", C.c_float), ("ozone_thickness_m", C.c_float),
        ("_pad_after_ozone", C.c_float * 2),
        ("ground_albedo", F4), ("haze_extinction_per_m", F4),
    ]


class CloudParams(C.Structure):
    """CloudParamsGpu (112 B)."""

    _fields_ = [
        ("sigma_s", C.c_float * 3), ("_pad_after_sigma_s", C.c_float),
        ("sigma_a", C.c_float * 3), ("droplet_diameter_bias", C.c_float),
        ("cone_light_sampling", C.c_uint32), ("forward_bias", C.c_float),
        ("detail_strength", C.c_float), ("curl_strength", C.c_float),
        ("powder_strength", C.c_float), ("multi_scatter_a", C.c_float),
        ("multi_scatter_b", C.c_float), ("multi_scatter_c", C.c_float),
        ("ambient_strength", C.c_float), ("base_scale_m", C.c_float),
        ("detail_scale_m", C.c_float), ("weather_scale_m", C.c_float),
        ("light_steps", C.c_uint32), ("cloud_steps", C.c_uint32),
        ("multi_scatter_octaves", C.c_uint32), ("cloud_layer_count", C.c_uint32),
        ("temporal_jitter", C.c_uint32), ("wind_drift_strength", C.c_float),
        ("wind_skew_strength", C.c_float), ("diurnal_strength", C.c_float),
    ]


class CloudLayer(C.Structure):
    """CloudLayerGpu (48 B)."""

    _fields_ = [
        ("base_m", C.c_float), ("top_m", C.c_float), ("coverage", C.c_float),
        ("density_scale", C.c_float), ("cloud_type", C.c_uint32),
        ("shape_bias", C.c_float), ("detail_bias", C.c_float), ("anvil_bias", C.c_float),
        ("droplet_diameter_um", C.c_float), ("_pad_after_droplets", C.c_float * 3),
    ]


class LutDims(C.Structure):
    _fields_ = [
        ("transmittance_w", C.c_int32), ("transmittance_h", C.c_int32),
        ("multiscatter_w", C.c_int32), ("multiscatter_h", C.c_int32),
        ("skyview_w", C.c_int32), ("skyview_h", C.c_int32),
        ("ap_w", C.c_int32), ("ap_h", C.c_int32), ("ap_d", C.c_int32),
    ]


class ImageOut(C.Structure):
    _fields_ = [
        ("composite", C.c_void_p), ("cloud_luminance", C.c_void_p),
        ("cloud_transmittance", C.c_void_p), ("format", C.c_int32), ("_reserved", C.c_int32),
    ]


class SceneDesc(C.Structure):
    """ps_scene_desc: one scene of a batch (host pointers to its uniform
    blocks, device pointers for depth and outputs)."""
    _fields_ = [
        ("frame", C.c_void_p), ("atmosphere", C.c_void_p), ("clouds", C.c_void_p), ("layers", C.c_void_p),
        ("n_layers", C.c_int32), ("_reserved", C.c_int32), ("layer_wind_xz", C.c_void_p),
        ("overcast_mask_128", C.c_void_p), ("depth", C.c_void_p), ("out", ImageOut),
    ]


PS_MAX_STEP_CLASSES = 8


class StepSchedule(C.Structure):
    """ps_step_schedule: primary / light march steps per resolution class
    of a batch (0 = each scene's own)."""
    _fields_ = [("n_classes", C.c_int32), ("width", C.c_int32 * PS_MAX_STEP_CLASSES),
                ("height", C.c_int32 * PS_MAX_STEP_CLASSES), ("cloud_steps", C.c_int32 * PS_MAX_STEP_CLASSES),
                ("light_steps", C.c_int32 * PS_MAX_STEP_CLASSES), ("batch_ms", C.c_float)]

    def as_dict(self) -> dict:
        n = int(self.n_classes)
        return {"classes": [{"size": f"{self.width[k]}x{self.height[k]}", "cloud_steps": int(self.cloud_steps[k]),
                             "light_steps": int(self.light_steps[k])} for k in range(n)],
                "batch_ms": round(float(self.batch_ms), 3)}


class Tuning(C.Structure):
    """ps_tuning: launch shape of a batch."""
    _fields_ = [("tile_w", C.c_int32), ("persistent", C.c_int32), ("workers_per_cu", C.c_int32),
                ("order", C.c_int32), ("grab", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class WorkCounters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "pixels", "layer_marches", "primary_steps", "covered_steps", "eroded_steps",
        "dense_samples", "light_steps", "light_covered", "octaves", "ap_taps")]


class FrameTiming(C.Structure):
    _fields_ = [
        ("transmittance_ms", C.c_float), ("multiscatter_ms", C.c_float),
        ("skyview_ms", C.c_float), ("aerial_perspective_ms", C.c_float),
        ("render_ms", C.c_float), ("frame_ms", C.c_float),
        ("frames", C.c_int32), ("_reserved", C.c_int32),
    ]


LAYOUT_PINS = {
    # sizes asserted by the reference's own tests
    "FrameUniforms": (FrameUniforms, 448),   # frame_uniforms.rs:195-201
    "AtmosphereParams": (AtmosphereParams, 144),  # weather.rs:544-555
    "CloudParams": (CloudParams, 112),       # params.rs:221-231
    "CloudLayer": (CloudLayer, 48),          # params.rs:228-231, weather.rs:544-555
}

# byte offsets the std140 WGSL declarations imply (uniforms.wgsl:11-58,
# cloud_uniforms.wgsl:23-106; the reference's naga layout linters check the same)
OFFSET_PINS = {
    FrameUniforms: {"view": 0, "proj": 64, "view_proj": 128, "inv_view_proj": 192,
                    "camera_position_world": 256, "camera_velocity_world": 272,
                    "sun_direction": 288, "sun_illuminance": 304,
                    "lightning_illuminance": 320, "lightning_origin_world": 336,
                    "viewport_size": 352, "time_seconds": 368, "simulated_seconds": 372,
                    "frame_index": 376, "ev100": 380, "prev_view_proj": 384},
    AtmosphereParams: {"planet_radius_m": 0, "atmosphere_top_m": 4, "rayleigh_scale_height_m": 8,
                       "mie_scale_height_m": 12, "rayleigh_scattering": 16, "mie_scattering": 32,
                       "mie_absorption": 48, "mie_g": 64, "ozone_absorption": 80,
                       "ozone_center_m": 96, "ozone_thickness_m": 100, "ground_albedo": 112,
                       "haze_extinction_per_m": 128},
    CloudParams: {"sigma_s": 0, "sigma_a": 16, "droplet_diameter_bias": 28,
                  "cone_light_sampling": 32, "forward_bias": 36, "detail_strength": 40,
                  "curl_strength": 44, "powder_strength": 48, "multi_scatter_a": 52,
                  "multi_scatter_b": 56, "multi_scatter_c": 60, "ambient_strength": 64,
                  "base_scale_m": 68, "detail_scale_m": 72, "weather_scale_m": 76,
                  "light_steps": 80, "cloud_steps": 84, "multi_scatter_octaves": 88,
                  "cloud_layer_count": 92, "temporal_jitter": 96, "wind_drift_strength": 100,
                  "wind_skew_strength": 104, "diurnal_strength": 108},
    CloudLayer: {"base_m": 0, "top_m": 4, "coverage": 8, "density_scale": 12, "cloud_type": 16,
                 "shape_bias": 20, "detail_bias": 24, "anvil_bias": 28, "droplet_diameter_um": 32},
}


def check_layouts() -> None:
    for name, (cls, size) in LAYOUT_PINS.items():
        if C.sizeof(cls) != size:
            raise AssertionError(f"{name}: size {C.sizeof(cls)} != {size}")
    for cls, pins in OFFSET_PINS.items():
        for field, off in pins.items():
            got = getattr(cls, field).offset
            if got != off:
                raise AssertionError(f"{cls.__name__}.{field}: offset {got} != {off}")


class CloudGrads(C.Structure):
    """ps_cloud_grads: device gradient accumulators."""

    _fields_ = [("layers", C.c_void_p), ("density_delta", C.c_void_p)]


# (name, restype, argtypes) of every exported entry point in include/pedalsky_hip.h
_P = C.c_void_p
_SIGNATURES = [
    ("ps_ctx_create", C.c_int32, [C.c_int32, C.POINTER(C.c_void_p)]),
    ("ps_ctx_destroy", None, [_P]),
    ("ps_last_error", C.c_char_p, []),
    ("ps_abi_version", C.c_int32, []),
    ("ps_ctx_set_stream", C.c_int32, [_P, _P]),
    ("ps_ctx_synchronize", C.c_int32, [_P]),
    ("ps_ctx_set_lut_dims", C.c_int32, [_P, C.POINTER(LutDims)]),
    ("ps_ctx_get_lut_dims", C.c_int32, [_P, C.POINTER(LutDims)]),
    ("ps_set_atmosphere", C.c_int32, [_P, C.POINTER(AtmosphereParams)]),
    ("ps_set_frame", C.c_int32, [_P, C.POINTER(FrameUniforms)]),
    ("ps_set_march_steps", C.c_int32, [_P, C.c_uint32, C.c_uint32]),
    ("ps_set_clouds", C.c_int32, [_P, C.POINTER(CloudParams), C.POINTER(CloudLayer), C.c_int32,
                                  C.POINTER(C.c_float)]),
    ("ps_generate_noise", C.c_int32, [_P, C.c_uint32, C.c_int32]),
    ("ps_upload_noise", C.c_int32, [_P, C.c_int32, _P, _P, _P]),
    ("ps_read_noise", C.c_int32, [_P, C.c_int32, _P, C.c_size_t]),
    ("ps_upload_blue_noise", C.c_int32, [_P, _P]),
    ("ps_synthesise_blue_noise", C.c_int32, [_P]),
    ("ps_read_blue_noise", C.c_int32, [_P, C.c_void_p]),
    ("ps_upload_weather", C.c_int32, [_P, _P, C.c_int32, C.c_int32, C.c_int32, _P, _P]),
    ("ps_bake_luts", C.c_int32, [_P, C.c_uint32]),
    ("ps_read_lut", C.c_int32, [_P, C.c_int32, _P, C.c_size_t]),
    ("ps_render", C.c_int32, [_P, _P, C.POINTER(ImageOut)]),
    ("ps_render_noalloc", C.c_int32, [_P, _P, C.POINTER(ImageOut)]),
    ("ps_frame", C.c_int32, [_P, C.c_uint32, _P, C.POINTER(ImageOut)]),
    ("ps_time_frames", C.c_int32, [_P, C.c_int32, C.c_uint32, _P, C.POINTER(ImageOut),
                                   C.POINTER(FrameTiming)]),
    ("ps_timing_begin", C.c_int32, [_P, C.c_int32]),
    ("ps_timing_end", C.c_int32, [_P, C.POINTER(FrameTiming)]),
    ("ps_count_work", C.c_int32, [_P, _P, C.POINTER(WorkCounters)]),
    ("ps_count_work_tiles", C.c_int32, [_P, _P, C.POINTER(WorkCounters), _P, C.c_int32]),
    ("ps_batch_create", C.c_int32, [_P, C.POINTER(SceneDesc), C.c_int32, C.POINTER(C.c_void_p)]),
    ("ps_batch_update", C.c_int32, [_P, C.POINTER(SceneDesc), C.c_int32]),
    ("ps_batch_set_tuning", C.c_int32, [_P, C.POINTER(Tuning)]),
    ("ps_batch_get_tuning", C.c_int32, [_P, C.POINTER(Tuning)]),
    ("ps_batch_autotune", C.c_int32, [_P, C.c_int32, C.POINTER(Tuning)]),
    ("ps_batch_run", C.c_int32, [_P]),
    ("ps_batch_capture", C.c_int32, [_P]),
    ("ps_batch_replay", C.c_int32, [_P]),
    ("ps_batch_time", C.c_int32, [_P, C.c_int32, C.POINTER(C.c_float)]),
    ("ps_batch_destroy", None, [_P]),
    ("ps_render_batch", C.c_int32, [_P, C.POINTER(SceneDesc), C.c_int32, C.POINTER(Tuning)]),
    ("ps_batch_get_step_schedule", C.c_int32, [_P, C.POINTER(StepSchedule)]),
    ("ps_batch_set_step_schedule", C.c_int32, [_P, C.POINTER(StepSchedule)]),
    ("ps_batch_autotune_steps", C.c_int32, [_P, C.c_float, C.c_int32, C.POINTER(StepSchedule)]),
    ("ps_set_cloud_extensions", C.c_int32, [_P, C.POINTER(C.c_float), C.c_int32, _P]),
    ("ps_set_checkpoint_interval", C.c_int32, [_P, C.c_int32]),
    ("ps_set_half_res", C.c_int32, [_P, C.c_int32]),
    ("ps_set_taa", C.c_int32, [_P, C.c_int32]),
    ("ps_invalidate_taa_history", C.c_int32, [_P]),
    ("ps_set_active_view", C.c_int32, [_P, C.c_int32]),
    ("ps_render_train", C.c_int32, [_P, _P, C.POINTER(ImageOut)]),
    ("ps_frame_train", C.c_int32, [_P, C.c_uint32, _P, C.POINTER(ImageOut)]),
    ("ps_backward", C.c_int32, [_P, _P, _P, _P, C.POINTER(CloudGrads)]),
    ("ps_sky_ap_image", C.c_int32, [_P, _P, C.c_float, _P]),
    ("ps_atmosphere_grad", C.c_int32, [_P, C.c_int32, _P, _P, C.c_float, _P, _P, _P, _P]),
    ("ps_default_atmosphere", None, [C.POINTER(AtmosphereParams)]),
    ("ps_default_cloud_params", None, [C.POINTER(CloudParams)]),
    ("ps_remap_coverage", C.c_float, [C.c_float]),
    ("ps_default_density_scale", C.c_float, [C.c_uint32]),
    ("ps_default_droplet_diameter_um", C.c_float, [C.c_uint32]),
    ("ps_synthesise_cloud_layer", None, [C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float,
                                         C.c_float, C.c_float, C.c_float, C.c_float,
                                         C.POINTER(CloudLayer)]),
    ("ps_check_layers_overlap", C.c_int32, [C.POINTER(CloudLayer), C.c_int32,
                                            C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("ps_layer_wind", None, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float * 2]),
    ("ps_synthesise_weather", C.c_int32, [_P, C.POINTER(WeatherSynth), C.c_void_p]),
    ("ps_read_weather", C.c_int32, [_P, C.c_void_p, C.c_void_p]),
    ("ps_present", C.c_int32, [_P, C.c_void_p, C.c_void_p, C.POINTER(ImageOut), C.POINTER(PresentParams),
                               C.c_void_p, C.c_void_p]),
    ("ps_default_present_params", None, [C.POINTER(PresentParams)]),
    ("ps_exposure_stats", C.c_int32, [_P, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    ("ps_ev100_from_exposure_stats", C.c_float, [C.c_float, C.c_float]),
    ("ps_julian_day", C.c_double, [C.POINTER(Utc)]),
    ("ps_delta_t_seconds", C.c_double, [C.c_int32]),
    ("ps_sun_position_precise", C.c_int32, [C.POINTER(Utc), C.c_double, C.c_double, C.c_double,
                                            C.POINTER(SunPosition)]),
    ("ps_world_sun", C.c_int32, [C.POINTER(Utc), C.c_double, C.c_double, C.c_double, C.POINTER(WorldSun)]),
    ("ps_frame_set_sun", C.c_int32, [C.POINTER(FrameUniforms), C.POINTER(WorldSun), C.c_float]),
    ("ps_overcast_mask_value", C.c_uint8, [C.POINTER(CloudLayer), C.c_int32]),
    ("ps_overcast_mask", None, [C.POINTER(CloudLayer), C.c_int32, _P, C.c_int32, C.c_int32, C.c_float, _P]),
    ("ps_generate_blue_noise", None, [_P]),
]

EXPORTED = [s[0] for s in _SIGNATURES]

_lib = None


class PedalskyError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"pedalsky_hip status {status}: {message}")
        self.status = status


def load(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load the in-tree gfx950 library. Raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    # torch ships its own HIP runtime under the same SONAME (libamdhip64.so.7);
    # whichever is loaded first serves the whole process. Load torch's first
    # so a Python process that also uses torch for device buffers runs on one
    # runtime (torch does not initialise on a runtime other than its own).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not p.exists():
        raise FileNotFoundError(
            f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the sky + cloud path has no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(status: int) -> None:
    if status != PS_OK:
        msg = load().ps_last_error().decode(errors="replace")
        raise PedalskyError(status, msg)
