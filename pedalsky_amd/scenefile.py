"""Reference scene files (schema 1 TOML: surface observations, cloud layers,
precipitation, optional coverage grid) rendered and presented the way the
reference's headless harness renders its golden frames:

* crates/ps-app/tests/golden.rs:80-140 — the engine config of pedalsky.toml
  (Dunblane, 60 m ground elevation, 192/6 cloud steps, 4 octaves, detail
  0.35), wind drift and diurnal modulation off, camera at (0, 1.7, 0) with
  the scene's pitch and yaw 0, 60 degree FlyCamera;
* crates/ps-app/src/test_harness.rs:335-400 — WeatherState synthesis, the
  SPA sun and TOA illuminance of WorldState, the exposure;
* crates/ps-synthesis/src/state.rs:50-110 — the atmosphere (physical
  defaults + Koschmieder haze from the visibility, koschmieder.rs:7-19) and
  the surface scalars;
* then the frame (ground, sky, clouds) and the post-process chain of
  ps_present (ground PBR pass with aerial perspective, godrays, bloom, ACES)
  into the harness's Rgba8Unorm target.

The reference's own run differs in its texture formats (Rgba16Float LUTs and
cloud targets, Rgba8Unorm noise — the noise here is the same UNORM8 bake);
its frame is single-shot (TAA history invalid), as here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path
from typing import Optional

import numpy as np

from . import abi, scene

# pedalsky.toml [world]: the observer the golden frames are rendered for
LATITUDE_DEG = 56.1922
LONGITUDE_DEG = -3.9645
GROUND_ELEVATION_M = 60.0
SUN_ANGULAR_RADIUS_DEG = 0.27        # pedalsky.toml [render.atmosphere]
KOSCHMIEDER = 3.912                  # ps-synthesis/src/koschmieder.rs:7

CLOUD_TYPES = {"Cumulus": 0, "Stratus": 1, "Stratocumulus": 2, "Altocumulus": 3, "Altostratus": 4,
               "Cirrus": 5, "Cirrostratus": 6, "Cumulonimbus": 7}          # ps-core/src/scene.rs:64-80
MATERIALS = {"Grass": 0, "BareSoil": 1, "Tarmac": 2, "Sand": 3, "WaterEdge": 4}  # scene.rs:256-270
PRECIP_KINDS = {"None": 0, "Rain": 1, "Snow": 2, "Sleet": 3, "Hail": 4}       # state.rs:93-99


@dataclass
class SceneFile:
    path: Path
    data: dict

    @property
    def surface(self) -> dict:
        return self.data.get("surface", {})

    @property
    def layers(self) -> list:
        return self.data.get("clouds", {}).get("layers", [])


def load(path) -> SceneFile:
    import tomli
    p = Path(path)
    with open(p, "rb") as f:
        data = tomli.load(f)
    if data.get("schema_version", 1) != 1:
        raise ValueError(f"{p}: unsupported scene schema {data.get('schema_version')}")
    return SceneFile(p, data)


def parse_utc(iso: str) -> abi.Utc:
    """'YYYY-MM-DDTHH:MM:SSZ' as a ps_utc."""
    d, t = iso.rstrip("Z").split("T")
    y, mo, da = (int(v) for v in d.split("-"))
    h, mi, s = (int(v) for v in t.split(":"))
    return abi.Utc(y, mo, da, h, mi, s)


def atmosphere(sf: SceneFile) -> abi.AtmosphereParams:
    """state.rs:62-77: the physical defaults, the config's planet and ground
    albedo (equal to the defaults) and the Koschmieder haze of the
    visibility, the same in R, G and B."""
    a = scene.earth_atmosphere(mie_extinction=None)
    v = float(sf.surface.get("visibility_m", 30000.0))
    haze = KOSCHMIEDER / v if v > 0.0 else 0.0
    for i in range(3):
        a.haze_extinction_per_m[i] = haze
    a.haze_extinction_per_m[3] = 0.0
    return a


def surface_params(sf: SceneFile, wet_surface: bool = False) -> abi.SurfaceParams:
    """state.rs:80-104, as the ground subsystem uploads them: with the
    wet_surface switch off (pedalsky.toml:47, the golden harness's config)
    it zeroes the wetness and the puddle coverage (ps-ground/src/lib.rs:
    742-749); snow stays."""
    s = sf.surface
    wet = s.get("wetness", {})
    pr = sf.data.get("precipitation", {})
    out = abi.SurfaceParams()
    out.visibility_m = s.get("visibility_m", 30000.0)
    out.temperature_c = s.get("temperature_c", 17.0)
    out.dewpoint_c = s.get("dewpoint_c", 7.0)
    out.pressure_hpa = s.get("pressure_hpa", 1018.0)
    out.wind_dir_deg = s.get("wind_dir_deg", 240.0)
    out.wind_speed_mps = s.get("wind_speed_mps", 5.0)
    out.ground_wetness = wet.get("ground_wetness", 0.0)
    out.puddle_coverage = wet.get("puddle_coverage", 0.0)
    out.snow_depth_m = wet.get("snow_depth_m", 0.0)
    out.puddle_start = wet.get("puddle_start", 0.6)
    out.precip_intensity_mm_per_h = pr.get("intensity_mm_per_h", 0.0)
    out.precip_kind = float(PRECIP_KINDS[pr.get("type", "None")])
    out.material = float(MATERIALS[s.get("material", "Grass")])
    if not wet_surface:
        out.ground_wetness = 0.0
        out.puddle_coverage = 0.0
    return out


def cloud_layers(sf: SceneFile) -> list:
    """synthesise_cloud_layers over the scene's layers (cloud_layers.rs:44-73)."""
    out = []
    for L in sf.layers:
        out.append(scene.synth_layer(CLOUD_TYPES[L["type"]], float(L["base_m"]), float(L["top_m"]),
                                     float(L["coverage"]), density_scale=L.get("density_scale"),
                                     shape_bias=float(L.get("shape_octave_bias", 0.0)),
                                     detail_bias=float(L.get("detail_octave_bias", 0.0)),
                                     anvil_bias=L.get("anvil_bias"),
                                     droplet_diameter_um=L.get("droplet_diameter_um")))
    return out


def coverage_grid(sf: SceneFile):
    """The scene's [clouds.coverage_grid] (coverage_grid.rs:31-61): a
    little-endian f32 grid beside the scene file, and its extent."""
    g = sf.data.get("clouds", {}).get("coverage_grid")
    if not g:
        return None, 32000.0
    w, h = (int(v) for v in g["size"])
    data = np.fromfile(sf.path.parent / Path(g["data_path"]).name, dtype="<f4")
    if data.size != w * h:
        raise ValueError(f"{sf.path}: coverage grid holds {data.size} values, expected {w}x{h}")
    return data.reshape(h, w), float(g.get("extent_m", 32000.0))


def world_sun(utc: abi.Utc) -> abi.WorldSun:
    lib = abi.load()
    sun = abi.WorldSun()
    abi.check(lib.ps_world_sun(utc, LATITUDE_DEG, LONGITUDE_DEG, GROUND_ELEVATION_M, sun))
    return sun


def frame(width: int, height: int, pitch_deg: float, sun: abi.WorldSun) -> abi.FrameUniforms:
    """FrameUniforms of test_harness.rs:357-387: FlyCamera at (0, 1.7, 0),
    yaw 0, fov 60 deg, near 0.1, with the world's sun."""
    f = scene.make_frame(width, height, camera=(0.0, 1.7, 0.0), pitch_deg=pitch_deg, yaw_deg=0.0)
    abi.check(abi.load().ps_frame_set_sun(f, sun, math.radians(SUN_ANGULAR_RADIUS_DEG)))
    return f


def setup(r, sf: SceneFile, utc: str, pitch_deg: float, width: int = 1280, height: int = 720,
          cloud_steps: int = 192, light_steps: int = 6) -> abi.FrameUniforms:
    """Load the scene into a Renderer: LUT sizes, atmosphere, frame, clouds,
    noise (reference hash, UNORM8), blue noise, weather map, ground depth."""
    layers = cloud_layers(sf)
    s = sf.surface
    sun = world_sun(parse_utc(utc))
    f = frame(width, height, pitch_deg, sun)
    r.set_lut_dims(scene.lut_dims(ap_depth=64))   # the reference's LUT sizes (atmosphere_luts.rs:15-38)
    r.set_atmosphere(atmosphere(sf))
    r.set_frame(f)
    clouds = scene.cloud_params(steps=cloud_steps, light_steps=light_steps, octaves=4, detail_strength=0.35,
                                drift=0.0, diurnal=0.0, n_layers=len(layers))
    winds = scene.layer_winds(layers, float(s.get("wind_speed_mps", 5.0)), float(s.get("wind_dir_deg", 240.0)))
    r.set_clouds(clouds, layers, winds)
    r.generate_noise(0, abi.PS_NOISE_UNORM8)
    r.synthesise_blue_noise()
    grid, extent = coverage_grid(sf)
    pr = sf.data.get("precipitation", {})
    r.synthesise_weather(has_layers=len(layers) > 0, precip_mm_per_h=float(pr.get("intensity_mm_per_h", 0.0)),
                         ground_wetness=float(s.get("wetness", {}).get("ground_wetness", 0.0)),
                         coverage_grid=grid, grid_extent_m=extent,
                         overcast=scene.overcast_mask(layers, grid, extent))   # state.rs:116
    r.set_depth(scene.flat_ground_depth(f))
    return f


def render_presented(r, sf: SceneFile, utc: str, ev100: float, pitch_deg: float, width: int = 1280,
                     height: int = 720, hdr: bool = False):
    """One golden-harness frame: (height, width, 4) uint8 as the harness's
    Rgba8Unorm target holds it, and the HDR image before the tone map when
    `hdr`."""
    import ctypes as C
    import torch
    setup(r, sf, utc, pitch_deg, width, height)
    bufs = r.alloc_outputs(width, height, abi.PS_PIXEL_F32, ("composite", "cloud_luminance", "cloud_transmittance"))
    out = r.image_out(bufs)
    r.frame(out)
    pp = abi.default_present_params()
    pp.ev100 = ev100
    pp.ground_model = abi.PS_GROUND_PBR
    pp.surface = surface_params(sf)
    pp.target = abi.PS_TARGET_UNORM8       # the harness's Rgba8Unorm output (test_harness.rs:56-58)
    dev = f"cuda:{r.device}"
    ldr = torch.zeros((height, width, 4), dtype=torch.uint8, device=dev)
    hdr_img = torch.zeros((height, width, 4), dtype=torch.float32, device=dev)
    r.synchronize()
    torch.cuda.synchronize()
    abi.check(r.lib.ps_present(r.ctx, r.depth_ptr, None, C.byref(out), C.byref(pp), ldr.data_ptr(),
                               hdr_img.data_ptr()))
    r.synchronize()
    torch.cuda.synchronize()
    return (ldr.cpu().numpy(), hdr_img.cpu().numpy()) if hdr else ldr.cpu().numpy()
