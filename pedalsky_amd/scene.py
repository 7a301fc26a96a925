"""Scene description for the sky + cloud path: camera matrices, sun, atmosphere,
cloud layers and the synthetic inputs of the benchmark workloads.

Camera conventions follow crates/ps-core/src/camera.rs:44-71 (FlyCamera:
forward = (cos p sin y, sin p, -cos p cos y), ``look_to_rh``,
``perspective_infinite_reverse_rh``) and the headless harness
(crates/ps-app/src/test_harness.rs:357-387). The matrices are built in
float64 and stored as float32 like the reference's uniform upload.
"""
from __future__ import annotations

import dataclasses
import functools
import math
from typing import Optional, Sequence

import numpy as np

from . import abi

TOA_ILLUMINANCE_LUX = 127_500.0          # crates/ps-core/src/world.rs:17
SUN_ANGULAR_RADIUS_DEG = 0.27            # crates/ps-core/src/config.rs:343
CLOUD_TYPES = ("cumulus", "stratus", "stratocumulus", "altocumulus", "altostratus",
               "cirrus", "cirrostratus", "cumulonimbus")
# altitude bands per type (docs/cloud_calibration.md:54-62)
TYPE_BANDS_M = {0: (800.0, 2500.0), 1: (200.0, 800.0), 2: (600.0, 2000.0), 3: (2500.0, 5500.0),
                4: (3000.0, 6000.0), 5: (7500.0, 11000.0), 6: (7000.0, 10000.0),
                7: (600.0, 11000.0)}


# ---------------------------------------------------------------------------
# glam-equivalent matrices (column-major, as Mat4::to_cols_array)
# ---------------------------------------------------------------------------
def perspective_infinite_reverse_rh(fov_y: float, aspect: float, z_near: float) -> np.ndarray:
    h = math.cos(0.5 * fov_y) / math.sin(0.5 * fov_y)
    w = h / aspect
    m = np.zeros((4, 4))  # m[col, row]
    m[0, 0] = w
    m[1, 1] = h
    m[2, 3] = -1.0
    m[3, 2] = z_near
    return m


def look_to_rh(eye: Sequence[float], direction: Sequence[float], up: Sequence[float]) -> np.ndarray:
    f = np.asarray(direction, dtype=np.float64)
    f = f / np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, dtype=np.float64))
    s = s / np.linalg.norm(s)
    u = np.cross(s, f)
    e = np.asarray(eye, dtype=np.float64)
    m = np.zeros((4, 4))
    m[0] = [s[0], u[0], -f[0], 0.0]
    m[1] = [s[1], u[1], -f[1], 0.0]
    m[2] = [s[2], u[2], -f[2], 0.0]
    m[3] = [-e.dot(s), -e.dot(u), e.dot(f), 1.0]
    return m


def fly_camera_forward(pitch: float, yaw: float) -> np.ndarray:
    sy, cy = math.sin(yaw), math.cos(yaw)
    sp, cp = math.sin(pitch), math.cos(pitch)
    v = np.array([cp * sy, sp, -cp * cy])
    return v / np.linalg.norm(v)


def _mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Column-major product a*b (arrays indexed [col, row])."""
    return (a.T @ b.T).T


def altaz_to_world(altitude: float, azimuth: float) -> np.ndarray:
    """crates/ps-core/src/world.rs:110-114 (+X east, +Y up, +Z south)."""
    return np.array([math.cos(altitude) * math.sin(azimuth), math.sin(altitude),
                     -math.cos(altitude) * math.cos(azimuth)])


def sun_from_zenith(zenith_deg: float, azimuth_deg: float = 180.0) -> np.ndarray:
    return altaz_to_world(math.radians(90.0 - zenith_deg), math.radians(azimuth_deg))


def make_frame(width: int, height: int, camera=(0.0, 500.0, 0.0), pitch_deg: float = 0.0,
               yaw_deg: float = 0.0, fov_deg: float = 60.0, near: float = 0.1,
               sun_dir: Optional[Sequence[float]] = None, sun_zenith_deg: float = 40.0,
               sun_azimuth_deg: float = 180.0, illuminance: float = TOA_ILLUMINANCE_LUX,
               frame_index: int = 0) -> abi.FrameUniforms:
    f = abi.FrameUniforms()
    eye = np.asarray(camera, dtype=np.float64)
    eye[1] = max(eye[1], 0.1)  # test_harness.rs:360-364
    view = look_to_rh(eye, fly_camera_forward(math.radians(pitch_deg), math.radians(yaw_deg)), (0, 1, 0))
    proj = perspective_infinite_reverse_rh(math.radians(fov_deg), width / height, near)
    vp = _mul(proj, view)
    inv = np.linalg.inv(vp.T).T
    f.view[:] = view.astype(np.float32).ravel().tolist()
    f.proj[:] = proj.astype(np.float32).ravel().tolist()
    f.view_proj[:] = vp.astype(np.float32).ravel().tolist()
    f.inv_view_proj[:] = inv.astype(np.float32).ravel().tolist()
    f.prev_view_proj[:] = f.view_proj[:]
    f.camera_position_world[:] = [eye[0], eye[1], eye[2], 0.0]
    s = np.asarray(sun_dir, dtype=np.float64) if sun_dir is not None else sun_from_zenith(
        sun_zenith_deg, sun_azimuth_deg)
    f.sun_direction[:] = [s[0], s[1], s[2], math.radians(SUN_ANGULAR_RADIUS_DEG)]
    f.sun_illuminance[:] = [illuminance, illuminance, illuminance, illuminance]
    f.viewport_size[:] = [width, height, 1.0 / width, 1.0 / height]
    f.frame_index = frame_index
    return f


def frame_view_proj(f: abi.FrameUniforms) -> np.ndarray:
    return np.array(f.view_proj[:], dtype=np.float32).reshape(4, 4)


def frame_inv_view_proj(f: abi.FrameUniforms) -> np.ndarray:
    return np.array(f.inv_view_proj[:], dtype=np.float32).reshape(4, 4)


# ---------------------------------------------------------------------------
# atmosphere and clouds
# ---------------------------------------------------------------------------
def earth_atmosphere(mie_extinction: Optional[float] = 4.44e-6, scale: Optional[dict] = None) -> abi.AtmosphereParams:
    """Earth defaults (weather.rs:60-84); `mie_extinction` sets the Mie
    absorption to extinction - scattering (the benchmark workloads quote an
    extinction of 4.44e-6/m). `scale` multiplies named coefficients."""
    lib = abi.load()
    a = abi.AtmosphereParams()
    lib.ps_default_atmosphere(a)
    if mie_extinction is not None:
        absorb = mie_extinction - a.mie_scattering[0]
        for i in range(3):
            a.mie_absorption[i] = absorb
    if scale:
        for key, k in scale.items():
            if key in ("rayleigh_scattering", "mie_scattering", "mie_absorption", "ozone_absorption"):
                arr = getattr(a, key)
                for i in range(3):
                    arr[i] = arr[i] * k
            else:
                setattr(a, key, getattr(a, key) * k)
    return a


def cloud_params(steps: int = 128, light_steps: int = 6, octaves: int = 4, detail_strength: float = 0.35,
                 drift: float = 0.0, skew: float = 0.5, diurnal: float = 0.0, n_layers: int = 0,
                 **overrides) -> abi.CloudParams:
    """CloudParamsGpu seeded like CloudsSubsystem::new from the shipped config
    (ps-clouds/src/lib.rs:561-600, pedalsky.toml:62-74)."""
    lib = abi.load()
    p = abi.CloudParams()
    lib.ps_default_cloud_params(p)
    p.cloud_steps = steps
    p.light_steps = light_steps
    p.multi_scatter_octaves = octaves
    p.detail_strength = detail_strength
    p.wind_drift_strength = drift
    p.wind_skew_strength = skew
    p.diurnal_strength = diurnal
    p.cloud_layer_count = n_layers
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def synth_layer(cloud_type: int, base_m: float, top_m: float, coverage: float,
                density_scale: Optional[float] = None, shape_bias: float = 0.0, detail_bias: float = 0.0,
                anvil_bias: Optional[float] = None, droplet_diameter_um: Optional[float] = None) -> abi.CloudLayer:
    """synthesise_cloud_layers for one scene layer (ps-synthesis/src/cloud_layers.rs:44-73)."""
    lib = abi.load()
    out = abi.CloudLayer()
    nan = float("nan")
    lib.ps_synthesise_cloud_layer(cloud_type, base_m, top_m, coverage,
                                  nan if density_scale is None else density_scale, shape_bias, detail_bias,
                                  nan if anvil_bias is None else anvil_bias,
                                  nan if droplet_diameter_um is None else droplet_diameter_um, out)
    return out


def layer_winds(layers: Sequence[abi.CloudLayer], speed_mps: float = 5.0, dir_deg: float = 240.0) -> np.ndarray:
    lib = abi.load()
    out = np.zeros((len(layers), 2), dtype=np.float32)
    for i, L in enumerate(layers):
        v = (abi.C.c_float * 2)()
        lib.ps_layer_wind(speed_mps, dir_deg, L.base_m, L.top_m, v)
        out[i] = [v[0], v[1]]
    return out


# ---------------------------------------------------------------------------
# synthetic inputs
# ---------------------------------------------------------------------------
class Pcg32:
    """PCG-XSH-RR 64/32 (O'Neill 2014), pcg32_srandom seeding."""

    MULT = 6364136223846793005
    MASK = (1 << 64) - 1

    def __init__(self, seed: int, stream: int = 54):
        self.inc = ((stream << 1) | 1) & self.MASK
        self.state = 0
        self.next_u32()
        self.state = (self.state + seed) & self.MASK
        self.next_u32()

    def next_u32(self) -> int:
        old = self.state
        self.state = (old * self.MULT + self.inc) & self.MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self, n: int) -> np.ndarray:
        """n floats in [0, 1): (u32 >> 8) * 2^-24."""
        out = np.empty(n, dtype=np.float64)
        for i in range(n):
            out[i] = (self.next_u32() >> 8) * (1.0 / 16777216.0)
        return out


def pcg_uniforms(seed: int, n: int) -> np.ndarray:
    return Pcg32(seed).uniform(n)


@functools.lru_cache(maxsize=4)
def weather_map(size: int = 512, seed: int = 0x5EED, sigma: float = 8.0) -> np.ndarray:
    """Synthetic weather map: PCG32 U[0,1] per texel, circular Gaussian blur
    (sigma in texels, wrap), then a rank transform so the smoothed coverage is
    again uniform on [0, 1]. Returns (size, size, 2) binary16 bits: R =
    coverage, G = 0."""
    u = pcg_uniforms(seed, size * size).reshape(size, size)
    k = np.arange(size)
    k = np.minimum(k, size - k).astype(np.float64)
    g1 = np.exp(-0.5 * (k / sigma) ** 2)
    g1 /= g1.sum()
    blurred = np.real(np.fft.ifft2(np.fft.fft2(u) * np.outer(np.fft.fft(g1), np.fft.fft(g1))))
    order = np.argsort(blurred.ravel(), kind="stable")
    ranks = np.empty(size * size, dtype=np.float64)
    ranks[order] = (np.arange(size * size) + 0.5) / (size * size)
    cov = ranks.reshape(size, size).astype(np.float16)
    out = np.zeros((size, size, 2), dtype=np.float16)
    out[..., 0] = cov
    return out.view(np.uint16)


@functools.lru_cache(maxsize=1)
def blue_noise() -> np.ndarray:
    """generate_blue_noise(64) (crates/ps-clouds/src/noise.rs:646-704)."""
    out = np.zeros(64 * 64, dtype=np.uint8)
    abi.load().ps_generate_blue_noise(out.ctypes.data)
    return out


def overcast_mask(layers: Sequence[abi.CloudLayer], grid: Optional[np.ndarray] = None,
                  grid_extent_m: float = 32000.0) -> np.ndarray:
    """The 128x128 top-down overcast mask: uniform, or over a (h, w)
    coverage grid (ps_overcast_mask)."""
    arr = (abi.CloudLayer * max(len(layers), 1))(*layers)
    if grid is None:
        v = abi.load().ps_overcast_mask_value(arr, len(layers))
        return np.full(128 * 128, v, dtype=np.uint8)
    g = np.ascontiguousarray(grid, dtype=np.float32)
    out = np.zeros(128 * 128, dtype=np.uint8)
    abi.load().ps_overcast_mask(arr, len(layers), g.ctypes.data, g.shape[1], g.shape[0], float(grid_extent_m),
                                out.ctypes.data)
    return out


def ground_distance(frame: abi.FrameUniforms, planet_radius: float = 6_360_000.0) -> np.ndarray:
    """Per pixel, the view distance to the first hit of the planet sphere of
    radius `planet_radius` centred (0, -R, 0); inf where the ray misses.
    float64 (H, W)."""
    cam, d = pixel_rays(frame)
    oc = cam + np.array([0.0, planet_radius, 0.0])
    b = d @ oc
    r_cam = np.linalg.norm(oc)
    c = (r_cam - planet_radius) * (r_cam + planet_radius)
    disc = b * b - c
    hit = disc >= 0
    t0 = -b - np.sqrt(np.where(hit, disc, 0.0))
    hit &= t0 > 0
    return np.where(hit, t0, np.inf)


def ground_depth(frame: abi.FrameUniforms, planet_radius: float = 6_360_000.0) -> np.ndarray:
    """Reverse-Z NDC depth an opaque planet-sized ground writes: per pixel, the
    view ray's first hit of the sphere of radius `planet_radius` centred
    (0, -R, 0); 0 (far) where the ray misses. float32 (H, W)."""
    return distance_depth(frame, ground_distance(frame, planet_radius))


def flat_ground_distance(frame: abi.FrameUniforms, half_extent: float = 100_000.0) -> np.ndarray:
    """Per pixel, the view distance to the reference's startup ground mesh:
    the plane y = 0 within |x|, |z| <= half_extent (QUAD_HALF_EXTENT_M,
    crates/ps-ground/src/lib.rs:151, 365-407); inf where the ray misses it.
    float64 (H, W)."""
    cam, d = pixel_rays(frame)
    down = d[..., 1] < 0
    t = np.where(down, -cam[1] / np.where(down, d[..., 1], -1.0), np.inf)
    p = cam + d * np.where(np.isfinite(t), t, 0.0)[..., None]
    inside = (np.abs(p[..., 0]) <= half_extent) & (np.abs(p[..., 2]) <= half_extent)
    return np.where(down & inside, t, np.inf)


def flat_ground_depth(frame: abi.FrameUniforms, half_extent: float = 100_000.0) -> np.ndarray:
    """Reverse-Z NDC depth the startup ground quad writes. float32 (H, W)."""
    return distance_depth(frame, flat_ground_distance(frame, half_extent))


def pixel_rays(frame: abi.FrameUniforms):
    """Camera position and the unit view ray through each pixel centre
    (math.wgsl:23-33 in FP64): (cam (3,), d (H, W, 3))."""
    W, H = int(frame.viewport_size[0]), int(frame.viewport_size[1])
    inv = frame_inv_view_proj(frame).astype(np.float64)  # [col, row]
    cam = np.array(frame.camera_position_world[:3], dtype=np.float64)
    xs = (np.arange(W) + 0.5) / W * 2 - 1
    ys = 1 - (np.arange(H) + 0.5) / H * 2
    nx, ny = np.meshgrid(xs, ys)
    hom = np.stack([nx, ny, np.ones_like(nx), np.ones_like(nx)], axis=-1)
    near = hom @ inv  # row vector times [col,row] matrix = M @ v
    p = near[..., :3] / near[..., 3:4]
    d = p - cam
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return cam, d


def distance_depth(frame: abi.FrameUniforms, t: np.ndarray) -> np.ndarray:
    """Reverse-Z NDC depth of opaque surfaces at view distance t (H, W) along
    each pixel's ray; non-finite t is the far plane (0). float32 (H, W)."""
    cam, d = pixel_rays(frame)
    vp = frame_view_proj(frame).astype(np.float64)
    t = np.asarray(t, dtype=np.float64)
    hit = np.isfinite(t)
    wp = cam + d * np.where(hit, t, 1.0)[..., None]
    clip = np.concatenate([wp, np.ones(t.shape)[..., None]], axis=-1) @ vp
    return np.where(hit, clip[..., 2] / clip[..., 3], 0.0).astype(np.float32)


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Workload:
    name: str
    frame: abi.FrameUniforms
    atmosphere: abi.AtmosphereParams
    clouds: abi.CloudParams
    layers: list
    winds: np.ndarray
    noise_seed: int
    noise_format: int
    lut_dims: abi.LutDims
    weather: np.ndarray            # (h, w, 2) uint16 binary16 bits
    overcast: np.ndarray           # 128*128 uint8
    blue: np.ndarray               # 64*64 uint8
    depth: Optional[np.ndarray]    # (H, W) float32 or None (all far)
    type_grid: Optional[np.ndarray] = None  # 128*128 uint8 cloud types (255 = layer type) or None

    @property
    def width(self) -> int:
        return int(self.frame.viewport_size[0])

    @property
    def height(self) -> int:
        return int(self.frame.viewport_size[1])


def lut_dims(ap_depth: int = 32) -> abi.LutDims:
    return abi.LutDims(256, 64, 32, 32, 192, 108, 32, 32, ap_depth)


def _workload(name, width, height, layers, steps, sun_zenith, pitch, yaw, noise_format, ground=True,
              weather_size=512, noise_seed=0) -> Workload:
    frame = make_frame(width, height, camera=(0.0, 500.0, 0.0), pitch_deg=pitch, yaw_deg=yaw,
                       sun_zenith_deg=sun_zenith)
    atm = earth_atmosphere()
    clouds = cloud_params(steps=steps, light_steps=6, n_layers=len(layers))
    winds = layer_winds(layers)
    return Workload(name=name, frame=frame, atmosphere=atm, clouds=clouds, layers=list(layers), winds=winds,
                    noise_seed=noise_seed, noise_format=noise_format, lut_dims=lut_dims(32),
                    weather=weather_map(weather_size), overcast=overcast_mask(layers), blue=blue_noise(),
                    depth=ground_depth(frame, atm.planet_radius_m) if ground else None)


def workload_small(**kw) -> Workload:
    """320x180, one cumulus layer 1500-2300 m (coverage 0.5, density 1), sun
    zenith 40 deg, camera at 500 m, pitch 15 deg, yaw 180 deg, 64/6 steps,
    reference-format (UNORM8) noise."""
    layers = [synth_layer(0, 1500.0, 2300.0, 0.5, density_scale=1.0)]
    args = dict(width=320, height=180, layers=layers, steps=64, sun_zenith=40.0, pitch=15.0, yaw=180.0,
                noise_format=abi.PS_NOISE_UNORM8)
    args.update(kw)
    return _workload("sky+clouds 320x180, 1 cumulus layer, 64/6 steps", **args)


def full_frame_camera(index: int) -> tuple:
    """Seeded camera of frame `index` of the full-frame sweep: pitch U[-10, 30]
    deg, yaw U[0, 360) deg from PCG32 seed 0xC2."""
    u = pcg_uniforms(0xC2, 2 * (index + 1))
    return -10.0 + 40.0 * u[2 * index], 360.0 * u[2 * index + 1]


def workload_full_frame(index: int = 0, width: int = 1280, height: int = 720, **kw) -> Workload:
    """1280x720, cumulus 1500-2300 m (coverage 0.5) + cirrus 8000-9000 m
    (coverage 0.3, density 0.3), 128/6 steps, FP16 noise, sun zenith 65 deg,
    camera at 500 m with the seeded pitch/yaw of sweep frame `index`."""
    pitch, yaw = full_frame_camera(index)
    layers = [synth_layer(0, 1500.0, 2300.0, 0.5, density_scale=1.0),
              synth_layer(5, 8000.0, 9000.0, 0.3, density_scale=0.3)]
    args = dict(width=width, height=height, layers=layers, steps=128, sun_zenith=65.0, pitch=pitch, yaw=yaw,
                noise_format=abi.PS_NOISE_FP16)
    args.update(kw)
    return _workload(f"sky+clouds {width}x{height}, cumulus + cirrus, 128/6 steps, frame {index}", **args)


BATCH_RESOLUTIONS = ((256, 144), (640, 360), (1280, 720), (1920, 1080))


def workload_batch(n: int = 64, seed: int = 0xC3, steps: int = 128, ground: bool = True) -> list:
    """`n` uneven scenes for one batched launch, drawn from one PCG32 stream
    (seed 0xC3): resolution uniform over 256x144 / 640x360 / 1280x720 /
    1920x1080; 1-3 vertically disjoint layers, each type uniform over the 8
    with base/top inside the type's altitude band, scene coverage U[0.1, 1]
    and density_scale U[0.3, 2]; sun zenith U[0, 95] deg (azimuth 180);
    Rayleigh, Mie scattering, Mie absorption and ozone coefficients each
    scaled by U[0.5, 1.5]; camera at 500 m with pitch U[-10, 30] and yaw
    U[0, 360) deg; 128/6 steps, FP16 noise shared by the batch."""
    rng = Pcg32(seed)
    u = lambda: float(rng.uniform(1)[0])  # noqa: E731
    out = []
    for i in range(n):
        W, H = BATCH_RESOLUTIONS[min(int(u() * 4), 3)]
        k = 1 + min(int(u() * 3), 2)
        layers = []
        for _ in range(k):
            t = min(int(u() * 8), 7)
            lo, hi = TYPE_BANDS_M[t]
            thick = (0.25 + 0.5 * u()) * (hi - lo)
            base = lo + u() * (hi - lo - thick)
            cov, ds = 0.1 + 0.9 * u(), 0.3 + 1.7 * u()
            cand = synth_layer(t, base, base + thick, cov, density_scale=ds)
            trial = (abi.CloudLayer * (len(layers) + 1))(*(layers + [cand]))
            if abi.load().ps_check_layers_overlap(trial, len(layers) + 1, None, None) == 0:
                layers.append(cand)
        zen = 95.0 * u()
        pitch, yaw = -10.0 + 40.0 * u(), 360.0 * u()
        scale = {key: 0.5 + u() for key in ("rayleigh_scattering", "mie_scattering", "mie_absorption",
                                             "ozone_absorption")}
        frame = make_frame(W, H, camera=(0.0, 500.0, 0.0), pitch_deg=pitch, yaw_deg=yaw, sun_zenith_deg=zen)
        atm = earth_atmosphere(scale=scale)
        clouds = cloud_params(steps=steps, light_steps=6, n_layers=len(layers))
        out.append(Workload(
            name=f"batch scene {i}: {W}x{H}, {len(layers)} layer(s), sun zenith {zen:.1f} deg", frame=frame,
            atmosphere=atm, clouds=clouds, layers=layers, winds=layer_winds(layers), noise_seed=0,
            noise_format=abi.PS_NOISE_FP16, lut_dims=lut_dims(32), weather=weather_map(512),
            overcast=overcast_mask(layers), blue=blue_noise(),
            depth=ground_depth(frame, atm.planet_radius_m) if ground else None))
    return out


# ---------------------------------------------------------------------------
# differentiation workload
# ---------------------------------------------------------------------------
def workload_adjoint(index: int = 0, width: int = 640, height: int = 360) -> Workload:
    """The cloud-differentiation workload: the full-frame two-layer stack
    (cumulus 1500-2300 m + cirrus 8000-9000 m), 128/6 steps, FP16 noise,
    sun zenith 65 deg, camera of sweep frame `index`, at 640x360."""
    w = workload_full_frame(index, width=width, height=height)
    w.name = f"cloud adjoint {width}x{height}, cumulus + cirrus, 128/6 steps, frame {index}"
    return w


@dataclasses.dataclass
class AdjointTarget:
    """Parameters of the target render of the differentiation workload."""
    density_scale: np.ndarray      # per layer
    extinction_scale: np.ndarray   # per layer
    coverage: np.ndarray           # per layer
    delta: np.ndarray              # (128, 128, 128) float32, (z, h, x)


def adjoint_target(layers, seed: int = 0xC4) -> AdjointTarget:
    """Perturbed parameters the target image is rendered with: per layer
    density_scale x U[0.7, 1.3], extinction scale U[0.7, 1.3], coverage +
    U[-0.15, 0.15] (clamped to [0, 1]) from PCG32 `seed`; the density
    perturbation volume is i.i.d. N(0, 0.05^2) from numpy's PCG64 with the
    same seed."""
    rng = Pcg32(seed)
    n = len(layers)
    u = rng.uniform(3 * n).reshape(3, n)
    ds = np.array([L.density_scale for L in layers]) * (0.7 + 0.6 * u[0])
    kap = 0.7 + 0.6 * u[1]
    cov = np.clip(np.array([L.coverage for L in layers]) + (-0.15 + 0.3 * u[2]), 0.0, 1.0)
    delta = np.random.default_rng(seed).normal(0.0, 0.05, (128, 128, 128)).astype(np.float32)
    return AdjointTarget(ds.astype(np.float32), kap.astype(np.float32), cov.astype(np.float32), delta)
