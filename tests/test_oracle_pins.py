"""Independent pins of the oracle functions that the closed forms of
test_oracle.py leave open (CPU only), each with a mutation check.

Every pin is a function that raises AssertionError when the oracle disagrees
with an expectation fixed by mathematics or by the algorithm's definition:

  sun in-scatter chain     a numpy step sum over a constant-density column
  (march_to_light, Beer-    (closed-form NDF, FP64 geometry, the pinned
  Powder, octaves, sun      phase function) at the configured steps, and the
  transmittance, ambient,   continuous radiative-transfer integral the march
  AP on luminance)          converges to at 2048/256 steps
  LUT-sampled light terms   a LUT linear in (u, v) read at the LUT
                            parametrisations pinned by their round trips
  multi-scatter LUT         constant density and T = 1: F / (1 - F) with F
                            from numpy ray-sphere geometry
  sky-view / AP LUTs        constant density, T = 1, constant MS: the march
                            telescopes to in / sigma_t (1 - e^(-sigma_t t))
  sun disk                  the limb-darkened disk integrates to 0.8 E
  noise                     Perlin: 0.5 and the hashed gradient at lattice
                            points (all eight cells around them); Worley:
                            brute-force nearest feature; FBM channel
                            assembly; curl map divergence-free
  blue noise                the greedy placement's first nine texels, the
                            exact rank histogram, suppressed low frequencies

The mutation checks build the oracle with a planted slip (-DORACLE_SLIP=k,
oracle/Makefile `slip` target) and require the pin of that function to fail,
so each pin is shown to catch the mistake it is meant to catch.
"""
import concurrent.futures as cf
import contextlib
import functools
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from pedalsky_amd import abi, scene

ORACLE_DIR = Path(oracle.__file__).resolve().parent
LW = np.array([0.2126, 0.7152, 0.0722])
M32 = 0xFFFFFFFF


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def _smoothstep(e0, e1, x):
    t = np.clip((x - e0) / (e1 - e0), 0.0, 1.0)
    return t * t * (3 - 2 * t)


def _ndf_cumulus(h):
    """Cumulus vertical profile: rising edge over [0, 0.07], falling over
    [0.2, 1] (cloud_density_common.wgsl:49-91, reversed-edge smoothstep)."""
    return _smoothstep(0.0, 0.07, h) * _smoothstep(1.0, 0.2, h)


def _transparent(atm):
    for k in ("rayleigh_scattering", "mie_scattering", "mie_absorption", "ozone_absorption"):
        arr = getattr(atm, k)
        for i in range(3):
            arr[i] = 0.0


def _constant_density_atmosphere():
    """Earth coefficients with scale heights of 1e12 m and no ozone: the
    density factors are 1 to within 1e-7 inside the atmosphere."""
    a = scene.earth_atmosphere()
    a.rayleigh_scale_height_m = 1e12
    a.mie_scale_height_m = 1e12
    for i in range(3):
        a.ozone_absorption[i] = 0.0
    return a


def _coeffs(a):
    br = np.array(a.rayleigh_scattering[:3], dtype=np.float64)
    bm = np.array(a.mie_scattering[:3], dtype=np.float64)
    ba = np.array(a.mie_absorption[:3], dtype=np.float64)
    return br, bm, br + bm + ba


def _sphere_hits(o, d, radius):
    """Both roots of |o + t d| = radius (d unit), NaN where the ray misses."""
    b = o @ d if o.ndim == 1 else np.einsum("...i,...i", o, d)
    r = np.linalg.norm(o, axis=-1)
    disc = b * b - (r - radius) * (r + radius)
    sq = np.sqrt(np.where(disc >= 0, disc, np.nan))
    return -b - sq, -b + sq


def _boundary_distance(p, d, R, top):
    """Distance from p along unit d to the top of the atmosphere, or to the
    ground when the ray meets it first (atmosphere.wgsl:117-132)."""
    _, t_top = _sphere_hits(p, d, top)
    g0, _ = _sphere_hits(p, d, R)
    ground = np.isfinite(g0) & (g0 > 0)
    return np.where(ground, g0, t_top), ground


def _hash(x, y, z=0, with_z=True):
    x, y, z = np.uint64(x), np.uint64(y), np.uint64(z)
    v = (x * np.uint64(0x27d4eb2d) + y * np.uint64(0x165667b1) +
         (z * np.uint64(0x9e3779b9) if with_z else np.uint64(0))) & np.uint64(M32)
    v = ((v ^ (v >> np.uint64(15))) * np.uint64(0x2c1b3c6d)) & np.uint64(M32)
    v = ((v ^ (v >> np.uint64(12))) * np.uint64(0x297a2d39)) & np.uint64(M32)
    return v ^ (v >> np.uint64(15))


def _rand3(x, y, z):
    """The hashed jitter of a lattice cell: three 16-bit fractions."""
    return np.array([float(_hash(x, y, z) & np.uint64(0xFFFF)), float(_hash(x + 1, y, z) & np.uint64(0xFFFF)),
                     float(_hash(x, y + 1, z) & np.uint64(0xFFFF))]) / 65535.0


def _feature_points(freq, offset=(0, 0, 0)):
    """Feature point of every cell (cell index + hashed jitter), cell units,
    shape (freq^3, 3); the jitter hash input is the cell index plus `offset`."""
    k = np.arange(freq)
    cz, cy, cx = np.meshgrid(k, k, k, indexing="ij")
    cells = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=-1)
    jit = np.array([_rand3(int(c[0]) + offset[0], int(c[1]) + offset[1], int(c[2]) + offset[2]) for c in cells])
    return cells + jit


@functools.lru_cache(maxsize=16)
def _features_cached(freq, offset):
    return _feature_points(freq, offset)


def _worley_brute(p, freq, offset=(0, 0, 0)):
    """1 - distance (cell units) to the nearest feature point over the whole
    periodic lattice (minimal image), clamped to [0, 1]."""
    fp = _features_cached(freq, tuple(offset))
    q = np.asarray(p, dtype=np.float64) * freq
    dv = fp[None, :, :] - q[:, None, :]
    dv -= freq * np.round(dv / freq)
    d = np.sqrt((dv * dv).sum(-1)).min(axis=1)
    return np.clip(1.0 - d, 0.0, 1.0)


def _luts(frame, atm, dims):
    return oracle.bake_luts(frame, atm, dims)


# ---------------------------------------------------------------------------
# planted-slip builds
# ---------------------------------------------------------------------------
SLIPS = {
    1: "octave n attenuates with od * b instead of od * b^n",
    2: "powder term exp(-sigma od) instead of exp(-2 sigma od)",
    3: "light-march samples at step ends instead of midpoints",
    4: "multi-scatter LUT without the 1 / (1 - f_ms) series",
    5: "aerial perspective without the sun illuminance",
    6: "ambient height gradient reversed",
    7: "Perlin corner index transposed (ix <-> iy)",
    8: "Worley neighbourhood missing the +z cells",
    9: "detail Worley hash offset added before the wrap",
    10: "sun-disk limb darkening 0.4 instead of 0.6",
    11: "sky-view multi-scatter weighted by extinction instead of scattering",
    12: "aerial-perspective froxel rows not flipped",
    13: "curl components from swapped derivatives",
    14: "Beer-Powder blend reversed",
    15: "sun transmittance LUT read with u and v swapped",
    16: "sky ambient sampled at the world position instead of the local up",
    17: "base G channel Worley FBM at frequency 4 instead of 2",
}


@functools.lru_cache(maxsize=1)
def _build_slips():
    def one(k):
        res = subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "slip", f"K={k}"], capture_output=True,
                             text=True)
        if res.returncode != 0:
            raise RuntimeError(res.stdout + res.stderr)
        return k, ORACLE_DIR / "_build" / f"liboracle_slip{k}.so"
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        return dict(ex.map(one, SLIPS))


@contextlib.contextmanager
def slipped(k):
    path = _build_slips()[k]
    saved = oracle._lib
    oracle._lib = oracle.open_library(path)
    try:
        yield
    finally:
        oracle._lib = saved


# ---------------------------------------------------------------------------
# 1. sun in-scatter chain over a constant-density column
# ---------------------------------------------------------------------------
SV_CONST = np.array([2e-3, 3e-3, 5e-3])
AP_CONST = np.array([40.0, 60.0, 90.0, 0.7])   # in-scatter rgb, transmittance a


def _column(sun_zenith, octaves, powder, steps=64, light_steps=6, ds=0.1, W=6, H=6):
    """A near-vertical view through a cumulus layer whose density is exactly
    ds * NDF(h): base noise R = 1 / GBA = 0, coverage 1, no erosion, no skew.
    Transparent atmosphere (T-LUT = 1), constant sky-view and AP LUTs."""
    w = scene.workload_small(width=W, height=H)
    w.frame = scene.make_frame(W, H, pitch_deg=88.0, yaw_deg=180.0, fov_deg=2.0, sun_zenith_deg=sun_zenith)
    _transparent(w.atmosphere)
    L = scene.synth_layer(0, 1500.0, 2300.0, 1.0, density_scale=ds, droplet_diameter_um=20.0)
    L.coverage = 1.0
    w.layers = [L]
    w.clouds = scene.cloud_params(steps=steps, light_steps=light_steps, octaves=octaves, detail_strength=0.0,
                                  skew=0.0, n_layers=1, powder_strength=powder, forward_bias=0.0)
    w.depth = None
    w.overcast = np.zeros(128 * 128, dtype=np.uint8)
    base = np.zeros((128, 128, 128, 4), dtype=np.float32)
    base[..., 0] = 1.0
    det = np.zeros((32, 32, 32, 4), dtype=np.float32)
    curl = np.full((128, 128, 2), 0.5, dtype=np.float32)
    wm = np.zeros((8, 8, 2), dtype=np.float16)
    wm[..., 0] = 1.0
    w.weather = wm.view(np.uint16)
    luts = _luts(w.frame, w.atmosphere, abi.LutDims(64, 16, 16, 16, 24, 16, 8, 8, 8))
    assert np.all(luts["T"][..., :3] == 1.0)
    luts["SV"] = np.zeros_like(luts["SV"])
    luts["SV"][..., :3] = SV_CONST
    luts["AP"] = np.zeros_like(luts["AP"])
    luts["AP"][...] = AP_CONST
    return w, oracle.Frame(w, base, det, curl, luts=luts)


def _column_expected(w, light_steps_override=None):
    """Numpy step sum of the march for every pixel of a _column workload."""
    lib = oracle.load()
    c = w.clouds
    L = w.layers[0]
    a = w.atmosphere
    R = float(a.planet_radius_m)
    centre = np.array([0.0, -R, 0.0])
    cam, dirs = scene.pixel_rays(w.frame)
    sun = np.array(w.frame.sun_direction[:3], dtype=np.float64)
    E = np.array(w.frame.sun_illuminance[:3], dtype=np.float64)
    sig_s = np.array(c.sigma_s[:3], dtype=np.float64)
    sig_t = sig_s + np.array(c.sigma_a[:3], dtype=np.float64)
    ds, base_m, top_m = float(L.density_scale), float(L.base_m), float(L.top_m)
    thick = top_m - base_m
    steps, nl = int(c.cloud_steps), int(c.light_steps)
    H, W = dirs.shape[:2]
    out = np.zeros((H, W, 3))
    for iy in range(H):
        for ix in range(W):
            d = dirs[iy, ix]
            jitter = w.blue[(iy & 63) * 64 + (ix & 63)] / 255.0
            oc = cam - centre
            ta = _sphere_hits(oc, d, R + base_m)[1]
            tb = _sphere_hits(oc, d, R + top_m)[1]
            step = (tb - ta) / steps
            cos_th = float(d @ sun)
            lum = np.zeros(3)
            T = np.ones(3)
            tw = twn = 0.0
            for k in range(steps):
                t = ta + (jitter + k) * step
                p = cam + d * t
                alt = np.linalg.norm(p - centre) - R
                h = (alt - base_m) / thick
                rho = ds * _ndf_cumulus(h) if 0.0 <= h <= 1.0 + 1e-5 else 0.0
                if rho > 1e-3:
                    dist = max((top_m - alt) / max(sun[1], 0.05), 1.0)
                    dl = dist / nl
                    q = p[None, :] + sun[None, :] * ((np.arange(nl) + 0.5) * dl)[:, None]
                    hq = (np.linalg.norm(q - centre, axis=1) - R - base_m) / thick
                    rq = np.where((hq >= 0) & (hq <= 1), ds * _ndf_cumulus(hq), 0.0)
                    od = rq.sum() * dl
                    beer = np.exp(-sig_t * od)
                    powder = 1.0 - np.exp(-2.0 * sig_t * od)
                    energy = (1 - c.powder_strength) * beer + c.powder_strength * 2.0 * beer * powder
                    hn = min(max(h, 0.0), 1.0)
                    sun_in = np.zeros(3)
                    for n in range(int(c.multi_scatter_octaves)):
                        ph = lib.or_cloud_phase(cos_th, float(c.multi_scatter_c) ** n, float(L.droplet_diameter_um),
                                                float(c.droplet_diameter_bias), 0, hn)
                        sun_in += (float(c.multi_scatter_a) ** n) * ph * np.exp(
                            -sig_t * od * float(c.multi_scatter_b) ** n)
                    sun_in *= E  # transparent atmosphere: sun transmittance 1
                    amb = SV_CONST * E * float(c.ambient_strength) * (0.3 + 0.7 * hn)
                    S = sig_s * rho * (sun_in * energy + amb)
                    sig = np.maximum(rho * sig_t, 1e-6)
                    st = np.exp(-sig * step)
                    dl_ = T * S * (1 - st) / sig
                    lum += dl_
                    tw += (dl_ @ LW) * t
                    twn += dl_ @ LW
                    T *= st
                if T.max() < 0.01:
                    break
            opacity = 1.0 - T @ LW
            if opacity > 1e-4:
                lum = lum * AP_CONST[3] + AP_CONST[:3] * opacity
            out[iy, ix] = lum
    return out


def check_column_step_sum(sun_zenith, octaves, powder):
    w, fr = _column(sun_zenith, octaves, powder)
    _, lum, trn, _ = fr.render_region(0, 0, w.width, w.height)
    exp = _column_expected(w)
    assert np.all(trn[..., 1] < 0.9), "the column must be optically thick enough to matter"
    assert np.allclose(lum[..., :3], exp, rtol=1e-9, atol=0), np.abs(lum[..., :3] / exp - 1).max()


def check_column_continuous_limit():
    """At 2048 primary and 256 light steps the march must agree with the
    continuous solution L = int T(t) sigma_s rho(t) J(t) dt of the same
    column (sun at the zenith, centre pixel, near-vertical ray), evaluated
    on a 400k-point grid: T from the cumulative optical depth, J from the
    optical depth to the layer top, both integrals of the closed-form NDF."""
    w, fr = _column(0.0, 4, 1.0, steps=2048, light_steps=256, W=1, H=1)
    w.frame = scene.make_frame(1, 1, pitch_deg=89.0, yaw_deg=180.0, fov_deg=0.5, sun_zenith_deg=0.0)
    fr = oracle.Frame(w, fr.base, fr.detail, fr.curl, luts=fr.luts)
    _, lum, _, _ = fr.render_region(0, 0, 1, 1)
    lib = oracle.load()
    c, L, a = w.clouds, w.layers[0], w.atmosphere
    R = float(a.planet_radius_m)
    centre = np.array([0.0, -R, 0.0])
    cam, dirs = scene.pixel_rays(w.frame)
    d = dirs[0, 0]
    sun = np.array(w.frame.sun_direction[:3], dtype=np.float64)
    E = np.array(w.frame.sun_illuminance[:3], dtype=np.float64)
    sig_s = np.array(c.sigma_s[:3], dtype=np.float64)
    sig_t = sig_s + np.array(c.sigma_a[:3], dtype=np.float64)
    ds, base_m, top_m = float(L.density_scale), float(L.base_m), float(L.top_m)
    oc = cam - centre
    ta = _sphere_hits(oc, d, R + base_m)[1]
    tb = _sphere_hits(oc, d, R + top_m)[1]
    t = np.linspace(ta, tb, 400001)
    alt = np.linalg.norm(cam[None, :] + t[:, None] * d[None, :] - centre, axis=1) - R
    h = (alt - base_m) / (top_m - base_m)
    rho = ds * _ndf_cumulus(h)
    rho_d = np.where(rho > 1e-3, rho, 0.0)      # only dense samples attenuate and scatter
    # optical depth to the layer top along the (vertical) sun ray: the
    # antiderivative of the profile in altitude on a fine grid
    hg = np.linspace(0.0, 1.0, 400001)
    fg = ds * _ndf_cumulus(hg) * (top_m - base_m)
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (fg[1:] + fg[:-1]) * np.diff(hg))])
    od = cum[-1] - np.interp(np.clip(h, 0.0, 1.0), hg, cum)
    cos_th = float(d @ sun)
    hn = np.clip(h, 0.0, 1.0)
    tau = np.concatenate([[0.0], np.cumsum(0.5 * (rho_d[1:] + rho_d[:-1]) * np.diff(t))])
    lum_c = np.zeros(3)
    for ch in range(3):
        sun_in = np.zeros_like(t)
        for n in range(int(c.multi_scatter_octaves)):
            ph = np.array([lib.or_cloud_phase(cos_th, float(c.multi_scatter_c) ** n, float(L.droplet_diameter_um),
                                              float(c.droplet_diameter_bias), 0, float(x)) for x in hn[::4000]])
            ph = np.interp(hn, hn[::4000], ph)
            sun_in += float(c.multi_scatter_a) ** n * ph * np.exp(-sig_t[ch] * od * float(c.multi_scatter_b) ** n)
        beer = np.exp(-sig_t[ch] * od)
        energy = 2.0 * beer * (1.0 - np.exp(-2.0 * sig_t[ch] * od))
        J = E[ch] * sun_in * energy + SV_CONST[ch] * E[ch] * float(c.ambient_strength) * (0.3 + 0.7 * hn)
        f = np.exp(-sig_t[ch] * tau) * sig_s[ch] * rho_d * J
        lum_c[ch] = np.sum(0.5 * (f[1:] + f[:-1]) * np.diff(t))
    T = np.exp(-sig_t * tau[-1])
    opacity = 1.0 - T @ LW
    exp = lum_c * AP_CONST[3] + AP_CONST[:3] * opacity
    rel = np.abs(lum[0, 0, :3] / exp - 1.0)
    assert rel.max() < 3e-4, rel


# ---------------------------------------------------------------------------
# 2. LUT-sampled light terms of a dense sample
# ---------------------------------------------------------------------------
def _linear_lut(h, w, a0, au, av):
    u = (np.arange(w) + 0.5) / w
    v = (np.arange(h) + 0.5) / h
    out = np.zeros((h, w, 4))
    for ch in range(3):
        out[..., ch] = a0[ch] + au[ch] * u[None, :] + av[ch] * v[:, None]
    out[..., 3] = 1.0
    return out


def _lookup_frame():
    w = scene.workload_small(width=4, height=4)
    w.frame = scene.make_frame(4, 4, sun_zenith_deg=40.0, sun_azimuth_deg=180.0)
    w.depth = None
    base = np.zeros((128, 128, 128, 4), dtype=np.float32)
    det = np.zeros((32, 32, 32, 4), dtype=np.float32)
    curl = np.zeros((128, 128, 2), dtype=np.float32)
    return w, base, det, curl


def check_sun_transmittance_lookup():
    """A transmittance LUT linear in (u, v) is read back exactly by bilinear
    filtering, so the sun transmittance at p must equal the linear function
    at the LUT parametrisation of (|p - c|, cos(up, sun)), which the uv round
    trip and the texel = exp(-OD) pins fix; below the horizon it is 0."""
    w, base, det, curl = _lookup_frame()
    a = w.atmosphere
    R = float(a.planet_radius_m)
    T = _linear_lut(64, 256, [0.3, 0.35, 0.4], [0.4, 0.3, 0.2], [0.2, 0.25, 0.1])
    luts = {"T": T, "SV": np.zeros((16, 24, 4)), "AP": np.zeros((8, 8, 8, 4))}
    fr = oracle.Frame(w, base, det, curl, luts=luts)
    rng = np.random.default_rng(11)
    pts = np.stack([rng.uniform(-2e5, 2e5, 64), rng.uniform(300.0, 9000.0, 64), rng.uniform(-2e5, 2e5, 64)], 1)
    got = np.zeros((64, 3))
    oracle.load().or_sun_transmittance_at(C_ref(fr), pts.ctypes.data, 64, got.ctypes.data)
    sun = np.array(w.frame.sun_direction[:3], dtype=np.float64)
    uv = np.zeros(2)
    n_in = 0
    for i, p in enumerate(pts):
        q = p + np.array([0.0, R, 0.0])
        r = np.linalg.norm(q)
        mu = float(q @ sun) / r
        oracle.load().or_transmittance_uv(oracle._ref(a), r, mu, uv.ctypes.data)
        if not (0.5 / 256 <= uv[0] <= 1 - 0.5 / 256 and 0.5 / 64 <= uv[1] <= 1 - 0.5 / 64):
            continue
        n_in += 1
        exp = [[0.3, 0.35, 0.4][c] + [0.4, 0.3, 0.2][c] * uv[0] + [0.2, 0.25, 0.1][c] * uv[1] for c in range(3)]
        assert np.allclose(got[i], exp, rtol=1e-9), (i, got[i], exp)
    assert n_in > 40
    # the sun 5 degrees below the horizon of a point at 300 m: Earth shadow
    w.frame = scene.make_frame(4, 4, sun_zenith_deg=95.0, sun_azimuth_deg=180.0)
    fr = oracle.Frame(w, base, det, curl, luts=luts)
    p = np.array([0.0, 300.0, 0.0])
    oracle.load().or_sun_transmittance_at(C_ref(fr), p.ctypes.data, 1, got.ctypes.data)
    assert np.all(got[0] == 0.0)


def check_sky_ambient_lookup():
    """The sky ambient of a dense sample is the sky-view LUT at the
    direction of the local up seen from the sample's radius, times E: a LUT
    linear in (u, v) gives the linear function at the sky-view
    parametrisation that its round-trip pin fixes."""
    w, base, det, curl = _lookup_frame()
    a = w.atmosphere
    R = float(a.planet_radius_m)
    SV = _linear_lut(108, 192, [0.1, 0.12, 0.15], [0.2, 0.1, 0.05], [0.3, 0.2, 0.1])
    luts = {"T": np.ones((64, 256, 4)), "SV": SV, "AP": np.zeros((8, 8, 8, 4))}
    fr = oracle.Frame(w, base, det, curl, luts=luts)
    rng = np.random.default_rng(12)
    pts = np.stack([rng.uniform(-4e5, 4e5, 64), rng.uniform(300.0, 9000.0, 64), rng.uniform(-4e5, 4e5, 64)], 1)
    got = np.zeros((64, 3))
    oracle.load().or_sky_ambient_at(C_ref(fr), pts.ctypes.data, 64, got.ctypes.data)
    sun = np.array(w.frame.sun_direction[:3], dtype=np.float64)
    E = np.array(w.frame.sun_illuminance[:3], dtype=np.float64)
    uv = np.zeros(2)
    n_in = 0
    for i, p in enumerate(pts):
        q = p + np.array([0.0, R, 0.0])
        r = np.linalg.norm(q)
        up = q / r
        oracle.load().or_skyview_uv(oracle._ref(a), up.ctypes.data, sun.ctypes.data, r, uv.ctypes.data)
        if not (0.5 / 192 <= uv[0] <= 1 - 0.5 / 192 and 0.5 / 108 <= uv[1] <= 1 - 0.5 / 108):
            continue
        n_in += 1
        exp = np.array([[0.1, 0.12, 0.15][c] + [0.2, 0.1, 0.05][c] * uv[0] + [0.3, 0.2, 0.1][c] * uv[1]
                        for c in range(3)]) * E
        assert np.allclose(got[i], exp, rtol=1e-9), (i, got[i], exp)
    assert n_in > 40


def C_ref(fr):
    import ctypes
    return ctypes.byref(fr.inputs)


# ---------------------------------------------------------------------------
# 3. LUT texel values under constant density
# ---------------------------------------------------------------------------
def _fibonacci(n=64):
    i = np.arange(n)
    z = 1.0 - 2.0 * (i + 0.5) / n
    r = np.sqrt(np.maximum(1.0 - z * z, 0.0))
    phi = 2.0 * math.pi * np.mod(i * 0.6180339887498949, 1.0)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=-1)


def check_multiscatter_texels():
    """Constant density and T = 1: every direction's second-order
    in-scatter and transfer telescope to sigma_s / sigma_t (1 -
    exp(-sigma_t t_max)); with the sun visible along every sample both
    equal F, so the texel is F / (1 - min(F, 0.95))."""
    a = _constant_density_atmosphere()
    br, bm, st = _coeffs(a)
    ss = br + bm
    R, top = float(a.planet_radius_m), float(a.atmosphere_top_m)
    W = H = 16
    T = np.ones((16, 64, 4))
    ms = oracle.bake_multiscatter(a, T, W, H)
    dirs = _fibonacci(64)
    checked = 0
    for y in range(H):
        r = R + (y + 0.5) / H * (top - R)
        p0 = np.array([0.0, r, 0.0])
        tmax, _ = _boundary_distance(np.broadcast_to(p0, dirs.shape), dirs, R, top)
        for x in range(W):
            sun_cos = (x + 0.5) / W * 2 - 1
            if sun_cos < 0.25:
                continue
            sun = np.array([math.sqrt(1 - sun_cos ** 2), sun_cos, 0.0])
            # precondition: the sun is above every sample's horizon
            ts = (np.arange(20) + 0.5)[None, :, None] / 20 * tmax[:, None, None]
            ps = p0 + dirs[:, None, :] * ts
            rp = np.linalg.norm(ps, axis=-1)
            assert np.all(ps @ sun / rp > -np.sqrt(np.maximum(1 - (R / rp) ** 2, 0)) + 1e-3)
            F = (ss[None, :] / st[None, :] * (1 - np.exp(-st[None, :] * tmax[:, None]))).mean(axis=0)
            exp = F / (1 - np.minimum(F, 0.95))
            assert np.allclose(ms[y, x, :3], exp, rtol=2e-6), (y, x, ms[y, x, :3], exp)
            checked += 1
    assert checked == H * 6


def check_skyview_texels():
    """Constant density, T = 1 and a constant multi-scatter LUT M: the
    in-scatter (beta_R p_R + beta_M p_M + (beta_R + beta_M) M) is constant
    along the ray, so each texel is in / sigma_t (1 - exp(-sigma_t t_max))
    plus, where the ray meets the ground, albedo / pi max(n.l, 0)
    exp(-sigma_t t_max)."""
    a = _constant_density_atmosphere()
    br, bm, st = _coeffs(a)
    R, top = float(a.planet_radius_m), float(a.atmosphere_top_m)
    f = scene.make_frame(16, 16, camera=(0.0, 500.0, 0.0), sun_zenith_deg=30.0, sun_azimuth_deg=150.0)
    M = np.array([0.1, 0.2, 0.3])
    MS = np.zeros((8, 8, 4))
    MS[..., :3] = M
    W, H = 48, 28
    sv = oracle.bake_skyview(f, a, np.ones((16, 64, 4)), MS, W, H)
    lib = oracle.load()
    sun = np.array(f.sun_direction[:3], dtype=np.float64)
    p = np.array(f.camera_position_world[:3], dtype=np.float64) + np.array([0.0, R, 0.0])
    r = np.linalg.norm(p)
    zha = math.pi - math.asin(R / r)
    sun_az = math.atan2(sun[0], sun[2])
    albedo = np.array(a.ground_albedo[:3], dtype=np.float64)
    for y in range(H):
        v = (y + 0.5) / H
        c = 2 * v if v < 0.5 else 2 * (v - 0.5)
        vza = c * c * zha if v < 0.5 else zha + c * c * (math.pi - zha)
        for x in range(W):
            az = sun_az + (x + 0.5) / W * 2 * math.pi
            d = np.array([math.sin(vza) * math.sin(az), math.cos(vza), math.sin(vza) * math.cos(az)])
            tmax, ground = _boundary_distance(p, d, R, top)
            tmax = float(tmax)
            cth = float(d @ sun)
            ins = br * lib.or_phase_rayleigh(cth) + bm * lib.or_phase_mie(cth, float(a.mie_g)) + (br + bm) * M
            exp = ins / st * (1 - np.exp(-st * tmax))
            if ground:
                n = (p + d * tmax) / np.linalg.norm(p + d * tmax)
                exp = exp + albedo / 3.14159265 * max(float(n @ sun), 0.0) * np.exp(-st * tmax)
            assert np.allclose(sv[y, x, :3], exp, rtol=2e-6), (y, x, sv[y, x, :3], exp)


def check_aerial_texels():
    """Constant density, T = 1 and constant multi-scatter: each froxel of a
    ray that stays above the ground holds E in / sigma_t (1 - exp(-sigma_t
    t)) and the luminance-weighted exp(-sigma_t t), t = 50 m 2000^((z +
    0.5) / D); the froxel directions come from a pinhole model of the camera
    (row 0 at the top of the screen)."""
    a = _constant_density_atmosphere()
    br, bm, st = _coeffs(a)
    pitch, yaw, fov = 10.0, 30.0, 60.0
    f = scene.make_frame(64, 36, camera=(0.0, 500.0, 0.0), pitch_deg=pitch, yaw_deg=yaw, fov_deg=fov,
                         sun_zenith_deg=40.0, sun_azimuth_deg=180.0)
    M = np.array([0.1, 0.2, 0.3])
    MS = np.zeros((8, 8, 4))
    MS[..., :3] = M
    W, H, D = 12, 10, 8
    ap = oracle.bake_ap(f, a, np.ones((16, 64, 4)), MS, W, H, D)
    lib = oracle.load()
    sun = np.array(f.sun_direction[:3], dtype=np.float64)
    E = np.array(f.sun_illuminance[:3], dtype=np.float64)
    fw = scene.fly_camera_forward(math.radians(pitch), math.radians(yaw))
    s = np.cross(fw, [0.0, 1.0, 0.0])
    s /= np.linalg.norm(s)
    u = np.cross(s, fw)
    th = math.tan(math.radians(fov) / 2)
    checked = 0
    for y in range(H):
        for x in range(W):
            nx = (x + 0.5) / W * 2 - 1
            ny = 1 - (y + 0.5) / H * 2
            d = fw + nx * th * (64 / 36) * s + ny * th * u
            d /= np.linalg.norm(d)
            if d[1] < 0.0:
                continue
            cth = float(d @ sun)
            ins = br * lib.or_phase_rayleigh(cth) + bm * lib.or_phase_mie(cth, float(a.mie_g)) + (br + bm) * M
            for z in range(D):
                t = 50.0 * 2000.0 ** ((z + 0.5) / D)
                exp = E * ins / st * (1 - np.exp(-st * t))
                assert np.allclose(ap[z, y, x, :3], exp, rtol=2e-6), (z, y, x, ap[z, y, x, :3], exp)
                assert ap[z, y, x, 3] == pytest.approx(float(np.exp(-st * t) @ LW), rel=1e-6)
                checked += 1
    assert checked >= W * H * D // 3


# ---------------------------------------------------------------------------
# 4. sun disk
# ---------------------------------------------------------------------------
def check_sun_disk_integral():
    """With a transparent atmosphere the sky is the sun disk alone, radiance
    E limb(theta / alpha) / (pi alpha^2) with limb(x) = 1 - 0.6 (1 -
    sqrt(1 - x^2)); over the disk, int limb dOmega = pi alpha^2 (1 - 0.6 /
    3), so the image's radiance summed over pixel solid angles is 0.8 E."""
    N, fov = 512, 0.7
    w = scene.workload_small(width=N, height=N)
    w.frame = scene.make_frame(N, N, pitch_deg=40.0, yaw_deg=180.0, fov_deg=fov, sun_zenith_deg=50.0)
    _transparent(w.atmosphere)
    w.clouds = scene.cloud_params(steps=4, n_layers=0)
    w.depth = None
    w.overcast = np.zeros(128 * 128, dtype=np.uint8)
    luts = _luts(w.frame, w.atmosphere, abi.LutDims(64, 16, 16, 16, 24, 16, 8, 8, 8))
    z = np.zeros((1, 1, 1, 4), dtype=np.float32)
    fr = oracle.Frame(w, np.zeros((128, 128, 128, 4), np.float32), np.zeros((32, 32, 32, 4), np.float32),
                      np.zeros((128, 128, 2), np.float32), luts=luts)
    del z
    comp, _, _, _ = fr.render_region(0, 0, N, N)
    t = math.tan(math.radians(fov) / 2)
    c = (np.arange(N) + 0.5) / N * 2 - 1
    nx, ny = np.meshgrid(c * t, c * t)
    domega = (2 * t / N) ** 2 / (1 + nx ** 2 + ny ** 2) ** 1.5
    E = float(w.frame.sun_illuminance[0])
    total = (comp[..., :3] * domega[..., None]).sum(axis=(0, 1)) / E
    assert np.allclose(total, 0.8, rtol=2e-3), total


# ---------------------------------------------------------------------------
# 5. noise
# ---------------------------------------------------------------------------
def check_perlin_lattice():
    """At a lattice point the Perlin tile is 0.5 and, in every one of the
    eight cells around it, its first-order behaviour is 0.5 + 0.5 freq g.dp
    with g the normalised hashed gradient of that lattice point (the fade
    weights of all other corners are O(dp^3))."""
    lib = oracle.load()
    eps = 1e-5
    for freq in (4, 8, 16):
        for c in ((0, 0, 0), (1, 2, 3), (freq - 1, 0, freq - 2), (3, freq - 1, 1)):
            w = [v % freq for v in c]
            g = 2.0 * _rand3(*w) - 1.0
            g /= np.linalg.norm(g)
            p0 = np.array(c, dtype=np.float64) / freq
            assert lib.or_perlin_tile(*p0, freq, 0) == pytest.approx(0.5, abs=1e-15)
            for sx in (-1, 1):
                for sy in (-1, 1):
                    for sz in (-1, 1):
                        dp = eps * np.array([sx, 0.7 * sy, 0.4 * sz]) / freq
                        got = lib.or_perlin_tile(*(p0 + dp), freq, 0)
                        exp = 0.5 + 0.5 * freq * float(g @ dp)
                        assert abs(got - exp) < 1e-12, (freq, c, (sx, sy, sz), got, exp)
            # tiling: the same lattice point one period away
            assert lib.or_perlin_tile(*(p0 + 1.0 + eps), freq, 0) == pytest.approx(
                lib.or_perlin_tile(*(p0 + eps), freq, 0), abs=1e-12)


def check_worley_brute_force():
    """Worley = 1 - distance to the nearest hashed feature point: compared
    with a brute-force search over the whole periodic lattice (the 27-cell
    neighbourhood finds it at nearly every point and can never find a nearer
    one); exactly 1 at the feature points. Base (no hash offset) and detail
    (offset 73, 19, 41 added to the wrapped cell) variants."""
    lib = oracle.load()
    rng = np.random.default_rng(21)
    for freq, off in ((2, (0, 0, 0)), (4, (0, 0, 0)), (14, (0, 0, 0)), (8, (73, 19, 41))):
        pts = rng.random((300, 3))
        got = np.array([lib.or_worley_tile(*p, freq, *off, 0) for p in pts])
        bf = _worley_brute(pts, freq, off)
        assert np.all(got <= bf + 1e-12)
        assert np.mean(np.abs(got - bf) < 1e-12) > 0.97, np.mean(np.abs(got - bf) < 1e-12)
        fp = _features_cached(freq, off)[rng.integers(0, freq ** 3, 10)] / freq
        for p in fp:
            assert lib.or_worley_tile(*np.mod(p, 1.0), freq, *off, 0) == pytest.approx(1.0, abs=1e-12)


def check_noise_channel_assembly():
    """The base volume texels are the published channel formulas over the
    pinned primitives: R = clamp(remap(Perlin FBM(4, 8, 16 at 1/2, 1/4,
    1/8), Worley FBM(4) - 1, 1, 0, 1)), G / B / A = Worley FBM at base
    frequency 2 / 8 / 14 with weights 0.625 / 0.25 / 0.125 of f, 2f, 4f.
    The detail texels are Worley FBM 2 / 8 / 16 with the hash offset."""
    lib = oracle.load()
    base = oracle.noise_base(0, 37, 38)[0]          # slice z = 37
    det = oracle.noise_detail(0)
    rng = np.random.default_rng(31)

    def wfbm(p, f, off=(0, 0, 0)):
        return (0.625 * lib.or_worley_tile(*p, f, *off, 0) + 0.25 * lib.or_worley_tile(*p, 2 * f, *off, 0) +
                0.125 * lib.or_worley_tile(*p, 4 * f, *off, 0))
    for _ in range(25):
        x, y = rng.integers(0, 128, 2)
        p = (np.array([x, y, 37]) + 0.5) / 128
        perlin = (0.5 * lib.or_perlin_tile(*p, 4, 0) + 0.25 * lib.or_perlin_tile(*p, 8, 0) +
                  0.125 * lib.or_perlin_tile(*p, 16, 0))
        wl = wfbm(p, 4)
        r = np.clip((perlin - (wl - 1.0)) / max(1.0 - (wl - 1.0), 1e-5), 0.0, 1.0)
        exp = [r, wfbm(p, 2), wfbm(p, 8), wfbm(p, 14)]
        assert np.allclose(base[y, x], exp, rtol=0, atol=1e-14), (x, y, base[y, x], exp)
    for _ in range(25):
        x, y, z = rng.integers(0, 32, 3)
        p = (np.array([x, y, z]) + 0.5) / 32
        exp = [wfbm(p, 2, (73, 19, 41)), wfbm(p, 8, (73, 19, 41)), wfbm(p, 16, (73, 19, 41)), 0.0]
        assert np.allclose(det[z, y, x], exp, rtol=0, atol=1e-14)


def check_curl_map():
    """The curl map is (d psi / dy, -d psi / dx) of a tileable 2-D Perlin
    potential by central differences at the texel spacing: those commute,
    so its discrete divergence vanishes wherever no component is clamped.
    The 2-D Perlin tile is 0 at lattice points with gradient freq g (g the
    normalised hashed 2-D gradient), and the first component grows with
    psi along +y."""
    lib = oracle.load()
    cm = oracle.noise_curl(0)
    vx, vy = 2.0 * cm[..., 0] - 1.0, 2.0 * cm[..., 1] - 1.0   # [y, x]
    ok = (np.abs(vx) < 0.999) & (np.abs(vy) < 0.999)
    ok = ok & np.roll(ok, 1, 0) & np.roll(ok, -1, 0) & np.roll(ok, 1, 1) & np.roll(ok, -1, 1)
    div = (np.roll(vx, -1, 1) - np.roll(vx, 1, 1)) + (np.roll(vy, -1, 0) - np.roll(vy, 1, 0))
    scale = np.abs(np.roll(vx, -1, 1)) + np.abs(np.roll(vy, -1, 0)) + 1e-3
    assert ok.sum() > 50, ok.sum()
    assert np.max(np.abs(div[ok]) / scale[ok]) < 1e-9, np.max(np.abs(div[ok]) / scale[ok])
    eps = 1e-6
    for freq in (4, 8):
        for c in ((0, 0), (1, 3), (freq - 1, 2)):
            a = float(_hash(c[0] % freq, c[1] % freq, with_z=False) & np.uint64(0xFFFF)) / 65535 * 2 - 1
            b = float(_hash(c[0] % freq + 1, c[1] % freq, with_z=False) & np.uint64(0xFFFF)) / 65535 * 2 - 1
            g = np.array([a, b]) / math.hypot(a, b)
            p0 = np.array(c, dtype=np.float64) / freq
            assert lib.or_perlin2(*p0, freq, 0) == pytest.approx(0.0, abs=1e-15)
            for sx in (-1, 1):
                for sy in (-1, 1):
                    dp = eps * np.array([sx, 0.6 * sy]) / freq
                    assert abs(lib.or_perlin2(*(p0 + dp), freq, 0) - freq * float(g @ dp)) < 1e-14
    # orientation: texel (x, y) first component from psi at (x, y +- 1)
    psi = lambda px, py: lib.or_perlin2(px, py, 4, 0) + 0.5 * lib.or_perlin2(px, py, 8, 0)  # noqa: E731
    x, y = 37, 90
    px, py = (x + 0.5) / 128, (y + 0.5) / 128
    dpdy = (psi(px, py + 1 / 128) - psi(px, py - 1 / 128)) * 64
    assert cm[y, x, 0] == pytest.approx(min(max(0.5 * dpdy + 0.5, 0.0), 1.0), abs=1e-12)


# ---------------------------------------------------------------------------
# 6. blue noise (crates/ps-clouds/src/noise.rs:646-704)
# ---------------------------------------------------------------------------
def check_blue_noise(bn):
    """The greedy energy placement starts at texel 0 and, with a splat of
    Chebyshev radius 6 and ties broken in scan order, places its next eight
    samples 7 texels apart along row 0; ranks 0..8 quantise to 0, rank 9 to
    1. The map is a permutation of the ranks 0..4095 (its histogram is the
    quantised ramp's exactly), and its low spatial frequencies carry a
    small fraction of the power a white-noise map carries."""
    bn = np.asarray(bn, dtype=np.uint8).reshape(64, 64)      # [y, x]
    zeros = sorted((int(x), int(y)) for y, x in zip(*np.nonzero(bn == 0)))
    assert zeros == [(7 * k, 0) for k in range(9)], zeros
    ramp = np.clip(np.round(np.arange(4096, dtype=np.float64) / 4095 * 255), 0, 255).astype(np.int64)
    assert np.array_equal(np.bincount(bn.ravel(), minlength=256), np.bincount(ramp, minlength=256))
    f = np.abs(np.fft.fft2(bn.astype(np.float64) - bn.mean())) ** 2
    k = np.fft.fftfreq(64) * 64
    kr = np.hypot(k[None, :], k[:, None])
    low = f[(kr > 0) & (kr <= 6)].mean()
    assert low < 0.1 * f[kr > 0].mean(), low / f[kr > 0].mean()


# ---------------------------------------------------------------------------
# tests
# ---------------------------------------------------------------------------
COLUMN_CASES = {"1 octave, no powder, sun at zenith": (0.0, 1, 0.0),
                "4 octaves, powder, sun 60 deg": (60.0, 4, 1.0),
                "4 octaves, half powder, sun 30 deg": (30.0, 4, 0.5)}

PINS = {
    "column step sum: " + k: (functools.partial(check_column_step_sum, *v), {1: "octaves", 2: "powder", 3: None,
                                                                           6: None, 14: "full powder"})
    for k, v in COLUMN_CASES.items()
}
PINS.update({
    "column continuous limit": (check_column_continuous_limit, {1: None, 2: None, 3: None, 6: None, 14: None}),
    "sun transmittance lookup": (check_sun_transmittance_lookup, {15: None}),
    "sky ambient lookup": (check_sky_ambient_lookup, {16: None}),
    "multi-scatter texels": (check_multiscatter_texels, {4: None}),
    "sky-view texels": (check_skyview_texels, {11: None}),
    "aerial-perspective texels": (check_aerial_texels, {5: None, 12: None}),
    "sun disk integral": (check_sun_disk_integral, {10: None}),
    "Perlin lattice": (check_perlin_lattice, {7: None}),
    "Worley brute force": (check_worley_brute_force, {8: None, 9: None}),
    "noise channel assembly": (check_noise_channel_assembly, {17: None}),
    "curl map": (check_curl_map, {13: None}),
})


def _requires(name, k, cond):
    """Whether a column case exercises slip k (octave slips need octaves > 1,
    Beer-Powder slips a non-zero powder strength, a reversed blend a strength
    other than 1/2, where both orders agree)."""
    if cond is None:
        return True
    sz, octv, pw = COLUMN_CASES[name.split(": ", 1)[1]]
    if cond == "octaves":
        return octv > 1
    return pw == 1.0 if cond == "full powder" else pw > 0


MUTATIONS = [(name, k) for name, (_, slips) in PINS.items() for k, cond in slips.items() if _requires(name, k, cond)]


@pytest.mark.parametrize("name", list(PINS))
def test_pin_holds(name):
    PINS[name][0]()


@pytest.mark.parametrize("name, k", MUTATIONS)
def test_pin_catches_planted_slip(name, k):
    with slipped(k):
        with pytest.raises(AssertionError):
            PINS[name][0]()


def test_every_planted_slip_is_caught_by_some_pin():
    assert sorted({k for _, k in MUTATIONS}) == sorted(SLIPS)


def test_blue_noise_generator():
    check_blue_noise(scene.blue_noise())


@pytest.mark.parametrize("slip", ["transposed", "shuffled"])
def test_blue_noise_pin_catches_slips(slip):
    bn = scene.blue_noise().reshape(64, 64)
    bad = bn.T.copy() if slip == "transposed" else np.random.default_rng(5).permutation(bn.ravel())
    with pytest.raises(AssertionError):
        check_blue_noise(bad)
