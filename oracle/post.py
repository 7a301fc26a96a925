"""Oracle: the ground pass, godrays, bloom and auto-exposure statistics,
restated in numpy. Test infrastructure only (tests/, smoke() and bench.py's
CPU baseline may use it; the product never does).

  * ground_lit: shaders/ground/pbr.wgsl fs_main (511-650) on a flat ground
    (normal +y) at the surface positions of a depth buffer: Voronoi albedo
    (101-194), Lagarde wet material (358-362), cloud-shadowed sun through the
    transmittance LUT with the Earth's shadow (atmosphere_lut_sampling.wgsl:
    6-22, atmosphere.wgsl:146-159), Cook-Torrance GGX/Smith/Schlick + Lambert
    (215-271), sky-view ambient blended to the overcast irradiance
    (368-497), puddles with ripples (281-351, 581-618) and snow (621-642).
  * godrays: radial.wgsl:51-92 and composite.wgsl:39-46, with the sun's
    screen position of crates/ps-godrays/src/lib.rs:358-386.
  * bloom: bright_pass.wgsl, downsample.wgsl, upsample.wgsl, composite.wgsl
    in the order of crates/ps-bloom/src/lib.rs:511-797.
  * exposure_stats: auto_exposure.comp.wgsl:31-73.

FP64, except two discrete choices made on float32 values exactly as an FP32
evaluation makes them: the Voronoi cell (float32 numpy, no fused operations)
and the godrays on-screen test (float32, glam's order). Images are sampled
bilinearly with ClampToEdge at texel centres (the reference's samplers).

Pinned by tests/test_post_oracle.py: BRDF energy and reciprocity closed
forms, the Voronoi cell against brute force over feature points, the ripple
field's ring geometry, the radial blur of a constant image, the pyramid's
partition of unity on constant images and its response to a single bright
texel, and the statistics against a direct sum.
"""
from __future__ import annotations

import numpy as np

PI = np.pi
f32 = np.float32
U32 = np.uint32


# ---------------------------------------------------------------------------
# sampling and geometry
# ---------------------------------------------------------------------------
def bilinear(img, u, v):
    """ClampToEdge bilinear tap of img (h, w, c) at uv (any shape) -> (..., c)."""
    img = np.asarray(img, np.float64)
    h, w = img.shape[:2]
    x = np.asarray(u, np.float64) * w - 0.5
    y = np.asarray(v, np.float64) * h - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    fx, fy = (x - x0)[..., None], (y - y0)[..., None]
    x0, y0 = x0.astype(np.int64), y0.astype(np.int64)
    xa, xb = np.clip(x0, 0, w - 1), np.clip(x0 + 1, 0, w - 1)
    ya, yb = np.clip(y0, 0, h - 1), np.clip(y0 + 1, 0, h - 1)
    top = img[ya, xa] * (1 - fx) + img[ya, xb] * fx
    bot = img[yb, xa] * (1 - fx) + img[yb, xb] * fx
    return top * (1 - fy) + bot * fy


def _mat(a):
    """Column-major 16 floats -> (4, 4) float64 matrix M with M @ v."""
    return np.asarray(a, np.float64).reshape(4, 4).T


def pixel_rays(frame):
    """Camera position (3,) and unit view rays (H, W, 3) through the pixel
    centres, unprojected at reverse-Z NDC 1."""
    W, H = int(frame.viewport_size[0]), int(frame.viewport_size[1])
    inv = _mat(frame.inv_view_proj)
    cam = np.array(frame.camera_position_world[:3], np.float64)
    nx = (np.arange(W) + 0.5) / W * 2 - 1
    ny = 1 - (np.arange(H) + 0.5) / H * 2
    X, Y = np.meshgrid(nx, ny)
    hom = np.stack([X, Y, np.ones_like(X), np.ones_like(X)], -1) @ inv.T
    d = hom[..., :3] / hom[..., 3:4] - cam
    return cam, d / np.linalg.norm(d, axis=-1, keepdims=True)


def surface_distance(frame, depth):
    """View distance of each pixel's opaque surface (inf where depth is 0)."""
    W, H = int(frame.viewport_size[0]), int(frame.viewport_size[1])
    inv = _mat(frame.inv_view_proj)
    cam = np.array(frame.camera_position_world[:3], np.float64)
    nx = (np.arange(W) + 0.5) / W * 2 - 1
    ny = 1 - (np.arange(H) + 0.5) / H * 2
    X, Y = np.meshgrid(nx, ny)
    dn = np.asarray(depth, np.float64)
    hom = np.stack([X, Y, dn, np.ones_like(X)], -1) @ inv.T
    w = np.where((dn > 0)[..., None], hom[..., 3:4], 1.0)
    q = hom[..., :3] / w - cam
    return np.where(dn > 0, np.linalg.norm(q, axis=-1), np.inf)


# ---------------------------------------------------------------------------
# the ground pass
# ---------------------------------------------------------------------------
PALETTE = np.array([
    [[0.10, 0.20, 0.06], [0.12, 0.24, 0.08], [0.08, 0.16, 0.05]],   # grass
    [[0.30, 0.22, 0.14], [0.36, 0.27, 0.18], [0.24, 0.18, 0.11]],   # bare soil
    [[0.05, 0.05, 0.06], [0.07, 0.07, 0.08], [0.04, 0.04, 0.05]],   # tarmac
    [[0.55, 0.48, 0.36], [0.62, 0.55, 0.41], [0.50, 0.43, 0.32]],   # sand
    [[0.08, 0.18, 0.28], [0.06, 0.15, 0.24], [0.10, 0.20, 0.30]]])  # water edge
ROUGHNESS = np.array([0.85, 0.88, 0.92, 0.78, 0.30])
F0 = np.array([0.04, 0.04, 0.04, 0.05, 0.08])


def hash21(x, y):
    """pbr.wgsl:94-99 on i32 coordinates reinterpreted as u32."""
    with np.errstate(over="ignore"):
        h = np.asarray(x, np.int64).astype(U32) * U32(0x27d4eb2d) + np.asarray(y, np.int64).astype(U32) * U32(0x165667b1)
        h = (h ^ (h >> U32(15))) * U32(0x2c1b3c6d)
        h = (h ^ (h >> U32(12))) * U32(0x297a2d39)
    return h ^ (h >> U32(15))


def voronoi_cell(px, pz):
    """Cell hash of the nearest jittered feature point of the 5 m tiling
    (pbr.wgsl:111-131), float32 without fused operations."""
    sx = np.asarray(px, f32) / f32(5.0)
    sz = np.asarray(pz, f32) / f32(5.0)
    flx, flz = np.floor(sx), np.floor(sz)
    cx, cz = flx.astype(np.int64), flz.astype(np.int64)
    frx, frz = sx - flx, sz - flz
    best_d2 = np.full(sx.shape, f32(4.0), f32)
    bx, bz = np.zeros(sx.shape, np.int64), np.zeros(sx.shape, np.int64)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            nx, nz = cx + dx, cz + dy
            jx = (hash21(nx, nz) & U32(0xffff)).astype(f32) / f32(65535.0)
            jz = (hash21(nx + 7919, nz + 1) & U32(0xffff)).astype(f32) / f32(65535.0)
            ex = (f32(dx) + jx) - frx
            ez = (f32(dy) + jz) - frz
            d2 = ex * ex + ez * ez
            take = d2 < best_d2
            best_d2 = np.where(take, d2, best_d2)
            bx, bz = np.where(take, nx, bx), np.where(take, nz, bz)
    return hash21(bx, bz)


def mix(a, b, t):
    """WGSL mix: a (1 - t) + b t."""
    return a * (1 - t) + b * t


def fade5(t):
    return t * t * t * (t * (t * 6 - 15) + 10)


def value_noise(x, y):
    """pbr.wgsl:200-209."""
    flx, fly = np.floor(x), np.floor(y)
    cx, cy = flx.astype(np.int64), fly.astype(np.int64)
    ux, uy = fade5(x - flx), fade5(y - fly)
    h = lambda a, b: (hash21(a, b) & U32(0xffff)).astype(np.float64) / 65535.0
    return mix(mix(h(cx, cy), h(cx + 1, cy), ux), mix(h(cx, cy + 1), h(cx + 1, cy + 1), ux), uy)


def _dot(a, b):
    return np.sum(a * b, axis=-1)


def ggx_d(ndh, rough):
    """GGX normal distribution (pbr.wgsl:215-220)."""
    a2 = (rough * rough) ** 2
    den = ndh * ndh * (a2 - 1) + 1
    return a2 / (PI * den * den + 1e-6)


def smith_g1(x, rough):
    """Schlick-Smith masking with k = (r + 1)^2 / 8 (pbr.wgsl:222-226)."""
    k = (rough + 1) ** 2 / 8
    return x / (x * (1 - k) + k)


def fresnel(cos_theta, f0):
    """Fresnel-Schlick (pbr.wgsl:233-237)."""
    x = np.clip(1 - cos_theta, 0, 1)
    return f0 + (1 - f0) * x ** 5


def brdf(n, v, l, albedo, rough, f0):
    """Cook-Torrance (pbr.wgsl:240-271): GGX D, Smith-Schlick G, Schlick F,
    Lambert with (1 - F); returns the BRDF (..., 3), not times n.l."""
    hv = v + l
    hv = hv / np.linalg.norm(hv, axis=-1, keepdims=True)
    ndv, ndl = np.maximum(_dot(n, v), 0), np.maximum(_dot(n, l), 0)
    ndh, vdh = np.maximum(_dot(n, hv), 0), np.maximum(_dot(v, hv), 0)
    d = ggx_d(ndh, rough)
    g = smith_g1(np.maximum(ndv, 1e-4), rough) * smith_g1(np.maximum(ndl, 1e-4), rough)
    f = fresnel(vdh, f0)
    spec = (d * g * f / np.maximum(4 * ndv * ndl, 1e-4))[..., None]
    return (1 - f)[..., None] * np.asarray(albedo) / PI + spec


def halton(i, base):
    """pbr.wgsl:279-289, elementwise over u32 arrays."""
    i = np.asarray(i, np.uint64).copy()
    f = np.ones(i.shape)
    r = np.zeros(i.shape)
    while np.any(i != 0):
        live = i != 0
        f = np.where(live, f / base, f)
        r = np.where(live, r + f * (i % base), r)
        i = i // base
    return r


def ripples(px, pz, time_s, intensity):
    """Ripple normal offset in the plane (pbr.wgsl:295-351): (..., 2)."""
    dn = np.zeros(np.shape(px) + (2,))
    if intensity <= 0:
        return dn
    cx, cz = np.floor(px).astype(np.int64), np.floor(pz).astype(np.int64)
    t_floor = np.floor(time_s)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            gx, gz = cx + dx, cz + dy
            h = hash21(gx, gz)
            for spawn in (0, 1):
                with np.errstate(over="ignore"):
                    salt = U32((spawn * 0x9e3779b9) & 0xFFFFFFFF)
                phase = halton(h ^ salt, 2)
                spawn_t = t_floor + phase
                spawn_t = np.where(spawn_t > time_s, spawn_t - 1, spawn_t)
                age = time_s - spawn_t
                ox, oz = halton(h, 3), halton(h ^ U32(0xa5a5a5a5), 5)
                tx, tz = px - (gx + ox), pz - (gz + oz)
                r = np.hypot(tx, tz)
                dr = r - age * 0.5
                ok = (age >= 0) & (age <= 0.6) & (r >= 1e-4) & (dr <= 0) & (dr >= -0.6)
                damp = np.clip(1 - age / 0.6, 0, 1)
                dh = np.sin(dr * 6 * 2 * PI / 0.5) * damp * damp * 0.4 * intensity
                rs = np.where(ok, r, 1.0)
                dn[..., 0] += np.where(ok, tx / rs * dh, 0)
                dn[..., 1] += np.where(ok, tz / rs * dh, 0)
    return dn


def transmittance_at(T, atm, p_atm, dirs):
    """sample_transmittance_lut (atmosphere_lut_sampling.wgsl:6-22) at
    atmosphere-frame points toward unit directions: 0 inside the Earth's
    shadow, else the T LUT at transmittance_lut_uv (atmosphere.wgsl:146-159)."""
    R, top = float(atm.planet_radius_m), float(atm.atmosphere_top_m)
    r = np.linalg.norm(p_atm, axis=-1)
    b = _dot(p_atm, dirs)
    disc = b * b - (r * r - R * R)
    shadow = (disc >= 0) & (-b - np.sqrt(np.maximum(disc, 0)) > 0)
    Hh = np.sqrt(max(top * top - R * R, 0.0))
    rho = np.sqrt(np.maximum(r * r - R * R, 0))
    disc_t = b * b - (r * r - top * top)
    d = np.where(disc_t >= 0, np.maximum(-b + np.sqrt(np.maximum(disc_t, 0)), 0), 0)
    d_min, d_max = top - r, rho + Hh
    u = np.clip((d - d_min) / np.maximum(d_max - d_min, 1.0), 0, 1)
    v = rho / max(Hh, 1.0)
    t = bilinear(T, u, v)[..., :3]
    return np.where(shadow[..., None], 0.0, t)


def skyview_uv_local(dirs, sun, cam_r, R):
    """pbr.wgsl:368-391."""
    vza = np.arccos(np.clip(dirs[..., 1], -1, 1))
    zha = PI - np.arcsin(np.clip(R / cam_r, 0, 1))
    below = np.maximum(PI - zha, 1e-6)
    v = np.where(vza < zha, np.sqrt(np.clip(vza / np.maximum(zha, 1e-6), 0, 1)) * 0.5,
                 0.5 + np.sqrt(np.clip((vza - zha) / below, 0, 1)) * 0.5)
    du = (np.arctan2(dirs[..., 0], dirs[..., 2]) - np.arctan2(sun[0], sun[2])) / (2 * PI)
    return du - np.floor(du), v


def overcast_irradiance(frame):
    """overcast_diffuse_irradiance (pbr.wgsl:452-476)."""
    sy = float(frame.sun_direction[1])
    t = np.clip((sy + 0.09) / 0.14, 0, 1)
    fade = t * t * (3 - 2 * t)
    return float(frame.sun_illuminance[3]) * np.sqrt(np.clip(sy, 0.05, 1.0)) * fade * 0.25


def ground_lit(frame, atm, luts, overcast_mask, surface, positions, view_dirs):
    """The ground pass's `lit` (before aerial perspective) at FP64 world
    positions (N, 3) seen along unit rays (N, 3); `surface` a SurfaceParams."""
    p = np.asarray(positions, np.float64)
    sun = np.array(frame.sun_direction[:3], np.float64)
    E = np.array(frame.sun_illuminance[:3], np.float64)
    R = float(atm.planet_radius_m)
    n = np.broadcast_to(np.array([0.0, 1.0, 0.0]), p.shape)
    v = -np.asarray(view_dirs, np.float64)
    l = np.broadcast_to(sun, p.shape)
    mf = float(surface.material)
    m = int(mf) if 0 <= mf <= 4 else 0
    bucket = voronoi_cell(p[:, 0].astype(f32), p[:, 2].astype(f32)) % U32(3)
    albedo = PALETTE[m][bucket.astype(np.int64)] * (np.array(atm.ground_albedo[:3], np.float64) / 0.18)
    wet = min(max(float(surface.ground_wetness), 0.0), 1.0)
    albedo = albedo ** mix(1.0, 3.0, wet)
    rough = mix(ROUGHNESS[m], max(ROUGHNESS[m] * 0.5, 0.05), wet)
    f0 = F0[m]
    # overcast blocking: the R8 mask over the 32 km square
    mask = np.asarray(overcast_mask, np.float64).reshape(128, 128, 1) / 255.0
    blocking = 1 - np.exp(-10 * bilinear(mask, p[:, 0] / 32000 + 0.5, p[:, 2] / 32000 + 0.5)[:, 0])
    p_atm = p + np.array([0.0, R, 0.0])
    cam_r = np.linalg.norm(p_atm, axis=-1)
    sun_t = transmittance_at(luts["T"], atm, p_atm, l)
    dE = E * sun_t * np.maximum(sun[1], 0) * (1 - blocking)[:, None]
    ov = overcast_irradiance(frame)

    def sky_at(dirs):
        u, vv = skyview_uv_local(dirs, sun, cam_r, R)
        clear = bilinear(luts["SV"], u, vv)[..., :3] * E
        return mix(clear, ov, blocking[:, None])

    sky = sky_at(n)
    lit = brdf(n, v, l, albedo, rough, f0) * dE + albedo / PI * sky
    cov_gate = min(max(float(surface.puddle_coverage), 0.0), 1.0)
    ps = float(surface.puddle_start)
    if wet > ps:
        pt = min(max((wet - ps) / max(1 - ps, 1e-3), 0.0), 1.0)
        pmask = value_noise(p[:, 0] * 0.15, p[:, 2] * 0.15) * cov_gate * pt
        nw = n
        precip = float(surface.precip_intensity_mm_per_h)
        if wet > 0.5 and precip > 0:
            dn = ripples(p[:, 0], p[:, 2], float(frame.simulated_seconds), min(max(precip / 25, 0.0), 1.0))
            nw = np.stack([-dn[:, 0], np.ones(len(p)), -dn[:, 1]], -1)
            nw = nw / np.linalg.norm(nw, axis=-1, keepdims=True)
        fw = brdf(nw, v, l, np.zeros(3), 0.05, 0.02)
        rf = -v - 2 * _dot(-v, nw)[:, None] * nw
        nv = np.maximum(_dot(nw, v), 1e-4)
        fv = 0.02 + 0.98 * np.clip(1 - nv, 0, 1) ** 5
        water = fw * dE + fv[:, None] * sky_at(rf)
        lit = np.where((pmask > 0.01)[:, None], mix(lit, water, pmask[:, None]), lit)
    if float(surface.temperature_c) < 0.5 and float(surface.snow_depth_m) > 0:
        inv_puddle = 1 - value_noise(p[:, 0] * 0.15, p[:, 2] * 0.15) * cov_gate
        amount = min(max(float(surface.snow_depth_m) / 0.05, 0.0), 1.0) * inv_puddle
        snow = brdf(n, v, l, np.full(3, 0.9), 0.85, f0) * dE + 0.9 / PI * sky
        lit = np.where((amount > 0.01)[:, None], mix(lit, snow, amount[:, None]), lit)
    return lit


def ground_image(frame, atm, luts, overcast_mask, surface, depth):
    """`lit` (H, W, 3) at every surface pixel of the depth buffer (0 elsewhere)."""
    cam, d = pixel_rays(frame)
    t = surface_distance(frame, depth)
    g = np.isfinite(t)
    pos = cam + d[g] * t[g][:, None]
    out = np.zeros(d.shape)
    out[g] = ground_lit(frame, atm, luts, overcast_mask, surface, pos, d[g])
    return out


# ---------------------------------------------------------------------------
# godrays, bloom, exposure statistics
# ---------------------------------------------------------------------------
def sun_screen_uv(frame):
    """The sun's uv on screen, or None when the godrays pass is off
    (ps-godrays lib.rs:358-386, radial.wgsl:71-74); float32 like the host."""
    sd = np.array(frame.sun_direction[:3], f32)
    if not sd[1] > 0:
        return None
    m = np.asarray(frame.view_proj, f32)   # column-major
    cx = ((m[0] * sd[0] + m[4] * sd[1]) + m[8] * sd[2]) + m[12] * f32(0)
    cy = ((m[1] * sd[0] + m[5] * sd[1]) + m[9] * sd[2]) + m[13] * f32(0)
    cw = ((m[3] * sd[0] + m[7] * sd[1]) + m[11] * sd[2]) + m[15] * f32(0)
    if not cw > f32(1e-6):
        return None
    nx, ny = cx / cw, cy / cw
    if abs(nx) > 2 or abs(ny) > 2:
        return None
    return float(nx) * 0.5 + 0.5, 1.0 - (float(ny) * 0.5 + 0.5)


def _centres(w, h):
    v, u = np.meshgrid((np.arange(h) + 0.5) / h, (np.arange(w) + 0.5) / w, indexing="ij")
    return u, v


def godrays(hdr, sun_uv, samples, decay, threshold):
    """Half-resolution radial accumulation (radial.wgsl:57-92): (rh, rw, 3)."""
    H, W = hdr.shape[:2]
    rw, rh = max(W // 2, 1), max(H // 2, 1)
    u0, v0 = _centres(rw, rh)
    du, dv = (sun_uv[0] - u0) / samples, (sun_uv[1] - v0) / samples
    acc = np.zeros((rh, rw, 3))
    for i in range(samples):
        c = bilinear(hdr, u0 + i * du, v0 + i * dv)[..., :3]
        lum = c @ np.array([0.2126, 0.7152, 0.0722])
        mask = np.clip((lum - threshold) / max(threshold, 1.0), 0, 1)
        acc += c * (mask * decay ** i)[..., None]
    return acc / samples


def bloom(hdr, threshold, knee):
    """The pyramid's half-resolution level after the upsample chain
    (crates/ps-bloom/src/lib.rs:587-747): (h0, w0, 3)."""
    H, W = hdr.shape[:2]
    sizes = []
    w, h = max(W // 2, 1), max(H // 2, 1)
    for _ in range(3):
        sizes.append((w, h))
        w, h = max(w // 2, 1), max(h // 2, 1)
    u, v = _centres(*sizes[0])
    c = bilinear(hdr, u, v)[..., :3]
    lum = c @ np.array([0.2126, 0.7152, 0.0722])
    lo = max(threshold - knee, 1e-6)
    t = np.clip((lum - lo) / (threshold - lo), 0, 1)
    levels = [c * (t * t * (3 - 2 * t))[..., None]]
    for l in (1, 2):
        src = levels[-1]
        sh, sw = src.shape[:2]
        u, v = _centres(*sizes[l])
        tap = lambda ox, oy: bilinear(src, u + ox / sw, v + oy / sh)
        inner = sum(tap(x, y) for x, y in ((-1, -1), (1, -1), (-1, 1), (1, 1))) * (0.5 / 4)
        box = sum(tap(x, y) for x, y in ((-2, 0), (2, 0), (0, -2), (0, 2))) * (0.125 / 4)
        corners = sum(tap(x, y) for x, y in ((-2, -2), (2, -2), (-2, 2), (2, 2))) * (0.125 / 4)
        levels.append(tap(0, 0) * 0.125 + inner + box + corners)
    for l in (2, 1):
        src = levels[l]
        sh, sw = src.shape[:2]
        u, v = _centres(*sizes[l - 1])
        tap = lambda ox, oy: bilinear(src, u + ox / sw, v + oy / sh)
        tent = (sum(tap(x, y) for x, y in ((-1, -1), (1, -1), (-1, 1), (1, 1))) / 16
                + sum(tap(x, y) for x, y in ((0, -1), (-1, 0), (1, 0), (0, 1))) * 2 / 16 + tap(0, 0) * 4 / 16)
        levels[l - 1] = levels[l - 1] + tent
    return levels[0]


def post_chain(hdr, frame, ev100, godrays_params=None, bloom_params=None):
    """HDR after godrays and bloom. godrays_params: (samples, decay,
    intensity, threshold) or None; bloom_params: (threshold_ev100,
    intensity, knee_ev) or None."""
    out = np.array(hdr[..., :3], np.float64)
    H, W = out.shape[:2]
    u, v = _centres(W, H)
    if godrays_params is not None:
        suv = sun_screen_uv(frame)
        if suv is not None:
            n, decay, inten, thr = godrays_params
            rays = godrays(out, suv, n, decay, thr)
            out = out + bilinear(rays, u, v) * inten
    if bloom_params is not None:
        thr_ev, inten, knee_ev = bloom_params
        threshold = 2.0 ** (ev100 + thr_ev)
        knee = max(knee_ev * np.log(2.0), 1e-6)
        l0 = bloom(out, threshold, knee)
        out = out + bilinear(l0, u, v) * inten
    return out


def exposure_stats(hdr):
    """(sum of log2 max(luminance, 1e-6), count) over the pixels on the 4-pixel
    grid (auto_exposure.comp.wgsl:31-73)."""
    s = np.asarray(hdr, np.float64)[::4, ::4, :3]
    lum = np.maximum(s @ np.array([0.2126, 0.7152, 0.0722]), 1e-6)
    return float(np.sum(np.log2(lum))), float(lum.size)


def ev100_from_stats(log_sum, count):
    """auto_exposure.rs:209-212."""
    return float(np.log2(max(2.0 ** (log_sum / count), 1e-6) / 0.216))
