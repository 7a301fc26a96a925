"""Pins of the FP64 oracle against mathematics and the reference's own test
numbers (CPU only). Each test is chosen so that a plausible restatement
slip (a dropped term, a wrong sign or index, a transposed operand) fails."""
import math

import numpy as np
import pytest
from scipy import integrate, ndimage

import oracle
from pedalsky_amd import abi, scene


@pytest.fixture(scope="module")
def atm():
    return scene.earth_atmosphere()


# ---------------------------------------------------------------------------
# atmosphere (atmosphere.wgsl)
# ---------------------------------------------------------------------------
def _vertical_od_closed_form(a, h0):
    """Exact vertical optical depth from altitude h0 to the top of the
    atmosphere: exponential profiles integrate to sigma*H*(e^{-h0/H} -
    e^{-htop/H}); the ozone tent to its piecewise-linear area."""
    top = a.atmosphere_top_m - a.planet_radius_m
    od = np.zeros(3)
    for i in range(3):
        r = a.rayleigh_scattering[i] * a.rayleigh_scale_height_m * (
            math.exp(-h0 / a.rayleigh_scale_height_m) - math.exp(-top / a.rayleigh_scale_height_m))
        m = (a.mie_scattering[i] + a.mie_absorption[i]) * a.mie_scale_height_m * (
            math.exp(-h0 / a.mie_scale_height_m) - math.exp(-top / a.mie_scale_height_m))
        half = a.ozone_thickness_m / 2
        oz, _ = integrate.quad(lambda h: max(0.0, 1.0 - abs(h - a.ozone_center_m) / half), h0, top,
                               points=[a.ozone_center_m - half, a.ozone_center_m, a.ozone_center_m + half],
                               epsabs=0, epsrel=1e-12, limit=200)
        od[i] = r + m + a.ozone_absorption[i] * oz
    return od


@pytest.mark.parametrize("h0", [0.0, 1500.0, 12000.0])
def test_vertical_optical_depth_closed_form(atm, h0):
    lib = oracle.load()
    r = atm.planet_radius_m + h0
    exact = _vertical_od_closed_form(atm, h0)
    od = np.zeros(3)
    lib.or_optical_depth(oracle._ref(atm), r, 1.0, 4000, od.ctypes.data)
    assert np.allclose(od, exact, rtol=1e-5)
    # second-order convergence of the trapezoid on the smooth (ozone-free)
    # integrand: halving the step quarters the error
    smooth = scene.earth_atmosphere()
    for i in range(3):
        smooth.ozone_absorption[i] = 0.0
    ex_s = _vertical_od_closed_form(smooth, h0)
    o1, o2 = np.zeros(3), np.zeros(3)
    lib.or_optical_depth(oracle._ref(smooth), r, 1.0, 4000, o1.ctypes.data)
    lib.or_optical_depth(oracle._ref(smooth), r, 1.0, 2000, o2.ctypes.data)
    ratio = (o2 - ex_s) / (o1 - ex_s)
    assert np.all(np.abs(ratio - 4.0) < 0.05)
    # the LUT's 40-step trapezoid (2.5 km steps vs a 1.2 km Mie scale height) is a few % high
    lib.or_optical_depth(oracle._ref(atm), r, 1.0, 40, od.ctypes.data)
    assert np.allclose(od, exact, rtol=5e-2)
    assert np.all(od > exact)  # convex integrand: trapezoid over-estimates


def test_transmittance_uv_round_trip(atm):
    lib = oracle.load()
    rm = np.zeros(2)
    uv = np.zeros(2)
    for u in (0.01, 0.3, 0.77, 0.99):
        for v in (0.02, 0.5, 0.98):
            lib.or_transmittance_pos_dir(oracle._ref(atm), u, v, rm.ctypes.data)
            lib.or_transmittance_uv(oracle._ref(atm), rm[0], rm[1], uv.ctypes.data)
            assert uv[0] == pytest.approx(u, abs=1e-9) and uv[1] == pytest.approx(v, abs=1e-9)


def test_transmittance_lut_texel_is_exp_of_optical_depth(atm):
    T = oracle.bake_transmittance(atm, 256, 64)
    lib = oracle.load()
    rm = np.zeros(2)
    od = np.zeros(3)
    for x, y in ((0, 0), (100, 10), (255, 63), (17, 40)):
        lib.or_transmittance_pos_dir(oracle._ref(atm), (x + 0.5) / 256, (y + 0.5) / 64, rm.ctypes.data)
        lib.or_optical_depth(oracle._ref(atm), rm[0], rm[1], 40, od.ctypes.data)
        assert np.allclose(T[y, x, :3], np.exp(-od), rtol=1e-12)
    # horizon rays are darker than zenith rays, blue attenuates most
    assert T[0, 255, 2] < T[0, 0, 2] and T[0, 0, 2] < T[0, 0, 0]


@pytest.mark.parametrize("g", [0.0, 0.5, 0.8, 0.95])
def test_phase_functions_normalise(g):
    lib = oracle.load()
    two_pi = 2 * math.pi
    pm, _ = integrate.quad(lambda c: lib.or_phase_mie(c, g), -1, 1, points=[1.0], limit=400, epsrel=1e-10)
    assert two_pi * pm == pytest.approx(1.0, rel=1e-6)
    hg, _ = integrate.quad(lambda c: lib.or_henyey_greenstein(c, g), -1, 1, limit=400, epsrel=1e-10)
    assert two_pi * hg == pytest.approx(1.0, rel=1e-6)
    for alpha in (0.0, 0.7, 4.0):
        dr, _ = integrate.quad(lambda c: lib.or_draine_phase(c, g, alpha), -1, 1, limit=400, epsrel=1e-10)
        assert two_pi * dr == pytest.approx(1.0, rel=1e-6)
    pr, _ = integrate.quad(lambda c: lib.or_phase_rayleigh(c), -1, 1)
    assert two_pi * pr == pytest.approx(1.0, rel=1e-12)


def test_draine_reduces_to_hg_and_rayleigh():
    lib = oracle.load()
    for c in np.linspace(-1, 1, 11):
        assert lib.or_draine_phase(c, 0.6, 0.0) == pytest.approx(lib.or_henyey_greenstein(c, 0.6), rel=1e-12)
        assert lib.or_draine_phase(c, 0.0, 1.0) == pytest.approx(lib.or_phase_rayleigh(c), rel=1e-12)


def test_cloud_phase_fit_values():
    # Jendersie & d'Eon 2023 Eqs. 4-7 at d = 20 um (cloud_march.wgsl:266-269)
    lib = oracle.load()
    d = 20.0
    g_hg = math.exp(-0.0990567 / (d - 1.67154))
    g_d = math.exp(-2.20679 / (d + 3.91029)) - 0.428934
    alpha = math.exp(3.62489 - 8.29288 / (d + 5.52825))
    w_d = math.exp(-0.599085 / (d - 0.641583) - 0.665888)
    for c in (-0.5, 0.3, 0.999):
        exp = (1 - w_d) * lib.or_henyey_greenstein(c, g_hg) + w_d * lib.or_draine_phase(c, g_d, alpha)
        assert lib.or_cloud_phase(c, 1.0, d, 1.0, 0, 0.5) == pytest.approx(exp, rel=1e-12)
    # water droplets: no halo; ice (50 um): halo lobe adds at 22 degrees
    c22 = 0.92718
    no_ice = (lib.or_cloud_phase(c22, 1.0, 30.0, 1.0, 0, 0.5))
    assert no_ice > 0
    ice = lib.or_cloud_phase(c22, 1.0, 50.0, 1.0, 5, 0.5)
    base = lib.or_cloud_phase(c22 + 0.02, 1.0, 50.0, 1.0, 5, 0.5)
    assert ice > base  # the 22-degree lobe peaks there


def test_skyview_uv_inverts_the_bake_parametrisation(atm):
    """sky.wgsl:73-102 maps the direction skyview.comp.wgsl:68-93 builds for
    texel (u, v) back to (u, v)."""
    lib = oracle.load()
    R = atm.planet_radius_m
    r = R + 500.0
    zha = math.pi - math.asin(R / r)
    sun = np.array([0.3, 0.6, -0.74])
    sun /= np.linalg.norm(sun)
    sun_az = math.atan2(sun[0], sun[2])
    uv = np.zeros(2)
    for u in (0.05, 0.4, 0.9):
        for v in (0.1, 0.45, 0.55, 0.93):
            c = 2 * v if v < 0.5 else 2 * (v - 0.5)
            vza = c * c * zha if v < 0.5 else zha + c * c * (math.pi - zha)
            az = sun_az + u * 2 * math.pi
            d = np.array([math.sin(vza) * math.sin(az), math.cos(vza), math.sin(vza) * math.cos(az)])
            lib.or_skyview_uv(oracle._ref(atm), d.ctypes.data, sun.ctypes.data, r, uv.ctypes.data)
            assert uv[0] == pytest.approx(u, abs=1e-9) and uv[1] == pytest.approx(v, abs=1e-9)


def test_lut_chain_invariants(atm):
    f = scene.make_frame(64, 32, sun_zenith_deg=50.0, sun_azimuth_deg=180.0)
    luts = oracle.bake_luts(f, atm, abi.LutDims(64, 32, 32, 32, 48, 28, 16, 16, 16))
    MS, SV, AP = luts["MS"], luts["SV"], luts["AP"]
    assert np.all(MS[..., :3] >= 0) and MS[..., :3].max() > 1e-4       # phase5_lut_diag.rs:95-110
    assert SV[..., :3].sum() > 1e-2
    # sky-view is symmetric about the sun's azimuth: column u == column 1 - u
    assert np.allclose(SV[..., :3], SV[:, ::-1, :3], rtol=1e-9, atol=1e-15)
    # sky above the horizon is blue-dominant (phase5_sky.rs:58-69)
    assert np.all(SV[2:12, :, 2] > SV[2:12, :, 0])
    # for froxel rows looking above the horizon, AP transmittance decreases and
    # in-scatter increases with depth slice (rows looking down stop at the
    # ground, aerialperspective.comp.wgsl:272)
    up = AP[:, :8]
    assert np.all(np.diff(up[..., 3], axis=0) <= 1e-12)
    assert np.all(np.diff(up[..., 1], axis=0) >= -1e-9)
    assert AP[0, ..., 3].min() > 0.99
    # no scattering: multi-scattering vanishes
    dark = scene.earth_atmosphere()
    for k in ("rayleigh_scattering", "mie_scattering"):
        arr = getattr(dark, k)
        for i in range(3):
            arr[i] = 0.0
    T0 = oracle.bake_transmittance(dark, 64, 32)
    assert np.all(oracle.bake_multiscatter(dark, T0, 16, 16)[..., :3] == 0.0)


# ---------------------------------------------------------------------------
# sampler (texel-centred linear filtering)
# ---------------------------------------------------------------------------
def test_trilinear_repeat_matches_scipy():
    rng = np.random.default_rng(7)
    vol = rng.random((5, 6, 7, 2)).astype(np.float32)  # (d, h, w, ch)
    lib = oracle.load()
    out = np.zeros(2)
    for _ in range(50):
        u, v, w = rng.uniform(-2, 3, 3)
        lib.or_sample3d_repeat(vol.ctypes.data, 7, 6, 5, 2, u, v, w, out.ctypes.data)
        coords = [[w * 5 - 0.5], [v * 6 - 0.5], [u * 7 - 0.5]]
        for c in range(2):
            exp = ndimage.map_coordinates(vol[..., c].astype(np.float64), coords, order=1, mode="grid-wrap")[0]
            assert out[c] == pytest.approx(exp, abs=1e-6)


def test_bilinear_clamp_matches_scipy():
    rng = np.random.default_rng(8)
    img = rng.random((9, 13, 1)).astype(np.float32)
    lib = oracle.load()
    out = np.zeros(1)
    for _ in range(50):
        u, v = rng.uniform(-0.2, 1.2, 2)
        lib.or_sample2d(img.ctypes.data, 13, 9, 1, u, v, 0, out.ctypes.data)
        exp = ndimage.map_coordinates(img[..., 0].astype(np.float64), [[v * 9 - 0.5], [u * 13 - 0.5]], order=1,
                                      mode="nearest")[0]
        assert out[0] == pytest.approx(exp, abs=1e-6)


# ---------------------------------------------------------------------------
# noise (noise_base/detail/curl.comp.wgsl)
# ---------------------------------------------------------------------------
def test_hash_is_the_reference_integer_mix():
    lib = oracle.load()
    M = 0xFFFFFFFF

    def h(x, y, z):
        v = (x * 0x27d4eb2d + y * 0x165667b1 + z * 0x9e3779b9) & M
        v = ((v ^ (v >> 15)) * 0x2c1b3c6d) & M
        v = ((v ^ (v >> 12)) * 0x297a2d39) & M
        return v ^ (v >> 15)
    for x, y, z in ((0, 0, 0), (1, 2, 3), (127, 5, 99), (200, 1000, 7)):
        assert lib.or_hash3(x, y, z, 0) == h(x, y, z)
    assert lib.or_hash3(1, 2, 3, 1) != h(1, 2, 3)


def test_noise_volumes_tile_and_vary():
    # phase6_noise_diag.rs:27-123: variation present, values in [0, 1]
    d = oracle.noise_detail(0)
    assert d.min() >= 0 and d.max() <= 1 and d[..., :3].std() > 0.05
    assert np.all(d[..., 3] == 0)
    b = oracle.noise_base(0, 0, 4)
    assert b.min() >= 0 and b.max() <= 1 and b.std() > 0.05
    c = oracle.noise_curl(0)
    assert c.min() >= 0 and c.max() <= 1 and c.std() > 0.01
    # tileable: the curl map's opposite edges are as close as interior neighbours
    edge = np.abs(c[0] - c[-1]).mean()
    inner = np.abs(np.diff(c, axis=0)).mean()
    assert edge < 3 * inner


# ---------------------------------------------------------------------------
# cloud march: closed forms on constructed inputs
# ---------------------------------------------------------------------------
def _uniform_inputs(w, base_rgba=(1.0, 0.0, 0.0, 0.0), coverage=1.0):
    base = np.zeros((128, 128, 128, 4), dtype=np.float32)
    base[...] = base_rgba
    det = np.zeros((32, 32, 32, 4), dtype=np.float32)
    curl = np.full((128, 128, 2), 0.5, dtype=np.float32)
    wm = np.zeros((8, 8, 2), dtype=np.float16)
    wm[..., 0] = coverage
    w.weather = wm.view(np.uint16)
    return base, det, curl


def _transparent(w):
    a = w.atmosphere
    for k in ("rayleigh_scattering", "mie_scattering", "mie_absorption", "ozone_absorption"):
        arr = getattr(a, k)
        for i in range(3):
            arr[i] = 0.0


def _smoothstep(e0, e1, x):
    t = np.clip((x - e0) / (e1 - e0), 0, 1)
    return t * t * (3 - 2 * t)


def test_constant_noise_column_transmittance_closed_form():
    """Base noise R=1, GBA=0, coverage 1 and no erosion give density =
    density_scale * NDF(h). Looking almost straight up through a cumulus
    layer the transmittance is exp(-sigma_t * ds * thickness/cos * int NDF)."""
    w = scene.workload_small(width=16, height=16)
    w.frame = scene.make_frame(16, 16, pitch_deg=89.0, yaw_deg=0.0, fov_deg=1.0, sun_zenith_deg=40.0)
    w.layers = [scene.synth_layer(0, 1500.0, 2300.0, 1.0, density_scale=0.05)]
    w.layers[0].coverage = 1.0
    w.clouds = scene.cloud_params(steps=4096, detail_strength=0.0, skew=0.0, n_layers=1)
    w.depth = None
    base, det, curl = _uniform_inputs(w)
    fr = oracle.Frame(w, base, det, curl)
    _, _, trn, _ = fr.render_region(7, 7, 2, 2)
    ndf_int, _ = integrate.quad(lambda h: _smoothstep(0, 0.07, h) * _smoothstep(1, 0.2, h), 0, 1,
                                points=[0.07, 0.2], epsrel=1e-12)
    # direction of the pixel-centre ray: pitch 89 deg within a 1-degree fov
    cosz = math.sin(math.radians(89.0))
    od_per_sigma = 0.05 * 800.0 / cosz * ndf_int
    for c, s in enumerate(w.clouds.sigma_s[:3]):
        exp_t = math.exp(-s * od_per_sigma)
        assert trn[..., c] == pytest.approx(exp_t, rel=2e-3)


def test_emission_only_march_telescopes():
    """With a transparent atmosphere and zero sun, the only source is the
    (flat, huge-radius) lightning term K: the energy-conserving integration
    telescopes to lum = K * sigma_s/sigma_t * (1 - T) per channel, whatever
    the density field."""
    w = scene.workload_small(width=48, height=32)
    _transparent(w)
    w.frame = scene.make_frame(48, 32, pitch_deg=20.0, yaw_deg=30.0, illuminance=0.0)
    w.frame.lightning_illuminance[:] = [3.0, 5.0, 7.0, 1e12]
    w.clouds = scene.cloud_params(steps=64, n_layers=1, sigma_a=(C3 := (0.01, 0.0, 0.02)))
    w.depth = None
    base = oracle.decode(oracle.quantize(oracle.noise_base(0), 0), 0)
    det = oracle.decode(oracle.quantize(oracle.noise_detail(0), 0), 0)
    curl = oracle.decode(oracle.quantize(oracle.noise_curl(0), 0), 0)
    fr = oracle.Frame(w, base, det, curl)
    comp, lum, trn, _ = fr.render_region(0, 0, 48, 32)
    assert (trn[..., 1] < 0.95).mean() > 0.05
    sig_s = np.array(w.clouds.sigma_s[:3], dtype=np.float64)
    sig_t = sig_s + np.array(C3)
    K = np.array([3.0, 5.0, 7.0])
    fall = 1.0  # radius 1e12 m: weight = 1 - 3(d/r)^2 >= 1 - 1e-9 even for far-side rays
    exp = K * sig_s / sig_t * (1.0 - trn[..., :3]) * fall
    assert np.allclose(lum[..., :3], exp, rtol=1e-6, atol=1e-9)
    # sky is black without sun, so the composite is the cloud emission alone
    assert np.allclose(comp[..., :3], lum[..., :3], atol=1e-9)


def test_zero_coverage_and_zero_density_are_identity():
    w = scene.workload_small(width=32, height=24)
    w.depth = None
    base = oracle.decode(oracle.quantize(oracle.noise_base(0, 0, 128), 0), 0)
    det = oracle.decode(oracle.quantize(oracle.noise_detail(0), 0), 0)
    curl = oracle.decode(oracle.quantize(oracle.noise_curl(0), 0), 0)
    w.layers[0].coverage = 0.0
    fr = oracle.Frame(w, base, det, curl)
    comp, lum, trn, _ = fr.render_region(0, 0, 32, 24)
    assert np.all(lum[..., :3] == 0) and np.all(trn[..., :3] == 1)
    w.layers[0].coverage = 0.5
    w.layers[0].density_scale = 0.0
    fr = oracle.Frame(w, base, det, curl)
    comp2, lum2, trn2, _ = fr.render_region(0, 0, 32, 24)
    assert np.all(lum2[..., :3] == 0) and np.all(trn2[..., :3] == 1)
    assert np.allclose(comp2, comp)


def test_depth_clip_removes_clouds_behind_ground():
    # phase9_composition.rs:84-134: nothing behind an opaque hit contributes
    w = scene.workload_small(width=40, height=30)
    w.frame = scene.make_frame(40, 30, pitch_deg=25.0, yaw_deg=180.0, sun_zenith_deg=40.0)
    base = oracle.decode(oracle.quantize(oracle.noise_base(0), 0), 0)
    det = oracle.decode(oracle.quantize(oracle.noise_detail(0), 0), 0)
    curl = oracle.decode(oracle.quantize(oracle.noise_curl(0), 0), 0)
    far = oracle.Frame(w, base, det, curl, depth=None)
    _, lum_far, trn_far, _ = far.render_region(0, 0, 40, 30)
    assert (trn_far[..., 1] < 0.99).mean() > 0.05
    # an opaque wall 500 m away (reverse-Z depth = near / 500) hides the clouds
    near = oracle.Frame(w, base, det, curl, depth=np.full((30, 40), 0.1 / 500.0, dtype=np.float32))
    _, lum_near, trn_near, _ = near.render_region(0, 0, 40, 30)
    assert np.all(trn_near[..., :3] == 1.0) and np.all(lum_near[..., :3] == 0.0)


def test_ndf_shapes():
    # WGSL NDF (cloud_density_common.wgsl:159-201): zero at base/top for cumulus,
    # cumulus peaks in the lower half, Cb anvil near the top, unknown type = 0
    lib = oracle.load()
    assert lib.or_ndf(0.0, 0, 0.0) < 1e-3 and lib.or_ndf(1.0, 0, 0.0) < 1e-3
    hs = np.arange(0, 1.0001, 0.05)
    vals = [lib.or_ndf(h, 0, 0.0) for h in hs]
    assert hs[int(np.argmax(vals))] < 0.55 and max(vals) > 0.5
    assert lib.or_ndf(0.92, 7, 1.0) > 0.05 and lib.or_ndf(0.4, 7, 1.0) > 0.05
    assert lib.or_ndf(0.92, 7, 0.0) < lib.or_ndf(0.92, 7, 1.0)
    assert lib.or_ndf(0.5, 99, 1.0) == 0.0
