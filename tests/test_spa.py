"""Sun position and TOA illuminance (host C++ behind the C ABI) against the
oracle restatement of crates/ps-core/src/astro/spa.rs and world.rs:95-111,
the reference's own tests, and the intermediate values NREL/TP-560-34302
publishes for its Appendix B example (Table A5.1). CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle.spa as S
from pedalsky_amd import abi

NREL_UTC = (2003, 10, 17, 19, 30, 30)          # local 12:30:30 at UTC-7
NREL_SITE = (39.742476, -105.1786, 1830.14)     # the paper's example site
REF_SITE = (39.7426, -105.1786, 1830.14)        # as spa.rs:352 rounds it


@pytest.fixture(scope="module")
def lib():
    return abi.load()


def _utc(t):
    return abi.Utc(*t)


def _product(lib, t, lat, lon, elev):
    p = abi.SunPosition()
    assert lib.ps_sun_position_precise(C.byref(_utc(t)), lat, lon, elev, C.byref(p)) == abi.PS_OK
    return p


# ---- the reference's own tests ------------------------------------------------------------

def test_julian_day_reference_cases(lib):
    # spa.rs:304-309 (J2000) and :311-323 (the NREL example)
    assert abs(lib.ps_julian_day(C.byref(_utc((2000, 1, 1, 12, 0, 0)))) - 2451545.0) < 1e-6
    assert abs(lib.ps_julian_day(C.byref(_utc(NREL_UTC))) - 2452930.312847) < 1e-5
    # the calendar switch and the epoch (Meeus ch. 7): 1582-10-04 (Julian) and
    # 1582-10-15 (Gregorian) are consecutive days; JD 0 is -4712-01-01 12:00
    assert lib.ps_julian_day(C.byref(_utc((1582, 10, 4, 0, 0, 0)))) == 2299159.5
    assert lib.ps_julian_day(C.byref(_utc((1582, 10, 15, 0, 0, 0)))) == 2299160.5
    assert lib.ps_julian_day(C.byref(_utc((-4712, 1, 1, 12, 0, 0)))) == 0.0


def test_delta_t_2026(lib):
    # spa.rs:326-333
    assert 70.0 <= lib.ps_delta_t_seconds(2026) <= 80.0
    for y in (1950, 1990, 2004, 2005, 2026, 2050, 2051, 2149, 2200):
        assert lib.ps_delta_t_seconds(y) == S.delta_t_seconds(y)


def test_nrel_example_within_reference_tolerance(lib):
    # spa.rs:347-360: zenith 50.11162 and azimuth 194.34024 (the paper's,
    # with refraction) within 0.05 deg
    p = _product(lib, NREL_UTC, *REF_SITE)
    assert abs(p.zenith_deg - 50.11162) < 0.05
    assert abs(p.azimuth_deg - 194.34024) < 0.05


# ---- the oracle against NREL's published intermediates --------------------------------------

# Table A5.1 of NREL/TP-560-34302 (Delta T = 67 s), with the tolerance its
# printed digits allow. e0 is the topocentric elevation before refraction:
# the reference omits the refraction term, so its zenith is 90 - e0.
NREL_A51 = {
    "julian_day": (2452930.312847, 1e-6),
    "l_deg": (24.0182616917, 1e-10),
    "b_deg": (-0.0001011219, 1e-10),
    "distance_au": (0.9965422974, 1e-10),
    "dpsi_deg": (-0.00399840, 1e-8),
    "deps_deg": (0.00166657, 1e-8),
    "eps_deg": (23.440465, 1e-6),
    "lambda_deg": (204.0085519281, 1e-10),
    "apparent_sidereal_time_deg": (318.5119, 1e-4),
    "right_ascension_deg": (202.22741, 1e-5),
    "declination_deg": (-9.31434, 1e-5),
    "hour_angle_deg": (11.105900, 1e-5),
    "topo_declination_deg": (-9.316179, 1e-6),
    "e0_deg": (39.872046, 1e-6),
    "azimuth_deg": (194.340241, 1e-6),
}


def nrel_errors(mod=S):
    """Deviation of each Table A5.1 quantity from the paper's value, in units
    of its tolerance (<= 1 passes)."""
    o = mod.sun_position_precise(NREL_UTC, *NREL_SITE, delta_t_override=67.0)
    got = dict(o.__dict__)
    got["e0_deg"] = 90.0 - o.zenith_deg
    return {k: abs(got[k] - v) / tol for k, (v, tol) in NREL_A51.items()}


def test_oracle_matches_nrel_table_a51():
    err = nrel_errors()
    bad = {k: e for k, e in err.items() if e > 1.0}
    assert not bad, bad


def _slip_no_l1(m):
    t = m.T.L1
    m.T.L1 = ()
    return lambda: setattr(m.T, "L1", t)


def _slip_nutation_swap(m):
    f = m._nutation

    def g(jce):
        a, b = f(jce)
        return b, a
    m._nutation = g
    return lambda: setattr(m, "_nutation", f)


def _slip_beta_sign(m):
    t = m.T.B0
    m.T.B0 = tuple((-a, b, c) for a, b, c in t)
    return lambda: setattr(m.T, "B0", t)


def _slip_series_in_jce(m):
    f = m._series
    m._series = lambda jme, tabs: f(jme * 10.0, tabs)
    return lambda: setattr(m, "_series", f)


def _slip_parallax_dropped(m):
    f = m.math.sin
    # the topocentric parallax is the only sin of a sub-arcsecond angle
    m.math.sin = lambda x: 0.0 if 0.0 < x < 1e-4 else f(x)
    return lambda: setattr(m.math, "sin", f)


@pytest.mark.parametrize("slip", [_slip_no_l1, _slip_nutation_swap, _slip_beta_sign, _slip_series_in_jce,
                                  _slip_parallax_dropped])
def test_table_a51_pins_catch_oracle_slips(slip):
    undo = slip(S)
    try:
        err = nrel_errors()
    finally:
        undo()
    assert max(err.values()) > 1.0, err
    assert max(nrel_errors().values()) <= 1.0


# ---- the product against the oracle -------------------------------------------------------------

def test_product_matches_oracle_over_dates_and_sites(lib):
    rng = np.random.default_rng(0x5A)
    for _ in range(300):
        y = int(rng.integers(1900, 2200))
        mo = int(rng.integers(1, 13))
        d = int(rng.integers(1, 29))
        t = (y, mo, d, int(rng.integers(0, 24)), int(rng.integers(0, 60)), int(rng.integers(0, 60)))
        lat, lon, el = float(rng.uniform(-89.9, 89.9)), float(rng.uniform(-180, 180)), float(rng.uniform(0, 5000))
        p = _product(lib, t, lat, lon, el)
        o = S.sun_position_precise(t, lat, lon, el)
        assert p.julian_day == o.julian_day
        assert abs(p.distance_au - o.distance_au) < 1e-13
        for k in ("zenith_deg", "altitude_deg", "right_ascension_deg", "declination_deg",
                  "apparent_sidereal_time_deg"):
            assert abs(getattr(p, k) - getattr(o, k)) < 1e-9, (t, k)
        daz = abs(p.azimuth_deg - o.azimuth_deg)
        assert min(daz, 360.0 - daz) < 1e-9 or p.zenith_deg < 1e-6


# ---- the world sun state and the frame fields ---------------------------------------------------

def test_altaz_axes():
    # world.rs:117-135: north horizon -> -Z, east horizon -> +X, zenith -> +Y
    assert np.allclose(S.altaz_to_world(0.0, 0.0), (0.0, 0.0, -1.0), atol=1e-12)
    assert np.allclose(S.altaz_to_world(0.0, math.pi / 2), (1.0, 0.0, 0.0), atol=1e-12)
    assert np.allclose(S.altaz_to_world(math.pi / 2, 1.0), (0.0, 1.0, 0.0), atol=1e-12)


def test_world_sun_matches_oracle(lib):
    for t, site in ((NREL_UTC, REF_SITE), ((2026, 6, 21, 12, 0, 0), (51.5, -0.12, 20.0)),
                    ((2026, 1, 3, 4, 0, 0), (-33.9, 151.2, 50.0)), ((2026, 7, 4, 18, 30, 0), (0.0, 0.0, 0.0))):
        w = abi.WorldSun()
        assert lib.ps_world_sun(C.byref(_utc(t)), *site, C.byref(w)) == abi.PS_OK
        alt, az, d, dirn, toa = S.world_sun(t, *site)
        assert (w.altitude_rad, w.azimuth_rad, w.distance_au) == (np.float32(alt), np.float32(az), np.float32(d))
        assert np.allclose(list(w.direction_world), dirn, atol=2e-7)
        assert abs(np.linalg.norm(list(w.direction_world)) - 1.0) < 1e-6
        assert abs(w.toa_illuminance_lux / toa - 1.0) < 1e-6


def test_toa_follows_inverse_square(lib):
    # world.rs:139-146 pins 127500 lux at 1 AU; perihelion (early January) and
    # aphelion (early July) differ by (Q / q)^2 ~ 1.069
    def toa(t):
        w = abi.WorldSun()
        assert lib.ps_world_sun(C.byref(_utc(t)), 0.0, 0.0, 0.0, C.byref(w)) == abi.PS_OK
        assert abs(w.toa_illuminance_lux * w.distance_au ** 2 / 127500.0 - 1.0) < 1e-6
        return w.toa_illuminance_lux
    ratio = toa((2026, 1, 3, 0, 0, 0)) / toa((2026, 7, 6, 0, 0, 0))
    assert abs(ratio - (1.01671 / 0.98329) ** 2) < 2e-3


def test_frame_set_sun_packs_w(lib):
    # frame_uniforms.rs:227-236 and the splat of state.rs:154
    w = abi.WorldSun()
    assert lib.ps_world_sun(C.byref(_utc((2026, 3, 20, 15, 0, 0))), 45.0, 7.0, 300.0, C.byref(w)) == abi.PS_OK
    f = abi.FrameUniforms()
    assert lib.ps_frame_set_sun(C.byref(f), C.byref(w), C.c_float(0.00471)) == abi.PS_OK
    assert list(f.sun_direction[:3]) == list(w.direction_world)
    assert f.sun_direction[3] == np.float32(0.00471)
    assert list(f.sun_illuminance) == [w.toa_illuminance_lux] * 4


def test_error_paths(lib):
    p = abi.SunPosition()
    for t in ((2026, 13, 1, 0, 0, 0), (2026, 2, 30, 0, 0, 0), (2026, 1, 1, 24, 0, 0)):
        assert lib.ps_sun_position_precise(C.byref(_utc(t)), 0.0, 0.0, 0.0, C.byref(p)) == abi.PS_ERR_INVALID_ARGUMENT
    assert lib.ps_sun_position_precise(C.byref(_utc(NREL_UTC)), 91.0, 0.0, 0.0, C.byref(p)) == \
        abi.PS_ERR_INVALID_ARGUMENT
    assert b"latitude" in lib.ps_last_error()
    assert lib.ps_sun_position_precise(None, 0.0, 0.0, 0.0, C.byref(p)) == abi.PS_ERR_INVALID_ARGUMENT
