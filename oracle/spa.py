"""Oracle: the sun position and TOA illuminance of the reference, restated.

Test infrastructure only (tests/, smoke() and bench.py's CPU baseline may use
it; the product never does). A plain-Python FP64 restatement of
crates/ps-core/src/astro/spa.rs (NREL SPA, Reda and Andreas 2003) and of the
sun half of WorldState::recompute (crates/ps-core/src/world.rs:95-111).
Each function cites the reference lines it follows.

Pinned by: the reference's own tests (J2000 and NREL example Julian days,
spa.rs:304-323; Delta T 2026, spa.rs:326-333; the Appendix B zenith and
azimuth, spa.rs:347-360; TOA at 1 AU, world.rs:139-146; altaz_to_world axes,
world.rs:117-135) and the intermediate values NREL/TP-560-34302 publishes for
its Appendix B example (Table A5.1), which `delta_t_override` reproduces with
the paper's Delta T of 67 s.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

from . import spa_terms as T

TOA_ILLUMINANCE_AT_1AU_LUX = 127_500.0   # world.rs:17


def julian_day(year, month, day, hour=0, minute=0, second=0) -> float:
    """spa.rs:194-216 (SPA eq. 4)."""
    y, m = float(year), float(month)
    frac = day + (hour + (minute + second / 60.0) / 60.0) / 24.0
    if m <= 2.0:
        y -= 1.0
        m += 12.0
    a = math.floor(y / 100.0)
    greg = year > 1582 or (year == 1582 and month > 10) or (year == 1582 and month == 10 and day >= 15)
    b = 2.0 - a + math.floor(a / 4.0) if greg else 0.0
    return math.floor(365.25 * (y + 4716.0)) + math.floor(30.6001 * (m + 1.0)) + frac + b - 1524.5


def delta_t_seconds(year: int) -> float:
    """spa.rs:222-249."""
    if 2005 <= year <= 2050:
        t = year - 2000.0
        return 62.92 + 0.32217 * t + 0.005589 * t * t
    if 2050 < year < 2150:
        u = (year - 1820.0) / 100.0
        return -20.0 + 32.0 * u * u - 0.5628 * (2150.0 - year)
    if 1986 <= year < 2005:
        t = year - 2000.0
        return (63.86 + 0.3345 * t - 0.060374 * t ** 2 + 0.0017275 * t ** 3 + 0.000651814 * t ** 4
                - 0.00002371599 * t ** 5)
    u = (year - 1820.0) / 100.0
    return -20.0 + 32.0 * u * u


def _wrap(x: float) -> float:
    r = math.fmod(x, 360.0)
    return r + 360.0 if r < 0.0 else r


def _series(jme: float, tables) -> float:
    """spa.rs:270-281: sum_i jme^i sum_rows A cos(B + C jme), / 1e8."""
    total = 0.0
    for i, tab in enumerate(tables):
        total += sum(a * math.cos(b + c * jme) for a, b, c in tab) * jme ** i
    return total / 1e8


def _nutation(jce: float):
    """spa.rs:289-302 with the arguments of tables.rs:310-337; degrees."""
    x = [math.radians(v) for v in (
        297.85036 + jce * (445267.111480 + jce * (-0.0019142 + jce / 189474.0)),
        357.52772 + jce * (35999.050340 + jce * (-0.0001603 - jce / 300000.0)),
        134.96298 + jce * (477198.867398 + jce * (0.0086972 + jce / 56250.0)),
        93.27191 + jce * (483202.017538 + jce * (-0.0036825 + jce / 327270.0)),
        125.04452 + jce * (-1934.136261 + jce * (0.0020708 + jce / 450000.0)))]
    dpsi = deps = 0.0
    for y, a, b, c, d in T.NUTATION:
        arg = sum(yi * xi for yi, xi in zip(y, x))
        dpsi += (a + b * jce) * math.sin(arg)
        deps += (c + d * jce) * math.cos(arg)
    return dpsi / 36_000_000.0, deps / 36_000_000.0


@dataclass
class SunPosition:
    altitude_deg: float
    azimuth_deg: float
    zenith_deg: float
    distance_au: float
    right_ascension_deg: float
    declination_deg: float
    apparent_sidereal_time_deg: float
    julian_day: float
    # intermediates (for the published-value pins)
    l_deg: float = 0.0
    b_deg: float = 0.0
    dpsi_deg: float = 0.0
    deps_deg: float = 0.0
    eps_deg: float = 0.0
    lambda_deg: float = 0.0
    hour_angle_deg: float = 0.0
    topo_declination_deg: float = 0.0
    topo_hour_angle_deg: float = 0.0


def sun_position_precise(utc, latitude_deg, longitude_deg, elevation_m, delta_t_override=None) -> SunPosition:
    """spa.rs:60-185. `utc` = (year, month, day, hour, minute, second)."""
    year = utc[0]
    jd = julian_day(*utc)
    dt = delta_t_seconds(year) if delta_t_override is None else delta_t_override
    jde = jd + dt / 86400.0
    jc = (jd - 2451545.0) / 36525.0
    jce = (jde - 2451545.0) / 36525.0
    jme = jce / 10.0
    l_deg = _wrap(math.degrees(_series(jme, (T.L0, T.L1, T.L2, T.L3, T.L4, T.L5))))
    b_deg = math.degrees(_series(jme, (T.B0, T.B1)))
    r_au = _series(jme, (T.R0, T.R1, T.R2, T.R3, T.R4))
    theta = _wrap(l_deg + 180.0)
    beta = -b_deg
    dpsi, deps = _nutation(jce)
    u = jme / 10.0
    eps0 = (84381.448 - 4680.93 * u - 1.55 * u ** 2 + 1999.25 * u ** 3 - 51.38 * u ** 4 - 249.67 * u ** 5
            - 39.05 * u ** 6 + 7.12 * u ** 7 + 27.87 * u ** 8 + 5.79 * u ** 9 + 2.45 * u ** 10)
    eps = eps0 / 3600.0 + deps
    lam = theta + dpsi + (-20.4898 / (3600.0 * r_au))
    v0 = _wrap(280.46061837 + 360.98564736629 * (jd - 2451545.0) + 0.000387933 * jc ** 2 - jc ** 3 / 38710000.0)
    v = _wrap(v0 + dpsi * math.cos(math.radians(eps)))
    lr, br, er = math.radians(lam), math.radians(beta), math.radians(eps)
    alpha = _wrap(math.degrees(math.atan2(math.sin(lr) * math.cos(er) - math.tan(br) * math.sin(er), math.cos(lr))))
    delta = math.degrees(math.asin(math.sin(br) * math.cos(er) + math.cos(br) * math.sin(er) * math.sin(lr)))
    h = _wrap(v + longitude_deg - alpha)
    xi = math.radians(8.794 / (3600.0 * r_au))
    phi = math.radians(latitude_deg)
    ut = math.atan(0.99664719 * math.tan(phi))
    x = math.cos(ut) + elevation_m / 6378140.0 * math.cos(phi)
    y = 0.99664719 * math.sin(ut) + elevation_m / 6378140.0 * math.sin(phi)
    hr, dr = math.radians(h), math.radians(delta)
    den = math.cos(dr) - x * math.sin(xi) * math.cos(hr)
    dalpha = math.atan2(-x * math.sin(xi) * math.sin(hr), den)
    dtopo = math.atan2((math.sin(dr) - y * math.sin(xi)) * math.cos(dalpha), den)
    htopo = hr - dalpha
    zen = math.acos(math.sin(phi) * math.sin(dtopo) + math.cos(phi) * math.cos(dtopo) * math.cos(htopo))
    az_s = math.atan2(math.sin(htopo), math.cos(htopo) * math.sin(phi) - math.tan(dtopo) * math.cos(phi))
    zen_deg = math.degrees(zen)
    return SunPosition(90.0 - zen_deg, _wrap(math.degrees(az_s) + 180.0), zen_deg, r_au, alpha, delta, v, jd,
                       l_deg, b_deg, dpsi, deps, eps, lam, h, math.degrees(dtopo), math.degrees(htopo))


def altaz_to_world(altitude_rad: float, azimuth_rad: float):
    """world.rs:106-111 (+X east, +Y up, +Z south)."""
    ca, sa = math.cos(altitude_rad), math.sin(altitude_rad)
    return (ca * math.sin(azimuth_rad), sa, -ca * math.cos(azimuth_rad))


def world_sun(utc, latitude_deg, longitude_deg, elevation_m):
    """world.rs:95-100 on sun_position's FP32 fields (spa.rs:53-58):
    (altitude_rad, azimuth_rad, distance_au, direction, toa_lux)."""
    import numpy as np
    p = sun_position_precise(utc, latitude_deg, longitude_deg, elevation_m)
    alt = float(np.float32(math.radians(p.altitude_deg)))
    az = float(np.float32(math.radians(p.azimuth_deg)))
    d = float(np.float32(p.distance_au))
    inv = 1.0 / max(d, 1e-6)
    return alt, az, d, altaz_to_world(alt, az), TOA_ILLUMINANCE_AT_1AU_LUX * inv * inv
