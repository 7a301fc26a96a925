// Sun position and top-of-atmosphere illuminance for the frame inputs (host
// code; a handful of FP64 series evaluations per frame does not belong on the
// GPU).
//
// The NREL Solar Position Algorithm (Reda and Andreas, NREL/TP-560-34302,
// 2003) in the form crates/ps-core/src/astro/spa.rs:60-185 evaluates it:
// Julian day from the UTC calendar date (no sub-second part), Delta T from
// the year's Espenak & Meeus 2006 polynomial segment, heliocentric L, B, R
// from the periodic series, nutation, true obliquity, aberration, apparent
// sidereal time, geocentric and then topocentric coordinates, and the
// topocentric zenith and azimuth without the atmospheric refraction term.
// Then WorldState::recompute's sun half (crates/ps-core/src/world.rs:95-100):
// the FP32 altitude / azimuth turned into the world direction (+X east, +Y
// up, +Z south, world.rs:106-111) and the TOA illuminance 127500 lux / d^2.
#include <cmath>
#include <string>

#include "ps_host.h"
#include "spa_terms.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kRadPerDeg = kPi / 180.0;
constexpr double kDegPerRad = 180.0 / kPi;
constexpr float kToaLuxAt1Au = 127500.f;  // world.rs:17

double wrap360(double x) {
    const double r = std::fmod(x, 360.0);
    return r < 0.0 ? r + 360.0 : r;
}

// Sum over the tables of one quantity: sum_i jme^i * sum_rows A cos(B + C jme),
// scaled by 1e-8 (spa.rs:270-281). jme^i by repeated multiplication.
template <size_t N>
double table_sum(const spa::Term (&t)[N], double jme) {
    double s = 0.0;
    for (size_t k = 0; k < N; ++k) s += t[k].a * std::cos(t[k].b + t[k].c * jme);
    return s;
}

double earth_l(double jme) {
    const double s[6] = {table_sum(spa::kL0, jme), table_sum(spa::kL1, jme), table_sum(spa::kL2, jme),
                         table_sum(spa::kL3, jme), table_sum(spa::kL4, jme), table_sum(spa::kL5, jme)};
    double total = 0.0, p = 1.0;
    for (int i = 0; i < 6; ++i) {
        total += s[i] * p;
        p *= jme;
    }
    return total / 1e8;
}

double earth_b(double jme) {
    return (table_sum(spa::kB0, jme) + table_sum(spa::kB1, jme) * jme) / 1e8;
}

double earth_r(double jme) {
    const double s[5] = {table_sum(spa::kR0, jme), table_sum(spa::kR1, jme), table_sum(spa::kR2, jme),
                         table_sum(spa::kR3, jme), table_sum(spa::kR4, jme)};
    double total = 0.0, p = 1.0;
    for (int i = 0; i < 5; ++i) {
        total += s[i] * p;
        p *= jme;
    }
    return total / 1e8;
}

// Nutation in longitude and obliquity, degrees (spa.rs:289-302 with the
// fundamental arguments of tables.rs:310-337: mean elongation of the Moon,
// mean anomalies of the Sun and the Moon, the Moon's argument of latitude
// and the longitude of its ascending node).
void nutation(double jce, double& dpsi_deg, double& deps_deg) {
    const double x[5] = {
        (297.85036 + jce * (445267.111480 + jce * (-0.0019142 + jce / 189474.0))) * kRadPerDeg,
        (357.52772 + jce * (35999.050340 + jce * (-0.0001603 - jce / 300000.0))) * kRadPerDeg,
        (134.96298 + jce * (477198.867398 + jce * (0.0086972 + jce / 56250.0))) * kRadPerDeg,
        (93.27191 + jce * (483202.017538 + jce * (-0.0036825 + jce / 327270.0))) * kRadPerDeg,
        (125.04452 + jce * (-1934.136261 + jce * (0.0020708 + jce / 450000.0))) * kRadPerDeg,
    };
    double psi = 0.0, eps = 0.0;
    for (const spa::NutTerm& t : spa::kNutation) {
        double arg = 0.0;
        for (int i = 0; i < 5; ++i) arg += (double)t.y[i] * x[i];
        psi += (t.a + t.b * jce) * std::sin(arg);
        eps += (t.c + t.d * jce) * std::cos(arg);
    }
    // 1e-4 arcsec -> degrees
    dpsi_deg = psi / 36000000.0;
    deps_deg = eps / 36000000.0;
}

bool valid_utc(const ps_utc* u) {
    static const int mdays[12] = {31, 29, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    if (u->month < 1 || u->month > 12 || u->day < 1 || u->day > mdays[u->month - 1]) return false;
    if (u->hour < 0 || u->hour > 23 || u->minute < 0 || u->minute > 59 || u->second < 0 || u->second > 60)
        return false;
    return true;
}

}  // namespace

extern "C" {

double ps_julian_day(const ps_utc* u) {
    // spa.rs:194-216: Meeus's calendar-to-JD with the Gregorian correction
    // from 1582-10-15 on
    double y = (double)u->year, m = (double)u->month;
    const double day = (double)u->day +
                       ((double)u->hour + ((double)u->minute + (double)u->second / 60.0) / 60.0) / 24.0;
    if (m <= 2.0) {
        y -= 1.0;
        m += 12.0;
    }
    const double a = std::floor(y / 100.0);
    const bool gregorian = u->year > 1582 || (u->year == 1582 && u->month > 10) ||
                           (u->year == 1582 && u->month == 10 && u->day >= 15);
    const double b = gregorian ? 2.0 - a + std::floor(a / 4.0) : 0.0;
    return std::floor(365.25 * (y + 4716.0)) + std::floor(30.6001 * (m + 1.0)) + day + b - 1524.5;
}

double ps_delta_t_seconds(int32_t year) {
    // spa.rs:222-249 (Espenak & Meeus 2006 segments)
    const double yr = (double)year;
    if (year >= 2005 && year <= 2050) {
        const double t = yr - 2000.0;
        return 62.92 + 0.32217 * t + 0.005589 * t * t;
    }
    if (year > 2050 && year < 2150) {
        const double u = (yr - 1820.0) / 100.0;
        return -20.0 + 32.0 * u * u - 0.5628 * (2150.0 - yr);
    }
    if (year >= 1986 && year < 2005) {
        const double t = yr - 2000.0, t2 = t * t, t3 = t2 * t;
        return 63.86 + 0.3345 * t - 0.060374 * t2 + 0.0017275 * t3 + 0.000651814 * (t2 * t2) -
               0.00002371599 * (t2 * t3);
    }
    const double u = (yr - 1820.0) / 100.0;
    return -20.0 + 32.0 * u * u;
}

ps_status ps_sun_position_precise(const ps_utc* utc, double latitude_deg, double longitude_deg,
                                  double elevation_m, ps_sun_position* out) {
    if (!utc || !out) return psi::fail(PS_ERR_INVALID_ARGUMENT, "ps_sun_position_precise: NULL argument");
    if (!valid_utc(utc)) return psi::fail(PS_ERR_INVALID_ARGUMENT, "ps_sun_position_precise: invalid UTC date");
    if (!(latitude_deg >= -90.0 && latitude_deg <= 90.0) || !std::isfinite(longitude_deg) ||
        !std::isfinite(elevation_m))
        return psi::fail(PS_ERR_INVALID_ARGUMENT, "ps_sun_position_precise: latitude must be in [-90, 90]");
    // time scales (SPA 3.1)
    const double jd = ps_julian_day(utc);
    const double jde = jd + ps_delta_t_seconds(utc->year) / 86400.0;
    const double jc = (jd - 2451545.0) / 36525.0;
    const double jce = (jde - 2451545.0) / 36525.0;
    const double jme = jce / 10.0;

    // heliocentric -> geocentric (3.2, 3.3)
    const double l_deg = wrap360(earth_l(jme) * kDegPerRad);
    const double b_deg = earth_b(jme) * kDegPerRad;
    const double r_au = earth_r(jme);
    const double theta = wrap360(l_deg + 180.0);
    const double beta = -b_deg;

    // nutation and the true obliquity (3.4, 3.5; Laskar's polynomial in U)
    double dpsi, deps;
    nutation(jce, dpsi, deps);
    const double u = jme / 10.0;
    const double u2 = u * u, u3 = u2 * u, u4 = u2 * u2, u5 = u4 * u, u6 = u3 * u3, u7 = u6 * u, u8 = u4 * u4,
                 u9 = u8 * u, u10 = u5 * u5;
    const double eps0 = 84381.448 - 4680.93 * u - 1.55 * u2 + 1999.25 * u3 - 51.38 * u4 - 249.67 * u5 -
                        39.05 * u6 + 7.12 * u7 + 27.87 * u8 + 5.79 * u9 + 2.45 * u10;
    const double eps = eps0 / 3600.0 + deps;

    // aberration, apparent longitude, sidereal time (3.6-3.8)
    const double lambda = theta + dpsi + (-20.4898 / (3600.0 * r_au));
    const double v0 = wrap360(280.46061837 + 360.98564736629 * (jd - 2451545.0) + 0.000387933 * (jc * jc) -
                              (jc * jc * jc) / 38710000.0);
    const double v = wrap360(v0 + dpsi * std::cos(eps * kRadPerDeg));

    // geocentric right ascension and declination (3.9)
    const double lr = lambda * kRadPerDeg, br = beta * kRadPerDeg, er = eps * kRadPerDeg;
    const double alpha =
        wrap360(std::atan2(std::sin(lr) * std::cos(er) - std::tan(br) * std::sin(er), std::cos(lr)) * kDegPerRad);
    const double delta =
        std::asin(std::sin(br) * std::cos(er) + std::cos(br) * std::sin(er) * std::sin(lr)) * kDegPerRad;

    // local hour angle, topocentric parallax (3.10, 3.11)
    const double h_deg = wrap360(v + longitude_deg - alpha);
    const double xi = (8.794 / (3600.0 * r_au)) * kRadPerDeg;
    const double phi = latitude_deg * kRadPerDeg;
    const double ut = std::atan(0.99664719 * std::tan(phi));
    const double x = std::cos(ut) + (elevation_m / 6378140.0) * std::cos(phi);
    const double y = 0.99664719 * std::sin(ut) + (elevation_m / 6378140.0) * std::sin(phi);
    const double hr = h_deg * kRadPerDeg, dr = delta * kRadPerDeg;
    const double den = std::cos(dr) - x * std::sin(xi) * std::cos(hr);
    const double dalpha = std::atan2(-x * std::sin(xi) * std::sin(hr), den);
    const double delta_topo = std::atan2((std::sin(dr) - y * std::sin(xi)) * std::cos(dalpha), den);
    const double h_topo = hr - dalpha;

    // topocentric zenith and azimuth (3.13, no refraction); azimuth from
    // north, clockwise
    const double cos_z = std::sin(phi) * std::sin(delta_topo) + std::cos(phi) * std::cos(delta_topo) * std::cos(h_topo);
    const double zen = std::acos(cos_z);
    const double az_south =
        std::atan2(std::sin(h_topo), std::cos(h_topo) * std::sin(phi) - std::tan(delta_topo) * std::cos(phi));

    out->zenith_deg = zen * kDegPerRad;
    out->altitude_deg = 90.0 - out->zenith_deg;
    out->azimuth_deg = wrap360(az_south * kDegPerRad + 180.0);
    out->distance_au = r_au;
    out->right_ascension_deg = alpha;
    out->declination_deg = delta;
    out->apparent_sidereal_time_deg = v;
    out->julian_day = jd;
    return PS_OK;
}

ps_status ps_world_sun(const ps_utc* utc, double latitude_deg, double longitude_deg, double elevation_m,
                       ps_world_sun_state* out) {
    if (!out) return psi::fail(PS_ERR_INVALID_ARGUMENT, "ps_world_sun: NULL argument");
    ps_sun_position p;
    ps_status s = ps_sun_position_precise(utc, latitude_deg, longitude_deg, elevation_m, &p);
    if (s != PS_OK) return s;
    // sun_position's FP32 fields (spa.rs:53-58), then world.rs:97-100
    const float alt = (float)(p.altitude_deg * kRadPerDeg), az = (float)(p.azimuth_deg * kRadPerDeg);
    const float d = (float)p.distance_au;
    out->altitude_rad = alt;
    out->azimuth_rad = az;
    out->distance_au = d;
    const float sa = std::sin(alt), ca = std::cos(alt), sz = std::sin(az), cz = std::cos(az);
    out->direction_world[0] = ca * sz;
    out->direction_world[1] = sa;
    out->direction_world[2] = -ca * cz;
    const float inv_d = 1.f / std::fmax(d, 1e-6f);
    out->toa_illuminance_lux = kToaLuxAt1Au * inv_d * inv_d;
    return PS_OK;
}

ps_status ps_frame_set_sun(ps_frame_uniforms* f, const ps_world_sun_state* sun, float angular_radius_rad) {
    if (!f || !sun) return psi::fail(PS_ERR_INVALID_ARGUMENT, "ps_frame_set_sun: NULL argument");
    // FrameUniforms::set_sun (frame_uniforms.rs:98-118) with the illuminance
    // WeatherState carries, the TOA lux on every channel (state.rs:154)
    for (int i = 0; i < 3; ++i) f->sun_direction[i] = sun->direction_world[i];
    f->sun_direction[3] = angular_radius_rad;
    for (int i = 0; i < 4; ++i) f->sun_illuminance[i] = sun->toa_illuminance_lux;
    return PS_OK;
}

}  // extern "C"
