"""Parallactic angles per time and antenna from the antennas' ITRF
positions, the field centre and the times (the E-beam rotation input).

Reference interface: montblanc/util/__init__.py:414-462,
``parallactic_angles(field_centre, antenna_positions, times)`` ->
(ntime, na), the position angle of the zenith seen from the field centre
(north through east), which the reference gets from casacore measures.
Restated here from spherical astronomy, without casacore:

* geodetic latitude and longitude of an ITRF position on the WGS84
  ellipsoid (a = 6378137 m, f = 1 / 298.257223563), by Bowring's method
  (B. R. Bowring, Survey Review 23, 1976) with two refinements;
* Greenwich mean sidereal time from the IAU 1982 expression (Aoki et al.,
  A&A 105, 1982) with UTC taken for UT1 (|UT1 - UTC| < 0.9 s, under
  0.004 degrees of hour angle);
* hour angle H = GMST + longitude - right ascension, and
  q = atan2(cos(lat) sin H, sin(lat) cos(dec) - cos(lat) sin(dec) cos H).

The field centre is taken as given (J2000), not precessed to the date as
casacore does; the difference is a fraction of a degree at the present
epoch. Times are MeasurementSet TIME values: UTC seconds since MJD 0.
"""

import numpy as np

WGS84_A = 6378137.0
WGS84_F = 1.0 / 298.257223563


def geodetic(positions):
    """ITRF (x, y, z) in metres, shape (..., 3) -> (latitude, longitude) in
    radians on the WGS84 ellipsoid."""
    p = np.asarray(positions, np.float64)
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    a, f = WGS84_A, WGS84_F
    b = a * (1.0 - f)
    e2 = f * (2.0 - f)
    ep2 = e2 / (1.0 - e2)
    r = np.hypot(x, y)
    lon = np.arctan2(y, x)
    beta = np.arctan2(z * a, r * b)                  # parametric latitude, first guess
    for _ in range(3):
        lat = np.arctan2(z + ep2 * b * np.sin(beta) ** 3, r - e2 * a * np.cos(beta) ** 3)
        beta = np.arctan2((1.0 - f) * np.sin(lat), np.cos(lat))
    return lat, lon


def itrf(lat, lon, height=0.0):
    """Inverse of geodetic: WGS84 latitude, longitude (radians) and height
    (metres) -> ITRF (x, y, z)."""
    a, f = WGS84_A, WGS84_F
    e2 = f * (2.0 - f)
    n = a / np.sqrt(1.0 - e2 * np.sin(lat) ** 2)
    return np.stack([(n + height) * np.cos(lat) * np.cos(lon),
                     (n + height) * np.cos(lat) * np.sin(lon),
                     (n * (1.0 - e2) + height) * np.sin(lat)], -1)


def gmst(times):
    """Greenwich mean sidereal time (radians, in [0, 2 pi)) at UTC seconds
    since MJD 0."""
    t = np.asarray(times, np.float64)
    jd = t / 86400.0 + 2400000.5
    T = (jd - 2451545.0) / 36525.0
    # seconds of sidereal time: 24110.54841 s at J2000.0 0h UT1 plus the
    # rotation; written from the 12h epoch (67310.54841 s)
    s = 67310.54841 + (876600.0 * 3600.0 + 8640184.812866) * T + 0.093104 * T ** 2 - 6.2e-6 * T ** 3
    return np.mod(s / 86400.0 * 2.0 * np.pi, 2.0 * np.pi)


def parallactic_angle(hour_angle, dec, lat):
    """Position angle of the zenith at a source (north through east)."""
    return np.arctan2(np.cos(lat) * np.sin(hour_angle),
                      np.sin(lat) * np.cos(dec) - np.cos(lat) * np.sin(dec) * np.cos(hour_angle))


def parallactic_angles(field_centre, antenna_positions, times):
    """(ntime, na) parallactic angles (see the module docstring)."""
    ra, dec = float(field_centre[0]), float(field_centre[1])
    if ra < 0:
        ra += 2.0 * np.pi            # as the reference's loader does
    lat, lon = geodetic(antenna_positions)
    H = gmst(times)[:, None] + lon[None, :] - ra
    return parallactic_angle(H, dec, lat[None, :])
