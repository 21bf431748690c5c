"""Parallactic angles per time and antenna from the antennas' ITRF
positions, the field centre and the times (the E-beam rotation input).

Reference: montblanc/util/__init__.py:414-462,
``parallactic_angles(field_centre, antenna_positions, times)`` ->
(ntime, na). For each antenna and time the reference sets the antenna's
ITRF position and the UTC epoch as the measures frame and asks casacore for
``posangle(J2000 field centre, AZEL zenith)``: casacore converts the second
direction into the first one's frame, so the result is the position angle
(north through east) at the field centre of the antenna's zenith
expressed in J2000 coordinates. This module restates that chain without
casacore:

* the AZEL zenith is the direction of the antenna's geocentric position
  vector (casacore's AZEL uses the geocentric latitude; AZELGEO would use
  the geodetic one); in the true equator and equinox of date it lies at
  right ascension = local apparent sidereal time, declination = geocentric
  latitude;
* apparent sidereal time = Greenwich mean sidereal time (IAU 1982, Aoki et
  al., A&A 105, 1982; UTC taken for UT1, as casacore does without IERS
  tables) plus the equation of the equinoxes dpsi cos(eps);
* true of date -> J2000: annual aberration removed (first order in v/c,
  the Earth's velocity from the Sun's low-accuracy true longitude and the
  orbit's eccentricity and perihelion, Meeus, Astronomical Algorithms,
  2nd ed., ch. 23 and 25), nutation removed (IAU 1980 in the truncated
  four-term form of Meeus ch. 22: 0.5" in dpsi, 0.1" in deps), then IAU
  1976 precession removed (Lieske et al., A&A 58, 1977: zeta, z, theta);
* position angle at the field centre: atan2(d . east, d . north) with the
  J2000 north and east unit vectors at the field centre.

Left out, each below an arcsecond of zenith direction (so below 5e-6 rad
of angle for sources more than 12 degrees from the zenith): the
nutation series' smaller terms, diurnal aberration (0.3"), gravitational
light deflection, polar motion and UT1 - UTC. Times are MeasurementSet
TIME values: UTC seconds since MJD 0; terrestrial time is taken as
UTC + 69.184 s (37 leap seconds, valid since 2017; an error of tens of
seconds moves the precession by 1e-9 rad).
"""

import numpy as np

WGS84_A = 6378137.0
WGS84_F = 1.0 / 298.257223563
ARCSEC = np.pi / (180.0 * 3600.0)
TT_MINUS_UTC = 69.184
# constant of aberration, arcseconds
KAPPA = 20.49552


def geodetic(positions):
    """ITRF (x, y, z) in metres, shape (..., 3) -> (latitude, longitude) in
    radians on the WGS84 ellipsoid (Bowring, Survey Review 23, 1976, with
    refinements)."""
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


def geocentric(positions):
    """ITRF (x, y, z) -> (geocentric latitude, longitude) in radians."""
    p = np.asarray(positions, np.float64)
    return np.arctan2(p[..., 2], np.hypot(p[..., 0], p[..., 1])), np.arctan2(p[..., 1], p[..., 0])


def itrf(lat, lon, height=0.0):
    """Inverse of geodetic: WGS84 latitude, longitude (radians) and height
    (metres) -> ITRF (x, y, z)."""
    a, f = WGS84_A, WGS84_F
    e2 = f * (2.0 - f)
    n = a / np.sqrt(1.0 - e2 * np.sin(lat) ** 2)
    return np.stack([(n + height) * np.cos(lat) * np.cos(lon),
                     (n + height) * np.cos(lat) * np.sin(lon),
                     (n * (1.0 - e2) + height) * np.sin(lat)], -1)


def julian_centuries(times, offset=0.0):
    """Julian centuries since J2000.0 of MJD seconds (plus offset seconds)."""
    return ((np.asarray(times, np.float64) + offset) / 86400.0 + 2400000.5 - 2451545.0) / 36525.0


def gmst(times):
    """Greenwich mean sidereal time (radians, in [0, 2 pi)) at UTC seconds
    since MJD 0."""
    T = julian_centuries(times)
    # seconds of sidereal time: 24110.54841 s at J2000.0 0h UT1 plus the
    # rotation; written from the 12h epoch (67310.54841 s)
    s = 67310.54841 + (876600.0 * 3600.0 + 8640184.812866) * T + 0.093104 * T ** 2 - 6.2e-6 * T ** 3
    return np.mod(s / 86400.0 * 2.0 * np.pi, 2.0 * np.pi)


def precession_angles(T):
    """IAU 1976 precession angles (zeta, z, theta) in radians from J2000.0
    to T Julian centuries (TT) after it (Lieske et al. 1977)."""
    T = np.asarray(T, np.float64)
    zeta = (2306.2181 * T + 0.30188 * T ** 2 + 0.017998 * T ** 3) * ARCSEC
    z = (2306.2181 * T + 1.09468 * T ** 2 + 0.018203 * T ** 3) * ARCSEC
    theta = (2004.3109 * T - 0.42665 * T ** 2 - 0.041833 * T ** 3) * ARCSEC
    return zeta, z, theta


def _r1(a):
    c, s = np.cos(a), np.sin(a)
    o, i = np.zeros_like(c), np.ones_like(c)
    return np.stack([np.stack([i, o, o], -1), np.stack([o, c, s], -1),
                     np.stack([o, -s, c], -1)], -2)


def _r2(a):
    c, s = np.cos(a), np.sin(a)
    o, i = np.zeros_like(c), np.ones_like(c)
    return np.stack([np.stack([c, o, -s], -1), np.stack([o, i, o], -1),
                     np.stack([s, o, c], -1)], -2)


def _r3(a):
    c, s = np.cos(a), np.sin(a)
    o, i = np.zeros_like(c), np.ones_like(c)
    return np.stack([np.stack([c, s, o], -1), np.stack([-s, c, o], -1),
                     np.stack([o, o, i], -1)], -2)


def precession_matrix(T):
    """(..., 3, 3) rotation taking J2000 mean equatorial vectors to the mean
    equator and equinox of date: R3(-z) R2(theta) R3(-zeta)."""
    zeta, z, theta = precession_angles(T)
    return _r3(-z) @ _r2(theta) @ _r3(-zeta)


def nutation(T):
    """(dpsi, deps, eps0) in radians at T Julian centuries (TT): nutation in
    longitude and obliquity (IAU 1980, four largest terms) and the mean
    obliquity of the ecliptic (IAU 1976)."""
    T = np.asarray(T, np.float64)
    L = np.radians(280.4665 + 36000.7698 * T)          # Sun's mean longitude
    Lm = np.radians(218.3165 + 481267.8813 * T)        # Moon's mean longitude
    om = np.radians(125.04452 - 1934.136261 * T + 0.0020708 * T ** 2 + T ** 3 / 450000.0)
    dpsi = -17.20 * np.sin(om) - 1.32 * np.sin(2 * L) - 0.23 * np.sin(2 * Lm) + 0.21 * np.sin(2 * om)
    deps = 9.20 * np.cos(om) + 0.57 * np.cos(2 * L) + 0.10 * np.cos(2 * Lm) - 0.09 * np.cos(2 * om)
    eps0 = 84381.448 - 46.8150 * T - 0.00059 * T ** 2 + 0.001813 * T ** 3
    return dpsi * ARCSEC, deps * ARCSEC, eps0 * ARCSEC


def nutation_matrix(T):
    """(..., 3, 3) rotation from the mean to the true equator and equinox of
    date: R1(-(eps0 + deps)) R3(-dpsi) R1(eps0)."""
    dpsi, deps, eps0 = nutation(T)
    return _r1(-(eps0 + deps)) @ _r3(-dpsi) @ _r1(eps0)


def gast(times):
    """Greenwich apparent sidereal time (radians): GMST plus the equation
    of the equinoxes dpsi cos(eps)."""
    dpsi, deps, eps0 = nutation(julian_centuries(times, TT_MINUS_UTC))
    return np.mod(gmst(times) + dpsi * np.cos(eps0 + deps), 2.0 * np.pi)


def earth_velocity(T):
    """The Earth's velocity over c, (..., 3), in the true equator and equinox
    of date, from the Sun's low-accuracy geometric longitude (Meeus ch. 25)
    and the orbit's eccentricity e and perihelion longitude pi (ch. 23):
    v / c = kappa (sin L - e sin pi, -(cos L - e cos pi), 0) in ecliptic
    coordinates."""
    T = np.asarray(T, np.float64)
    L0 = 280.46646 + 36000.76983 * T + 0.0003032 * T ** 2
    M = np.radians(357.52911 + 35999.05029 * T - 0.0001537 * T ** 2)
    C = ((1.914602 - 0.004817 * T - 0.000014 * T ** 2) * np.sin(M)
         + (0.019993 - 0.000101 * T) * np.sin(2 * M) + 0.000289 * np.sin(3 * M))
    sun = np.radians(L0 + C)
    e = 0.016708634 - 0.000042037 * T - 0.0000001267 * T ** 2
    peri = np.radians(102.93735 + 1.71946 * T + 0.00046 * T ** 2)
    k = KAPPA * ARCSEC
    vx = k * (np.sin(sun) - e * np.sin(peri))
    vy = -k * (np.cos(sun) - e * np.cos(peri))
    dpsi, deps, eps0 = nutation(T)
    eps = eps0 + deps
    return np.stack([vx, vy * np.cos(eps), vy * np.sin(eps)], -1)


def unit(ra, dec):
    return np.stack([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)], -1)


def radec(u):
    return np.mod(np.arctan2(u[..., 1], u[..., 0]), 2 * np.pi), np.arcsin(np.clip(u[..., 2], -1, 1))


def add_aberration(u, beta):
    """Apparent direction of a geometric unit vector u seen with velocity
    beta = v / c (first order)."""
    d = u + beta - u * (u * beta).sum(-1, keepdims=True)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def remove_aberration(u, beta):
    """Inverse of add_aberration to first order in beta."""
    d = u - beta + u * (u * beta).sum(-1, keepdims=True)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def zenith_j2000(antenna_positions, times):
    """(ntime, na, 3) J2000 unit vectors of each antenna's zenith (the
    direction of its geocentric position vector) at each time."""
    lat, lon = geocentric(antenna_positions)
    T = julian_centuries(times, TT_MINUS_UTC)
    last = gast(times)[:, None] + lon[None, :]
    z_true = unit(last, np.broadcast_to(lat[None, :], last.shape))
    z_true = remove_aberration(z_true, earth_velocity(T)[:, None, :])
    rot = nutation_matrix(T) @ precession_matrix(T)        # J2000 -> true of date
    return np.einsum('tji,taj->tai', rot, z_true)


def position_angle(centre, d):
    """Position angle (north through east) at the direction centre = (ra,
    dec) of the unit vectors d (..., 3) in the same frame."""
    ra, dec = float(centre[0]), float(centre[1])
    north = np.array([-np.sin(dec) * np.cos(ra), -np.sin(dec) * np.sin(ra), np.cos(dec)])
    east = np.array([-np.sin(ra), np.cos(ra), 0.0])
    return np.arctan2(d @ east, d @ north)


def parallactic_angle(hour_angle, dec, lat):
    """Position angle of the zenith at a source (north through east) from
    its hour angle and declination and the observer's latitude, all in one
    frame."""
    return np.arctan2(np.cos(lat) * np.sin(hour_angle),
                      np.sin(lat) * np.cos(dec) - np.cos(lat) * np.sin(dec) * np.cos(hour_angle))


def parallactic_angles(field_centre, antenna_positions, times):
    """(ntime, na) parallactic angles: the position angle at the J2000
    field centre of each antenna's zenith in J2000 (see the module
    docstring)."""
    ra, dec = float(field_centre[0]), float(field_centre[1])
    if ra < 0:
        ra += 2.0 * np.pi            # as the reference's loader does
    z = zenith_j2000(np.atleast_2d(antenna_positions), np.atleast_1d(times))
    return position_angle((ra, dec), z)
