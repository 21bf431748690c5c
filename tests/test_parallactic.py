"""Parallactic angles (montblanc_amd/parallactic.py) pinned by published
worked examples (Meeus, Astronomical Algorithms, 2nd ed.: sidereal time
12.a/12.b, precession 21.b, nutation 22.a, apparent place 23.a), by an
independent vector derivation of the position angle, and by a second route
through the frame of date (hour angle of the precessed and nutated field
centre, then the position angle of the J2000 pole there). The reference
computes them with casacore measures (montblanc/util/__init__.py:414-462),
absent here: agreement with casacore's own output is unpinned beyond these
published values."""

import numpy as np

from montblanc_amd import parallactic as pa

DAY = 86400.0
HOUR_RAD = 2.0 * np.pi / 24.0


def test_gmst_published_values():
    # J2000.0 = JD 2451545.0 = MJD 51544.5: 280.46061837 degrees
    assert abs(np.degrees(pa.gmst(51544.5 * DAY)) - 280.46061837) < 1e-7
    # Meeus, Astronomical Algorithms, examples 12.a / 12.b: 1987 April 10
    # (MJD 46895) at 0h UT -> 13h10m46.3668s, at 19h21m00s UT -> 128.7378734 degrees
    h = pa.gmst(46895.0 * DAY) / HOUR_RAD
    assert abs(h - (13 + 10 / 60 + 46.3668 / 3600)) < 2e-6
    d = np.degrees(pa.gmst(46895.0 * DAY + 19 * 3600 + 21 * 60))
    assert abs(d - 128.7378734) < 2e-5


def test_geodetic_round_trip_and_special_points():
    lat, lon = pa.geodetic([[pa.WGS84_A, 0.0, 0.0], [0.0, pa.WGS84_A, 0.0]])
    np.testing.assert_allclose(lat, [0.0, 0.0], atol=1e-15)
    np.testing.assert_allclose(lon, [0.0, np.pi / 2], atol=1e-15)
    b = pa.WGS84_A * (1 - pa.WGS84_F)
    lat, _ = pa.geodetic([[1e-3, 0.0, b], [1e-3, 0.0, -b]])
    np.testing.assert_allclose(lat, [np.pi / 2, -np.pi / 2], atol=1e-9)
    r = np.random.default_rng(4)
    la = r.uniform(-1.5, 1.5, 50)
    lo = r.uniform(-np.pi, np.pi, 50)
    hh = r.uniform(-100.0, 5000.0, 50)
    la2, lo2 = pa.geodetic(pa.itrf(la, lo, hh))
    np.testing.assert_allclose(la2, la, atol=1e-12)
    np.testing.assert_allclose(lo2, lo, atol=1e-12)


def test_angle_against_vector_geometry():
    """Position angle of the zenith at the source from the zenith's
    components along the source's north and east tangent vectors."""
    r = np.random.default_rng(5)
    lst, ra = r.uniform(0, 2 * np.pi, 200), r.uniform(0, 2 * np.pi, 200)
    dec, lat = r.uniform(-1.5, 1.5, 200), r.uniform(-1.4, 1.4, 200)
    zen = np.stack([np.cos(lat) * np.cos(lst), np.cos(lat) * np.sin(lst), np.sin(lat)], -1)
    north = np.stack([-np.sin(dec) * np.cos(ra), -np.sin(dec) * np.sin(ra), np.cos(dec)], -1)
    east = np.stack([-np.sin(ra), np.cos(ra), np.zeros_like(ra)], -1)
    want = np.arctan2((zen * east).sum(-1), (zen * north).sum(-1))
    got = pa.parallactic_angle(lst - ra, dec, lat)
    np.testing.assert_allclose(np.angle(np.exp(1j * (got - want))), 0, atol=1e-12)
    # on the meridian: 0 south of the zenith, pi north of it; antisymmetric in H
    assert abs(pa.parallactic_angle(0.0, 0.2, 0.5)) < 1e-15
    assert abs(abs(pa.parallactic_angle(0.0, 0.8, 0.5)) - np.pi) < 1e-15
    assert abs(pa.parallactic_angle(0.7, 0.2, 0.5) + pa.parallactic_angle(-0.7, 0.2, 0.5)) < 1e-15


ARC = pa.ARCSEC


def theta_persei():
    """Meeus example 21.b: theta Persei, J2000 place with its proper motion
    carried to 2028 November 13.19 TD (JD 2462088.69)."""
    T = (2462088.69 - 2451545.0) / 36525.0
    a0 = np.radians(15 * (2 + 44 / 60 + 11.986 / 3600)) + np.radians(15 * 0.03425 * 100 * T / 3600)
    d0 = np.radians(49 + 13 / 60 + 42.48 / 3600) - 0.0895 * 100 * T * ARC
    return T, pa.unit(a0, d0)


def test_precession_meeus_21b():
    T, u0 = theta_persei()
    a, d = pa.radec(pa.precession_matrix(T) @ u0)
    # alpha = 2h46m11.331s = 41.547214 deg, delta = +49 20' 54.54" = 49.348483 deg
    assert abs(np.degrees(a) - 41.547214) < 3e-6
    assert abs(np.degrees(d) - 49.348483) < 3e-6
    # orthogonal, and the identity at J2000.0
    P = pa.precession_matrix(T)
    np.testing.assert_allclose(P @ P.T, np.eye(3), atol=1e-15)
    np.testing.assert_allclose(pa.precession_matrix(0.0), np.eye(3), atol=1e-15)


def test_nutation_meeus_22a_and_apparent_sidereal_time_12a():
    T = (2446895.5 - 2451545.0) / 36525.0        # 1987 April 10, 0h TD
    dpsi, deps, eps0 = pa.nutation(T)
    # IAU 1980 series: dpsi = -3.788", deps = +9.443"; eps0 = 23 26' 27.407"
    assert abs(dpsi / ARC + 3.788) < 0.5
    assert abs(deps / ARC - 9.443) < 0.1
    assert abs(eps0 / ARC - (23 * 3600 + 26 * 60 + 27.407)) < 1e-3
    # apparent sidereal time at 0h UT that day: 13h10m46.1351s (mean 46.3668s)
    h = pa.gast(46895.0 * DAY) / HOUR_RAD
    assert abs((h - 13 - 10 / 60) * 3600 - 46.1351) < 0.04


def test_apparent_place_meeus_23a():
    """Nutation and annual aberration of theta Persei's mean place of date:
    aberration +30.045" in right ascension and +6.697" in declination; the
    apparent place 2h46m14.390s, +49 21' 07.45"."""
    T, u0 = theta_persei()
    um = pa.precession_matrix(T) @ u0
    ut = pa.nutation_matrix(T) @ um
    ua = pa.add_aberration(ut, pa.earth_velocity(T))
    a1, d1 = pa.radec(ut)
    a2, d2 = pa.radec(ua)
    assert abs((a2 - a1) / ARC - 30.045) < 0.02
    assert abs((d2 - d1) / ARC - 6.697) < 0.02
    am, dm = pa.radec(um)
    # nutation in position, IAU 1980 series: +15.843", +6.218" (the four-term
    # form is within 0.5" of dpsi)
    assert abs((a1 - am) / ARC - 15.843) < 0.4
    assert abs((d1 - dm) / ARC - 6.218) < 0.1
    assert abs(np.degrees(a2) / 15 * 3600 - (2 * 3600 + 46 * 60 + 14.390)) < 0.03
    assert abs(np.degrees(d2) * 3600 - (49 * 3600 + 21 * 60 + 7.45)) < 0.2
    np.testing.assert_allclose(pa.remove_aberration(ua, pa.earth_velocity(T)), ut, atol=1e-12)


def test_position_angle_against_hour_angle_formula():
    r = np.random.default_rng(6)
    lst, ra = r.uniform(0, 2 * np.pi, 100), r.uniform(0, 2 * np.pi, 100)
    dec, lat = r.uniform(-1.5, 1.5, 100), r.uniform(-1.4, 1.4, 100)
    for i in range(100):
        got = pa.position_angle((ra[i], dec[i]), pa.unit(lst[i], lat[i]))
        want = pa.parallactic_angle(lst[i] - ra[i], dec[i], lat[i])
        assert abs(np.angle(np.exp(1j * (got - want)))) < 1e-12


def _frame_of_date_route(fc, pos, times):
    """The same angle without aberration, through the frame of date: the
    field centre precessed and nutated to the date, its hour angle from the
    apparent sidereal time, the parallactic angle there by the hour-angle
    formula, minus the position angle (at that point) of the J2000 pole
    carried to the date."""
    lat, lon = pa.geocentric(pos)
    T = pa.julian_centuries(times, pa.TT_MINUS_UTC)
    rot = pa.nutation_matrix(T) @ pa.precession_matrix(T)
    fa, fd = pa.radec(rot @ pa.unit(fc[0], fc[1]))
    pa_, pd = pa.radec(rot @ np.array([0.0, 0.0, 1.0]))
    H = pa.gast(times)[:, None] + lon[None, :] - fa[:, None]
    q = pa.parallactic_angle(H, fd[:, None], lat[None, :])
    da = pa_ - fa
    chi = np.arctan2(np.sin(da) * np.cos(pd), np.cos(fd) * np.sin(pd) - np.sin(fd) * np.cos(pd) * np.cos(da))
    return q - chi[:, None]


def test_parallactic_angles_two_routes(monkeypatch):
    pos = pa.itrf(np.radians([-30.7, -30.72, 52.9]), np.radians([21.44, 21.45, 6.6]), 1050.0)
    times = np.array([4.0e9, 4.5e9, 5.0e9 + 600.0, 5.3e9])
    fc = np.array([1.1, -0.4])
    full = pa.parallactic_angles(fc, pos, times)
    assert full.shape == (4, 3)
    monkeypatch.setattr(pa, 'earth_velocity', lambda T: np.zeros(np.shape(T) + (3,)))
    rigid = pa.parallactic_angles(fc, pos, times)
    other = _frame_of_date_route(fc, pos, times)
    np.testing.assert_allclose(np.angle(np.exp(1j * (rigid - other))), 0, atol=1e-12)
    # aberration moves the zenith by at most ~20.5": well under 1e-3 rad here
    assert np.abs(np.angle(np.exp(1j * (full - rigid)))).max() < 1e-3
    # the J2000 / geodetic short cut is off by the frame and latitude terms
    lat, lon = pa.geodetic(pos)
    simple = pa.parallactic_angle(pa.gmst(times)[:, None] + lon[None, :] - fc[0], fc[1], lat[None, :])
    d = np.abs(np.angle(np.exp(1j * (full - simple))))
    assert 1e-4 < d.max() < 2e-2


def test_per_time_and_antenna_and_negative_ra():
    pos = pa.itrf(np.radians([-30.7, -30.72, -30.69]), np.radians([21.44, 21.45, 21.41]), 1050.0)
    times = 5.0e9 + np.arange(4) * 600.0
    q = pa.parallactic_angles(np.array([-0.5, -0.6]), pos, times)
    assert q.shape == (4, 3)
    np.testing.assert_allclose(q, pa.parallactic_angles(np.array([2 * np.pi - 0.5, -0.6]), pos, times),
                               atol=1e-14)
    for k in range(3):
        np.testing.assert_allclose(q[:, k], pa.parallactic_angles(np.array([-0.5, -0.6]), pos[k],
                                                                  times)[:, 0], atol=1e-15)
