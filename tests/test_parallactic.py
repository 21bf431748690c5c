"""Parallactic angles (montblanc_amd/parallactic.py) pinned by published
values and by an independent vector derivation. The reference computes them
with casacore measures (montblanc/util/__init__.py:414-462), absent here:
agreement with casacore itself is unpinned (it also precesses the J2000
field centre to the date)."""

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


def test_per_time_and_antenna():
    pos = pa.itrf(np.radians([-30.7, -30.72, -30.69]), np.radians([21.44, 21.45, 21.41]), 1050.0)
    times = 5.0e9 + np.arange(4) * 600.0
    fc = np.array([-0.5, -0.6])            # negative right ascension wraps
    q = pa.parallactic_angles(fc, pos, times)
    assert q.shape == (4, 3)
    lat, lon = pa.geodetic(pos)
    H = pa.gmst(times)[:, None] + lon[None, :] - (fc[0] + 2 * np.pi)
    np.testing.assert_allclose(q, pa.parallactic_angle(H, fc[1], lat[None, :]), atol=1e-15)
