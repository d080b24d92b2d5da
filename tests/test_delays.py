"""The phasor module's delay engine against closed forms (CPU, no GPU).

The reference computes the delays with radiointerferometryc99 + ERFA
(src/modules/phasor/ata.cc:105-173), neither of which is available, and its
own test compares with astropy at rtol = atol = 0.5
(tests/modules/phasor/advanced.py:373).  These checks pin the restated model
with facts that do not come from its code:

* the delay definition the reference documents (ata.cc:27-46):
  tau = (W_ant - W_ref)/c towards the beam minus the same towards the
  boresight, with W the baseline's component towards the source (Thompson,
  Moran & Swenson, eq. 4.1: w = X cos(d) cos(H) - Y cos(d) sin(H) + Z sin(d));
* Greenwich mean sidereal time at J2000.0, 18h41m50.54841s, and its daily
  advance 360.98564736629 deg (IAU 1982, in its 0h-UT1 form);
* the annual general precession of J2000 coordinates, m = 4612.4362"/cy in
  right ascension and n = 2004.3109"/cy in declination, to first order;
* one sidereal day (0.99726956633 solar days) returns the sky to its place;
* UT1 - UTC (dut1) shifts the Earth's rotation only.

A flipped hour-angle sign, a solar instead of sidereal rate, a precession
applied with the wrong sign or size, or a dut1 in the wrong unit fails one of
them.
"""

import math

import numpy as np
import pytest

from blade_amd.delays import ata_delays

C = 299792458.0
J2000 = 2451545.0
ARCSEC = math.pi / (180 * 3600)
LON = math.radians(-121.470733)
LAT = math.radians(40.815987)
REF_LLA = (LON, LAT, 1020.86)
P0 = np.array([-2524041.5388905862, -4123587.965024342, 4147646.4222955606])   # ATA antenna 1c
BASE = 1000.0   # metres


def local_axes(lon):
    """ECEF directions of the local X (meridian at the equator), Y (east), Z (pole)."""
    return (np.array([math.cos(lon), math.sin(lon), 0.0]),
            np.array([-math.sin(lon), math.cos(lon), 0.0]),
            np.array([0.0, 0.0, 1.0]))


def antennas(lon=LON):
    """Reference antenna at P0, then one antenna BASE metres along X, Y, Z."""
    ex, ey, ez = local_axes(lon)
    return np.stack([P0, P0 + BASE * ex, P0 + BASE * ey, P0 + BASE * ez])


def gmst_rad(jd_ut1):
    """GMST from the IAU 1982 expression at 0h UT1 plus the UT1 fraction of the day."""
    jd0 = math.floor(jd_ut1 - 0.5) + 0.5
    tu = (jd0 - J2000) / 36525.0
    sec0 = 24110.54841 + 8640184.812866 * tu + 0.093104 * tu ** 2 - 6.2e-6 * tu ** 3
    sec = sec0 + 1.00273790935 * (jd_ut1 - jd0) * 86400.0
    return (sec / 86400.0 * 2 * math.pi) % (2 * math.pi)


def w(base_xyz, ra, dec, H):
    X, Y, Z = base_xyz
    return X * math.cos(dec) * math.cos(H) - Y * math.cos(dec) * math.sin(H) + Z * math.sin(dec)


def test_gmst_at_j2000_is_the_published_value():
    g = gmst_rad(J2000)
    assert abs(g - math.radians(280.46061837504)) < 1e-9
    g1 = gmst_rad(J2000 + 1.0)
    adv = math.degrees(g1 - g) % 360.0
    assert abs(adv - 0.98564736629) < 1e-8


def test_x_y_z_baselines_at_j2000():
    # At J2000.0 precession is the identity and GMST is 280.46061837504 deg.
    ra, dec = 1.0, 0.5
    d = ata_delays(antennas(), REF_LLA, (0.0, math.pi / 2), [(ra, dec)], J2000)
    H = math.radians(280.46061837504) + LON - ra
    # boresight at the pole: W_bore = Z only
    exp = [0.0,
           BASE * math.cos(dec) * math.cos(H) / C,
           -BASE * math.cos(dec) * math.sin(H) / C,
           BASE * (math.sin(dec) - 1.0) / C]
    np.testing.assert_allclose(d[0], exp, rtol=0, atol=1e-15)


def test_delays_do_not_depend_on_the_reference_position_beyond_longitude():
    # Baselines are differences of ECEF positions; latitude and altitude of the
    # array reference only translate them.
    beams = [(0.3, 0.2), (2.0, -0.4)]
    jd = 2459676.5 + 0.37
    a = ata_delays(antennas(), REF_LLA, (0.31, 0.25), beams, jd)
    b = ata_delays(antennas(), (LON, 0.1, -30.0), (0.31, 0.25), beams, jd)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-15)   # ECEF rounding at 4e6 m


def test_pole_of_date_is_hour_angle_independent():
    # Near J2000 the J2000 pole is the pole of date (precession moves it
    # 0.055"/day): delays of X and Y baselines vanish at every hour angle.
    for jd in (J2000, J2000 + 0.13, J2000 + 0.41):
        d = ata_delays(antennas(), REF_LLA, (0.0, math.pi / 2), [(0.7, math.pi / 2)], jd)
        np.testing.assert_allclose(d[0], 0.0, atol=3e-15)


def test_one_sidereal_day_repeats_the_delays():
    jd = (1649366473.0 / 86400) + 2440587.5       # the reference test's observation date
    beams = [(0.63722, 1.07552424), (0.2, -0.3), (4.0, 0.9)]
    bore = (0.64169, 1.079896295)
    a = ata_delays(antennas(), REF_LLA, bore, beams, jd)
    b = ata_delays(antennas(), REF_LLA, bore, beams, jd + 0.99726956633)
    # one day of precession (0.14") bounds what may change
    assert np.max(np.abs(a - b)) < BASE / C * 0.3 * ARCSEC
    # a solar day is ~0.9856 deg more rotation: the X/Y delays change a lot
    c = ata_delays(antennas(), REF_LLA, bore, beams, jd + 1.0)
    assert np.max(np.abs(a - c)) > BASE / C * 1e-3


def test_precession_matches_the_general_precession_rates():
    # Beams at (0, 0) and (pi, 0) J2000, 20 years later.  To first order the
    # mean place of date is alpha = m t, delta = +n t for the first and
    # alpha = pi + m t, delta = -n t for the second.  Differences between the
    # two beams cancel the boresight and the reference antenna.
    t_cy = 0.2
    jd = J2000 + 36525.0 * t_cy
    m = 4612.4362 * t_cy * ARCSEC
    n = 2004.3109 * t_cy * ARCSEC
    d = ata_delays(antennas(), REF_LLA, (0.0, 0.0), [(0.0, 0.0), (math.pi, 0.0)], jd)
    H0 = gmst_rad(jd) + LON - m
    H1 = gmst_rad(jd) + LON - (math.pi + m)
    exp = np.array([[w(v, 0, n, H0) for v in np.eye(3) * BASE],
                    [w(v, 0, -n, H1) for v in np.eye(3) * BASE]]) / C
    got = d[:, 1:] - d[:, :1]          # per baseline, relative to the reference antenna
    tol = BASE / C * 0.5 * ARCSEC      # second-order terms are below 0.1"
    np.testing.assert_allclose(got[0] - got[1], exp[0] - exp[1], rtol=0, atol=tol)
    # the precession itself is ~1000" here: without it the check fails
    H0n = gmst_rad(jd) + LON
    exp_none = np.array([w(v, 0, 0, H0n) - w(v, 0, 0, H0n - math.pi) for v in np.eye(3) * BASE]) / C
    assert np.max(np.abs((got[0] - got[1]) - exp_none)) > 50 * tol


def test_dut1_shifts_only_the_earth_rotation():
    jd = 2459676.5 + 0.2
    beams = [(0.63722, 1.07552424), (1.5, 0.1)]
    bore = (0.64169, 1.079896295)
    a = ata_delays(antennas(), REF_LLA, bore, beams, jd, dut1=0.4)
    b = ata_delays(antennas(), REF_LLA, bore, beams, jd + 0.4 / 86400.0, dut1=0.0)
    # the two differ only by 0.4 s of precession, far below a picosecond
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-15)
    c = ata_delays(antennas(), REF_LLA, bore, beams, jd, dut1=0.0)
    assert np.max(np.abs(a - c)) > 1e-12


def test_reference_antenna_and_boresight_rows():
    pos = antennas()
    bore = (0.64169, 1.079896295)
    d = ata_delays(pos, REF_LLA, bore, [bore, (0.63722, 1.07552424)], 2459676.8, reference_antenna=2)
    np.testing.assert_array_equal(d[0], 0.0)       # the boresight as a beam
    assert d[1, 2] == 0.0                          # the reference antenna


def test_rejects_reference_antenna_out_of_range():
    from blade_amd import _lib
    with pytest.raises(_lib.BladeError, match="reference antenna"):
        ata_delays(antennas(), REF_LLA, (0.0, 0.0), [(0.0, 0.0)], J2000, reference_antenna=4)
