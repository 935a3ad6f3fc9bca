"""GPU RA/DEC conversions (src/coordmath.c:38-80) and the angular path.

The device sin/cos are correctly rounded (checked here against 60-digit
Decimal evaluations); the reference calls the host libm, and glibc 2.35 is off
by one ulp for ~0.3% of arguments, so the comparison with the reference's
fixture allows one-ulp differences at a bounded rate."""
import os
from decimal import Decimal, getcontext

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

getcontext().prec = 60
_PI = Decimal("3.14159265358979323846264338327950288419716939937510582097494459230781640628620899862803")


def _dec_sincos(x):
    x = Decimal(x)
    k = (x / (_PI / 2)).to_integral_value()
    r = x - k * (_PI / 2)
    eps = Decimal(10) ** -58
    s, c, t, n = Decimal(0), Decimal(0), r, 1
    while t != 0 and (s == 0 or abs(t) > abs(s) * eps):
        s += t
        n += 2
        t = -t * r * r / ((n - 1) * n)
    t, n = Decimal(1), 0
    while t != 0 and (c == 0 or abs(t) > abs(c) * eps):
        c += t
        n += 2
        t = -t * r * r / ((n - 1) * n)
    q = int(k) % 4
    return [(s, c), (c, -s), (-s, -c), (-c, s)][q]


def _ulps(a, b):
    a = np.asarray(a).view(np.int64)
    b = np.asarray(b).view(np.int64)
    return np.abs(a - b)


def test_radec2cart_correctly_rounded():
    from corruscant_amd.twopoint import angular
    rng = np.random.default_rng(4)
    ra = np.concatenate([rng.uniform(0, 360, 1500), [0, 90, 180, 270, 360, 1e-300, -0.0]])
    dec = np.concatenate([np.degrees(np.arcsin(rng.uniform(-1, 1, 1500))), [0, 90, -90, 45, -45, -0.0, 1e-300]])
    xyz = angular.cartesian(ra, dec)
    assert angular.coordlib.coords_last_uncertain() == 0
    p180 = np.pi / 180.0
    for i in range(ra.size):
        sl, cl = _dec_sincos(dec[i] * p180)
        so, co = _dec_sincos(ra[i] * p180)
        cl_d, co_d, so_d, sl_d = float(cl), float(co), float(so), float(sl)
        assert xyz[i, 2] == sl_d, i
        assert xyz[i, 0] == cl_d * co_d, i
        assert xyz[i, 1] == cl_d * so_d, i
    assert np.signbit(xyz[-2, 2])  # dec = -0.0 keeps its sign


def test_radec2cart_against_reference_fixture():
    from corruscant_amd.twopoint import angular
    z = np.load(os.path.join(GOLD, "ref_coords.npz"), allow_pickle=False)
    xyz = angular.cartesian(z["ra"], z["dec"])
    u = _ulps(xyz, z["xyz"])
    # z = sin(dec) is one libm value; x and y are products of two, each of
    # which glibc may round one ulp away from the correctly rounded value
    assert u[:, 2].max() <= 1 and u.max() <= 2
    assert (u > 0).mean() < 0.01  # glibc's own rounding errors (about 0.3% of sin/cos values)


def test_radecdist2cart_and_radec2sph(oracle):
    from corruscant_amd.twopoint import threedim
    import corruscant_amd
    from ctypes import POINTER, c_double, c_int
    z = np.load(os.path.join(GOLD, "ref_coords.npz"), allow_pickle=False)
    xyz = threedim.cartesian(z["ra"], z["dec"], distance=z["dist"])
    assert _ulps(xyz, z["xyzd"]).max() <= 3
    lib = corruscant_amd.coordlib
    lib.radec2sph64.restype = None
    lib.radec2sph64.argtypes = [POINTER(c_double)] * 4 + [c_int]
    th, ph = np.empty_like(z["ra"]), np.empty_like(z["ra"])
    dp = lambda a: a.ctypes.data_as(POINTER(c_double))
    ra, dec = np.ascontiguousarray(z["ra"]), np.ascontiguousarray(z["dec"])
    lib.radec2sph64(dp(ra), dp(dec), dp(th), dp(ph), ra.size)
    assert np.array_equal(th, z["theta"]) and np.array_equal(ph, z["phi"])


def test_angular_autocorr_counts_match_oracle(oracle):
    """A sky patch (RA 0-60 deg, DEC -30..30 deg with cos DEC density), xyz
    from the GPU conversion, chord radii as angular.py:40 computes them."""
    from corruscant_amd import clustering
    from corruscant_amd.twopoint import angular
    rng = np.random.default_rng(3)

    def patch(n):
        return rng.uniform(0, 60, n), np.degrees(np.arcsin(rng.uniform(-0.5, 0.5, n)))

    dra, ddec = patch(8000)
    rra, rdec = patch(40000)
    D, R = angular.cartesian(dra, ddec), angular.cartesian(rra, rdec)
    df = (dra // 15).astype(np.int32) * 2 + (ddec > 0)
    rf = (rra // 15).astype(np.int32) * 2 + (rdec > 0)
    deg = np.logspace(-2, 0, 10)
    res = angular.autocorr(clustering.tree(D, df), clustering.tree(R, rf), deg)
    chord = 2.0 * np.sin(deg * (np.pi / 180.0) / 2.0)
    od, orr = oracle.OracleTree(D, df, 8), oracle.OracleTree(R, rf, 8)
    assert np.array_equal(res.dd, np.diff(od.cumulative(D, df, 8, chord), axis=0))
    assert np.array_equal(res.dr, np.diff(orr.cumulative(D, df, 8, chord), axis=0))
    assert np.array_equal(res.rr, np.diff(orr.cumulative(R, rf, 8, chord), axis=0))
    assert res.radii_units == "degrees"
    assert np.all(np.isfinite(res.jackknife_error()))
