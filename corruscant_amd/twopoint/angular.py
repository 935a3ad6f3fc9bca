"""Angular correlation functions (corruscant/twopoint/angular.py)."""
from __future__ import division, print_function

import ctypes
from ctypes import POINTER, c_double, c_int

import numpy as np

import corruscant_amd
from corruscant_amd import clustering

coordlib = corruscant_amd.coordlib
coordlib.radec2cart64.restype = None
coordlib.radec2cart64.argtypes = [POINTER(c_double), POINTER(c_double), POINTER(c_double), c_int]
coordlib.coords_last_error.restype = ctypes.c_char_p
coordlib.coords_last_uncertain.restype = ctypes.c_ulonglong


def chord(radii_deg):
    """Angular separation in degrees -> unit-sphere chord length, in FP64
    numpy exactly as angular.py:40 and :54 compute it."""
    radii_deg = np.array(radii_deg)
    return 2. * np.sin(radii_deg * (np.pi / 180.) / 2.)


def autocorr(data_tree, rand_tree, radii_deg, **kwargs):
    """angular.py:37-50."""
    radii_deg = np.array(radii_deg)
    results = clustering._autocorr(data_tree, rand_tree, chord(radii_deg), **kwargs)
    results.radii_nominal = radii_deg
    results.radii_units = "degrees"
    return results


def crosscorr(data_tree_1, data_tree_2, rand_tree_1, rand_tree_2, radii_deg, **kwargs):
    """angular.py:52-62."""
    radii_deg = np.array(radii_deg)
    results = clustering._crosscorr(data_tree_1, data_tree_2, rand_tree_1, rand_tree_2, chord(radii_deg),
                                    **kwargs)
    results.radii_nominal = radii_deg
    results.radii_units = "degrees"
    return results


def cartesian(ra, dec):
    """RA/DEC in degrees -> (N, 3) points on the unit sphere, converted on the
    GPU (angular.py:64-107 -> coordmath.c:38-51)."""
    ra = np.asarray(ra)
    N = ra.size
    ra = np.require(ra, requirements="AC", dtype="float64")
    dec = np.require(dec, requirements="AC", dtype="float64")
    cart = np.empty((N, 3), dtype="float64")
    coordlib.radec2cart64(ra.ctypes.data_as(POINTER(c_double)), dec.ctypes.data_as(POINTER(c_double)),
                          cart.ctypes.data_as(POINTER(c_double)), c_int(N))
    err = coordlib.coords_last_error()
    if err:
        raise RuntimeError("radec2cart64 failed: " + err.decode())
    return cart
