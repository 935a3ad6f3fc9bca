"""3-D correlation functions (corruscant/twopoint/threedim.py)."""
from __future__ import division, print_function

from ctypes import POINTER, c_double, c_int

import numpy as np

import corruscant_amd
from corruscant_amd import clustering

coordlib = corruscant_amd.coordlib
coordlib.radecdist2cart64.restype = None
coordlib.radecdist2cart64.argtypes = [POINTER(c_double)] * 4 + [c_int]
coordlib.coords_last_error.restype = __import__("ctypes").c_char_p


def autocorr(data_tree, rand_tree, radii, **kwargs):
    """threedim.py:39-46: radii are Euclidean; units "Data units"."""
    results = clustering._autocorr(data_tree, rand_tree, radii, **kwargs)
    results.radii_nominal = results.radii_euclidean
    results.radii_units = "Data units"
    return results


def crosscorr(data_tree_1, data_tree_2, rand_tree_1, rand_tree_2, radii, **kwargs):
    """threedim.py:48-56."""
    results = clustering._crosscorr(data_tree_1, data_tree_2, rand_tree_1, rand_tree_2, radii, **kwargs)
    results.radii_nominal = results.radii_euclidean
    results.radii_units = "Data units"
    return results


def cartesian(ra, dec, distance=None, z=None, cosmology=None):
    """RA/DEC/distance (or redshift with a cosmology object providing
    ``comoving_distance``) to (N, 3) Cartesian points (threedim.py:58-113)."""
    if distance is None:
        z = np.array(z)
        distance = cosmology.comoving_distance(z)
    ra = np.require(ra, requirements="AC", dtype="float64")
    dec = np.require(dec, requirements="AC", dtype="float64")
    distance = np.require(distance, requirements="AC", dtype="float64")
    N = ra.size
    cart = np.empty((N, 3), dtype="float64")
    coordlib.radecdist2cart64(ra.ctypes.data_as(POINTER(c_double)), dec.ctypes.data_as(POINTER(c_double)),
                              distance.ctypes.data_as(POINTER(c_double)), cart.ctypes.data_as(POINTER(c_double)),
                              c_int(N))
    err = coordlib.coords_last_error()
    if err:
        raise RuntimeError("radecdist2cart64 failed: " + err.decode())
    return cart
