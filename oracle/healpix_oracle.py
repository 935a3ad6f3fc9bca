"""HEALPix RING pixel centres by direct enumeration -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ as the checker of corruscant_amd.healpix and of the GPU
ang2pix; the product never imports it.  healpy (which corruscant/fields.py:20-22
uses) is not installed, so the pixelisation is pinned against the closed forms
of the HEALPix paper (Gorski et al. 2005, ApJ 622, 759, eqs. 4-9), a
formulation independent of the product's (index arithmetic on edge lines and
base faces): rings i = 1 .. 4 Nside - 1 from north to south, pixels j = 1 .. n_i
in each ring in order of increasing phi,

  north cap   i <  Nside:       n_i = 4 i,      z = 1 - i^2 / (3 Nside^2),  phi = pi/(2 i) (j - 1/2)
  equator     Nside <= i <= 3 Nside: n_i = 4 Nside, z = 4/3 - 2 i / (3 Nside),
                                phi = pi/(2 Nside) (j - 1)      if i + Nside is odd,
                                phi = pi/(2 Nside) (j - 1/2)    if i + Nside is even
  south cap   i >  3 Nside:     mirror of the north cap with i' = 4 Nside - i.

(The equatorial rings alternate between a pixel centred on phi = 0 and a
half-pixel offset; for Nside = 1 the equatorial base pixels 4..7 sit at
phi = 0, pi/2, pi, 3 pi/2.)

The RING index of a pixel is its position in this enumeration.  The
enumeration is pinned against the worked example of healpy's pix2ang
documentation (five pixels at Nside 16 covering both caps and both equatorial
phi offsets; healpy is unversioned in the reference's setup.py:50) and the
known base pixels; equal pixel areas are checked statistically in the tests.
"""
from __future__ import division

import numpy as np


def ring_centres(nside):
    """(z, phi) of every pixel centre in RING order."""
    N = int(nside)
    zs, phis = [], []
    for i in range(1, 4 * N):
        if i < N:
            n, z, step, off = 4 * i, 1.0 - i * i / (3.0 * N * N), np.pi / (2 * i), 0.5
        elif i <= 3 * N:
            off = 1.0 if (i + N) % 2 else 0.5
            n, z, step = 4 * N, 4.0 / 3.0 - 2.0 * i / (3.0 * N), np.pi / (2 * N)
        else:
            k = 4 * N - i
            n, z, step, off = 4 * k, -(1.0 - k * k / (3.0 * N * N)), np.pi / (2 * k), 0.5
        j = np.arange(1, n + 1)
        zs.append(np.full(n, z))
        phis.append(step * (j - off))
    return np.concatenate(zs), np.concatenate(phis)
