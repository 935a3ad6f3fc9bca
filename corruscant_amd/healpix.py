"""HEALPix RING-scheme geometry for the automatic field division.

corruscant/fields.py:30-31, :67, :153 take pixel indices, neighbours and
boundaries from healpy, which is not available here.  The per-point work
(ang2pix of every catalogue point) runs on the GPU (`ang2pix_ring64` in
libcoords); the per-pixel geometry below (a few thousand pixels) runs on the
host.  Formulas follow the HEALPix paper (Gorski et al. 2005, ApJ 622, 759):
the twelve base faces with their ring/phi offsets, a pixel's face coordinates
(ix, iy), and the map from face coordinates to (z, phi).
"""
from __future__ import division

from ctypes import POINTER, c_char_p, c_double, c_int, c_longlong

import numpy as np

import corruscant_amd

coordlib = corruscant_amd.coordlib
coordlib.ang2pix_ring64.restype = None
coordlib.ang2pix_ring64.argtypes = [POINTER(c_double), POINTER(c_double), POINTER(c_longlong), c_int, c_longlong]
coordlib.coords_last_error.restype = c_char_p
coordlib.coords_last_error.argtypes = []

# ring (in units of nside) and phi (in units of pi/4) of each base face's
# southern corner
_JRLL = np.array([2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4], dtype=np.int64)
_JPLL = np.array([1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7], dtype=np.int64)


def ang2pix(nside, theta, phi):
    """RING pixel of each (theta, phi) on the GPU (healpy.ang2pix(nside,
    theta, phi) with nest=False)."""
    theta = np.ascontiguousarray(theta, dtype=np.float64).ravel()
    phi = np.ascontiguousarray(phi, dtype=np.float64).ravel()
    if theta.shape != phi.shape:
        raise ValueError("theta and phi must have the same size")
    out = np.empty(theta.size, dtype=np.int64)
    if theta.size:
        coordlib.ang2pix_ring64(theta.ctypes.data_as(POINTER(c_double)), phi.ctypes.data_as(POINTER(c_double)),
                                out.ctypes.data_as(POINTER(c_longlong)), theta.size, int(nside))
        err = coordlib.coords_last_error()
        if err:
            raise ValueError(err.decode("utf-8", "replace"))
    return out


def _isqrt(v):
    r = np.floor(np.sqrt(v.astype(np.float64))).astype(np.int64)
    r = np.where((r + 1) * (r + 1) <= v, r + 1, r)
    return np.where(r * r > v, r - 1, r)


def ring2xyf(nside, pix):
    """(ix, iy, face) of RING pixels: ring index and position in the ring,
    then the offsets of the pixel's base face."""
    pix = np.asarray(pix, dtype=np.int64)
    N = int(nside)
    ncap = 2 * N * (N - 1)
    npix = 12 * N * N
    nl2 = 2 * N
    iring = np.empty_like(pix)
    iphi = np.empty_like(pix)
    kshift = np.zeros_like(pix)
    nr = np.empty_like(pix)
    face = np.empty_like(pix)

    north = pix < ncap
    south = pix >= npix - ncap
    equ = ~north & ~south
    if north.any():
        p = pix[north]
        ir = (1 + _isqrt(1 + 2 * p)) >> 1
        ip = (p + 1) - 2 * ir * (ir - 1)
        iring[north], iphi[north], nr[north] = ir, ip, ir
        face[north] = (ip - 1) // ir
    if equ.any():
        p = pix[equ] - ncap
        tmp = p // (4 * N)
        ir = tmp + N
        ip = p - tmp * 4 * N + 1
        iring[equ], iphi[equ], nr[equ] = ir, ip, N
        kshift[equ] = (ir + N) & 1
        ire = tmp + 1
        irm = nl2 + 2 - ire
        ifm = (ip - ire // 2 + N - 1) // N
        ifp = (ip - irm // 2 + N - 1) // N
        face[equ] = np.where(ifp == ifm, ifp | 4, np.where(ifp < ifm, ifp, ifm + 8))
    if south.any():
        p = npix - pix[south]
        ir = (1 + _isqrt(2 * p - 1)) >> 1
        ip = 4 * ir + 1 - (p - 2 * ir * (ir - 1))
        iphi[south], nr[south] = ip, ir
        iring[south] = 2 * nl2 - ir
        face[south] = 8 + (ip - 1) // ir
    irt = iring - _JRLL[face] * N + 1
    ipt = 2 * iphi - _JPLL[face] * nr - kshift - 1
    ipt = np.where(ipt >= nl2, ipt - 8 * N, ipt)
    ix = (ipt - irt) >> 1
    iy = (-ipt - irt) >> 1
    return ix, iy, face


def xyf2ang(x, y, face):
    """(theta, phi) of face coordinates (x, y) in units of the face side
    (0..1 inside the face; slightly outside continues onto the neighbouring
    face)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    face = np.asarray(face, dtype=np.int64)
    jr = _JRLL[face] - x - y
    nr = np.where(jr < 1, jr, np.where(jr > 3, 4 - jr, 1.0))
    t = nr * nr / 3.0
    z = np.where(jr < 1, 1 - t, np.where(jr > 3, t - 1, (2 - jr) * (2.0 / 3.0)))
    # phi in units of pi/4 is jpll + (x - y) / nr; wrapped after the division
    # (one turn is 8 nr in the polar caps), so that points continued past a
    # face edge land on the neighbouring face
    tmp = _JPLL[face] * nr + x - y
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = np.where(nr < 1e-15, 0.0, np.mod((np.pi / 4) * tmp / nr, 2 * np.pi))
    # theta from z, or from sin(theta) where 1 - |z| loses digits
    polar = np.abs(z) > 0.99
    sth = np.sqrt(np.maximum(t * (2.0 - t), 0.0))
    theta = np.where(polar, np.where(z > 0, np.arcsin(np.minimum(sth, 1.0)), np.pi - np.arcsin(np.minimum(sth, 1.0))),
                     np.arccos(np.clip(z, -1.0, 1.0)))
    return theta, phi


def pix2ang(nside, pix):
    """Centre (theta, phi) of RING pixels."""
    ix, iy, face = ring2xyf(nside, pix)
    return xyf2ang((ix + 0.5) / nside, (iy + 0.5) / nside, face)


def edge_neighbours(nside):
    """(npix, 4) RING indices of the pixels sharing an edge with each pixel
    (healpy's get_all_neighbours(...)[::2], corruscant/fields.py:30): probe
    points just outside the midpoint of each of the four edges, located with
    the face map continued across the face boundary, then ang2pix."""
    N = int(nside)
    pix = np.arange(12 * N * N, dtype=np.int64)
    ix, iy, face = ring2xyf(N, pix)
    d = 0.55  # half a pixel plus a margin, in pixel units
    cx, cy = ix + 0.5, iy + 0.5
    px = np.stack([cx + d, cx - d, cx, cx], axis=1) / N
    py = np.stack([cy, cy, cy + d, cy - d], axis=1) / N
    th, ph = xyf2ang(px, py, np.repeat(face[:, None], 4, axis=1))
    nb = ang2pix(N, th.ravel(), ph.ravel()).reshape(-1, 4)
    # every pixel has four distinct edge neighbours and the relation is symmetric
    if np.any(nb == pix[:, None]) or np.any(np.sort(nb, axis=1)[:, 1:] == np.sort(nb, axis=1)[:, :-1]):
        raise RuntimeError("HEALPix edge neighbours: probe landed in the pixel itself or twice in one neighbour")
    back = (nb[nb.ravel()].reshape(-1, 4, 4) == pix[:, None, None]).any(axis=2)
    if not back.all():
        raise RuntimeError("HEALPix edge neighbours: relation not symmetric")
    return nb
