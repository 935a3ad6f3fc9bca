"""Automatic jackknife field division on the sky (corruscant/fields.py).

Same API as the reference: ``template(N_fields, theta=, phi=)`` or
``template(N_fields, ra=, dec=)``, then ``create_fields(Nside, min_ratio)``
returns a field id per point, ``assign_fields(...)`` labels other points
(randoms) with the same division, ``assign_colors()`` colours the field
adjacency graph.

The reference bins the points into HEALPix pixels (healpy) and partitions the
pixel adjacency graph with METIS (pymetis), doubling Nside until every field
is populated and the smallest field holds at least ``min_ratio`` times the
average (fields.py:62-148).  Neither library is available here: the pixel of
every point is computed on the GPU (corruscant_amd.healpix), and the
partition is a deterministic recursive bisection of the pixels, weighted by
their point counts, along the principal axis of each part's pixel centres
(contiguous, count-balanced fields; METIS's edge-cut optimum is not
reproduced, so field ids differ from the reference's).

Difference on purpose: the reference's default min_ratio = 1.3 can never be
met (the smallest field cannot exceed the average), so create_fields recursed
until it failed; here min_ratio must be in (0, 1] and defaults to 0.8.
"""
from __future__ import division, print_function

from ctypes import POINTER, c_double

import numpy as np

import corruscant_amd
from corruscant_amd import healpix

_MAX_NSIDE = 1 << 13


def _corners(Nside, ids):
    """(n, 4, 3) unit vectors of the corners of RING pixels."""
    ix, iy, f = healpix.ring2xyf(Nside, ids)
    corners = []
    for dx, dy in ((0, 0), (1, 0), (1, 1), (0, 1)):
        th, ph = healpix.xyf2ang((ix + dx) / Nside, (iy + dy) / Nside, f)
        corners.append(_unit(th, ph))
    return np.stack(corners, axis=1)


class pixel:
    """One HEALPix pixel (corruscant/fields.py:27-33): id, point count, the
    four edge neighbours and the corners (unit vectors).  ``neighbors`` and
    ``vertecies`` are computed when not given."""

    def __init__(self, Nside, id, count, neighbors=None, vertecies=None):
        self.id = id
        self.count = count
        self.Nside = Nside
        if neighbors is None:
            neighbors = healpix.edge_neighbours(Nside)[id]
        if vertecies is None:
            vertecies = _corners(Nside, np.array([id]))[0]
        self.neighbors = set(int(n) for n in neighbors)
        self.vertecies = set(map(tuple, vertecies))


def _sph(ra, dec):
    """radec2sph64 on the GPU (the reference calls it through ctypes,
    fields.py:49-55)."""
    ra = np.require(ra, requirements="AC", dtype="float64")
    dec = np.require(dec, requirements="AC", dtype="float64")
    theta = np.empty(dec.size)
    phi = np.empty(ra.size)
    lib = corruscant_amd.coordlib
    lib.radec2sph64.restype = None
    lib.radec2sph64(ra.ctypes.data_as(POINTER(c_double)), dec.ctypes.data_as(POINTER(c_double)),
                    theta.ctypes.data_as(POINTER(c_double)), phi.ctypes.data_as(POINTER(c_double)), ra.size)
    return theta, phi


def _unit(theta, phi):
    st = np.sin(theta)
    return np.stack([st * np.cos(phi), st * np.sin(phi), np.cos(theta)], axis=1)


def bisect_fields(counts, centres, n_fields):
    """Field id per pixel: recursive bisection of the pixel set along the
    principal axis of its count-weighted centres, at the pixel where the
    cumulative count reaches the share of the left part (floor(k/2) of k
    fields), keeping at least one populated pixel per field on each side."""
    counts = np.asarray(counts, dtype=np.float64)
    out = np.full(counts.size, -1, dtype=np.int64)

    def axis_of(ids):
        w = counts[ids]
        if w.sum() <= 0:
            w = np.ones_like(w)
        c = centres[ids]
        m = (c * w[:, None]).sum(0) / w.sum()
        d = c - m
        cov = (d * w[:, None]).T @ d
        vals, vecs = np.linalg.eigh(cov)
        a = vecs[:, -1]
        return a if a[np.argmax(np.abs(a))] > 0 else -a

    stack = [(np.arange(counts.size), n_fields, 0)]
    while stack:
        ids, k, first = stack.pop()
        if k == 1 or ids.size == 1:
            out[ids] = first
            continue
        kl = k // 2
        proj = centres[ids] @ axis_of(ids)
        order = ids[np.lexsort((ids, proj))]
        w = counts[order]
        cum = np.cumsum(w)
        pop = np.cumsum(w > 0)
        target = cum[-1] * kl / k
        # cut = number of pixels on the left, 1..n-1
        cuts = np.arange(1, order.size)
        ok = (pop[cuts - 1] >= kl) & (pop[-1] - pop[cuts - 1] >= k - kl)
        if not ok.any():
            ok = np.ones_like(cuts, dtype=bool)
        err = np.abs(cum[cuts - 1] - target)
        err[~ok] = np.inf
        cut = int(cuts[np.argmin(err)])
        stack.append((order[cut:], k - kl, first + kl))
        stack.append((order[:cut], kl, first))
    return out


class template:
    """corruscant/fields.py:35-149."""

    def __init__(self, N_fields, theta=None, phi=None, ra=None, dec=None):
        if theta is not None and phi is not None:
            self.theta = np.array(theta, dtype=np.float64)
            self.phi = np.array(phi, dtype=np.float64)
        else:
            self.theta, self.phi = _sph(ra, dec)
        if not (self.phi.shape == self.theta.shape):
            raise ValueError("Theta and phi arrays must be same size")
        self.N_fields = N_fields
        self.Nside = None
        self.color_dict = None
        self.filled_pix = None
        self.fld_ids = None
        self.adj_dict = None
        self._nb = None
        self._counts = None

    @property
    def pix_list(self):
        """The pixels of the last division as ``pixel`` objects (built on
        demand: one per HEALPix pixel)."""
        N = self.Nside
        ids = np.arange(12 * N * N)
        corners = _corners(N, ids)
        return [pixel(N, int(i), int(self._counts[i]), self._nb[i], corners[i]) for i in ids]

    def create_fields(self, Nside=1, min_ratio=0.8):
        """Field id (0..N_fields-1) of every point (fields.py:62-148)."""
        if not (0 < min_ratio <= 1):
            raise ValueError("min_ratio must be in (0, 1]: the smallest field cannot hold more than the average")
        while True:
            if Nside > _MAX_NSIDE:
                raise RuntimeError("no balanced division into {:d} fields up to Nside = {:d}".format(
                    self.N_fields, _MAX_NSIDE))
            pixels = healpix.ang2pix(Nside, self.theta, self.phi)
            self.Nside = Nside
            counts = np.bincount(pixels, minlength=12 * Nside * Nside)
            self.filled_pix = [int(i) for i in np.flatnonzero(counts)]
            if len(self.filled_pix) < self.N_fields:
                print("Nside = {:d} too small for {:d} fields ({:d} pix with data). "
                      "Trying Nside = {:d}...".format(Nside, self.N_fields, len(self.filled_pix), 2 * Nside))
                Nside *= 2
                continue
            nb = healpix.edge_neighbours(Nside)
            th, ph = healpix.pix2ang(Nside, np.arange(counts.size))
            fld_ids = bisect_fields(counts, _unit(th, ph), self.N_fields)
            pix_ids = fld_ids[pixels]
            hist = np.bincount(pix_ids, minlength=self.N_fields)
            if hist.min() == 0:
                print("At least one field has no population. Trying Nside = {:d}...".format(2 * Nside))
                Nside *= 2
                continue
            if float(hist.min()) / np.average(hist) < min_ratio:
                print("The smallest field has less than {:.3f} ".format(min_ratio) +
                      "times the average population. Trying Nside = {:d}...".format(2 * Nside))
                Nside *= 2
                continue
            self.fld_ids = fld_ids
            self._nb = nb
            self._counts = counts
            self.adj_dict = {int(p): [int(n) for n in nb[p]] for p in range(counts.size)}
            return pix_ids

    def assign_fields(self, theta=None, phi=None, ra=None, dec=None):
        """Field ids of other points with the division made by
        create_fields (fields.py:151-172)."""
        if theta is not None and phi is not None:
            theta = np.array(theta, dtype=np.float64)
            phi = np.array(phi, dtype=np.float64)
        else:
            theta, phi = _sph(ra, dec)
        pix = healpix.ang2pix(self.Nside, theta, phi)
        return self.fld_ids[pix]

    def assign_colors(self):
        """Colour classes of the field adjacency graph with the Welsh-Powell
        greedy method (fields.py:174-205): fields in order of decreasing
        degree, each class takes every field not adjacent to the class."""
        graph = {}
        for p, nbs in self.adj_dict.items():
            f = int(self.fld_ids[p])
            graph.setdefault(f, set()).update(int(self.fld_ids[n]) for n in nbs)
        for f in graph:
            graph[f].discard(f)
        order = sorted(graph, key=lambda f: len(graph[f]))
        color_dict = {}
        color = 0
        while len(color_dict) < len(graph):
            members = []
            for f in reversed(order):
                if not any(n in members for n in graph[f]):
                    color_dict[f] = color
                    members.append(f)
            order = [f for f in order if f not in color_dict]
            color += 1
        self.color_dict = color_dict
        return color_dict
