"""GPU HEALPix pixelisation and the automatic field division
(corruscant/fields.py): ang2pix against the paper's ring enumeration, equal
pixel areas, edge neighbours, and balanced, contiguous fields.  Field ids are
not compared with the reference's (they come from METIS there)."""
import numpy as np
import pytest

from oracle import healpix_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nside", [1, 2, 3, 4, 5, 16, 64, 256])
def test_ang2pix_of_every_pixel_centre(nside):
    from corruscant_amd import healpix
    z, phi = healpix_oracle.ring_centres(nside)
    pix = healpix.ang2pix(nside, np.arccos(z), phi)
    assert np.array_equal(pix, np.arange(z.size))


def test_ang2pix_near_the_poles_and_phi_wrap():
    from corruscant_amd import healpix
    th = np.array([0.0, 1e-12, np.pi, np.pi - 1e-12, np.pi / 2, np.pi / 2])
    ph = np.array([0.3, 5.0, 0.3, 5.0, -1e-9, 2 * np.pi + 1e-9])
    pix = healpix.ang2pix(8, th, ph)
    npix = 12 * 64
    assert pix[0] < 4 and pix[1] < 4 and pix[2] >= npix - 4 and pix[3] >= npix - 4
    # phi just below 0 is the last pixel of its ring, just above 2 pi the first
    assert healpix.ang2pix(8, [np.pi / 2], [2 * np.pi - 1e-9])[0] == pix[4]
    assert healpix.ang2pix(8, [np.pi / 2], [1e-9])[0] == pix[5]
    with pytest.raises(ValueError):
        healpix.ang2pix(8, [4.0], [0.0])


def test_equal_area_pixels():
    """Uniform points on the sphere: counts per pixel are Poisson with one mean."""
    from corruscant_amd import healpix
    rng = np.random.default_rng(11)
    n, nside = 3_000_000, 8
    th = np.arccos(rng.uniform(-1, 1, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    c = np.bincount(healpix.ang2pix(nside, th, ph), minlength=12 * nside * nside)
    mean = n / c.size
    chi2 = ((c - mean) ** 2 / mean).sum()
    dof = c.size - 1
    assert abs(chi2 - dof) < 6 * np.sqrt(2 * dof)


def test_base_pixel_adjacency():
    """Edges of the twelve base pixels: each polar face touches its two polar
    neighbours and the two equatorial faces below it."""
    from corruscant_amd import healpix
    nb = [sorted(int(v) for v in r) for r in healpix.edge_neighbours(1)]
    assert nb == [[1, 3, 4, 5], [0, 2, 5, 6], [1, 3, 6, 7], [0, 2, 4, 7], [0, 3, 8, 11], [0, 1, 8, 9],
                  [1, 2, 9, 10], [2, 3, 10, 11], [4, 5, 9, 11], [5, 6, 8, 10], [6, 7, 9, 11], [4, 7, 8, 10]]


def test_edge_neighbours_are_adjacent_pixels():
    from corruscant_amd import healpix
    for nside in (1, 2, 4, 16):
        nb = healpix.edge_neighbours(nside)  # symmetric, four distinct, never the pixel itself
        th, ph = healpix.pix2ang(nside, np.arange(nb.shape[0]))
        v = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
        ang = np.arccos(np.clip((v[:, None, :] * v[nb]).sum(2), -1, 1))
        size = np.sqrt(4 * np.pi / nb.shape[0])
        assert ang.max() < 1.6 * size and ang.min() > 0.3 * size


def _patch(rng, n):
    ra = rng.uniform(0, 60, n)
    dec = np.degrees(np.arcsin(rng.uniform(-0.5, 0.5, n)))
    return ra, dec


def test_create_fields_balanced_contiguous_and_assignable():
    from corruscant_amd import fields
    rng = np.random.default_rng(12)
    ra, dec = _patch(rng, 200_000)
    t = fields.template(16, ra=ra, dec=dec)
    f = t.create_fields(min_ratio=0.8)
    assert f.shape == ra.shape and f.min() == 0 and f.max() == 15
    hist = np.bincount(f, minlength=16)
    assert hist.min() / hist.mean() >= 0.8
    # the same division labels the same points identically, and randoms too
    assert np.array_equal(t.assign_fields(ra=ra, dec=dec), f)
    rra, rdec = _patch(rng, 400_000)
    fr = t.assign_fields(ra=rra, dec=rdec)
    hr = np.bincount(fr, minlength=16)
    assert hr.min() / hr.mean() > 0.6
    # every field is one connected set of populated pixels
    populated = set(t.filled_pix)
    for k in range(16):
        pix = [p for p in populated if t.fld_ids[p] == k]
        seen, todo = {pix[0]}, [pix[0]]
        while todo:
            p = todo.pop()
            for q in t.adj_dict[p]:
                if q in populated and t.fld_ids[q] == k and q not in seen:
                    seen.add(q)
                    todo.append(q)
        assert len(seen) == len(pix), k
    # a proper colouring of the field adjacency graph
    col = t.assign_colors()
    for p, nbs in t.adj_dict.items():
        for q in nbs:
            if t.fld_ids[p] != t.fld_ids[q]:
                assert col[int(t.fld_ids[p])] != col[int(t.fld_ids[q])]
    assert len(t.pix_list) == 12 * t.Nside ** 2


def test_create_fields_from_theta_phi_matches_radec():
    from corruscant_amd import fields
    rng = np.random.default_rng(13)
    ra, dec = _patch(rng, 50_000)
    a = fields.template(8, ra=ra, dec=dec).create_fields()
    th, ph = np.radians(90.0 - dec), np.radians(ra)
    b = fields.template(8, theta=th, phi=ph).create_fields()
    assert np.array_equal(a, b)


def test_full_sky_demo_case():
    """The reference's demo (corruscant/fields.py:210-240): 50,000 points
    uniform on the sphere into 20 fields with min_ratio 0.8."""
    from corruscant_amd import fields
    rng = np.random.default_rng(0)
    th = np.arccos(rng.uniform(-1, 1, 50_000))
    ph = rng.uniform(0, 2 * np.pi, 50_000)
    t = fields.template(20, th, ph)
    f = t.create_fields(min_ratio=0.8)
    hist = np.bincount(f, minlength=20)
    assert f.min() == 0 and f.max() == 19 and hist.min() / hist.mean() >= 0.8
    col = t.assign_colors()
    assert sorted(col) == list(range(20))
    for a, nbs in t.adj_dict.items():
        for b in nbs:
            fa, fb = int(t.fld_ids[a]), int(t.fld_ids[b])
            assert fa == fb or col[fa] != col[fb]
    # a pixel object built on its own agrees with the division's list
    p = t.pix_list[t.filled_pix[0]]
    q = fields.pixel(t.Nside, p.id, p.count)
    assert q.neighbors == p.neighbors and len(q.vertecies) == 4
    assert q.vertecies == p.vertecies
