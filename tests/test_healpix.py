"""HEALPix geometry of the field division (host side) against the paper's
closed-form ring enumeration, and the count-balanced bisection."""
import numpy as np
import pytest

from oracle import healpix_oracle


def test_oracle_against_published_healpy_values():
    """The oracle's enumeration against the example in healpy's pix2ang
    documentation (healpy is unversioned in the reference's setup.py:50):
    pix2ang(16, [1440, 427, 1520, 0, 3068]) gives theta
    [1.52911759, 0.78550497, 1.57079633, 0.05103658, 3.09055608] and phi
    [0, 0.78539816, 1.61988371, 0.78539816, 0.78539816].  These cover a
    north-cap ring, both equatorial phi offsets and the two polar rings."""
    z, phi = healpix_oracle.ring_centres(16)
    ids = [1440, 427, 1520, 0, 3068]
    assert np.allclose(np.arccos(z[ids]), [1.52911759, 0.78550497, 1.57079633, 0.05103658, 3.09055608],
                       rtol=0, atol=5e-9)
    assert np.allclose(phi[ids], [0.0, 0.78539816, 1.61988371, 0.78539816, 0.78539816], rtol=0, atol=5e-9)
    # twelve base pixels at z = 2/3, 0, -2/3 and the ring sizes sum to 12 Nside^2
    z1, phi1 = healpix_oracle.ring_centres(1)
    assert np.allclose(z1, [2 / 3.] * 4 + [0.] * 4 + [-2 / 3.] * 4)
    assert np.allclose(phi1[4:8], np.arange(4) * np.pi / 2)
    for n in (2, 5, 16):
        assert healpix_oracle.ring_centres(n)[0].size == 12 * n * n


@pytest.mark.parametrize("nside", [1, 2, 3, 4, 8, 16, 32])
def test_pixel_centres_match_ring_enumeration(nside):
    from corruscant_amd import healpix
    z, phi = healpix_oracle.ring_centres(nside)
    assert z.size == 12 * nside * nside
    th, ph = healpix.pix2ang(nside, np.arange(z.size))
    assert np.allclose(np.cos(th), z, rtol=0, atol=1e-13)
    assert np.allclose(ph, phi, rtol=0, atol=1e-12)


def test_nside1_base_pixels():
    """The twelve base pixels: four at z = 2/3 (phi = pi/4 + k pi/2), four on
    the equator (phi = k pi/2), four at z = -2/3."""
    from corruscant_amd import healpix
    th, ph = healpix.pix2ang(1, np.arange(12))
    z = np.cos(th)
    assert np.allclose(z, [2 / 3.] * 4 + [0.] * 4 + [-2 / 3.] * 4, atol=1e-15)
    k = np.arange(4)
    assert np.allclose(ph, np.concatenate([np.pi / 4 + k * np.pi / 2, k * np.pi / 2, np.pi / 4 + k * np.pi / 2]))


def test_face_coordinates_cover_each_face_once():
    from corruscant_amd import healpix
    for nside in (1, 2, 4, 8):
        ix, iy, f = healpix.ring2xyf(nside, np.arange(12 * nside * nside))
        assert ix.min() == 0 and iy.min() == 0 and ix.max() == nside - 1 and iy.max() == nside - 1
        key = (f * nside + ix) * nside + iy
        assert np.unique(key).size == key.size


def test_bisection_balances_counts_and_is_deterministic():
    from corruscant_amd import fields
    rng = np.random.default_rng(7)
    v = rng.normal(size=(3000, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    counts = rng.integers(0, 50, 3000)
    for k in (2, 3, 7, 16, 20):
        f = fields.bisect_fields(counts, v, k)
        assert f.min() == 0 and f.max() == k - 1
        w = np.bincount(f, weights=counts, minlength=k)
        assert w.min() / w.mean() > 0.9, (k, w)
        assert np.array_equal(f, fields.bisect_fields(counts, v, k))


def test_min_ratio_above_one_is_an_error():
    from corruscant_amd import fields
    t = fields.template(4, theta=np.array([1.0, 2.0]), phi=np.array([0.5, 1.0]))
    with pytest.raises(ValueError, match="min_ratio"):
        t.create_fields(min_ratio=1.3)
