"""End-to-end runs of the reference's own scripts, with assertions.

test/test_alltypes.py builds trees of 4,000 data and 80,000 random points
(np.random.seed(3), 6 radii, 8 fields) and prints every estimator with every
error type for 3-D auto-correlations without and with fields, a 3-D
cross-correlation and angular auto- and cross-correlations.  Here each of
those runs is checked: the count arrays bit-exact against the oracle (same
tree/query orientation), the estimate and the jackknife, field-to-field,
Poisson and bootstrap errors against the results oracle on those counts, and
the printed table.  test/test_duplicates.py (30,000 points, each three times)
checks the deduplicated size."""
import numpy as np
import pytest

from oracle import results_oracle as RO

pytestmark = pytest.mark.gpu

DSIZE, RSIZE, NF = 4000, 80000, 8
RADII = np.logspace(-2.5, -1., 6)
RADII_ANG = np.logspace(-1.5, 0.5, 6)
EST = {"standard": "est_standard", "hamilton": "est_hamilton", "landy-szalay": "est_landy_szalay"}


def random_cube(n, nf=None):
    xyz = np.random.rand(n, 3)
    if nf is None:
        return (xyz,)
    return xyz, (xyz[:, 0] * nf).astype("int32")


def random_sphere(n, nf=None):
    from corruscant_amd.twopoint import angular
    ra = np.random.rand(n) * 360.
    dec = np.arcsin(np.random.rand(n) * 2. - 1.) * 180. / np.pi
    xyz = angular.cartesian(ra, dec)
    if nf is None:
        return (xyz,)
    return xyz, (ra / 360. * nf).astype("int32")


def _oracle_counts(oracle, pairs, radii):
    """np.diff of the oracle's cumulative counts for (tree, query) pairs."""
    out = []
    for t, q in pairs:
        o = oracle.OracleTree(t.points, t.fields, t.N_fields)
        cum = o.cumulative(q.points, q.fields, q.N_fields, radii)
        out.append(np.diff(cum, axis=0))
    return out


def _check(res, counts, name_prefix):
    from corruscant_amd import clustering as C
    assert len(res.count_arrays) == len(counts)
    for a, b in zip(res.count_arrays, counts):
        assert np.array_equal(a, b), name_prefix
    # a tree without fields has 0-d field sizes; meshgrid treats them as length 1
    sizes = [(np.atleast_1d(a), np.atleast_1d(b)) for a, b in res.count_sizes]
    for name, fn in EST.items():
        res.estimator = getattr(C, fn)
        np.testing.assert_array_equal(res.estimate(), RO.estimate(name, counts, sizes))
        if res.N_fields > 1:
            # exact integer totals and the reference's multiplication order:
            # the errors equal the oracle's bit for bit
            for et in ("jackknife", "ftf", "poisson"):
                np.testing.assert_array_equal(res.error(et), RO.error(name, counts, sizes, et))
            np.random.seed(11)
            got = res.error("bootstrap")
            np.random.seed(11)
            draws = np.random.randint(0, res.N_fields, 1000 * res.N_fields).reshape(1000, res.N_fields)
            np.testing.assert_allclose(got, RO.bootstrap(name, counts, sizes, draws), rtol=1e-12)
        for et in ([None, "poisson", "ftf", "jackknife", "bootstrap"] if res.N_fields > 1 else [None]):
            res.error_type = et
            text = str(res)
            assert len(text.splitlines()) >= res.nbins


def test_reference_alltypes_script(oracle):
    from corruscant_amd import clustering, twopoint
    np.random.seed(3)
    print("alltypes: 3-D auto without fields", flush=True)
    d, r = clustering.tree(*random_cube(DSIZE)), clustering.tree(*random_cube(RSIZE))
    res = twopoint.threedim.autocorr(d, r, RADII, num_threads=4, err_type=None)
    _check(res, _oracle_counts(oracle, [(d, d), (r, d), (r, r)], RADII), "auto")

    print("alltypes: 3-D auto with fields", flush=True)
    d, r = clustering.tree(*random_cube(DSIZE, NF)), clustering.tree(*random_cube(RSIZE, NF))
    res = twopoint.threedim.autocorr(d, r, RADII, num_threads=4)
    _check(res, _oracle_counts(oracle, [(d, d), (r, d), (r, r)], RADII), "auto fields")

    print("alltypes: 3-D cross", flush=True)
    d1, d2 = clustering.tree(*random_cube(DSIZE, NF)), clustering.tree(*random_cube(DSIZE, NF))
    r1, r2 = clustering.tree(*random_cube(RSIZE, NF)), clustering.tree(*random_cube(RSIZE, NF))
    res = twopoint.threedim.crosscorr(d1, d2, r1, r2, RADII, num_threads=4)
    dd = (d1, d2) if d1.size > d2.size else (d2, d1)
    rr = (r1, r2) if r1.size > r2.size else (r2, r1)
    _check(res, _oracle_counts(oracle, [dd, (r2, d1), (r1, d2), rr], RADII), "cross")

    print("alltypes: angular auto", flush=True)
    d, r = clustering.tree(*random_sphere(DSIZE, NF)), clustering.tree(*random_sphere(RSIZE, NF))
    res = twopoint.angular.autocorr(d, r, RADII_ANG, num_threads=4)
    assert res.radii_units == "degrees" and np.array_equal(res.radii_nominal, RADII_ANG)
    _check(res, _oracle_counts(oracle, [(d, d), (r, d), (r, r)], res.radii_euclidean), "angular auto")

    print("alltypes: angular cross", flush=True)
    d1, d2 = clustering.tree(*random_sphere(DSIZE, NF)), clustering.tree(*random_sphere(DSIZE, NF))
    r1, r2 = clustering.tree(*random_sphere(RSIZE, NF)), clustering.tree(*random_sphere(RSIZE, NF))
    res = twopoint.angular.crosscorr(d1, d2, r1, r2, RADII_ANG, num_threads=4)
    dd = (d1, d2) if d1.size > d2.size else (d2, d1)
    rr = (r1, r2) if r1.size > r2.size else (r2, r1)
    _check(res, _oracle_counts(oracle, [dd, (r2, d1), (r1, d2), rr], res.radii_euclidean), "angular cross")


def test_reference_duplicates_script():
    """test/test_duplicates.py: 30,000 random points, each repeated three
    times, fields from x: 30,000 unique points; field sizes count every copy."""
    from corruscant_amd import clustering
    X = np.repeat(np.random.default_rng(30).random((30000, 3)), 3, axis=0)
    f = (X[:, 0] * 4).astype("int32")
    t = clustering.tree(X, f)
    assert t.size == 30000
    assert np.array_equal(t.field_sizes, np.bincount(f, minlength=4))
