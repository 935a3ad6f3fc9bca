"""GPU pair counts vs the oracle, bit-exact int64 (src/kdquery.c:165-249 per
radius, corruscant/clustering.py:218-230 over radii)."""
import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

PRECISIONS = ["fp64", "fp32_guarded"]


def _trees(X, f):
    from corruscant_amd import clustering
    return clustering.tree(X, f)


def _gpu_cum(tree, query, radii, prec):
    from corruscant_amd import clustering
    return clustering._count_jobs([(tree, query)], radii, prec)[0]


@pytest.mark.parametrize("prec", PRECISIONS)
def test_config1_dd_dr_rr(oracle, prec):
    D, df, R, rf, radii = cases.config1()
    td, tr = _trees(D, df), _trees(R, rf)
    od, orr = oracle.OracleTree(D, df, 4), oracle.OracleTree(R, rf, 4)
    for name, (t, q, o, qp, qf) in {"DD": (td, td, od, D, df), "DR": (tr, td, orr, D, df),
                                    "RR": (tr, tr, orr, R, rf)}.items():
        print("config1", name, prec, flush=True)
        g = _gpu_cum(t, q, radii, prec)
        ref = o.cumulative(qp, qf, 4, radii)
        assert np.array_equal(g, ref), name


@pytest.mark.parametrize("prec", PRECISIONS)
def test_lattice_edges_on_lattice_distances(oracle, prec):
    """Dyadic lattice, radii exactly equal to lattice distances: the strict
    n < r*r and the reference's pruning decide every boundary pair."""
    X = cases.lattice(12, spacing=0.125)
    f = (np.arange(X.shape[0]) % 3).astype(np.int32)
    radii = np.array([0.125, 0.125 * np.sqrt(2), 0.25, 0.375, 0.125 * np.sqrt(3), 0.5, 1.0])
    t = _trees(X, f)
    o = oracle.OracleTree(X, f, 3)
    print("lattice", prec, flush=True)
    g = _gpu_cum(t, t, radii, prec)
    ref = o.cumulative(X, f, 3, radii)
    assert np.array_equal(g, ref)


@pytest.mark.parametrize("prec", PRECISIONS)
def test_duplicates_weights(oracle, prec):
    X = cases.duplicates(2000, 3, seed=3)
    f = (np.arange(X.shape[0]) % 4).astype(np.int32)
    t = _trees(X, f)
    o = oracle.OracleTree(X, f, 4)
    radii = np.logspace(-2.5, -0.5, 8)
    g = _gpu_cum(t, t, radii, prec)
    assert np.array_equal(g, o.cumulative(X, f, 4, radii))


@pytest.mark.parametrize("prec", PRECISIONS)
def test_clustered_crossed_fields(oracle, prec):
    X = cases.clustered(6000, 20, 200, 0.004, seed=4)
    f = (4 * np.floor(4 * X[:, 0]) + np.floor(4 * X[:, 1])).astype(np.int32)
    Y, g_ = cases.cube(30000, seed=44)
    gf = (4 * np.floor(4 * Y[:, 0]) + np.floor(4 * Y[:, 1])).astype(np.int32)
    radii = np.logspace(-3, -1, 20)
    ty, tx = _trees(Y, gf), _trees(X, f)
    oy = oracle.OracleTree(Y, gf, 16)
    g = _gpu_cum(ty, tx, radii, prec)
    assert np.array_equal(g, oy.cumulative(X, f, 16, radii))


def test_reference_single_radius_api(oracle):
    """pair_count with the reference's argument list (arbitrary query points)."""
    from corruscant_amd import clustering
    X, f = cases.cube(3000, seed=21, nfields=5)
    Q, qf = cases.cube(1000, seed=22, nfields=5)
    t = _trees(X, f)
    o = oracle.OracleTree(X, f, 5)
    for r in (0.0, 0.01, 0.05, 0.2):
        got = clustering._query_tree(t, Q, r, 4, qf, 5)
        assert np.array_equal(got, o.pair_count(Q, qf, 5, r))


def test_unsorted_and_repeated_radii(oracle):
    X, f = cases.cube(5000, seed=31, nfields=4)
    t = _trees(X, f)
    o = oracle.OracleTree(X, f, 4)
    radii = np.array([0.05, 0.01, 0.05, 0.0, 0.2, 0.02])
    g = _gpu_cum(t, t, radii, "fp64")
    assert np.array_equal(g, o.cumulative(X, f, 4, radii))


def test_field_index_outside_matrix_is_an_error():
    from corruscant_amd import clustering
    X, f = cases.cube(2000, seed=1, nfields=4)
    t = _trees(X, f)
    with pytest.raises(RuntimeError):
        clustering._query_tree(t, X, 0.1, 4, f, 2)  # query fields 0..3 with a stride of 2
