"""GPU tree build vs the oracle: node arrays field by field, permutation,
error paths (src/kdbuild.c:451-537)."""
import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


class _RawTree:
    """Tree built through the C ABI directly (clustering.tree, like the
    reference's validate_points, cannot tell (N,3) from (3,N) for N < 3)."""

    def __init__(self, X, f):
        from corruscant_amd import clustering as C
        X = np.ascontiguousarray(X, dtype=np.float64)
        self.ctree = C.kdlib.tree_construct(C._dp(X), C._ip(f), X.shape[0], 1)
        if not self.ctree.dev:
            raise RuntimeError(C._err())
        self.size = self.ctree.size

    node_array = lambda self: __import__("corruscant_amd.clustering").clustering.tree.node_array(self)
    heap_permutation = lambda self: __import__("corruscant_amd.clustering").clustering.tree.heap_permutation(self)


def _gpu_tree(X, f=None):
    from corruscant_amd import clustering
    if X.shape[0] < 3:
        return _RawTree(X, f)
    return clustering.tree(X, f)


def _check(oracle, X, f=None, nf=1):
    from oracle.oracle import nodes_equal
    print("build case n=%d" % X.shape[0], flush=True)
    t = _gpu_tree(X, f)
    o = oracle.OracleTree(X, f, nf if f is not None else 1)
    assert t.size == o.size
    assert t.ctree.memsize == o.memsize
    nodes = t.node_array()
    assert nodes_equal(nodes, o.nodes), "node arrays differ"
    perm, ino = t.heap_permutation()
    assert np.array_equal(perm, o.perm)
    assert np.array_equal(ino, o.inorder)
    return t, o


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 15, 16, 17, 31, 32, 33, 100, 257, 1000, 4097])
def test_small_sizes(oracle, n):
    X, f = cases.cube(n, seed=n, nfields=3)
    if n < 3:
        f = None
    _check(oracle, X, f, 3)


def test_config1_trees(oracle):
    D, df, R, rf, _ = cases.config1()
    _check(oracle, D, df, 4)
    _check(oracle, R, rf, 4)


def test_duplicates_and_signed_zero(oracle):
    X = cases.duplicates(3000, 3, seed=7)
    f = (np.arange(X.shape[0]) % 7).astype(np.int32)
    t, o = _check(oracle, X, f, 7)
    # the +/-0 edits in cases.duplicates split some runs; -0.0 and +0.0 merge
    assert t.size == np.unique(X + 0.0, axis=0).shape[0]


def test_many_duplicates_up_to_255(oracle):
    rng = np.random.default_rng(11)
    base = rng.random((50, 3))
    X = np.repeat(base, 255, axis=0)
    _check(oracle, X, None)


def test_256_copies_is_an_error():
    X = np.repeat(np.random.default_rng(1).random((3, 3)), 256, axis=0)
    with pytest.raises(RuntimeError, match="Invalid weights"):
        _gpu_tree(X)


def test_nan_is_an_error():
    X = np.random.default_rng(1).random((100, 3))
    X[17, 2] = np.nan
    with pytest.raises(RuntimeError, match="NaN"):
        _gpu_tree(X)


def test_lattice_variance_ties(oracle):
    """Equal extents on every axis: exact variance ties at every level, so the
    split dimension comes from the reference's rounding (sequential replay)."""
    X = cases.lattice(24, spacing=1.0 / 24)
    _check(oracle, X)


def test_wide_fields_truncate_to_8_bits(oracle):
    X, _ = cases.cube(5000, seed=5)
    f = (np.arange(5000) % 300).astype(np.int32)
    _check(oracle, X, f, 300)


def test_clustered_medium(oracle):
    X = cases.clustered(20000, 40, 250, 0.005, seed=2)
    f = (4 * np.floor(4 * X[:, 0]) + np.floor(4 * X[:, 1])).astype(np.int32)
    _check(oracle, X, f, 16)


def test_large_segments_certified(oracle):
    """Above the sequential-replay size the split dimension is decided by the
    certified parallel sums."""
    X, f = cases.cube(300000, seed=9, nfields=8)
    _check(oracle, X, f, 8)


def test_tree_attributes_match_reference_expressions():
    """N_fields, field_sizes (clustering.py:336-338: per-id sizes of the raw
    fields) and size (the unique count) of the Python tree object."""
    from corruscant_amd import clustering
    D, df, _, _, _ = cases.config1()
    t = clustering.tree(D, df)
    assert t.N_fields == 4
    assert np.array_equal(t.field_sizes, np.array([df[df == i].size for i in range(4)]))
    X = np.repeat(D[:100], 3, axis=0)
    t2 = clustering.tree(X)
    assert t2.size == 100 and int(t2.field_sizes) == 100 and t2.N_fields == 1


def test_large_set_with_shared_coordinates(oracle):
    """More than 2^20 points, so the presort first tries one radix sort by the
    leading coordinate: x and y on coarse grids (shared by many points, so
    those sorts fall back to the full stable passes), z all distinct (its
    single sort stands), exact duplicates and a signed zero."""
    rng = np.random.default_rng(13)
    X = rng.random((1_200_000, 3))
    X[:, 0] = np.floor(X[:, 0] * 1000) / 1000
    X[:, 1] = np.floor(X[:, 1] * 4096) / 4096
    X[1000:1100] = X[0:100]
    X[5, 0], X[6, 0] = -0.0, 0.0
    f = (4 * np.floor(4 * X[:, 2]) + np.floor(4 * np.abs(X[:, 1]))).astype(np.int32)
    _check(oracle, X, f, 16)


def test_large_set_without_shared_coordinates(oracle):
    """More than 2^20 distinct real-valued points: every presort is the single
    radix sort by its leading coordinate."""
    X, f = cases.cube(1_100_000, seed=17, nfields=8)
    _check(oracle, X, f, 8)
