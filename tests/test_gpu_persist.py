"""Tree files: tree.save / tree.load round trips.  The saved node array is the
reference's record layout (src/kdtree.h:13-25); it is compared with the
oracle's, and a file whose node array disagrees with its points is refused."""
import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


def _roundtrip(oracle, tmp_path, X, f, nf):
    from corruscant_amd import clustering
    from oracle.oracle import NODE_DTYPE, nodes_equal
    print("persist case n=%d fields=%s" % (X.shape[0], f is not None), flush=True)
    t = clustering.tree(X, f)
    path = str(tmp_path / "tree.npz")
    t.save(path)
    with np.load(path, allow_pickle=False) as z:
        saved = np.ascontiguousarray(z["nodes"]).view(NODE_DTYPE).reshape(-1)
        perm = z["perm"]
    o = oracle.OracleTree(X, f, nf if f is not None else 1)
    assert nodes_equal(saved, o.nodes) and np.array_equal(perm, o.perm)
    u = clustering.tree.load(path)
    assert u.size == t.size and u.N_fields == t.N_fields
    assert np.array_equal(u.points, t.points)
    assert np.array_equal(u.field_sizes, t.field_sizes)
    assert (u.fields is None) == (t.fields is None)
    # the loaded tree counts like the saved one
    r = [0.05, 0.1, 0.2]
    a = clustering._count_jobs([(t, t)], r)[0]
    b = clustering._count_jobs([(u, u)], r)[0]
    assert np.array_equal(a, b)
    return path


def test_roundtrip_with_fields(oracle, tmp_path):
    X, f = cases.cube(5000, seed=21, nfields=4)
    _roundtrip(oracle, tmp_path, X, f, 4)


def test_roundtrip_without_fields_and_with_duplicates(oracle, tmp_path):
    rng = np.random.default_rng(22)
    X = rng.random((3000, 3))
    X = np.concatenate([X, X[:700], X[:50]])
    _roundtrip(oracle, tmp_path, X, None, 1)


def test_tampered_node_array_is_refused(oracle, tmp_path):
    from corruscant_amd import clustering
    X, f = cases.cube(2000, seed=23, nfields=3)
    path = _roundtrip(oracle, tmp_path, X, f, 3)
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    arrays["nodes"] = arrays["nodes"].copy()
    arrays["nodes"][7, 0] ^= 1  # one bit of one coordinate
    bad = str(tmp_path / "bad.npz")
    np.savez(bad, **arrays)
    with pytest.raises(ValueError, match="differs"):
        clustering.tree.load(bad)
    assert clustering.tree.load(bad, verify=False).size == 2000
