"""CPU tests: the oracle pinned against the reference's own outputs (golden
fixtures written by tests/golden/make_golden.py from the reference's C code),
against the compiled reference where it is present, and against what the
mathematics fixes (brute force, closed forms, invariants)."""
import os

import numpy as np
import pytest

from tests import cases

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _npz(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _nodes_from_fixture(raw):
    from oracle.oracle import NODE_DTYPE
    return np.ascontiguousarray(raw).view(NODE_DTYPE).reshape(-1)


# ---- golden fixtures ------------------------------------------------------

def test_small_trees_match_reference_fixture(oracle):
    z = _npz("ref_nodes_small.npz")
    names = sorted({k.split("/")[0] for k in z.files})
    assert len(names) >= 10
    for name in names:
        X = z[name + "/xyz"]
        f = z[name + "/fields"]
        f = f if f.size else None
        nf = int(z[name + "/num_fields"])
        t = oracle.OracleTree(X, f, nf)
        ref = _nodes_from_fixture(z[name + "/nodes"])
        assert t.size == int(z[name + "/size"]), name
        assert oracle.nodes_equal(t.nodes, ref), name


def test_counts_match_reference_fixture(oracle):
    z = _npz("ref_counts.npz")
    for name, nf in (("lattice", 3), ("dups", 4)):
        X, f, radii = z[name + "/xyz"], z[name + "/fields"], z[name + "/radii"]
        t = oracle.OracleTree(X, f, nf)
        assert np.array_equal(t.cumulative(X, f, nf, radii), z[name + "/counts"]), name
    t = oracle.OracleTree(z["cross/xyz"], z["cross/fields"], 6)
    got = t.cumulative(z["cross/q"], z["cross/qf"], 6, z["cross/radii"])
    assert np.array_equal(got, z["cross/counts"])


def test_config1_matches_reference_fixture(oracle):
    import hashlib
    z = _npz("ref_config1.npz")
    D, df, R, rf, radii = cases.config1()
    assert np.array_equal(radii, z["radii"])
    td, tr = oracle.OracleTree(D, df, 4), oracle.OracleTree(R, rf, 4)

    def digest(nodes):
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(nodes["value"]).view(np.uint64).tobytes())
        h.update(np.ascontiguousarray(nodes["field"]).tobytes())
        h.update(np.ascontiguousarray(nodes["weight"]).tobytes())
        h.update(np.ascontiguousarray(nodes["bits"] & 0x3FF).tobytes())
        return h.hexdigest()

    assert digest(td.nodes) == str(z["d_digest"])
    assert digest(tr.nodes) == str(z["r_digest"])
    assert np.array_equal(td.cumulative(D, df, 4, radii), z["dd"])
    assert np.array_equal(tr.cumulative(D, df, 4, radii), z["dr"])
    assert np.array_equal(tr.cumulative(R, rf, 4, radii), z["rr"])


def test_coords_match_reference_fixture(oracle):
    z = _npz("ref_coords.npz")
    assert np.array_equal(oracle.radec2cart64(z["ra"], z["dec"]).view(np.uint64), z["xyz"].view(np.uint64))
    assert np.array_equal(oracle.radecdist2cart64(z["ra"], z["dec"], z["dist"]).view(np.uint64),
                          z["xyzd"].view(np.uint64))
    th, ph = oracle.radec2sph64(z["ra"], z["dec"])
    assert np.array_equal(th, z["theta"]) and np.array_equal(ph, z["phi"])


# ---- compiled reference, where present (this container) -------------------

needs_ref = pytest.mark.skipif(not __import__("oracle.oracle").oracle.ref_available(),
                               reason="oracle/_ref not built (reference sources absent)")


@needs_ref
@pytest.mark.parametrize("seed", range(6))
def test_oracle_equals_compiled_reference(oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(50, 3000))
    if seed % 3 == 0:
        X = np.round(rng.random((n, 3)) * 16) / 16  # many coincident coordinates and duplicates
    elif seed % 3 == 1:
        X = cases.clustered(n, 5, 40, 0.01, seed)
    else:
        X = rng.random((n, 3)) * [1.0, 1e-3, 100.0]  # very unequal extents
    f = rng.integers(0, 5, X.shape[0]).astype(np.int32)
    a = oracle.OracleTree(X, f, 5)
    b = oracle.RefTree(X, f, 5)
    assert oracle.nodes_equal(a.nodes, b.nodes)
    radii = np.sort(rng.random(5) * 0.3)
    assert np.array_equal(a.cumulative(X, f, 5, radii), b.cumulative(X, f, 5, radii, nthreads=2))


# ---- mathematics ------------------------------------------------------------

def test_build_closed_forms_for_tiny_n(oracle):
    """n <= 3: the split dimension is 0 and the node order is fixed by the
    (x,y,z) order (src/kdbuild.c:396-417)."""
    X = np.array([[0.3, 0.0, 0.0], [0.1, 0.0, 0.0], [0.2, 0.0, 0.0]])
    t = oracle.OracleTree(X)
    assert t.memsize == 4
    assert [t.nodes[i]["value"][0] for i in (1, 2, 3)] == [0.2, 0.1, 0.3]
    assert [int(t.nodes[i]["bits"]) for i in (1, 2, 3)] == [3, 0, 0]
    t2 = oracle.OracleTree(X[:2])
    assert t2.memsize == 4 and t2.nodes[1]["value"][0] == 0.3 and int(t2.nodes[1]["bits"]) == 1
    assert t2.nodes[2]["value"][0] == 0.1


def test_heap_shape_depends_only_on_n(oracle):
    for n in (4, 9, 100, 777):
        a = oracle.OracleTree(cases.cube(n, 1)[0])
        b = oracle.OracleTree(cases.cube(n, 2)[0] * 7 - 3)
        assert np.array_equal(a.perm >= 0, b.perm >= 0)
        assert np.array_equal(a.inorder, b.inorder)
        assert np.array_equal(a.nodes["bits"] & 3, b.nodes["bits"] & 3)


def test_tree_invariants(oracle):
    """Children lie on the correct side of every ancestor's split (the
    property src/kdtest.c:73-99 checks for immediate children), the in-order
    positions are a permutation, and weights sum to N."""
    X = cases.clustered(3000, 10, 100, 0.02, 3)
    X = np.vstack([X, X[:500]])
    t = oracle.OracleTree(X)
    assert t.nodes["weight"].sum() == X.shape[0]
    live = np.nonzero(t.perm >= 0)[0]
    assert np.array_equal(np.sort(t.inorder[live]), np.arange(t.size))
    for h in live[live > 1][:4000]:
        c = h
        while c > 1:
            a = c >> 1
            k = (int(t.nodes[a]["bits"]) >> 2) & 0xFF
            s = t.nodes[a]["value"][k]
            v = t.nodes[h]["value"][k]
            if c % 2 == 0:
                assert v <= s
            else:
                assert v >= s
            c = a


def test_pair_count_equals_brute_force(oracle):
    """The traversal equals the ancestor-walk formulation of the reference's
    pruning, and on generic inputs also the plain exact count."""
    rng = np.random.default_rng(4)
    X = rng.random((600, 3))
    f = rng.integers(0, 3, 600).astype(np.int32)
    Q = rng.random((300, 3))
    qf = rng.integers(0, 3, 300).astype(np.int32)
    t = oracle.OracleTree(X, f, 3)
    for r in (0.0, 0.05, 0.1, 0.3, 1.8):
        a = t.pair_count(Q, qf, 3, r)
        assert np.array_equal(a, t.pair_count_brute(Q, qf, 3, r, use_reach=True))
        assert np.array_equal(a, t.pair_count_brute(Q, qf, 3, r, use_reach=False))


def test_pruning_never_drops_a_counted_pair(oracle):
    """The reference's pruning (src/kdquery.c:96-109) cannot drop a pair with
    n < r*r: if fl(q_k - r) > s then q_k - r > s exactly (rounding is
    monotone and s is a double), every point p of the left subtree has
    p_k <= s, so |p_k - q_k| > r, fl(p_k - q_k) >= r and the FMA chain gives
    n >= fl(r*r).  The right side is symmetric.  So the traversal count is the
    plain predicate count, which the GPU kernels compute.  Adversarial
    queries: every node, every axis, both sides, at distance r and the next
    ulps inside it."""
    rng = np.random.default_rng(1)
    checked = 0
    for trial in range(60):
        X = rng.random((9, 3)) if trial % 2 else np.round(rng.random((9, 3)) * 8) / 8
        t = oracle.OracleTree(X)
        r = float(rng.random() * 0.2 + 0.01) if trial % 3 else 0.125
        for h in np.nonzero(t.perm >= 0)[0]:
            p = t.nodes[h]["value"].copy()
            for k in range(3):
                for sgn in (1.0, -1.0):
                    q = p.copy()
                    q[k] = p[k] + sgn * r
                    for _ in range(4):
                        Q = q.reshape(1, 3)
                        a = t.pair_count(Q, None, 1, r)
                        assert np.array_equal(a, t.pair_count_brute(Q, None, 1, r, use_reach=True))
                        assert np.array_equal(a, t.pair_count_brute(Q, None, 1, r, use_reach=False))
                        checked += 1
                        q[k] = np.nextafter(q[k], p[k])
    assert checked > 5000


def test_lattice_closed_form_counts(oracle):
    """Dyadic lattice (exact distances): the count at radius r is the number
    of ordered lattice pairs with squared distance < r*r, computable exactly."""
    m = 6
    h = 0.125
    X = cases.lattice(m, spacing=h)
    t = oracle.OracleTree(X)
    d = X[:, None, :] - X[None, :, :]
    n2 = (d ** 2).sum(-1)
    for r in (h, h * np.sqrt(2), 2 * h, 0.3, 1.0):
        expected = int((n2 < r * r).sum())
        assert int(t.pair_count(X, None, 1, r).sum()) == expected


def test_symmetry_field_sum_and_weights(oracle):
    X, f = cases.cube(1500, 12, nfields=4)
    t = oracle.OracleTree(X, f, 4)
    c = t.pair_count(X, f, 4, 0.1)
    assert np.array_equal(c, c.T)  # DD without duplicates is symmetric
    t1 = oracle.OracleTree(X)
    assert c.sum() == t1.pair_count(X, None, 1, 0.1).sum()
    # repeating the tree points k times == node weight k
    X3 = np.repeat(X[:300], 3, axis=0)
    t3 = oracle.OracleTree(X3)
    Q = cases.cube(200, 13)[0]
    assert t3.pair_count(Q, None, 1, 0.2).sum() == 3 * oracle.OracleTree(X[:300]).pair_count(Q, None, 1, 0.2).sum()


def test_256_copies_error(oracle):
    base = np.random.default_rng(0).random((2, 3))
    with pytest.raises(ValueError, match="Invalid weights"):
        oracle.OracleTree(np.repeat(base, 256, axis=0))
    t = oracle.OracleTree(np.repeat(base, 255, axis=0))  # 255 copies each is fine
    assert t.size == 2 and t.nodes["weight"].sum() == 510
