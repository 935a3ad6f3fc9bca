"""Edge cases of the GPU pair counts against the oracle: trees and queries of
1-4 points (the build's base cases, src/kdbuild.c:384-433), radius 0, a
radius beyond every pair, offset and negative coordinates in both precision
forms, the radius-count limit and invalid radii."""
import ctypes

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


class _Handle:
    """A tree (or query set) built through the C ABI directly: clustering.tree,
    like the reference's validate_points, cannot tell (N,3) from (3,N) for
    N < 3."""

    def __init__(self, X, f, nf, query_only=False):
        from corruscant_amd import clustering as C
        self.points = np.ascontiguousarray(X, dtype=np.float64)
        self.fields = None if f is None else np.ascontiguousarray(f, dtype=np.int32)
        self.N_fields = nf
        make = C.kdlib.query_construct if query_only else C.kdlib.tree_construct
        self.ctree = make(C._dp(self.points), C._ip(self.fields), self.points.shape[0], nf)
        if not self.ctree.dev:
            raise RuntimeError(C._err())

    def __del__(self):
        from corruscant_amd import clustering as C
        if getattr(self, "ctree", None) is not None and self.ctree.dev:
            C.kdlib.destroy(self.ctree)


def _cum(tree, query, radii, prec="fp64"):
    from corruscant_amd import clustering as C
    return C._count_jobs([(tree, query)], radii, prec)[0]


@pytest.mark.parametrize("nt", [1, 2, 3, 4])
@pytest.mark.parametrize("nq", [1, 2, 5])
def test_tiny_trees_and_queries(oracle, nt, nq):
    print("tiny nt=%d nq=%d" % (nt, nq), flush=True)
    rng = np.random.default_rng(100 * nt + nq)
    T = rng.random((nt, 3)) * 0.3
    Q = rng.random((nq, 3)) * 0.3
    tf = rng.integers(0, 2, nt).astype(np.int32)
    qf = rng.integers(0, 2, nq).astype(np.int32)
    radii = np.array([0.0, 0.05, 0.1, 0.2, 0.3, 1.0])
    t = _Handle(T, tf, 2)
    for query_only in (False, True):
        q = _Handle(Q, qf, 2, query_only=query_only)
        ref = oracle.OracleTree(T, tf, 2).cumulative(Q, qf, 2, radii)
        for prec in ("fp64", "fp32_guarded"):
            assert np.array_equal(_cum(t, q, radii, prec), ref), (query_only, prec)


def test_radius_zero_and_radius_beyond_every_pair(oracle):
    """r = 0 counts nothing (n < 0 never holds); r = 2 in the unit cube counts
    every (query point, tree point) pair, copies included: the matrix is the
    outer product of the raw field sizes."""
    from corruscant_amd import clustering
    X, f = cases.cube(20000, seed=5, nfields=4)
    X = np.concatenate([X, X[:300]])
    f = np.concatenate([f, f[:300]])
    Y, g = cases.cube(7000, seed=6, nfields=4)
    t, q = clustering.tree(X, f), clustering.tree(Y, g)
    for prec in ("fp64", "fp32_guarded"):
        c = _cum(t, q, [0.0, 2.0], prec)
        assert not c[0].any()
        assert np.array_equal(c[1], np.outer(np.bincount(g, minlength=4), np.bincount(f, minlength=4)))


@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_offset_and_negative_coordinates(oracle, prec):
    """Points near (-1000.5, 1000, -3): FP64 differences lose digits and the
    FP32 form must fall back to its FP64 recheck; counts stay exact."""
    rng = np.random.default_rng(7)
    X = rng.random((15000, 3)) * 0.05 + np.array([-1000.5, 1000.0, -3.0])
    Y = rng.random((5000, 3)) * 0.05 + np.array([-1000.5, 1000.0, -3.0])
    fx = (X[:, 0] > -1000.475).astype(np.int32)
    fy = (Y[:, 1] > 1000.025).astype(np.int32)
    from corruscant_amd import clustering
    t, q = clustering.tree(X, fx), clustering.tree(Y, fy)
    radii = np.logspace(-4, -1.5, 9)
    ref = oracle.OracleTree(X, fx, 2).cumulative(Y, fy, 2, radii)
    assert np.array_equal(_cum(t, q, radii, prec), ref)


def test_many_radii_and_invalid_radii(oracle):
    """Any number of radii, as the reference takes (one pair_count per radius,
    clustering.py:218-230): 64 distinct radii go to the library in one call,
    150 (unsorted, with repeats) in several, and every count equals the
    oracle's.  Negative, NaN and infinite radii are refused."""
    from corruscant_amd import clustering
    X, f = cases.cube(3000, seed=8, nfields=2)
    t = clustering.tree(X, f)
    ot = oracle.OracleTree(X, f, 2)
    r64 = np.linspace(0.001, 0.2, 64)
    c = _cum(t, t, r64)
    assert c.shape == (64, 2, 2) and np.array_equal(c, ot.cumulative(X, f, 2, r64))
    rng = np.random.default_rng(80)
    r150 = rng.permutation(np.concatenate([np.linspace(0.001, 0.25, 140), np.linspace(0.001, 0.25, 140)[:10]]))
    for prec in ("fp64", "fp32_guarded"):
        c = _cum(t, t, r150, prec)
        assert c.shape == (150, 2, 2) and np.array_equal(c, ot.cumulative(X, f, 2, r150)), prec
    assert clustering.last_stats["evaluated_pairs"] > 0
    w = np.ones((2, X.shape[0]), np.uint8)
    w[1, ::3] = 2
    r40 = np.linspace(0.002, 0.2, 40)  # beyond one weighted call's 32 radii
    (cw,) = clustering._count_weighted([(t, t, w, w)], r40, 2)
    assert np.array_equal(cw[0], ot.cumulative(X, f, 2, r40))
    Xw, fw = np.repeat(X, w[1], axis=0), np.repeat(f, w[1])
    assert np.array_equal(cw[1], oracle.OracleTree(Xw, fw, 2).cumulative(Xw, fw, 2, r40))
    for bad in ([-0.1], [np.nan], [np.inf], np.append(np.linspace(0.001, 0.2, 70), np.nan)):
        with pytest.raises(RuntimeError, match="finite and >= 0"):
            _cum(t, t, bad)


def test_empty_catalogue_is_refused_before_device_work():
    from corruscant_amd import clustering
    with pytest.raises(ValueError, match="shape"):
        clustering.tree(np.zeros((0, 3)))


def test_counts_do_not_depend_on_the_tile_size(oracle):
    """Trees laid out in tiles of 64, 48, 33 and 1 points give the same
    counts (the tile size only changes how the pairs are grouped)."""
    from corruscant_amd import clustering as C
    X, f = cases.cube(6000, seed=12, nfields=3)
    Y, g = cases.cube(2500, seed=13, nfields=3)
    radii = np.array([0.01, 0.03, 0.07, 0.15])
    ref = oracle.OracleTree(X, f, 3).cumulative(Y, g, 3, radii)
    try:
        for tile in (64, 48, 33, 1):
            assert C.kdlib.kd_set_tile_size(tile) == 0
            t, q = C.tree(X, f), C.tree(Y, g)
            for prec in ("fp64", "fp32_guarded"):
                assert np.array_equal(_cum(t, q, radii, prec), ref), (tile, prec)
    finally:
        C.kdlib.kd_set_tile_size(64)


def test_counts_do_not_depend_on_the_staging_pool(oracle):
    """The enumeration stages each query tile's items in a pool so that it
    walks the tile hierarchy once; tiles that do not fit (a third chunk, or a
    used-up pool) are walked again.  Staging off, a pool of 3000 items (most
    tiles walked twice, some items across two chunks) and the default give
    the oracle's counts, for auto- and cross-pairs in both precision forms."""
    from corruscant_amd import clustering as C
    X, f = cases.cube(12000, seed=14, nfields=3)
    Y, g = cases.cube(5000, seed=15, nfields=3)
    radii = np.logspace(-2.2, -0.9, 8)
    t, q = C.tree(X, f), C.tree(Y, g)
    ref_tq = oracle.OracleTree(X, f, 3).cumulative(Y, g, 3, radii)
    ref_tt = oracle.OracleTree(X, f, 3).cumulative(X, f, 3, radii)
    try:
        for items in (-1, 3000, 0):
            assert C.kdlib.kd_set_stage_pool(items) == 0
            for prec in ("fp64", "fp32_guarded"):
                tq, tt = C._count_jobs([(t, q), (t, t)], radii, prec)
                assert np.array_equal(tq, ref_tq), (items, prec)
                assert np.array_equal(tt, ref_tt), (items, prec)
    finally:
        C.kdlib.kd_set_stage_pool(0)


def test_trimmed_items_match_the_oracle(oracle):
    """The trim pass (FP32 form) cuts off the end quads of items whose radius
    window reaches past the largest radius when their boxes lie beyond it; the
    FP64 loop skips every such quad through its per-quad radius windows.  With
    the largest radius close to the tile size most such items are cut: the
    counts stay the oracle's (cross, auto with mirrored tile pairs, and a tree
    with duplicates) in both forms, the FP32 form's trim pass cuts items, and
    no more pairs are evaluated than with an extra radius beyond every pair,
    which leaves every window short of the largest radius (nothing is cut;
    the quad windows skip the same quads either way)."""
    from corruscant_amd import clustering as C
    X, f = cases.cube(20000, seed=16, nfields=4)
    Y, g = cases.cube(8000, seed=17, nfields=4)
    Xd = np.concatenate([X, X[:3000]])
    fd = np.concatenate([f, f[:3000]])
    t, q, td = C.tree(X, f), C.tree(Y, g), C.tree(Xd, fd)
    for radii in (np.array([0.01, 0.02, 0.04, 0.06]), np.array([0.005, 0.05, 0.12])):
        wide = np.append(radii, 10.0)
        ref = [oracle.OracleTree(X, f, 4).cumulative(Y, g, 4, wide),
               oracle.OracleTree(X, f, 4).cumulative(X, f, 4, wide),
               oracle.OracleTree(Xd, fd, 4).cumulative(Y, g, 4, wide)]
        for prec in ("fp64", "fp32_guarded"):
            evaluated, trimmed = [], []
            for rr in (radii, wide):
                got = C._count_jobs([(t, q), (t, t), (td, q)], rr, prec)
                for k in range(3):
                    assert np.array_equal(got[k], ref[k][: rr.size]), (rr, prec, k)
                evaluated.append(C.last_stats["evaluated_pairs"])
                trimmed.append(C.last_stats["trimmed_items"])
            assert evaluated[0] <= evaluated[1], (radii, prec, evaluated)
            assert trimmed[1] == 0, (radii, prec, trimmed)
            if prec == "fp32_guarded":
                assert trimmed[0] > 0, (radii, prec, trimmed)
