"""Weighted pair counts for S weight sets (per-point multiplicities, e.g.
bootstrap draws), bit-exact against the oracle run on the expanded catalogues
np.repeat(points, w_s): tree-side repeats merge into node weights and
query-side repeats count one by one, as in the reference's own dedup
(src/kdbuild.c:451-516) and query loop (src/kdquery.c:165-234)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _expanded(oracle, T, tf, Q, qf, nf, wt, wq, radii):
    """Oracle cumulative counts (nradii, nf, nf) for one weight set."""
    if wt is not None:
        T, tf = np.repeat(T, wt, axis=0), np.repeat(tf, wt)
    if wq is not None:
        Q, qf = np.repeat(Q, wq, axis=0), np.repeat(qf, wq)
    return oracle.OracleTree(T, tf, nf).cumulative(Q, qf, nf, radii)


def _fields(X, nf):
    return np.minimum((X[:, 0] * nf).astype(np.int32), nf - 1)


@pytest.mark.parametrize("nsets", [1, 11])
def test_weighted_jobs_match_expanded_catalogues(oracle, nsets):
    from corruscant_amd import clustering
    rng = np.random.default_rng(21)
    nf = 4
    A, B, R = rng.random((2500, 3)), rng.random((3200, 3)), rng.random((6000, 3))
    fa, fb, fr = _fields(A, nf), _fields(B, nf), _fields(R, nf)
    ta, tb, tr = clustering.tree(A, fa), clustering.tree(B, fb), clustering.tree(R, fr)
    wa = rng.poisson(1.0, (nsets, A.shape[0])).astype(np.uint8)
    wb = rng.poisson(1.0, (nsets, B.shape[0])).astype(np.uint8)
    radii = np.logspace(-2.5, -0.8, 12)  # more radii than one pass of the pair loop takes
    jobs = [(tb, ta, wb, wa),      # both sides weighted
            (tr, ta, None, wa),    # query side only
            (tb, tr, wb, None),    # tree side only
            (ta, ta, wa, wa),      # auto pair, both sides
            (tr, tr, None, None)]  # unweighted, through the weighted entry point
    out, fsz = clustering._count_weighted(jobs, radii, nsets, sizes=True)
    assert clustering.last_stats["evaluated_pairs"] > 0
    # per-set field sizes of the resampled catalogues (sum of multiplicities per field)
    wsz = lambda f, w: np.stack([np.bincount(f, weights=w[k], minlength=nf) for k in range(nsets)]).astype(np.int64)
    ones = lambda f: np.tile(np.bincount(f, minlength=nf), (nsets, 1))
    assert np.array_equal(fsz[0][0], wsz(fb, wb)) and np.array_equal(fsz[0][1], wsz(fa, wa))
    assert np.array_equal(fsz[1][0], ones(fr)) and np.array_equal(fsz[2][1], ones(fr))
    for s in range(nsets):
        assert np.array_equal(out[0][s], _expanded(oracle, B, fb, A, fa, nf, wb[s], wa[s], radii)), s
        assert np.array_equal(out[1][s], _expanded(oracle, R, fr, A, fa, nf, None, wa[s], radii)), s
        assert np.array_equal(out[2][s], _expanded(oracle, B, fb, R, fr, nf, wb[s], None, radii)), s
        assert np.array_equal(out[3][s], _expanded(oracle, A, fa, A, fa, nf, wa[s], wa[s], radii)), s
        assert np.array_equal(out[4][s], oracle.OracleTree(R, fr, nf).cumulative(R, fr, nf, radii)), s


def test_many_sets_and_unit_weights(oracle):
    """70 sets (more than one wave's lanes take at once), with weight 0 and
    large multiplicities; all-ones sets equal the unweighted counts."""
    from corruscant_amd import clustering
    rng = np.random.default_rng(22)
    D, R = rng.random((1500, 3)), rng.random((4000, 3))
    td, tr = clustering.tree(D), clustering.tree(R)
    nsets = 70
    wd = rng.integers(0, 4, (nsets, D.shape[0])).astype(np.uint8)
    wd[3] = 1
    wd[5, ::7] = 200
    radii = np.array([0.01, 0.03, 0.05, 0.1, 0.2])
    dd, dr = clustering._count_weighted([(td, td, wd, wd), (tr, td, None, wd)], radii, nsets)
    (dd1, dr1) = clustering._count_jobs([(td, td), (tr, td)], radii)
    assert np.array_equal(dd[3], dd1) and np.array_equal(dr[3], dr1)
    one = np.zeros(D.shape[0], np.int32)
    for s in [0, 5, 37, 69]:
        assert np.array_equal(dd[s], _expanded(oracle, D, one, D, one, 1, wd[s], wd[s], radii)), s
        assert np.array_equal(dr[s], _expanded(oracle, R, np.zeros(R.shape[0], np.int32), D, one, 1, None, wd[s],
                                               radii)), s


def test_query_weights_against_tree_with_duplicates(oracle):
    """Unweighted tree side with duplicate points (node weights > 1) and
    duplicate query points with their own multiplicities."""
    from corruscant_amd import clustering
    rng = np.random.default_rng(23)
    base = np.round(rng.random((1200, 3)) * 64) / 64.0  # a lattice: many exact repeats
    T = np.concatenate([base, base[:400], -0.0 + base[:50]])
    Q = np.concatenate([rng.random((900, 3)), base[:300]])
    nf = 2
    tf, qf = _fields(T, nf), _fields(Q, nf)
    tt, tq = clustering.tree(T, tf), clustering.tree(Q, qf)
    w = rng.poisson(1.5, (3, Q.shape[0])).astype(np.uint8)
    radii = np.array([0.0, 1 / 64.0, 0.05, 0.15])
    (c,) = clustering._count_weighted([(tt, tq, None, w)], radii, 3)
    for s in range(3):
        assert np.array_equal(c[s], _expanded(oracle, T, tf, Q, qf, nf, None, w[s], radii)), s


def test_tree_weights_need_a_tree_without_duplicates():
    from corruscant_amd import clustering
    rng = np.random.default_rng(24)
    X = rng.random((500, 3))
    X[10] = X[20]
    t = clustering.tree(X)
    with pytest.raises(RuntimeError, match="without duplicate points"):
        clustering._count_weighted([(t, t, np.ones((2, 500), np.uint8), None)], [0.1], 2)
    with pytest.raises(ValueError):
        clustering._count_weighted([(t, t, None, np.ones((3, 500), np.uint8))], [0.1], 2)


def test_autocorr_and_crosscorr_weighted(oracle):
    """The resampled-catalogue correlation functions against the unweighted
    path run on the expanded catalogues, counts and field sizes alike."""
    from corruscant_amd import clustering
    rng = np.random.default_rng(25)
    nf = 4
    D1, D2 = rng.random((1800, 3)), rng.random((2600, 3))
    R1, R2 = rng.random((5000, 3)), rng.random((4000, 3))
    f1, f2, g1, g2 = _fields(D1, nf), _fields(D2, nf), _fields(R1, nf), _fields(R2, nf)
    t1, t2 = clustering.tree(D1, f1), clustering.tree(D2, f2)
    r1, r2 = clustering.tree(R1, g1), clustering.tree(R2, g2)
    S = 3
    w1 = rng.poisson(1.0, (S, D1.shape[0])).astype(np.uint8)
    w2 = rng.poisson(1.0, (S, D2.shape[0])).astype(np.uint8)
    radii = np.logspace(-2, -0.9, 7)
    auto = clustering.autocorr_weighted(t1, r1, radii, w1)
    cross = clustering.crosscorr_weighted(t1, t2, r1, r2, radii, w1, w2)
    for s in range(S):
        E1, e1 = np.repeat(D1, w1[s], axis=0), np.repeat(f1, w1[s])
        E2, e2 = np.repeat(D2, w2[s], axis=0), np.repeat(f2, w2[s])
        ref_a = clustering._autocorr(clustering.tree(E1, e1), r1, radii)
        assert np.array_equal(auto[s].dd, ref_a.dd) and np.array_equal(auto[s].dr, ref_a.dr)
        assert np.array_equal(auto[s].rr, ref_a.rr)
        for (a, b) in zip(auto[s].count_sizes, ref_a.count_sizes):
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.allclose(auto[s].estimate(), ref_a.estimate(), equal_nan=True)
        ref_c = clustering._crosscorr(clustering.tree(E1, e1), clustering.tree(E2, e2), r1, r2, radii)
        for k in ["d1d2", "d1r2", "d2r1", "r1r2"]:
            assert np.array_equal(getattr(cross[s], k), getattr(ref_c, k)), (s, k)
        assert np.allclose(cross[s].jackknife_error(), ref_c.jackknife_error(), equal_nan=True)
        # and the unweighted path itself against the oracle once
        if s == 0:
            oe = oracle.OracleTree(E1, e1, nf)
            assert np.array_equal(ref_a.dd, np.diff(oe.cumulative(E1, e1, nf, radii), axis=0))


def test_large_radii_whole_subtrees():
    """Radii up to 2 (every pair of the unit cube): query-weighted jobs take
    whole subtrees as bulk items, tree-weighted jobs tiles; every set equals
    the unweighted count of its expanded catalogues, and at radius 2 the counts
    are the outer product of the per-field multiplicity sums."""
    from corruscant_amd import clustering
    rng = np.random.default_rng(61)
    T = rng.random((40000, 3))
    Q = rng.random((12000, 3))
    ft, fq = (T[:, 0] * 3).astype(np.int32), (Q[:, 1] * 3).astype(np.int32)
    t, q = clustering.tree(T, ft), clustering.tree(Q, fq)
    radii = np.array([0.02, 0.1, 0.4, 2.0])
    S = 3
    wt = rng.poisson(1.0, (S, T.shape[0])).astype(np.uint8)
    wq = rng.poisson(1.0, (S, Q.shape[0])).astype(np.uint8)
    print("weighted large radii", flush=True)
    cq, ctq = clustering._count_weighted([(t, q, None, wq), (t, q, wt, wq)], radii, S)
    for s in range(S):
        Qs, fqs = np.repeat(Q, wq[s], axis=0), np.repeat(fq, wq[s])
        Ts, fts = np.repeat(T, wt[s], axis=0), np.repeat(ft, wt[s])
        qs_tree = clustering.tree(Qs, fqs)
        ref_q = clustering._count_jobs([(t, qs_tree)], radii)[0]
        ref_tq = clustering._count_jobs([(clustering.tree(Ts, fts), qs_tree)], radii)[0]
        assert np.array_equal(cq[s], ref_q), s
        assert np.array_equal(ctq[s], ref_tq), s
        sq = np.bincount(fq, weights=wq[s], minlength=3).astype(np.int64)
        st = np.bincount(ft, weights=wt[s], minlength=3).astype(np.int64)
        assert np.array_equal(cq[s][-1], np.outer(sq, np.bincount(ft, minlength=3)))
        assert np.array_equal(ctq[s][-1], np.outer(sq, st))


@pytest.mark.parametrize("fields", ["slabs", "per_point"])
def test_query_weights_past_sixteen_bit_histograms(fields):
    """Every query point has 200,000 tree neighbours under the largest radius,
    so the query-weighted kernel's 16-bit per-lane histograms must be flushed
    several times per query tile (before any lane's total could pass 65,535).
    Thirty radii closer than a tile's extent keep every pair in tile items
    (no whole-subtree bulk).  Slab fields give field-pure tiles (the quad
    loop), per-point fields mixed tiles (the per-point loop).  Checks: all-one
    multiplicities give the unweighted counts, doubled multiplicities double
    them, and set a+b equals set a plus set b."""
    from corruscant_amd import clustering
    rng = np.random.default_rng(23)
    nf = 3
    R, Q = rng.random((200_000, 3)), rng.random((1500, 3))
    if fields == "slabs":
        fr, fq = _fields(R, nf), _fields(Q, nf)
    else:
        fr, fq = rng.integers(0, nf, R.shape[0]).astype(np.int32), rng.integers(0, nf, Q.shape[0]).astype(np.int32)
    tr, tq = clustering.tree(R, fr), clustering.tree(Q, fq)
    radii = np.linspace(0.06, 1.8, 30)
    a = rng.integers(0, 4, Q.shape[0]).astype(np.uint8)
    b = rng.integers(0, 3, Q.shape[0]).astype(np.uint8)
    w = np.stack([np.ones_like(a), a, 2 * a, b, a + b])
    print("16-bit histograms, %s fields" % fields, flush=True)
    (c,) = clustering._count_weighted([(tr, tq, None, w)], radii, w.shape[0])
    (u,) = clustering._count_jobs([(tr, tq)], radii)
    assert u[-1].sum() == R.shape[0] * Q.shape[0]  # the largest radius holds every pair
    assert np.array_equal(c[0], u)
    assert np.array_equal(c[2], 2 * c[1])
    assert np.array_equal(c[4], c[1] + c[3])
