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
@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_weighted_jobs_match_expanded_catalogues(oracle, nsets, prec):
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
    out, fsz = clustering._count_weighted(jobs, radii, nsets, sizes=True, prec=prec)
    assert clustering.last_stats["evaluated_pairs"] > 0
    if prec == "fp32_guarded":  # the query-weighted jobs took the matrix-core loop
        assert clustering.last_stats["fp32_pairs"] > 0
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


@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_many_sets_and_unit_weights(oracle, prec):
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
    dd, dr = clustering._count_weighted([(td, td, wd, wd), (tr, td, None, wd)], radii, nsets, prec=prec)
    (dd1, dr1) = clustering._count_jobs([(td, td), (tr, td)], radii)
    assert np.array_equal(dd[3], dd1) and np.array_equal(dr[3], dr1)
    one = np.zeros(D.shape[0], np.int32)
    for s in [0, 5, 37, 69]:
        assert np.array_equal(dd[s], _expanded(oracle, D, one, D, one, 1, wd[s], wd[s], radii)), s
        assert np.array_equal(dr[s], _expanded(oracle, R, np.zeros(R.shape[0], np.int32), D, one, 1, None, wd[s],
                                               radii)), s


@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_tree_multiplicities_past_127(oracle, prec):
    """The tree-weighted kernel takes its multiplicities as signed int8
    matrix-core operands while all are below 128 (127 included) and as bf16
    once any tree multiplicity of the call reaches 128: the same job gives the
    same counts alone (int8) and beside a job with multiplicities up to 255
    (bf16), and both calls equal the oracle's counts."""
    from corruscant_amd import clustering
    rng = np.random.default_rng(31)
    T, Q = rng.random((2000, 3)), rng.random((700, 3))
    tt, tq = clustering.tree(T), clustering.tree(Q)
    w = rng.poisson(1.0, (3, T.shape[0])).astype(np.uint8)
    w[1, ::5] = 127
    big = w.copy()
    big[2, ::11] = 255
    big[0, 3] = 128
    radii = np.array([0.02, 0.05, 0.1, 0.2])
    (a,) = clustering._count_weighted([(tt, tq, w, None)], radii, 3, prec=prec)
    b, c = clustering._count_weighted([(tt, tq, w, None), (tt, tq, big, None)], radii, 3, prec=prec)
    assert np.array_equal(a, b)
    ft, fq = np.zeros(T.shape[0], np.int32), np.zeros(Q.shape[0], np.int32)
    for s in range(3):
        assert np.array_equal(a[s], _expanded(oracle, T, ft, Q, fq, 1, w[s], None, radii)), s
        assert np.array_equal(c[s], _expanded(oracle, T, ft, Q, fq, 1, big[s], None, radii)), s


@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_query_weights_against_tree_with_duplicates(oracle, prec):
    """Unweighted tree side with duplicate points (node weights > 1) and
    duplicate query points with their own multiplicities; lattice radii put
    many pairs exactly on a radius (the FP32 form's guard band)."""
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
    (c,) = clustering._count_weighted([(tt, tq, None, w)], radii, 3, prec=prec)
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


@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_autocorr_and_crosscorr_weighted(oracle, prec):
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
    auto = clustering.autocorr_weighted(t1, r1, radii, w1, precision=prec)
    cross = clustering.crosscorr_weighted(t1, t2, r1, r2, radii, w1, w2, precision=prec)
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


@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_large_radii_whole_subtrees(prec):
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
    cq, ctq = clustering._count_weighted([(t, q, None, wq), (t, q, wt, wq)], radii, S, prec=prec)
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
@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_query_weights_past_sixteen_bit_histograms(fields, prec):
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
    (c,) = clustering._count_weighted([(tr, tq, None, w)], radii, w.shape[0], prec=prec)
    (u,) = clustering._count_jobs([(tr, tq)], radii)
    assert u[-1].sum() == R.shape[0] * Q.shape[0]  # the largest radius holds every pair
    assert np.array_equal(c[0], u)
    assert np.array_equal(c[2], 2 * c[1])
    assert np.array_equal(c[4], c[1] + c[3])


@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_bootstrap_workload_at_200_sets(prec):
    """The bootstrap workload's sizes and set count: A 1M clustered, B 4M,
    R 10M, 16 fields, 20 radii, 200 weight sets in one call (D1D2 weighted on
    both sides with B as the tree, D1R2 by A's sets, D2R1 by B's sets), so the
    tree-weighted kernel runs over seven blocks of 32 sets.  Set 0 has all
    multiplicities one and equals the unweighted counts; counts are linear in
    each side's multiplicities (sets 3 = 1 + 2 and 199 = 197 + 198 in the tree
    weights with the query weights held, 6 = 4 + 5 in the query weights), and
    the per-set field sizes add the same way."""
    import bench
    from corruscant_amd import clustering as C
    print("c5 at 200 sets: generating", flush=True)
    A, fa, B, fb, R, fr, _, _, radii = bench.c5_workload(1)
    ta, tb, tr = C.trees([(A, fa), (B, fb), (R, fr)])
    assert tb.size == B.shape[0]
    S = 200
    rng = np.random.default_rng(78)
    wq = rng.integers(0, 3, (S, A.shape[0]), dtype=np.uint8)
    wt = rng.integers(0, 3, (S, B.shape[0]), dtype=np.uint8)
    wq[0:4] = 1
    wt[0] = 1
    wt[3] = wt[1] + wt[2]
    wt[4:7] = 1
    wq[6] = wq[4] + wq[5]
    wq[198] = wq[197]
    wq[199] = wq[197]
    wt[199] = wt[197] + wt[198]
    print("c5 at 200 sets: weighted call", flush=True)
    (d12, d1r2, d2r1), fsz = C._count_weighted([(tb, ta, wt, wq), (tr, ta, None, wq), (tr, tb, None, wt)],
                                                radii, S, sizes=True, prec=prec)
    u12, u1r2, u2r1 = C._count_jobs([(tb, ta), (tr, ta), (tr, tb)], radii)
    assert np.array_equal(d12[0], u12) and np.array_equal(d1r2[0], u1r2) and np.array_equal(d2r1[0], u2r1)
    for c in (d12, d2r1):  # linear in the tree side's (B's) multiplicities
        assert np.array_equal(c[3], c[1] + c[2]) and np.array_equal(c[199], c[197] + c[198])
    assert np.array_equal(d12[6], d12[4] + d12[5]) and np.array_equal(d1r2[6], d1r2[4] + d1r2[5])
    assert np.array_equal(d1r2[199], d1r2[197]) and np.array_equal(d1r2[1], u1r2)
    for c in (d12, d1r2, d2r1):
        assert np.all(np.diff(c, axis=1) >= 0)
    ts, qs = fsz[0]
    assert np.array_equal(ts[3], ts[1] + ts[2]) and np.array_equal(qs[6], qs[4] + qs[5])
    assert np.array_equal(ts[0], np.bincount(fb, minlength=16)) and np.array_equal(qs[0], np.bincount(fa, minlength=16))
    assert np.array_equal(ts[150], np.bincount(fb, weights=wt[150], minlength=16).astype(np.int64))


@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_200_sets_match_expanded_catalogues(oracle, prec):
    """200 weight sets in one call on a reduced bootstrap workload (A 3,000
    clustered, B 5,000, R 8,000, 16 fields, 20 radii): sets 0, 77 and 199
    equal the oracle counted on np.repeat-expanded catalogues."""
    from corruscant_amd import clustering as C
    from tests import cases
    rng = np.random.default_rng(79)
    A = cases.clustered(1500, 30, 50, 0.005, 79)
    B, R = rng.random((5000, 3)), rng.random((8000, 3))
    f = lambda X: (4 * np.floor(4 * X[:, 0]) + np.floor(4 * X[:, 1])).astype(np.int32)
    fa, fb, fr = f(A), f(B), f(R)
    ta, tb, tr = C.tree(A, fa), C.tree(B, fb), C.tree(R, fr)
    S = 200
    wa = rng.poisson(1.0, (S, A.shape[0])).astype(np.uint8)
    wb = rng.poisson(1.0, (S, B.shape[0])).astype(np.uint8)
    radii = np.logspace(-3, -1, 20)
    d12, d1r2, d2r1 = C._count_weighted([(tb, ta, wb, wa), (tr, ta, None, wa), (tr, tb, None, wb)], radii, S,
                                        prec=prec)
    for s in (0, 77, 199):
        assert np.array_equal(d12[s], _expanded(oracle, B, fb, A, fa, 16, wb[s], wa[s], radii)), s
        assert np.array_equal(d1r2[s], _expanded(oracle, R, fr, A, fa, 16, None, wa[s], radii)), s
        assert np.array_equal(d2r1[s], _expanded(oracle, R, fr, B, fb, 16, None, wb[s], radii)), s


@pytest.mark.parametrize("case", ["one tile", "few tiles", "many tiles", "four fields", "twelve radii"])
def test_tree_weighted_fp32_masks_match_the_oracle(oracle, case):
    """Tree-weighted jobs (the slot-mask kernel) in the FP32 form against the
    oracle, from a single tile up to runs of tiles with several fields: every
    radius of a pass, including the first one compared after the matrix
    cores' write (the compare reads the MFMA's result registers, which must
    hold the finished values)."""
    from corruscant_amd import clustering
    rng = np.random.default_rng(5)
    nt, nq, nf, radii = {"one tile": (50, 50, 1, np.array([0.05, 0.1, 0.3])),
                         "few tiles": (300, 64, 1, np.array([0.05, 0.1, 0.3])),
                         "many tiles": (3000, 500, 1, np.array([0.02, 0.05, 0.1])),
                         "four fields": (3000, 500, 4, np.array([0.02, 0.05, 0.1])),
                         "twelve radii": (3000, 2500, 4, np.logspace(-2.5, -0.8, 12))}[case]
    T, Q = rng.random((nt, 3)), rng.random((nq, 3))
    ft, fq = _fields(T, nf), _fields(Q, nf)
    tt, tq = clustering.tree(T, ft if nf > 1 else None), clustering.tree(Q, fq if nf > 1 else None)
    w = rng.poisson(1.0, (2, nt)).astype(np.uint8)
    w[0] = 1
    (c,) = clustering._count_weighted([(tt, tq, w, None)], radii, 2, prec="fp32_guarded")
    assert clustering.last_stats["fp32_pairs"] > 0
    for s in range(2):
        assert np.array_equal(c[s], _expanded(oracle, T, ft, Q, fq, nf, w[s], None, radii)), s


@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_all_zero_and_all_one_sets(prec):
    """A set whose multiplicities are all zero counts nothing (every pair
    type, both weighted sides); a set of ones counts what the unweighted call
    counts; a set of 127s on the tree side (the largest int8 operand) counts
    127 times that."""
    from corruscant_amd import clustering
    rng = np.random.default_rng(41)
    T, Q = rng.random((3000, 3)), rng.random((900, 3))
    tt, tq = clustering.tree(T), clustering.tree(Q)
    wt = np.stack([np.zeros(T.shape[0], np.uint8), np.ones(T.shape[0], np.uint8),
                   np.full(T.shape[0], 127, np.uint8)])
    wq = np.stack([np.zeros(Q.shape[0], np.uint8), np.ones(Q.shape[0], np.uint8), np.ones(Q.shape[0], np.uint8)])
    radii = np.array([0.03, 0.07, 0.15])
    a, b = clustering._count_weighted([(tt, tq, wt, wq), (tt, tq, None, wq)], radii, 3, prec=prec)
    (plain,) = clustering._count_jobs([(tt, tq)], radii)
    assert not a[0].any() and not b[0].any()
    assert np.array_equal(a[1], plain) and np.array_equal(b[1], plain)
    assert np.array_equal(a[2], 127 * plain) and np.array_equal(b[2], plain)
