"""The bench workloads at full size, checked through properties that need no
oracle.  The 1M-data / 10M-random 3-D workload (seeded, the bench's default):
the FP64 and FP32 guarded forms agree bit for bit, auto-correlation matrices
are symmetric in the fields, counts never decrease with the radius, a radius
of 2 (beyond the unit cube's diagonal) counts every pair (the outer product of
the field sizes), and a radius of 1e-150 (r^2 = 1e-300 > 0) counts only
coincident points: each point with itself in DD and RR, nothing in DR.  The
weighted and angular workloads: see their tests."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_full_size_properties():
    import bench
    from corruscant_amd import clustering
    print("full size: generating 1M/10M", flush=True)
    D, fd, R, fr, radii = bench.c2_workload(1.0)
    td, tr = clustering.tree(D, fd), clustering.tree(R, fr)
    nf = 16
    sd, sr = np.bincount(fd, minlength=nf), np.bincount(fr, minlength=nf)
    # the workload has no coincident points (checked through the tree sizes)
    assert td.size == D.shape[0] and tr.size == R.shape[0]
    rad = np.concatenate([[1e-150], radii, [2.0]])
    pairs = [(td, td), (tr, td), (tr, tr)]
    print("full size: fp64", flush=True)
    c64 = clustering._count_jobs(pairs, rad, "fp64")
    print("full size: fp32_guarded", flush=True)
    c32 = clustering._count_jobs(pairs, rad, "fp32_guarded")
    for a, b in zip(c64, c32):
        assert np.array_equal(a, b)
    dd, dr, rr = c64
    for c in (dd, rr):
        assert np.array_equal(c, c.transpose(0, 2, 1))
    for c in c64:
        assert np.all(np.diff(c, axis=0) >= 0)
    assert np.array_equal(dd[-1], np.outer(sd, sd))
    assert np.array_equal(dr[-1], np.outer(sd, sr))
    assert np.array_equal(rr[-1], np.outer(sr, sr))
    assert np.array_equal(dd[0], np.diag(sd)) and np.array_equal(rr[0], np.diag(sr))
    assert not dr[0].any()


def test_full_size_weighted_linearity():
    """Weighted counts at the bootstrap workload's sizes (A 1M clustered, B 4M,
    R 10M, 20 radii, 16 fields): all-one multiplicities give the unweighted
    counts, and counts are linear in either side's multiplicities (set a+b =
    set a + set b, the other side held at one)."""
    import bench
    from corruscant_amd import clustering
    print("full size weighted: generating", flush=True)
    A, fa, B, fb, R, fr, _, _, radii = bench.c5_workload(1)
    ta, tb, tr = clustering.trees([(A, fa), (B, fb), (R, fr)])
    assert tb.size == B.shape[0]  # tree-side multiplicities need a tree without duplicates
    rng = np.random.default_rng(77)
    aA, bA = rng.poisson(1.0, (2, A.shape[0])).astype(np.uint8)
    aB, bB = rng.poisson(1.0, (2, B.shape[0])).astype(np.uint8)
    oA, oB = np.ones(A.shape[0], np.uint8), np.ones(B.shape[0], np.uint8)
    wt = np.stack([oB, aB, bB, aB + bB, oB, oB, oB])
    wq = np.stack([oA, oA, oA, oA, aA, bA, aA + bA])
    print("full size weighted: 7 sets", flush=True)
    (d12, d1r2), fsz = clustering._count_weighted([(tb, ta, wt, wq), (tr, ta, None, wq)], radii, 7, sizes=True)
    u12, u1r2 = clustering._count_jobs([(tb, ta), (tr, ta)], radii)
    assert np.array_equal(d12[0], u12) and np.array_equal(d1r2[0], u1r2)
    assert np.array_equal(d12[3], d12[1] + d12[2])
    assert np.array_equal(d12[6], d12[4] + d12[5])
    assert np.array_equal(d1r2[6], d1r2[4] + d1r2[5])
    assert np.array_equal(d1r2[1], u1r2) and np.array_equal(d1r2[3], u1r2)
    ts, qs = fsz[0]
    assert np.array_equal(ts[3], ts[1] + ts[2]) and np.array_equal(qs[6], qs[4] + qs[5])
    assert np.array_equal(ts[0], np.bincount(fb, minlength=16)) and np.array_equal(qs[0], np.bincount(fa, minlength=16))


def test_full_size_angular_with_automatic_fields():
    """The angular workload (2M data, 20M randoms on a 60 x 60 deg patch, 64
    fields from the automatic division, 15 angles as chord radii): FP64 and
    FP32 guarded agree, DD and RR are symmetric, counts never decrease, and a
    chord of 2 (antipodal) counts every pair."""
    import bench
    from corruscant_amd import clustering, fields
    from corruscant_amd.twopoint import angular
    print("full size angular: generating", flush=True)
    dra, ddec, _, rra, rdec, _, deg = bench.c3_workload()
    t = fields.template(64, ra=dra, dec=ddec)
    fd = t.create_fields(min_ratio=0.8).astype(np.int32)
    fr = t.assign_fields(ra=rra, dec=rdec).astype(np.int32)
    td, tr = clustering.trees([(angular.cartesian(dra, ddec), fd), (angular.cartesian(rra, rdec), fr)])
    chord = 2.0 * np.sin(deg * (np.pi / 180.0) / 2.0)
    rad = np.concatenate([chord, [2.0]])
    pairs = [(td, td), (tr, td), (tr, tr)]
    print("full size angular: fp64 and fp32_guarded", flush=True)
    c64 = clustering._count_jobs(pairs, rad, "fp64")
    c32 = clustering._count_jobs(pairs, rad, "fp32_guarded")
    sd, sr = np.bincount(fd, minlength=64), np.bincount(fr, minlength=64)
    for a, b in zip(c64, c32):
        assert np.array_equal(a, b)
        assert np.all(np.diff(a, axis=0) >= 0)
    dd, dr, rr = c64
    assert np.array_equal(dd, dd.transpose(0, 2, 1)) and np.array_equal(rr, rr.transpose(0, 2, 1))
    assert np.array_equal(dd[-1], np.outer(sd, sd)) and np.array_equal(dr[-1], np.outer(sd, sr))
    assert np.array_equal(rr[-1], np.outer(sr, sr))
