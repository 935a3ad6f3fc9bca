"""Many uneven auto-correlation jobs in one GPU call (DD, DR, RR of all jobs
through one enumeration and one persistent kernel launch), per-job counts
bit-exact against the oracle run job by job."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_batch_of_uneven_jobs(oracle, prec):
    from corruscant_amd import clustering
    rng = np.random.default_rng(4)
    radii = np.logspace(np.log10(2e-3), np.log10(2e-1), 12)
    pairs, data = [], []
    for j in range(24):
        n = int(round(np.exp(rng.uniform(np.log(300), np.log(6000)))))
        D = rng.random((n, 3))
        R = rng.random((5 * n, 3))
        fd = np.minimum((D[:, 0] * 8).astype(np.int32), 7)
        fr = np.minimum((R[:, 0] * 8).astype(np.int32), 7)
        pairs.append((clustering.tree(D, fd), clustering.tree(R, fr)))
        data.append((D, fd, R, fr))
    res = clustering.autocorr_batch(pairs, radii, precision=prec)
    assert clustering.last_stats["work_items"] > 0
    for (D, fd, R, fr), r in zip(data, res):
        od, orr = oracle.OracleTree(D, fd, 8), oracle.OracleTree(R, fr, 8)
        assert np.array_equal(r.dd, np.diff(od.cumulative(D, fd, 8, radii), axis=0))
        assert np.array_equal(r.dr, np.diff(orr.cumulative(D, fd, 8, radii), axis=0))
        assert np.array_equal(r.rr, np.diff(orr.cumulative(R, fr, 8, radii), axis=0))


def test_crosscorr_orientation(oracle):
    """crosscorr queries the larger tree (clustering.py:256-263, :275-284)."""
    from corruscant_amd import clustering
    from corruscant_amd.twopoint import threedim
    rng = np.random.default_rng(9)
    A, B = rng.random((3000, 3)), rng.random((5000, 3))
    RA, RB = rng.random((9000, 3)), rng.random((7000, 3))
    f = lambda X: np.minimum((X[:, 1] * 4).astype(np.int32), 3)
    radii = np.logspace(-2, -0.7, 6)
    res = threedim.crosscorr(clustering.tree(A, f(A)), clustering.tree(B, f(B)), clustering.tree(RA, f(RA)),
                             clustering.tree(RB, f(RB)), radii)
    ob, ora, orb = oracle.OracleTree(B, f(B), 4), oracle.OracleTree(RA, f(RA), 4), oracle.OracleTree(RB, f(RB), 4)
    assert np.array_equal(res.d1d2, np.diff(ob.cumulative(A, f(A), 4, radii), axis=0))  # B larger: query A
    assert np.array_equal(res.d1r2, np.diff(orb.cumulative(A, f(A), 4, radii), axis=0))
    assert np.array_equal(res.d2r1, np.diff(ora.cumulative(B, f(B), 4, radii), axis=0))
    assert np.array_equal(res.r1r2, np.diff(ora.cumulative(RB, f(RB), 4, radii), axis=0))  # RA larger
    assert res.radii_units == "Data units"
