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


def test_batched_workload_at_config_size(oracle):
    """The batched workload as configured: 256 auto-correlation jobs with
    data sizes log-uniform in [1e3, 2e5] (randoms 5x, 8 fields, 12 radii),
    5.9e7 points in 512 trees, every DD/DR/RR in one call.  Against the
    oracle: the 8 smallest jobs in full, the 8 median and the 8 largest on
    2,000 seeded query rows per pair type (and their trees record for
    record).  Every job: the FP64 and FP32 guarded calls agree, and so do two
    replays of the FP32 guarded graph plan (the whole batch in one graph
    launch); DD and RR are symmetric in the fields, counts never decrease with
    the radius, and each
    job's counts equal its own single-job call (no cross-talk between the
    jobs of one launch) on a spread of 16 jobs."""
    import bench
    from corruscant_amd import clustering as C
    print("c4: generating 256 jobs", flush=True)
    jobs, radii = bench.c4_workload()
    built = C.trees([c for D, fd, R, fr in jobs for c in ((D, fd), (R, fr))])
    pairs = list(zip(built[0::2], built[1::2]))
    print("c4: batched calls", flush=True)
    res64 = C.autocorr_batch(pairs, radii, precision="fp64")
    res32 = C.autocorr_batch(pairs, radii, precision="fp32_guarded")
    assert C.last_stats["trimmed_items"] > 0
    for a, b in zip(res64, res32):
        assert np.array_equal(a.dd, b.dd) and np.array_equal(a.dr, b.dr) and np.array_equal(a.rr, b.rr)
    print("c4: graph plan", flush=True)
    plan = C.autocorr_plan(pairs, radii, precision="fp32_guarded")
    for _ in range(2):
        for a, b in zip(res64, plan.run()):
            assert np.array_equal(a.dd, b.dd) and np.array_equal(a.dr, b.dr) and np.array_equal(a.rr, b.rr)
    plan.close()
    for r in res64:
        assert np.array_equal(r.dd, r.dd.transpose(0, 2, 1)) and np.array_equal(r.rr, r.rr.transpose(0, 2, 1))
        assert (r.dd >= 0).all() and (r.dr >= 0).all() and (r.rr >= 0).all()
    order = np.argsort([D.shape[0] for D, _, _, _ in jobs])
    mid = len(order) // 2
    for j in order[::16]:
        (d, r), res = pairs[j], res64[j]
        dd, dr, rr = C._count_jobs([(d, d), (r, d), (r, r)], radii, "fp64")
        assert np.array_equal(np.diff(dd, axis=0), res.dd) and np.array_equal(np.diff(rr, axis=0), res.rr), j
        assert np.array_equal(np.diff(dr, axis=0), res.dr), j
    rng = np.random.default_rng(44)
    for group, sample in ((order[:8], None), (order[mid - 4:mid + 4], 2000), (order[-8:], 2000)):
        for j in group:
            D, fd, R, fr = jobs[j]
            d, r = pairs[j]
            print("c4: oracle on job %d (%d data)" % (j, D.shape[0]), flush=True)
            od, orr = oracle.OracleTree(D, fd, 8), oracle.OracleTree(R, fr, 8)
            assert oracle.nodes_equal(d.node_array(), od.nodes) and oracle.nodes_equal(r.node_array(), orr.nodes)
            for name, gt, ot, Q, qf in (("dd", d, od, D, fd), ("dr", r, orr, D, fd), ("rr", r, orr, R, fr)):
                if sample is None:
                    ref = np.diff(ot.cumulative(Q, qf, 8, radii), axis=0)
                    assert np.array_equal(getattr(res64[j], name), ref), (j, name)
                else:
                    idx = np.sort(rng.choice(Q.shape[0], sample, replace=False))
                    qs = C.query_set(Q[idx], qf[idx], 8)
                    got = C._count_jobs([(gt, qs)], radii, "fp64")[0]
                    assert np.array_equal(got, ot.cumulative(Q[idx], qf[idx], 8, radii)), (j, name)
