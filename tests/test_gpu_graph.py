"""Pair-count plans: the batched count recorded once into a HIP graph and
replayed (pair_count_plan / pair_count_plan_run through clustering.count_plan),
and the autotuner that times such replays.

Every replay must equal the eager call (pair_count_jobs) on the same jobs, and
the eager call is pinned to the oracle here too.  Replays are interleaved with
other library work on the device (eager counts of other trees, tree builds),
so a replay that depended on what an earlier one left in its buffers, or on
memory state a clear recorded into the graph failed to reset, would differ."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RADII = np.logspace(np.log10(2e-3), np.log10(2e-1), 12)


def _catalogues(seed, njobs, lo, hi, nf=8):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(njobs):
        n = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        D, R = rng.random((n, 3)), rng.random((5 * n, 3))
        out.append((D, np.minimum((D[:, 0] * nf).astype(np.int32), nf - 1), R,
                    np.minimum((R[:, 0] * nf).astype(np.int32), nf - 1)))
    return out


def _pairs(C, cats):
    flat = C.trees([c for D, fd, R, fr in cats for c in ((D, fd), (R, fr))])
    return list(zip(flat[0::2], flat[1::2]))


def _jobs(pairs):
    return [j for d, r in pairs for j in ((d, d), (r, d), (r, r))]


@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_replays_equal_the_eager_call_and_the_oracle(oracle, prec):
    from corruscant_amd import clustering as C
    cats = _catalogues(31, 10, 300, 5000)
    pairs = _pairs(C, cats)
    jobs = _jobs(pairs)
    eager = C._count_jobs(jobs, RADII, prec)
    # the eager call against the oracle on the two smallest jobs
    for k in np.argsort([D.shape[0] for D, _, _, _ in cats])[:2]:
        D, fd, R, fr = cats[k]
        od, orr = oracle.OracleTree(D, fd, 8), oracle.OracleTree(R, fr, 8)
        assert np.array_equal(eager[3 * k], od.cumulative(D, fd, 8, RADII))
        assert np.array_equal(eager[3 * k + 1], orr.cumulative(D, fd, 8, RADII))
        assert np.array_equal(eager[3 * k + 2], orr.cumulative(R, fr, 8, RADII))
    plan = C.count_plan(jobs, RADII, prec)
    assert plan.items > 0
    other = _pairs(C, _catalogues(32, 3, 2000, 4000))
    for rep in range(4):
        print("replay %d" % rep, flush=True)
        got = plan.run()
        st = plan.last_stats
        assert st["work_items"] == plan.items and st["count_ms"] > 0 and st["evaluated_pairs"] > 0
        for j, (a, b) in enumerate(zip(got, eager)):
            assert np.array_equal(a, b), (rep, j)
        # eager work on the device between replays
        C._count_jobs(_jobs(other), RADII[::2], prec, record=False)
    plan.close()


def test_plan_statistics_match_the_eager_call():
    """A replay does the eager call's work: the same items, evaluated pairs,
    trimmed items and (FP32 form) FP32 and rechecked pairs."""
    from corruscant_amd import clustering as C
    pairs = _pairs(C, _catalogues(33, 6, 1000, 20000))
    jobs = _jobs(pairs)
    for prec in ("fp64", "fp32_guarded"):
        C._count_jobs(jobs, RADII, prec)
        want = dict(C.last_stats)
        plan = C.count_plan(jobs, RADII, prec)
        plan.run()
        got = plan.last_stats
        for k in ("work_items", "evaluated_pairs", "trimmed_items", "fp32_pairs", "rechecked_pairs"):
            assert got[k] == want[k], (prec, k, got[k], want[k])
        plan.close()


def test_tuning_changes_nothing_but_the_time():
    from corruscant_amd import clustering as C
    pairs = _pairs(C, _catalogues(34, 8, 500, 8000))
    jobs = _jobs(pairs)
    eager = C._count_jobs(jobs, RADII, "fp64")
    for tuning in ({"blocks_per_cu": 1}, {"blocks_per_cu": 16, "chunk": 1}, {"chunk": 4096},
                   {"enum_blocks_per_cu": 1}, {"blocks_per_cu": 3, "chunk": 7, "enum_blocks_per_cu": 64}):
        plan = C.count_plan(jobs, RADII, "fp64", tuning)
        got = plan.run()
        assert all(np.array_equal(a, b) for a, b in zip(got, eager)), tuning
        plan.close()
    with pytest.raises(RuntimeError, match="tuning"):
        C.count_plan(jobs, RADII, "fp64", {"chunk": 5000})


def test_autocorr_plan_equals_autocorr_batch():
    from corruscant_amd import clustering as C
    pairs = _pairs(C, _catalogues(35, 5, 1000, 10000))
    want = C.autocorr_batch(pairs, RADII, precision="fp32_guarded")
    plan = C.autocorr_plan(pairs, RADII, precision="fp32_guarded")
    for _ in range(2):
        got = plan.run()
        for a, b in zip(got, want):
            assert np.array_equal(a.dd, b.dd) and np.array_equal(a.dr, b.dr) and np.array_equal(a.rr, b.rr)
            np.testing.assert_array_equal(a.estimate(), b.estimate())
            np.testing.assert_array_equal(a.jackknife_error(), b.jackknife_error())
    plan.close()


def test_a_plan_outlives_other_trees_and_keeps_its_own():
    """Trees built and destroyed after the plan was made leave its replays
    unchanged (the plan holds its trees)."""
    from corruscant_amd import clustering as C
    pairs = _pairs(C, _catalogues(36, 4, 1000, 6000))
    jobs = _jobs(pairs)
    plan = C.count_plan(jobs, RADII, "fp64")
    first = plan.run()
    del jobs, pairs
    for seed in range(3):
        tmp = _pairs(C, _catalogues(40 + seed, 2, 3000, 9000))
        C._count_jobs(_jobs(tmp), RADII, "fp64", record=False)
        del tmp
        again = plan.run()
        assert all(np.array_equal(a, b) for a, b in zip(again, first))
    plan.close()


def test_autotuner_picks_a_candidate_and_keeps_counts():
    from corruscant_amd import autotune, clustering as C
    cats = _catalogues(37, 6, 500, 8000)
    try:
        tuned = autotune.tune_autocorr(cats, RADII, "fp64", tile_sizes=(32, 64), blocks_per_cu=(4, 8),
                                       chunks=(0, 8), enum_blocks_per_cu=(16,), reps=1)
    finally:
        C.set_tile_size(64)
    assert tuned.tile_size in (32, 64)
    assert any(t == tuned.tile_size and u == tuned.tuning and m == tuned.ms for t, u, m, n in tuned.table)
    assert len(tuned.table) >= 4
    want = C.autocorr_batch(_pairs(C, cats), RADII, precision="fp64")
    got = C.autocorr_batch(tuned.trees, RADII, precision="fp64")
    for a, b in zip(got, want):
        assert np.array_equal(a.dd, b.dd) and np.array_equal(a.dr, b.dr) and np.array_equal(a.rr, b.rr)
