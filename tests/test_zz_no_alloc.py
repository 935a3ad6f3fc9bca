"""pair_count_jobs_noalloc (the entry point that allocates nothing) against
pair_count_jobs, called through the C ABI: two calls in a row, the second
with a smaller input, in both precision forms, give identical counts; a call
larger than the fixed scratch fails with a clear error and leaves the next
call working."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(pairs, radii, prec):
    from corruscant_amd import clustering as C
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    outs = [np.full((radii.size, q.N_fields, t.N_fields), -7, dtype=np.int64) for t, q in pairs]
    jobs = (C.kd_job * len(pairs))()
    for i, (t, q) in enumerate(pairs):
        jobs[i].tree = t.ctree
        jobs[i].query = q.ctree
        jobs[i].num_fields_query = q.N_fields
        jobs[i].counts = outs[i].ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
    st = C.kd_stats()
    rc = C.kdlib.pair_count_jobs_noalloc(jobs, len(pairs), C._dp(radii), radii.size, C._PRECISIONS[prec],
                                         ctypes.byref(st))
    return rc, outs, st


def _setup():
    from corruscant_amd import clustering as C
    f = C.kdlib.pair_count_jobs_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.POINTER(C.kd_job), ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                  ctypes.c_int, ctypes.POINTER(C.kd_stats)]
    return C


def _catalogue(n, seed, nf):
    rng = np.random.default_rng(seed)
    X = rng.random((n, 3))
    return X, np.minimum((X[:, 0] * nf).astype(np.int32), nf - 1)


@pytest.mark.parametrize("prec", ["fp64", "fp32_guarded"])
def test_noalloc_equals_pair_count_jobs(prec):
    C = _setup()
    radii = np.logspace(-2.5, -1.0, 6)
    for n_d, n_r, seed in ((4000, 80000, 3), (900, 7000, 4)):  # the second call is smaller
        print("noalloc %s: %d / %d points" % (prec, n_d, n_r), flush=True)
        D, fd = _catalogue(n_d, seed, 4)
        R, fr = _catalogue(n_r, seed + 100, 4)
        td, tr = C.tree(D, fd), C.tree(R, fr)
        pairs = [(td, td), (tr, td), (tr, tr)]
        ref = C._count_jobs(pairs, radii, prec)
        rc, got, st = _noalloc(pairs, radii, prec)
        assert rc == 0, C._err()
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
        assert st.evaluated_pairs == C.last_stats["evaluated_pairs"] and st.work_items > 0
        assert st.enumerate_ms == 0.0 and st.count_ms == 0.0


def test_noalloc_refuses_a_call_beyond_its_scratch():
    """A call whose histograms outgrow the fixed scratch (250 fields and 64
    radii: 62,500 field slots x 129 int64 per plane, two planes per
    auto-correlation job, three jobs: 369 MiB) fails after its enumeration with
    an error naming the scratch, and a small call afterwards still matches
    pair_count_jobs."""
    C = _setup()
    rng = np.random.default_rng(5)
    X = rng.random((3000, 3))
    f = (np.arange(3000) % 250).astype(np.int32)
    t = C.tree(X, f)
    assert t.N_fields == 250
    rc, _, _ = _noalloc([(t, t)] * 3, np.linspace(0.01, 0.3, 64), "fp64")
    assert rc == -1 and "scratch" in C._err(), C._err()
    D, fd = _catalogue(3000, 6, 4)
    td = C.tree(D, fd)
    radii = np.array([0.01, 0.05, 0.2])
    rc, got, _ = _noalloc([(td, td)], radii, "fp64")
    assert rc == 0, C._err()
    assert np.array_equal(got[0], C._count_jobs([(td, td)], radii, "fp64")[0])
