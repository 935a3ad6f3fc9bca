"""Library calls from several host threads: each leases its own HIP stream, so
tree builds and pair counts overlap on the GPU.  Results must equal those of
the same calls made one after another (and the oracle's)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


def test_concurrent_tree_builds_match_sequential(oracle):
    from corruscant_amd import clustering
    from oracle.oracle import nodes_equal
    rng = np.random.default_rng(40)
    cats = []
    for k in range(24):
        n = int(np.exp(rng.uniform(np.log(500), np.log(60000))))
        X = rng.random((n, 3))
        cats.append((X, (X[:, 0] * 4).astype(np.int32)))
    print("concurrent builds of %d catalogues" % len(cats), flush=True)
    par = clustering.trees(cats, max_workers=8)
    for (X, f), t in zip(cats, par):
        s = clustering.tree(X, f)
        assert t.size == s.size
        assert np.array_equal(t.node_array().view(np.uint8), s.node_array().view(np.uint8))
        pa, ia = t.heap_permutation()
        pb, ib = s.heap_permutation()
        assert np.array_equal(pa, pb) and np.array_equal(ia, ib)
    X, f = cats[3]
    assert nodes_equal(par[3].node_array(), oracle.OracleTree(X, f, 4).nodes)


def test_concurrent_pair_counts_match_sequential():
    from corruscant_amd import clustering
    radii = np.logspace(-2.5, -1, 7)
    cats = [cases.cube(n, seed=50 + n % 97, nfields=4) for n in (3000, 9000, 20000, 40000)]
    ts = clustering.trees(cats)
    jobs = [(a, b) for a in ts for b in ts]
    seq = [clustering._count_jobs([j], radii)[0] for j in jobs]
    print("concurrent counts of %d jobs" % len(jobs), flush=True)
    for prec in ("fp64", "fp32_guarded"):
        with ThreadPoolExecutor(max_workers=6) as ex:
            par = list(ex.map(lambda j: clustering._count_jobs([j], radii, prec)[0], jobs))
        for a, b in zip(seq, par):
            assert np.array_equal(a, b), prec
