"""Parity at the benchmark sizes, against the oracle.

1M-data / 10M-random 3-D workload (the bench's default) and the 2M / 20M
angular workload:

* trees: the GPU node arrays and heap permutations of all four catalogues
  equal the oracle's restatement of tree_construct (src/kdbuild.c:451-537)
  record for record;
* counts: 20,000 seeded query rows per pair type (DD, DR, RR) counted by the
  GPU against the full trees, at the workload's own radii only (so the trim
  pass cuts items, which a radius beyond every pair would prevent), equal the
  oracle's traversal (src/kdquery.c:62-113) in both precision forms;
* the full auto-correlation step: DD through the symmetric (mirrored) path
  equals DD of the same points as a plain query set, which pins the mirror
  path to the row-checked direct path at full size.

The coordinate conversion (radec2cart64, src/coordmath.c:38-51) is correctly
rounded on the GPU and differs from glibc in ~0.3% of components by an ulp;
test_angular_counts_with_reference_coordinates counts the angular workload
from both coordinate sets and requires every count cell to agree."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROWS = 20000


def _rows(n, seed):
    return np.sort(np.random.default_rng(seed).choice(n, min(ROWS, n), replace=False))


def _check_rows(oracle, C, gtree, otree, Q, qf, nf, radii, seed, label):
    """GPU counts of sampled query rows against the full tree, both forms,
    equal the oracle's; returns the trimmed-item count of the FP32 call (the
    trim pass serves the FP32 form's matrix-core loop)."""
    idx = _rows(Q.shape[0], seed)
    ref = otree.cumulative(Q[idx], qf[idx], nf, radii)
    qs = C.query_set(Q[idx], qf[idx], nf)
    trimmed = None
    for prec in ("fp64", "fp32_guarded"):
        got = C._count_jobs([(gtree, qs)], radii, prec)[0]
        if prec == "fp32_guarded":
            trimmed = C.last_stats["trimmed_items"]
        bad = np.argwhere(got != ref)
        assert bad.size == 0, "%s %s: %d cells differ, first %s" % (label, prec, len(bad), bad[:3].tolist())
    return trimmed


@pytest.fixture(scope="module")
def c2(oracle):
    import bench
    from corruscant_amd import clustering
    print("c2 parity: generating 1M/10M", flush=True)
    D, fd, R, fr, radii = bench.c2_workload(1.0)
    td, tr = clustering.trees([(D, fd), (R, fr)])
    print("c2 parity: oracle trees", flush=True)
    od, orr = oracle.OracleTree(D, fd, 16), oracle.OracleTree(R, fr, 16)
    return D, fd, R, fr, radii, td, tr, od, orr


def test_c2_trees_equal_the_oracle(oracle, c2):
    D, fd, R, fr, radii, td, tr, od, orr = c2
    for g, o, label in ((td, od, "1M data"), (tr, orr, "10M randoms")):
        assert g.size == o.size and g.ctree.memsize == o.memsize, label
        assert oracle.nodes_equal(g.node_array(), o.nodes), label
        perm, ino = g.heap_permutation()
        assert np.array_equal(perm, o.perm) and np.array_equal(ino, o.inorder), label


def test_c2_sampled_rows_equal_the_oracle(oracle, c2):
    from corruscant_amd import clustering as C
    D, fd, R, fr, radii, td, tr, od, orr = c2
    trimmed = []
    for k, (gt, ot, Q, qf, label) in enumerate(((td, od, D, fd, "DD"), (tr, orr, D, fd, "DR"),
                                               (tr, orr, R, fr, "RR"))):
        print("c2 parity: rows of %s" % label, flush=True)
        trimmed.append(_check_rows(oracle, C, gt, ot, Q, qf, 16, radii, 100 + k, label))
    assert all(t > 0 for t in trimmed), trimmed  # the trim pass ran on every pair type


def test_c2_mirrored_dd_equals_direct_dd(c2):
    """The step's DD (each unordered tile pair once, mirrored) equals DD of
    the same points counted as a query set (every ordered pair evaluated),
    at the workload's radii in both forms; RR and DR keep their symmetry and
    monotonicity."""
    from corruscant_amd import clustering as C
    D, fd, R, fr, radii, td, tr, od, orr = c2
    print("c2 parity: full step", flush=True)
    dd, dr, rr = C._count_jobs([(td, td), (tr, td), (tr, tr)], radii, "fp64")
    qd = C.query_set(D, fd, 16)
    direct = C._count_jobs([(td, qd)], radii, "fp64")[0]
    assert np.array_equal(dd, direct)
    dd32, dr32, rr32 = C._count_jobs([(td, td), (tr, td), (tr, tr)], radii, "fp32_guarded")
    assert C.last_stats["trimmed_items"] > 0
    assert np.array_equal(dd, dd32) and np.array_equal(dr, dr32) and np.array_equal(rr, rr32)
    for c in (dd, rr):
        assert np.array_equal(c, c.transpose(0, 2, 1))
    for c in (dd, dr, rr):
        assert np.all(np.diff(c, axis=0) >= 0)


@pytest.fixture(scope="module")
def c3(oracle):
    import bench
    from corruscant_amd import clustering, fields
    from corruscant_amd.twopoint import angular
    print("c3 parity: generating 2M/20M", flush=True)
    dra, ddec, _, rra, rdec, _, deg = bench.c3_workload()
    t = fields.template(64, ra=dra, dec=ddec)
    fd = t.create_fields(min_ratio=0.8).astype(np.int32)
    fr = t.assign_fields(ra=rra, dec=rdec).astype(np.int32)
    D, R = angular.cartesian(dra, ddec), angular.cartesian(rra, rdec)
    td, tr = clustering.trees([(D, fd), (R, fr)])
    chord = 2.0 * np.sin(deg * (np.pi / 180.0) / 2.0)
    return dra, ddec, rra, rdec, D, fd, R, fr, chord, td, tr


def test_c3_trees_and_sampled_rows_equal_the_oracle(oracle, c3):
    from corruscant_amd import clustering as C
    dra, ddec, rra, rdec, D, fd, R, fr, chord, td, tr = c3
    print("c3 parity: oracle trees", flush=True)
    od = oracle.OracleTree(D, fd, 64)
    assert oracle.nodes_equal(td.node_array(), od.nodes), "2M data"
    assert np.array_equal(td.heap_permutation()[0], od.perm)
    orr = oracle.OracleTree(R, fr, 64)
    assert oracle.nodes_equal(tr.node_array(), orr.nodes), "20M randoms"
    assert np.array_equal(tr.heap_permutation()[0], orr.perm)
    for k, (gt, ot, Q, qf, label) in enumerate(((td, od, D, fd, "DD"), (tr, orr, D, fd, "DR"),
                                               (tr, orr, R, fr, "RR"))):
        print("c3 parity: rows of %s" % label, flush=True)
        _check_rows(oracle, C, gt, ot, Q, qf, 64, chord, 200 + k, label)


def test_angular_counts_with_reference_coordinates(oracle, c3):
    """The reference converts RA/DEC with glibc's sin/cos (src/coordmath.c:38-51),
    which is not correctly rounded; the GPU conversion is.  The angular
    workload counted from the oracle's glibc coordinates and from the GPU's
    gives the same DD, DR and RR in every (radius, field, field) cell."""
    from corruscant_amd import clustering as C
    dra, ddec, rra, rdec, D, fd, R, fr, chord, td, tr = c3
    print("c3 coordinates: host libm conversion", flush=True)
    Dh, Rh = oracle.radec2cart64(dra, ddec), oracle.radec2cart64(rra, rdec)
    differ = int(np.count_nonzero(Dh != D)) + int(np.count_nonzero(Rh != R))
    print("c3 coordinates: %d of %d components differ from glibc's" % (differ, D.size + R.size), flush=True)
    assert differ > 0  # the two conversions do differ at the ulp level
    th, trh = C.trees([(Dh, fd), (Rh, fr)])
    got = C._count_jobs([(td, td), (tr, td), (tr, tr)], chord, "fp64")
    ref = C._count_jobs([(th, th), (trh, th), (trh, trh)], chord, "fp64")
    cells = [int(np.count_nonzero(a != b)) for a, b in zip(got, ref)]
    print("c3 coordinates: differing count cells DD/DR/RR = %s" % cells, flush=True)
    assert cells == [0, 0, 0]
