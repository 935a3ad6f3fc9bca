"""Tiles that hold several fields, against the oracle.  Tree tiles are
grouped by field into runs of whole quads, and query tiles whose points
belong to several fields go to a second item list counted per lane.  Fields
here are drawn per point (every tile mixed), in a checkerboard (tiles mixed
along the cell faces) or in slabs; the radius windows reach past the seven
radii of one pass, and duplicated points put weights into some tiles."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(X, kind, nf, rng):
    if kind == "random":
        return rng.integers(0, nf, X.shape[0]).astype(np.int32)
    if kind == "checker":
        c = np.floor(X * 6).astype(np.int64)
        return ((c[:, 0] + 3 * c[:, 1] + 5 * c[:, 2]) % nf).astype(np.int32)
    return np.minimum((X[:, 2] * nf).astype(np.int32), nf - 1)


@pytest.mark.parametrize("kind,nf", [("random", 3), ("random", 37), ("checker", 5), ("slab", 4)])
def test_mixed_field_tiles_match_the_oracle(oracle, kind, nf):
    from corruscant_amd import clustering
    rng = np.random.default_rng(300 + nf)
    X = rng.random((9000, 3))
    Y = rng.random((4000, 3))
    fx, fy = _fields(X, kind, nf, rng), _fields(Y, kind, nf, rng)
    radii = np.logspace(-2.2, -0.7, 12)
    print("mixed fields %s nf=%d" % (kind, nf), flush=True)
    t, q = clustering.tree(X, fx), clustering.tree(Y, fy)
    ref_tq = oracle.OracleTree(X, fx, nf).cumulative(Y, fy, nf, radii)
    ref_tt = oracle.OracleTree(X, fx, nf).cumulative(X, fx, nf, radii)
    for prec in ("fp64", "fp32_guarded"):
        tq, tt = clustering._count_jobs([(t, q), (t, t)], radii, prec)
        assert np.array_equal(tq, ref_tq), prec
        assert np.array_equal(tt, ref_tt), prec


def test_mixed_field_tiles_with_duplicates(oracle):
    """Copies of points give weights above one: those tiles take the general
    path, beside mixed unit-weight tiles of the same call."""
    from corruscant_amd import clustering
    rng = np.random.default_rng(321)
    X = rng.random((7000, 3))
    X = np.concatenate([X, X[:1500], X[:400]])
    fx = rng.integers(0, 6, X.shape[0]).astype(np.int32)
    # a copy keeps its own field: the oracle's dedup merges points with equal
    # coordinates, so make the fields of the copies equal to the originals'
    fx[7000:8500] = fx[:1500]
    fx[8500:] = fx[:400]
    Y = rng.random((3000, 3))
    fy = rng.integers(0, 6, Y.shape[0]).astype(np.int32)
    radii = np.logspace(-2, -0.8, 9)
    t, q = clustering.tree(X, fx), clustering.tree(Y, fy)
    ref = oracle.OracleTree(X, fx, 6).cumulative(Y, fy, 6, radii)
    for prec in ("fp64", "fp32_guarded"):
        assert np.array_equal(clustering._count_jobs([(t, q)], radii, prec)[0], ref), prec
