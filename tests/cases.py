"""Seeded inputs shared by the CPU and GPU tests."""
import numpy as np


def cube(n, seed, nfields=None, axis=0):
    rng = np.random.default_rng(seed)
    X = rng.random((n, 3))
    if nfields is None:
        return X, None
    return X, np.minimum((X[:, axis] * nfields).astype(np.int32), nfields - 1)


def config1():
    """The reference's own example (test/example.py:11-24): seed 3, 4000 data,
    80000 randoms in the unit cube, fields floor(4x), 6 log-spaced edges."""
    np.random.seed(3)
    D = np.random.rand(4000, 3)
    R = np.random.rand(80000, 3)
    radii = np.logspace(-2.5, -1., 6)
    return D, (D[:, 0] * 4).astype("int32"), R, (R[:, 0] * 4).astype("int32"), radii


def lattice(m, spacing=0.125, offset=0.0):
    """Points on an m x m x m lattice with a dyadic spacing: every coordinate
    difference and squared distance is exact in FP64."""
    g = np.arange(m) * spacing + offset
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    return X


def duplicates(n, copies, seed):
    rng = np.random.default_rng(seed)
    X = np.repeat(rng.random((n, 3)), copies, axis=0)
    perm = rng.permutation(X.shape[0])
    X = X[perm]
    X[::5, 1] = 0.0
    X[1::5, 1] = -0.0
    return X


def clustered(n_uniform, n_clusters, per_cluster, sigma, seed):
    rng = np.random.default_rng(seed)
    U = rng.random((n_uniform, 3))
    C = rng.random((n_clusters, 3))
    K = np.repeat(C, per_cluster, 0) + rng.normal(0, sigma, (n_clusters * per_cluster, 3))
    K -= np.floor(K)
    K[K == 1.0] = 0.0
    return np.vstack([U, K])
