"""Writes the golden fixtures in tests/golden/ from the reference's own C code.

Run in a container where /root/reference exists and `make -C oracle` has
compiled oracle/_ref (the reference's kdbuild.c, kdquery.c and coordmath.c,
compiled from source).  Nothing here runs the reference's Python.

Fixtures (numpy .npz, no pickles):
  ref_nodes_small.npz   inputs and full node arrays of small trees
  ref_counts.npz        inputs and cumulative [radius][qf][tf] counts
  ref_config1.npz       node-array digests and DD/DR/RR counts of the
                        reference example (test/example.py:11-24)
  ref_coords.npz        radec2cart64 / radecdist2cart64 / radec2sph64 outputs
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests import cases  # noqa: E402


def node_digest(nodes):
    """sha256 over the meaningful node fields (value bits, field, weight,
    child flags and dim); padding is excluded."""
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(nodes["value"]).view(np.uint64).tobytes())
    h.update(np.ascontiguousarray(nodes["field"]).tobytes())
    h.update(np.ascontiguousarray(nodes["weight"]).tobytes())
    h.update(np.ascontiguousarray(nodes["bits"] & 0x3FF).tobytes())
    return h.hexdigest()


def small_tree_inputs():
    out = {}
    for n in (3, 4, 5, 7, 8, 17, 33, 100, 1000):
        X, f = cases.cube(n, seed=100 + n, nfields=3)
        out["cube%d" % n] = (X, f, 3)
    X = cases.duplicates(200, 3, seed=5)
    out["dups"] = (X, (np.arange(X.shape[0]) % 5).astype(np.int32), 5)
    X = cases.lattice(8, spacing=0.125)
    out["lattice8"] = (X, None, 1)
    X = np.zeros((40, 3))
    X[:, 0] = np.repeat([0.0, -0.0, 0.5, -0.5], 10)
    X[:, 1] = np.tile(np.arange(10) * 0.25, 4)
    out["signed_zero"] = (X, (np.arange(40) % 2).astype(np.int32), 2)
    X, _ = cases.cube(600, seed=9)
    out["fields300"] = (X, (np.arange(600) % 300).astype(np.int32), 300)
    return out


def main():
    assert O.ref_available(), "build oracle/_ref first (make -C oracle)"
    # 1. small trees
    z = {}
    for name, (X, f, nf) in small_tree_inputs().items():
        t = O.RefTree(X, f, nf)
        nd = t.nodes.copy()
        nd["pad0"] = 0
        nd["pad1"] = 0
        nd["bits"] &= 0x3FF
        z[name + "/xyz"] = X
        z[name + "/fields"] = f if f is not None else np.zeros(0, np.int32)
        z[name + "/num_fields"] = np.array(nf)
        z[name + "/nodes"] = nd.view(np.uint8).reshape(-1, 40)
        z[name + "/size"] = np.array(t.size)
    np.savez_compressed(os.path.join(HERE, "ref_nodes_small.npz"), **z)

    # 2. counts on small inputs (every radius, reference one call per radius)
    z = {}
    X = cases.lattice(8, spacing=0.125)
    f = (np.arange(X.shape[0]) % 3).astype(np.int32)
    radii = np.array([0.125, 0.125 * np.sqrt(2), 0.25, 0.375, 0.125 * np.sqrt(3), 0.5])
    t = O.RefTree(X, f, 3)
    z["lattice/xyz"], z["lattice/fields"], z["lattice/radii"] = X, f, radii
    z["lattice/counts"] = t.cumulative(X, f, 3, radii, nthreads=4)
    X = cases.duplicates(300, 3, seed=6)
    f = (np.arange(X.shape[0]) % 4).astype(np.int32)
    radii = np.logspace(-2, -0.5, 7)
    t = O.RefTree(X, f, 4)
    z["dups/xyz"], z["dups/fields"], z["dups/radii"] = X, f, radii
    z["dups/counts"] = t.cumulative(X, f, 4, radii, nthreads=4)
    X, f = cases.cube(2000, seed=77, nfields=6)
    Q, qf = cases.cube(500, seed=78, nfields=6)
    radii = np.array([0.0, 0.02, 0.05, 0.1, 0.3, 2.0])
    t = O.RefTree(X, f, 6)
    z["cross/xyz"], z["cross/fields"], z["cross/q"], z["cross/qf"], z["cross/radii"] = X, f, Q, qf, radii
    z["cross/counts"] = t.cumulative(Q, qf, 6, radii, nthreads=4)
    np.savez_compressed(os.path.join(HERE, "ref_counts.npz"), **z)

    # 3. the reference example (config 1)
    D, df, R, rf, radii = cases.config1()
    td, tr = O.RefTree(D, df, 4), O.RefTree(R, rf, 4)
    z = {
        "radii": radii,
        "d_digest": np.array(node_digest(td.nodes)),
        "r_digest": np.array(node_digest(tr.nodes)),
        "d_size": np.array(td.size),
        "r_size": np.array(tr.size),
        "dd": td.cumulative(D, df, 4, radii, nthreads=8),
        "dr": tr.cumulative(D, df, 4, radii, nthreads=8),
        "rr": tr.cumulative(R, rf, 4, radii, nthreads=8),
    }
    np.savez_compressed(os.path.join(HERE, "ref_config1.npz"), **z)

    # 4. coordinate conversions (glibc sincos of this container, via the
    #    reference's coordmath.c)
    rng = np.random.default_rng(8)
    ra = np.concatenate([rng.uniform(0, 360, 4000), [0.0, 90.0, 180.0, 270.0, 360.0, 45.0, 1e-300, 359.999999]])
    dec = np.concatenate([np.degrees(np.arcsin(rng.uniform(-1, 1, 4000))), [0.0, 90.0, -90.0, 30.0, -30.0, 45.0,
                                                                            -1e-300, 89.9999999]])
    dist = rng.uniform(10, 3000, ra.size)
    _, co, _ = O.ref_lib()
    xyz = O.ref_radec2cart64(ra, dec)
    xyzd = np.empty_like(xyz)
    co.radecdist2cart64(O._dp(ra), O._dp(dec), O._dp(dist), O._dp(xyzd), ra.size)
    th, ph = np.empty_like(ra), np.empty_like(ra)
    co.radec2sph64(O._dp(ra), O._dp(dec), O._dp(th), O._dp(ph), ra.size)
    np.savez_compressed(os.path.join(HERE, "ref_coords.npz"), ra=ra, dec=dec, dist=dist, xyz=xyz, xyzd=xyzd,
                        theta=th, phi=ph)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
