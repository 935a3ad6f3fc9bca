"""The 1M-data / 10M-random 3-D workload (seeded, the bench's default) at full
size, checked through properties that need no oracle: the FP64 and FP32
guarded forms agree bit for bit, auto-correlation matrices are symmetric in
the fields, counts never decrease with the radius, a radius of 2 (beyond the
unit cube's diagonal) counts every pair (the outer product of the field
sizes), and a radius of 1e-150 (r^2 = 1e-300 > 0) counts only coincident
points: each point with itself in DD and RR, nothing in DR."""
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
