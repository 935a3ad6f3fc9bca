"""Bootstrap resampling on the GPU (results.bootstrap_pair_counts(device=True),
corruscant/clustering.py:416-464) against the plain-loop oracle and the host
tensordot with the same draws.  Every term of the contraction is >= 0, so the
two summation orders agree to a few ulps: rtol 1e-12 on the counts; the
estimator's spread (LS cancels dd - 2 dr + rr) to rtol 1e-9."""
import ctypes
import types

import numpy as np
import pytest

from oracle import results_oracle as RO

pytestmark = pytest.mark.gpu


def _tree(sizes):
    return types.SimpleNamespace(field_sizes=np.asarray(sizes), N_fields=len(sizes))


def _auto(seed, nb, nf):
    from corruscant_amd import clustering as C
    rng = np.random.default_rng(seed)
    dd = rng.integers(100, 10000, (nb, nf, nf))
    dr = rng.integers(1000, 50000, (nb, nf, nf))
    rr = rng.integers(5000, 90000, (nb, nf, nf))
    ds, rs = rng.integers(50, 150, nf), rng.integers(300, 900, nf)
    radii = np.linspace(0.01, 0.1, nb + 1)
    res = C.autocorr_results(dd, dr, rr, _tree(ds), _tree(rs), C.est_landy_szalay, radii)
    return res, [dd, dr, rr], [(ds, ds), (ds, rs), (rs, rs)]


def test_device_bootstrap_matches_oracle():
    print("bootstrap nf=4 nb=5 trials=300", flush=True)
    res, counts, sizes = _auto(3, 5, 4)
    np.random.seed(77)
    got = res.bootstrap_error(N_trials=300, device=True)
    np.random.seed(77)
    resample = np.random.randint(0, 4, 300 * 4).reshape(300, 4)
    np.testing.assert_allclose(got, RO.bootstrap("landy-szalay", counts, sizes, resample), rtol=1e-9)


@pytest.mark.parametrize("nb,nf,trials", [(14, 64, 2000), (3, 7, 129), (2, 256, 65), (19, 16, 1)])
def test_device_counts_match_host_tensordot(nb, nf, trials):
    print("bootstrap nf=%d nb=%d trials=%d" % (nf, nb, trials), flush=True)
    res, _, _ = _auto(nf, nb, nf)
    np.random.seed(5)
    dev = res.bootstrap_pair_counts(trials, device=True)
    np.random.seed(5)
    host = res.bootstrap_pair_counts(trials)
    for a, b in zip(dev, host):
        assert a.shape == b.shape == (nb, trials)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=0)
    np.random.seed(5)
    e_dev = res.bootstrap_error(trials, device=True)
    np.random.seed(5)
    np.testing.assert_allclose(e_dev, res.bootstrap_error(trials), rtol=1e-9, atol=1e-15)


def test_every_field_once_gives_the_total():
    """Draws 0..Nf-1 in every trial: h = 1, each form is the sum of the
    normalised counts (the closed form behind clustering.py:437-438)."""
    from corruscant_amd import clustering as C
    res, counts, sizes = _auto(9, 6, 12)
    # normalised by the number of point pairs, (sum of a) x (sum of b)
    cts = [arr / np.float64(np.sum(a) * np.sum(b)) for arr, (a, b) in zip(counts, sizes)]
    resample = np.tile(np.arange(12), 100).reshape(100, 12)
    out = C._bootstrap_forms(cts, resample)
    for o, c in zip(out, cts):
        np.testing.assert_allclose(o, np.repeat(c.sum(axis=(1, 2))[:, None], 100, axis=1), rtol=1e-13)


def test_out_of_range_draw_is_an_error():
    from corruscant_amd import clustering as C
    mats = np.ones((2, 3, 3))
    draws = np.array([[0, 1, 2], [0, 3, 1]], dtype=np.uint8)
    out = np.empty((2, 2))
    rc = C.kdlib.bootstrap_quadratic_forms(C._dp(mats), 2, 3, draws.ctypes.data, 2, C._dp(out))
    assert rc == -1 and b"not below" in C.kdlib.kd_last_error()
    assert C.kdlib.bootstrap_quadratic_forms(C._dp(mats), 2, 300, draws.ctypes.data, 0, C._dp(out)) == -1
    assert C.kdlib.bootstrap_quadratic_forms(C._dp(mats), 2, 3, draws.ctypes.data, 0, C._dp(out)) == 0
