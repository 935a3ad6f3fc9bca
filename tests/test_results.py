"""CPU tests of the results object (estimators, jackknife / field-to-field /
bootstrap / Poisson errors, table) against the plain-loop oracle and closed
forms (corruscant/clustering.py:132-159, :354-672)."""
import types

import numpy as np
import pytest

from oracle import results_oracle as RO

EST = ["landy-szalay", "hamilton", "standard"]


def _tree(sizes):
    return types.SimpleNamespace(field_sizes=np.asarray(sizes), N_fields=len(sizes))


def _auto(seed, nb=5, nf=4, homogeneous=False):
    from corruscant_amd import clustering as C
    rng = np.random.default_rng(seed)
    if homogeneous:
        base = rng.integers(1000, 5000, (nb, 1, 1)) * np.ones((1, nf, nf), dtype=np.int64)
        dd, dr, rr = base * 1, base * 3, base * 9
        ds, rs = [100] * nf, [300] * nf
    else:
        dd = rng.integers(100, 10000, (nb, nf, nf))
        dr = rng.integers(1000, 50000, (nb, nf, nf))
        rr = rng.integers(5000, 90000, (nb, nf, nf))
        ds = rng.integers(50, 150, nf)
        rs = rng.integers(300, 900, nf)
    radii = np.linspace(0.01, 0.1, nb + 1)
    res = C.autocorr_results(dd, dr, rr, _tree(ds), _tree(rs), C.est_landy_szalay, radii)
    res.radii_nominal = radii
    res.radii_units = "Data units"
    sizes = [(ds, ds), (ds, rs), (rs, rs)]
    return res, [dd, dr, rr], sizes


def _cross(seed, nb=4, nf=3):
    from corruscant_amd import clustering as C
    rng = np.random.default_rng(seed)
    arrs = [rng.integers(100, 60000, (nb, nf, nf)) for _ in range(4)]
    s = [rng.integers(40, 900, nf) for _ in range(4)]
    trees = [_tree(x) for x in s]
    radii = np.linspace(0.02, 0.2, nb + 1)
    res = C.crosscorr_results(*arrs, *trees, C.est_landy_szalay, radii)
    res.radii_nominal, res.radii_units = radii, "degrees"
    sizes = [(s[0], s[1]), (s[0], s[3]), (s[1], s[2]), (s[2], s[3])]
    return res, arrs, sizes


def _set_est(res, name):
    from corruscant_amd import clustering as C
    res.estimator = {"landy-szalay": C.est_landy_szalay, "hamilton": C.est_hamilton,
                     "standard": C.est_standard}[name]


@pytest.mark.parametrize("name", EST)
@pytest.mark.parametrize("kind", ["auto", "cross"])
def test_estimates_and_errors_match_oracle(name, kind):
    res, counts, sizes = (_auto(1) if kind == "auto" else _cross(2))
    _set_est(res, name)
    # integer totals make every normalised count exact, and the covariance
    # terms are multiplied in the same order: equal bit for bit
    np.testing.assert_array_equal(res.estimate(), RO.estimate(name, counts, sizes))
    for et in ("jackknife", "ftf", "poisson"):
        np.testing.assert_array_equal(res.error(et), RO.error(name, counts, sizes, et))
    for et in ("jackknife", "ftf"):
        np.testing.assert_array_equal(res.covariance(et), np.array(RO.covariance(name, counts, sizes, et)))
        corr = res.covariance(et, normalized=True)
        np.testing.assert_allclose(np.diagonal(corr), 1.0, rtol=1e-15)


@pytest.mark.parametrize("name", EST)
def test_bootstrap_matches_oracle_with_the_same_draws(name):
    res, counts, sizes = _auto(3)
    _set_est(res, name)
    np.random.seed(77)
    got = res.bootstrap_error(N_trials=300)
    np.random.seed(77)
    resample = np.random.randint(0, res.N_fields, 300 * res.N_fields).reshape(300, res.N_fields)
    np.testing.assert_allclose(got, RO.bootstrap(name, counts, sizes, resample), rtol=1e-12)


def test_closed_forms_homogeneous_fields():
    """Every field pair carries the same counts: every jackknife and
    field-to-field sub-estimate equals the full estimate, so both covariances
    and the bootstrap spread are zero; LS = (1 - 2*3*(1/3) + 9*(1/9)) ..."""
    res, counts, sizes = _auto(5, homogeneous=True)
    assert np.all(res.covariance("jackknife") == 0.0) or np.allclose(res.covariance("jackknife"), 0, atol=1e-30)
    assert np.allclose(res.covariance("ftf"), 0, atol=1e-30)
    assert np.allclose(res.bootstrap_error(50), 0, atol=1e-15)
    # dd:dr:rr = 1:3:9 with data:random sizes 1:3 gives normalised dd = dr = rr
    np.testing.assert_allclose(res.estimate(), 0.0, atol=1e-15)


def test_closed_form_estimators():
    from corruscant_amd import clustering as C
    assert C.est_landy_szalay(4.0, 1.0, 2.0) == (4 - 1 - 1 + 2) / 2
    assert C.est_landy_szalay(4.0, 1.0, 3.0, 2.0) == (4 - 1 - 3 + 2) / 2
    assert C.est_hamilton(4.0, 2.0, 2.0) == 4 * 2 / (2 * 2) - 1
    assert C.est_standard(3.0, 2.0, 7.0) == 0.5


def test_poisson_closed_form_and_zero_counts():
    res, counts, sizes = _auto(6)
    tot = [np.sum(a, axis=(1, 2)) for a in counts]
    np.testing.assert_allclose(res.poisson_error(), (1 + res.estimate()) / np.sqrt(tot[0]), rtol=1e-15)
    res.dd[0] = 0  # a bin without data pairs: -1/0 and 1/0 become NaN
    e = res.poisson_error()
    assert np.isnan(e[0]) and np.all(np.isfinite(e[1:]))


def test_jackknife_two_fields_by_hand():
    """Two fields, one bin: leaving field 0 out keeps only the [1][1] element."""
    from corruscant_amd import clustering as C
    dd = np.array([[[10, 2], [3, 20]]])
    dr = np.array([[[30, 6], [7, 50]]])
    rr = np.array([[[90, 9], [8, 120]]])
    s = np.array([5, 7])
    r = np.array([15, 20])
    res = C.autocorr_results(dd, dr, rr, _tree(s), _tree(r), C.est_standard, np.array([0.0, 1.0]))
    tot = (10 + 2 + 3 + 20) / 144.0 / ((30 + 6 + 7 + 50) / float(12 * 35)) - 1
    jk0 = (20 / 49.0) / (50 / 140.0) - 1
    jk1 = (10 / 25.0) / (30 / 75.0) - 1
    q0 = np.sqrt(120 / 227.0)
    q1 = np.sqrt(90 / 227.0)
    expect = (q0 * (jk0 - tot)) ** 2 + (q1 * (jk1 - tot)) ** 2
    np.testing.assert_allclose(res.covariance("jackknife")[0, 0], expect, rtol=1e-13)


def test_table_format():
    res, _, _ = _auto(7)
    res.error_type = "jackknife"
    text = str(res)
    lines = text.split("\n")
    assert lines[2].startswith("Bin L    Bin R    DD")
    assert lines[3].strip() == "(Data units)"
    assert lines[4] == "-" * 79
    assert len(lines) == 5 + res.nbins
    res.error_type = None
    assert str(res).split("\n")[-1].rstrip().endswith("None")


def test_error_without_fields_raises():
    from corruscant_amd import clustering as C
    a = np.ones((2, 1, 1), dtype=np.int64)
    res = C.autocorr_results(a, a, a, _tree([10]), _tree([10]), C.est_standard, np.array([0.1, 0.2, 0.3]))
    for f in (res.jackknife_error, res.ftf_error, res.bootstrap_error):
        with pytest.raises(ValueError):
            f()


@pytest.mark.parametrize("seed", range(6))
def test_wide_cases_bit_identical(seed):
    """Up to 40 fields and 14 bins, auto and cross, every estimator: the
    estimate, the jackknife / field-to-field covariances and the Poisson
    errors equal the plain-loop oracle bit for bit."""
    from corruscant_amd import clustering as C
    rng = np.random.default_rng(1000 + seed)
    nf, nb = int(rng.integers(2, 41)), int(rng.integers(1, 15))
    kind = "auto" if seed % 2 == 0 else "cross"
    ntypes = 3 if kind == "auto" else 4
    counts = [rng.integers(1, 10 ** int(rng.integers(2, 9)), (nb, nf, nf)) for _ in range(ntypes)]
    cat = [rng.integers(1, 10 ** 6, nf) for _ in range(4)]
    radii = np.linspace(0.01, 0.2, nb + 1)
    if kind == "auto":
        res = C.autocorr_results(*counts, _tree(cat[0]), _tree(cat[1]), C.est_landy_szalay, radii)
        sizes = [(cat[0], cat[0]), (cat[0], cat[1]), (cat[1], cat[1])]
    else:
        res = C.crosscorr_results(*counts, *[_tree(c) for c in cat], C.est_landy_szalay, radii)
        sizes = [(cat[0], cat[1]), (cat[0], cat[3]), (cat[1], cat[2]), (cat[2], cat[3])]
    for name in EST:
        _set_est(res, name)
        with np.errstate(divide="ignore", invalid="ignore"):
            np.testing.assert_array_equal(res.estimate(), RO.estimate(name, counts, sizes))
            for et in ("jackknife", "ftf"):
                np.testing.assert_array_equal(res.covariance(et), np.array(RO.covariance(name, counts, sizes, et)))
            np.testing.assert_array_equal(res.error("poisson"), RO.error(name, counts, sizes, "poisson"))


def test_float_counts_take_the_masked_sums():
    """Count arrays of a float dtype (not what the library returns) give the
    same jackknife totals as the integer path, through masked sums."""
    res, counts, sizes = _auto(8)
    ints = res.jackknife_pair_counts(1, normalized=True)
    res.count_arrays = [c.astype(np.float64) for c in res.count_arrays]
    np.testing.assert_array_equal(res.jackknife_pair_counts(1, normalized=True), ints)


def test_unknown_covariance_type_raises():
    res, _, _ = _auto(9)
    with pytest.raises(ValueError):
        res.covariance("poisson")
