"""Correlation-function results and their error estimates.

The behaviour is that of the reference's result objects
(corruscant/clustering.py:354-672): per-bin pair-count totals, the estimate
of a chosen estimator, jackknife, field-to-field, bootstrap and Poisson
errors, the jackknife / field-to-field covariance and the printed table, with
the same class names, attributes and method signatures.

The arithmetic is organised around integer summaries of each pair type's
count array ``C[bin][i][j]`` (i: first catalogue's field, j: second's):

* total over both field indices           T   = sum_ij C_ij
* leave-one-out (jackknife) total         T_f = T - row_f - col_f + C_ff
* field-to-field sample                   C_ff

and their normalisations, the products of the summed field sizes
``A * B``, ``(A - a_f) * (B - b_f)`` and ``a_f * b_f``.  All of these are
exact in int64, so every normalised count, estimate and covariance term is
the same floating-point number whichever way the integers were summed.
"""
import numpy as np

JACKKNIFE = ("jackknife", "jk")
FIELD_TO_FIELD = ("field-to-field", "ftf")


class _PairType(object):
    """One pair type: its count array and the field sizes of its two
    catalogues (a, b), summarised once per call."""

    def __init__(self, counts, sizes):
        self.counts = np.asarray(counts)
        self.a = np.asarray(sizes[0])
        self.b = np.asarray(sizes[1])
        self.total = self.counts.sum(axis=(1, 2))

    # -- integer totals ----------------------------------------------------
    def leave_out(self, f):
        """Totals of the pairs that involve neither catalogue's field f."""
        c = self.counts
        if np.issubdtype(c.dtype, np.integer):
            return self.total - c[:, f, :].sum(axis=1) - c[:, :, f].sum(axis=1) + c[:, f, f]
        keep = np.arange(c.shape[1]) != f
        return c[:, keep][:, :, keep].sum(axis=(1, 2))

    def same_field(self, f):
        return np.array(self.counts[:, f, f])

    # -- normalisations: the number of (a, b) point pairs -----------------
    def pairs_all(self):
        return np.float64(self.a.sum() * self.b.sum())

    def pairs_leave_out(self, f):
        return np.float64((self.a.sum() - self.a[f]) * (self.b.sum() - self.b[f]))

    def pairs_same_field(self, f):
        return float(self.a[f] * self.b[f])


def _finite_or_nan(x):
    x = np.array(x, dtype=np.float64)
    x[np.isinf(x)] = np.nan
    return x


class results(object):
    """Counts of one correlation measurement with its estimator and radii
    (corruscant/clustering.py:354-634).  Subclasses set ``count_arrays`` (one
    (nbins, N_fields, N_fields) array per pair type, the random-random type
    last), ``count_sizes`` (per type, the two catalogues' field sizes) and
    ``N_fields``."""

    def __init__(self, estimator, radii):
        self.estimator = estimator
        self.nbins = len(radii) - 1
        self.radii_euclidean = radii
        self.radii_nominal = None  # set by the threedim / angular wrappers
        self.radii_units = None
        self.error_type = None

    def _types(self):
        return [_PairType(c, s) for c, s in zip(self.count_arrays, self.count_sizes)]

    def _need_fields(self, what):
        if self.N_fields <= 1:
            raise ValueError("{} cannot be calculated without data fields".format(what))

    # -- counts ------------------------------------------------------------
    def total_pair_counts(self, normalized):
        """Per-bin totals of each pair type; normalised by the number of
        point pairs of its two catalogues."""
        types = self._types()
        if not normalized:
            return [t.total for t in types]
        return [t.total / t.pairs_all() for t in types]

    def ftf_pair_counts(self, fid, normalized):
        """Per-bin counts of the pairs within field fid."""
        types = self._types()
        if not normalized:
            return [t.same_field(fid) for t in types]
        return [t.same_field(fid) / t.pairs_same_field(fid) for t in types]

    def jackknife_pair_counts(self, fid, normalized):
        """Per-bin totals without the pairs that touch field fid."""
        types = self._types()
        if not normalized:
            return [t.leave_out(fid) for t in types]
        return [t.leave_out(fid) / t.pairs_leave_out(fid) for t in types]

    def bootstrap_pair_counts(self, N_trials, device=False):
        """Normalised counts of N_trials field resamplings, per type an array
        (nbins, N_trials).  Each trial draws N_fields field ids with
        replacement from numpy's global random state (as the reference does),
        and weights the count between fields i and j by h_i h_j, h the
        trial's histogram of draws.  With ``device`` the weighted sums run on
        the GPU (``bootstrap_quadratic_forms``); the draws are the same."""
        nf = self.N_fields
        draws = np.random.randint(0, nf, N_trials * nf).reshape(N_trials, nf)
        normed = [t.counts / t.pairs_all() for t in self._types()]
        if device:
            from .clustering import _bootstrap_forms
            return _bootstrap_forms(normed, draws)
        offsets = (nf * np.arange(N_trials))[:, None]
        hist = np.bincount((draws + offsets).ravel(), minlength=N_trials * nf).reshape(N_trials, nf)
        weights = hist[:, :, None] * hist[:, None, :]
        return [np.tensordot(c, weights, axes=((1, 2), (1, 2))) for c in normed]

    # -- estimates and errors ---------------------------------------------
    def estimate(self):
        return self.estimator(*self.total_pair_counts(normalized=True))

    def covariance(self, error_type, normalized=False):
        """Covariance of the estimate over the fields: per field f, the
        deviation of the leave-f-out (jackknife) or within-f (field-to-field)
        estimate from the full one, weighted by the square root of that
        sample's share of the random-random pairs; summed over fields (and
        divided by N_fields - 1 for field-to-field).  ``normalized`` gives
        the correlation matrix instead."""
        self._need_fields("Covariance matrix")
        jackknife = error_type in JACKKNIFE
        if not jackknife and error_type not in FIELD_TO_FIELD:
            raise ValueError("covariance: error type must be jackknife or field-to-field, not {!r}"
                             .format(error_type))
        types = self._types()
        full = self.estimator(*[t.total / t.pairs_all() for t in types])
        rr_total = types[-1].total.astype(np.float64)
        cov = np.zeros((self.nbins, self.nbins))
        for f in range(self.N_fields):
            if jackknife:
                sample = [t.leave_out(f) for t in types]
                pairs = [t.pairs_leave_out(f) for t in types]
            else:
                sample = [t.same_field(f) for t in types]
                pairs = [t.pairs_same_field(f) for t in types]
            dev = self.estimator(*[s / p for s, p in zip(sample, pairs)]) - full
            share = np.sqrt(sample[-1] / rr_total)
            # term (i, j) = share_j dev_j share_i dev_i, multiplied in that order
            cov += np.multiply.outer(share, share * dev) * dev[:, None]
        if not jackknife:
            cov /= float(self.N_fields - 1)
        if normalized:
            sigma = np.sqrt(np.diagonal(cov))
            return cov / np.multiply.outer(sigma, sigma)
        return cov

    def jackknife_error(self):
        self._need_fields("Jackknife error")
        return np.sqrt(np.diagonal(self.covariance("jackknife")))

    def ftf_error(self):
        self._need_fields("field-to-field error")
        return np.sqrt(np.diagonal(self.covariance("ftf")))

    def bootstrap_error(self, N_trials=1000, device=False):
        self._need_fields("Bootstrap error")
        return np.std(self.estimator(*self.bootstrap_pair_counts(N_trials, device=device)), axis=1)

    def poisson_error(self):
        """(1 + estimate) / sqrt(first type's total); NaN where that is not
        finite (bins without pairs)."""
        first = self._types()[0].total
        with np.errstate(divide="ignore", invalid="ignore"):
            return _finite_or_nan((1 + self.estimate()) / np.sqrt(first))

    def error(self, error_type):
        if error_type == "poisson":
            return self.poisson_error()
        if error_type in JACKKNIFE:
            return self.jackknife_error()
        if error_type in FIELD_TO_FIELD:
            return self.ftf_error()
        if error_type == "bootstrap":
            return self.bootstrap_error()
        return None

    # -- the table (output format of corruscant/clustering.py:572-634) -----
    _LABELS = {3: ["DD", "DR", "RR"], 4: ["D1D2", "D1R2", "D2R1", "R1R2"]}

    def __str__(self):
        totals = self.total_pair_counts(normalized=False)
        est = self.estimate()
        err = self.error(self.error_type)
        if err is None:
            err = [None] * len(est)
        widths = [2 + max([4] + [len(str(v)) for v in t]) for t in totals]
        columns = ["Bin L", "Bin R"] + self._LABELS[len(totals)] + ["Estimator", "Error"]
        header = "".join(c.ljust(w) for c, w in zip(columns, [9, 9] + widths + [11, 10]))
        lines = ["\n", header, "{:^14}".format("({:s})".format(self.radii_units)), "-" * 79]
        edges = self.radii_nominal
        for k, (e, s) in enumerate(zip(est, err)):
            if k + 1 >= len(edges):
                break
            cells = ["{:<9}".format("{:.2E}".format(edges[k])), "{:<9}".format("{:.2E}".format(edges[k + 1]))]
            cells += ["{:<{:d}}".format(t[k], w) for t, w in zip(totals, widths)]
            cells.append("{:<11}".format("{:+.2E}".format(e)))
            cells.append("{:<10}".format("None" if s is None else "{:.2E}".format(s)))
            lines.append("".join(cells))
        return "\n".join(lines)


class autocorr_results(results):
    """DD, DR and RR of one catalogue and its randoms
    (corruscant/clustering.py:636-651)."""

    def __init__(self, dd, dr, rr, dtree, rtree, estimator, radii):
        self.dd, self.dr, self.rr = dd, dr, rr
        self.N_fields = dtree.N_fields
        self.count_arrays = [dd, dr, rr]
        d, r = dtree.field_sizes, rtree.field_sizes
        self.count_sizes = [(np.copy(a), np.copy(b)) for a, b in ((d, d), (d, r), (r, r))]
        super(autocorr_results, self).__init__(estimator, radii)


class crosscorr_results(results):
    """D1D2, D1R2, D2R1 and R1R2 of two catalogues and their randoms
    (corruscant/clustering.py:653-672)."""

    def __init__(self, d1d2, d1r2, d2r1, r1r2, dtree_1, dtree_2, rtree_1, rtree_2, estimator, radii):
        self.d1d2, self.d1r2, self.d2r1, self.r1r2 = d1d2, d1r2, d2r1, r1r2
        self.N_fields = dtree_1.N_fields
        self.count_arrays = [d1d2, d1r2, d2r1, r1r2]
        d1, d2 = dtree_1.field_sizes, dtree_2.field_sizes
        r1, r2 = rtree_1.field_sizes, rtree_2.field_sizes
        self.count_sizes = [(np.copy(a), np.copy(b)) for a, b in ((d1, d2), (d1, r2), (d2, r1), (r1, r2))]
        super(crosscorr_results, self).__init__(estimator, radii)
