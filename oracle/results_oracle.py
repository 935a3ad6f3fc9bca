"""Plain-Python restatement of corruscant's correlation results and error
estimates (corruscant/clustering.py:132-159 and :354-558).

TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker).  Loops over
bins and fields with Python floats, no numpy linear algebra, so that a
mistake in the product's numpy expressions shows up as a difference.

Inputs: ``counts`` = list of per-type arrays indexed [bin][i][j] (DD, DR, RR
or D1D2, D1R2, D2R1, R1R2), ``sizes`` = list of (asizes, bsizes) per type,
``estimator`` in {"landy-szalay", "hamilton", "standard"}.

Square roots use math.sqrt (IEEE correctly rounded, as numpy's np.sqrt in
the reference), and every covariance term is multiplied in the reference's
order (share_j dev_j share_i dev_i, clustering.py:552-553), so jackknife,
field-to-field and Poisson errors are reproduced bit for bit; bootstrap sums
run in loop order and agree to rounding.
"""
import math


def estimator(name, cts):
    """clustering.py:132-159 on one bin: cts = (d1d2, d1r2, r1r2) or four."""
    if len(cts) == 3:
        dd, dr, rr = cts
        rd = dr
    else:
        dd, dr, rd, rr = cts
    if name == "landy-szalay":
        return (dd - dr - rd + rr) / rr
    if name == "hamilton":
        return dd * rr / (dr * rd) - 1.0
    if name == "standard":
        return dd / dr - 1.0
    raise ValueError(name)


def _nbins(counts):
    return len(counts[0])


def _nf(counts):
    return len(counts[0][0])


def _outer_sum(a, b, skip=None):
    s = 0
    for i in range(len(a)):
        for j in range(len(b)):
            if skip is not None and (i == skip or j == skip):
                continue
            s += int(a[i]) * int(b[j])
    return s


def totals(counts, sizes, normalized):
    """clustering.py:363-375: per-bin sums over both field indices."""
    out = []
    for arr, (a, b) in zip(counts, sizes):
        t = [sum(int(arr[k][i][j]) for i in range(len(arr[k])) for j in range(len(arr[k][i])))
             for k in range(len(arr))]
        if normalized:
            norm = float(_outer_sum(a, b))
            t = [float(x) / norm for x in t]
        out.append(t)
    return out


def ftf(counts, sizes, fid, normalized):
    """clustering.py:377-386: the [fid][fid] element."""
    out = []
    for arr, (a, b) in zip(counts, sizes):
        t = [int(arr[k][fid][fid]) for k in range(len(arr))]
        if normalized:
            t = [float(x) / float(int(a[fid]) * int(b[fid])) for x in t]
        out.append(t)
    return out


def jackknife(counts, sizes, fid, normalized):
    """clustering.py:388-414: totals without row and column fid."""
    out = []
    for arr, (a, b) in zip(counts, sizes):
        nf = len(arr[0])
        t = [sum(int(arr[k][i][j]) for i in range(nf) for j in range(nf) if i != fid and j != fid)
             for k in range(len(arr))]
        if normalized:
            norm = float(_outer_sum(a, b, skip=fid))
            t = [float(x) / norm for x in t]
        out.append(t)
    return out


def estimate(name, counts, sizes):
    tot = totals(counts, sizes, True)
    return [estimator(name, [t[k] for t in tot]) for k in range(_nbins(counts))]


def covariance(name, counts, sizes, error_type):
    """clustering.py:520-558 (no (N-1)/N factor for jackknife; / (N-1) for ftf)."""
    nb, nf = _nbins(counts), _nf(counts)
    est = estimate(name, counts, sizes)
    tot = totals(counts, sizes, False)
    cov = [[0.0] * nb for _ in range(nb)]
    for fid in range(nf):
        if error_type in ("jackknife", "jk"):
            norm = jackknife(counts, sizes, fid, True)
            samp = jackknife(counts, sizes, fid, False)
        else:
            norm = ftf(counts, sizes, fid, True)
            samp = ftf(counts, sizes, fid, False)
        diff = [estimator(name, [t[k] for t in norm]) - est[k] for k in range(nb)]
        q = [math.sqrt(samp[-1][k] / float(tot[-1][k])) for k in range(nb)]
        for i in range(nb):
            for j in range(nb):
                cov[i][j] += q[j] * diff[j] * q[i] * diff[i]
    if error_type in ("field-to-field", "ftf"):
        cov = [[c / float(nf - 1) for c in row] for row in cov]
    return cov


def error(name, counts, sizes, error_type):
    if error_type == "poisson":
        tot = totals(counts, sizes, False)
        est = estimate(name, counts, sizes)
        out = []
        for k in range(_nbins(counts)):
            d = tot[0][k]
            if d == 0:
                out.append(float("nan"))
            else:
                out.append((1.0 + est[k]) / math.sqrt(d))
        return out
    cov = covariance(name, counts, sizes, error_type)
    return [math.sqrt(cov[k][k]) for k in range(len(cov))]


def bootstrap(name, counts, sizes, resample):
    """clustering.py:416-464 for a given resample matrix (N_trials x Nf field
    ids): per trial the count arrays weighted by h_i h_j, h the field
    histogram of the trial; returns the per-bin population standard deviation
    of the estimator over trials."""
    nb, nf = _nbins(counts), _nf(counts)
    ests = []
    for row in resample:
        h = [0] * nf
        for f in row:
            h[int(f)] += 1
        cts = []
        for arr, (a, b) in zip(counts, sizes):
            norm = float(_outer_sum(a, b))
            cts.append([sum(float(arr[k][i][j]) / norm * h[i] * h[j] for i in range(nf) for j in range(nf))
                        for k in range(nb)])
        ests.append([estimator(name, [c[k] for c in cts]) for k in range(nb)])
    out = []
    for k in range(nb):
        vals = [e[k] for e in ests]
        m = sum(vals) / len(vals)
        out.append(math.sqrt(sum((v - m) ** 2 for v in vals) / len(vals)))
    return out
