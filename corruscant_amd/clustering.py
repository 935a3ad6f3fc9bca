"""Trees, pair counts and correlation-function results (GPU back end).

Mirrors corruscant/clustering.py: ``tree`` (:293-352), ``_autocorr`` /
``_crosscorr`` (:161-291), the estimators (:132-159) and the ``results``
classes (:354-672), with the same names, arguments and error behaviour.
The difference is underneath: every radius of every pair type goes to the GPU
library in one call (``pair_count_jobs``) instead of one ``pair_count`` call
per radius, and the counts come back cumulative per radius as before.
"""
from __future__ import division, print_function

import ctypes
from ctypes import POINTER, Structure, c_double, c_int, c_longlong, c_void_p
import warnings

import numpy as np

import corruscant_amd
from .results import autocorr_results, crosscorr_results, results  # noqa: F401  (reference names)

kdlib = corruscant_amd.kdlib


class datum(Structure):
    """src/kdtree.h:13-17 (32 bytes; field and weight are 8-bit)."""
    _fields_ = [("value", c_double * 3), ("bits", ctypes.c_uint32), ("_pad", ctypes.c_uint32)]


class node(Structure):
    """src/kdtree.h:19-25 (40 bytes)."""
    _fields_ = [("data", datum), ("bits", ctypes.c_uint32), ("_pad", ctypes.c_uint32)]


class kdtree(Structure):
    """Tree handle, 64 bytes like src/kdtree.h:33-41 (the reference's own ctypes
    mirror declares 48, clustering.py:36-45)."""
    _fields_ = [
        ("node_data", POINTER(node)),
        ("size", c_int),
        ("memsize", c_int),
        ("num_fields", c_int),
        ("dev", c_void_p),
        ("fields", POINTER(c_int)),
        ("args", POINTER(c_int) * 3),
    ]


assert ctypes.sizeof(kdtree) == 64


class kd_job(Structure):
    _fields_ = [("tree", kdtree), ("query", kdtree), ("num_fields_query", c_int),
                ("counts", POINTER(c_longlong))]


class kd_wjob(Structure):
    _fields_ = [("tree", kdtree), ("query", kdtree), ("num_fields_query", c_int),
                ("tree_weights", c_void_p), ("query_weights", c_void_p),
                ("counts", POINTER(c_longlong)),
                ("tree_field_sizes", POINTER(c_longlong)), ("query_field_sizes", POINTER(c_longlong))]


class kd_stats(Structure):
    _fields_ = [("evaluated_pairs", c_longlong), ("rechecked_pairs", c_longlong),
                ("reach_checks", c_longlong), ("work_items", c_longlong),
                ("enumerate_ms", c_double), ("count_ms", c_double), ("total_ms", c_double),
                ("trimmed_items", c_longlong), ("fp32_pairs", c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class kd_tuning(Structure):
    _fields_ = [("blocks_per_cu", c_int), ("chunk", c_int), ("enum_blocks_per_cu", c_int)]


kdlib.tree_construct.restype = kdtree
kdlib.tree_construct.argtypes = [POINTER(c_double), POINTER(c_int), c_int, c_int]
kdlib.query_construct.restype = kdtree
kdlib.query_construct.argtypes = [POINTER(c_double), POINTER(c_int), c_int, c_int]
kdlib.destroy.restype = None
kdlib.destroy.argtypes = [kdtree]
kdlib.pair_count.restype = POINTER(c_longlong)
kdlib.pair_count.argtypes = [kdtree, POINTER(c_double), POINTER(c_int), c_int, c_int, c_double, c_int]
kdlib.pair_count_jobs.restype = c_int
kdlib.pair_count_jobs.argtypes = [POINTER(kd_job), c_int, POINTER(c_double), c_int, c_int,
                                  POINTER(kd_stats)]
kdlib.pair_count_plan.restype = c_void_p
kdlib.pair_count_plan.argtypes = [POINTER(kd_job), c_int, POINTER(c_double), c_int, c_int, POINTER(kd_tuning)]
kdlib.pair_count_plan_run.restype = c_int
kdlib.pair_count_plan_run.argtypes = [c_void_p, POINTER(kd_stats)]
kdlib.pair_count_plan_items.restype = c_longlong
kdlib.pair_count_plan_items.argtypes = [c_void_p]
kdlib.pair_count_plan_free.restype = None
kdlib.pair_count_plan_free.argtypes = [c_void_p]
kdlib.pair_count_weighted.restype = c_int
kdlib.pair_count_weighted.argtypes = [c_void_p, c_int, c_int, POINTER(c_double), c_int, c_int, POINTER(kd_stats)]
kdlib.kd_last_error.restype = ctypes.c_char_p
kdlib.kd_last_error.argtypes = []
kdlib.kd_free.restype = None
kdlib.kd_free.argtypes = [c_void_p]
kdlib.tree_fetch_nodes.restype = c_int
kdlib.tree_fetch_nodes.argtypes = [POINTER(kdtree)]
kdlib.tree_fetch_perm.restype = c_int
kdlib.tree_fetch_perm.argtypes = [kdtree, POINTER(c_int), POINTER(c_int)]
kdlib.bootstrap_quadratic_forms.restype = c_int
kdlib.bootstrap_quadratic_forms.argtypes = [POINTER(c_double), c_int, c_int, c_void_p, c_longlong,
                                            POINTER(c_double)]
kdlib.kd_set_tile_size.restype = c_int
kdlib.kd_set_tile_size.argtypes = [c_int]
kdlib.kd_set_stage_pool.restype = c_int
kdlib.kd_set_stage_pool.argtypes = [c_longlong]

PRECISION_FP64 = 0
PRECISION_FP32_GUARDED = 1
_PRECISIONS = {"fp64": PRECISION_FP64, "fp32_guarded": PRECISION_FP32_GUARDED}

#: precision form used by _autocorr / _crosscorr; both give exact counts
precision = "fp64"
#: kd_stats of the last batched pair count, as a dict
last_stats = None


def set_precision(name):
    """Select "fp64" (every distance in FP64) or "fp32_guarded" (FP32 with an
    FP64 recheck near every radius); counts are identical."""
    global precision
    if name not in _PRECISIONS:
        raise ValueError("precision must be one of {}".format(sorted(_PRECISIONS)))
    precision = name


def set_tile_size(points_per_tile):
    """Points per tile (1..64) of the trees built from now on (the leaf size
    of the tile hierarchy; trees already built keep theirs).  Counts do not
    depend on it."""
    if kdlib.kd_set_tile_size(int(points_per_tile)) != 0:
        raise ValueError(_err())


def _precision_code(name):
    """The library's code of a precision form name (None: the module's form)."""
    return _PRECISIONS[name or precision]


def _err():
    return kdlib.kd_last_error().decode("utf-8", "replace")


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(POINTER(c_int))


def _query_tree(tree, points, radius, num_threads, fields=None, N_fields=1):
    """One cumulative count matrix (N_fields, tree.N_fields) at one radius,
    through the reference's own entry point (clustering.py:73-90)."""
    fields_arg = None
    if fields is not None:
        fields_arg = fields.ctypes.data_as(POINTER(c_int))
    counts = kdlib.pair_count(tree.ctree, points.ctypes.data_as(POINTER(c_double)), fields_arg,
                              c_int(points.shape[0]), c_int(N_fields), c_double(radius),
                              c_int(num_threads))
    if not counts:
        raise RuntimeError("pair_count failed: " + _err())
    out = np.ctypeslib.as_array(counts, shape=(N_fields * tree.N_fields,)).copy()
    kdlib.kd_free(ctypes.cast(counts, c_void_p))
    return out.reshape(N_fields, tree.N_fields)


#: distinct radii one library call takes (pair_count_jobs, pair_count_weighted)
MAX_RADII_PER_CALL = 64
MAX_RADII_PER_WEIGHTED_CALL = 32


def _radius_groups(radii, limit):
    """Split the distinct radii into groups of at most ``limit``.  Returns
    [(group_radii, rows)], where ``rows`` are the positions in ``radii`` whose
    value belongs to the group.  The reference takes any number of radii (one
    pair_count per radius, clustering.py:218-230); the library's calls take a
    bounded number each, and cumulative counts at different radii do not
    depend on each other, so a longer list becomes several calls."""
    distinct = np.unique(radii)
    if distinct.size <= limit or not np.all(np.isfinite(radii)):
        # (a NaN or infinite radius goes to the library as it is, which
        # refuses it with its own message)
        return [(radii, np.arange(radii.size))]
    groups = []
    for k in range(0, distinct.size, limit):
        part = distinct[k:k + limit]
        rows = np.nonzero((radii >= part[0]) & (radii <= part[-1]))[0]
        groups.append((radii[rows], rows))
    return groups


def _merge_stats(dicts):
    if len(dicts) == 1:
        return dicts[0]
    return {k: sum(d[k] for d in dicts) for k in dicts[0]}


def _count_jobs(pairs, radii, prec=None, record=True):
    """Cumulative counts for several (tree, query-tree) pairs and all radii in
    one GPU call (several when there are more than MAX_RADII_PER_CALL distinct
    radii).  Returns a list of arrays (len(radii), Nq, Nt).  ``record`` keeps
    the call's kd_stats in ``last_stats``."""
    global last_stats
    radii = np.ascontiguousarray(radii, dtype=np.float64).ravel()
    outs = [np.zeros((radii.size, q.N_fields, t.N_fields), dtype=np.int64) for t, q in pairs]
    p = _PRECISIONS[prec or precision]
    stats = []
    for group, rows in _radius_groups(radii, MAX_RADII_PER_CALL):
        group = np.ascontiguousarray(group)
        parts = [np.zeros((group.size, q.N_fields, t.N_fields), dtype=np.int64) for t, q in pairs]
        jobs = (kd_job * len(pairs))()
        for i, (t, q) in enumerate(pairs):
            jobs[i].tree = t.ctree
            jobs[i].query = q.ctree
            jobs[i].num_fields_query = q.N_fields
            jobs[i].counts = parts[i].ctypes.data_as(POINTER(c_longlong))
        st = kd_stats()
        rc = kdlib.pair_count_jobs(jobs, len(pairs), _dp(group), group.size, p, ctypes.byref(st))
        if rc != 0:
            raise RuntimeError("pair_count_jobs failed: " + _err())
        for out, part in zip(outs, parts):
            out[rows] = part
        stats.append(st.as_dict())
    if record:
        last_stats = _merge_stats(stats)
    return outs


def _concurrently(background, foreground):
    """Run two library calls at once (each on its own HIP stream); returns
    (foreground(), background())."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=1) as ex:
        fut = ex.submit(background)
        return foreground(), fut.result()


def _bootstrap_forms(cts, resample):
    """out[k][b][t] = sum_ij cts[k][b][i][j] h_t[i] h_t[j] on the GPU, h_t the
    field histogram of trial t's draws (resample: N_trials x N_fields ids)."""
    ntr, nf = resample.shape
    if nf > 256:
        raise ValueError("bootstrap on the device supports at most 256 fields")
    mats = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.float64).reshape(-1, nf, nf) for c in cts]))
    draws = np.ascontiguousarray(resample, dtype=np.uint8)
    out = np.empty((mats.shape[0], ntr), dtype=np.float64)
    if kdlib.bootstrap_quadratic_forms(_dp(mats), mats.shape[0], nf, draws.ctypes.data, ntr, _dp(out)) != 0:
        raise RuntimeError("bootstrap_quadratic_forms failed: " + _err())
    res, k = [], 0
    for c in cts:
        res.append(out[k:k + c.shape[0]])
        k += c.shape[0]
    return res


def _weights(w, n, nsets):
    if w is None:
        return None
    w = np.asarray(w)
    if w.ndim == 1:
        w = w[None, :]
    if w.shape != (nsets, n):
        raise ValueError("weight sets must have shape (nsets, N) = ({}, {}); got {}".format(nsets, n, w.shape))
    if w.dtype != np.uint8:
        if np.any(w < 0) or np.any(w > 255):
            raise ValueError("point multiplicities must be integers in [0, 255]")
        w = w.astype(np.uint8)
    return np.ascontiguousarray(w)


def _count_weighted(jobs, radii, nsets, sizes=False, prec=None):
    """Cumulative counts of weighted jobs in one GPU call (several when there
    are more than MAX_RADII_PER_WEIGHTED_CALL distinct radii).  ``jobs`` is a
    list of (tree, query_tree, tree_weights, query_weights); weights are
    (nsets, N) integer multiplicities of the trees' input points, or None for
    all ones.  Returns a list of arrays (nsets, len(radii), Nq, Nt): set s
    holds the counts of the catalogues np.repeat(points, weights[s], axis=0).
    With ``sizes``, also a list of (tree, query) arrays (nsets, N_fields): the
    per-field sums of the multiplicities (the resampled field sizes).
    ``prec``: "fp64" or "fp32_guarded" (None: the module's form); both give
    the same counts."""
    global last_stats
    radii = np.ascontiguousarray(radii, dtype=np.float64).ravel()
    weights = [(_weights(tw, t.points.shape[0], nsets), _weights(qw, q.points.shape[0], nsets))
               for t, q, tw, qw in jobs]
    outs = [np.zeros((nsets, radii.size, q.N_fields, t.N_fields), dtype=np.int64) for t, q, _, _ in jobs]
    fsz = [(np.zeros((nsets, t.N_fields), dtype=np.int64), np.zeros((nsets, q.N_fields), dtype=np.int64))
           for t, q, _, _ in jobs] if sizes else []
    stats = []
    for k, (group, rows) in enumerate(_radius_groups(radii, MAX_RADII_PER_WEIGHTED_CALL)):
        group = np.ascontiguousarray(group)
        parts = [np.zeros((nsets, group.size, q.N_fields, t.N_fields), dtype=np.int64) for t, q, _, _ in jobs]
        cj = (kd_wjob * len(jobs))()
        for i, ((t, q, _, _), (tw, qw)) in enumerate(zip(jobs, weights)):
            cj[i].tree = t.ctree
            cj[i].query = q.ctree
            cj[i].num_fields_query = q.N_fields
            cj[i].tree_weights = None if tw is None else tw.ctypes.data
            cj[i].query_weights = None if qw is None else qw.ctypes.data
            cj[i].counts = parts[i].ctypes.data_as(POINTER(c_longlong))
            if sizes and k == 0:  # the field sizes do not depend on the radii
                cj[i].tree_field_sizes = fsz[i][0].ctypes.data_as(POINTER(c_longlong))
                cj[i].query_field_sizes = fsz[i][1].ctypes.data_as(POINTER(c_longlong))
        st = kd_stats()
        rc = kdlib.pair_count_weighted(ctypes.addressof(cj), len(jobs), nsets, _dp(group), group.size,
                                       _precision_code(prec), ctypes.byref(st))
        if rc != 0:
            raise RuntimeError("pair_count_weighted failed: " + _err())
        for out, part in zip(outs, parts):
            out[:, rows] = part
        stats.append(st.as_dict())
    last_stats = _merge_stats(stats)
    return (outs, fsz) if sizes else outs


def validate_points(points):
    """clustering.py:92-114, plus a float64 conversion (the reference hands any
    dtype's buffer to C as double, which misreads float32 input)."""
    points = np.array(points)
    N_points = max(points.shape)
    if not points.shape == (N_points, 3):
        points = points.T
    if not points.shape == (N_points, 3):
        raise ValueError("Array must be of shape (N, 3) "
                         "or (3, N). The provided array has "
                         "shape {:s}".format(str(points.shape)))
    newpoints = np.require(points, dtype="float64", requirements="AC")
    if newpoints is not points:
        warnings.warn("Data array is being copied "
                      "to have shape {:s}".format(str(newpoints.shape)), RuntimeWarning)
    return newpoints


def validate_fields(fields, points):
    """clustering.py:116-130."""
    fields = np.array(fields)
    N_points = max(points.shape)
    if not fields.shape == (N_points,):
        raise ValueError("Field IDs must be a 1-d array of length equal to "
                         "the number of points")
    return np.require(fields, dtype="int32", requirements="AC")


def est_landy_szalay(*args):
    """(DD - DR - RD + RR) / RR (clustering.py:132-139)."""
    if len(args) == 3:
        d1d2, d1r2, r1r2 = args
        d2r1 = d1r2
    elif len(args) == 4:
        d1d2, d1r2, d2r1, r1r2 = args
    return (d1d2 - d1r2 - d2r1 + r1r2) / r1r2


def est_hamilton(*args):
    """DD RR / (DR RD) - 1 (clustering.py:141-149)."""
    if len(args) == 3:
        d1d2, d1r2, r1r2 = args
        d2r1 = d1r2
    elif len(args) == 4:
        d1d2, d1r2, d2r1, r1r2 = args
    return (d1d2 * r1r2 / (d1r2 * d2r1)) - 1.


def est_standard(*args):
    """DD / DR - 1 (clustering.py:151-159)."""
    if len(args) == 3:
        d1d2, d1r2, r1r2 = args
    elif len(args) == 4:
        d1d2, d1r2, d2r1, r1r2 = args
    return d1d2 / d1r2 - 1.


_ESTIMATORS = {"landy-szalay": est_landy_szalay, "hamilton": est_hamilton, "standard": est_standard}


def _estimator(est_type):
    try:
        return _ESTIMATORS[est_type]
    except (KeyError, TypeError):
        raise ValueError("Estimator type for Xi \"{:s}\" "
                         "not valid. Try \"standard\", \"hamilton\", or "
                         "\"landy-szalay\"".format(str(est_type)))


def _autocorr(data_tree, rand_tree, radii, est_type="landy-szalay", err_type="jackknife",
              num_threads=4):
    """DD, DR and RR at every radius (clustering.py:161-235), all in one GPU
    call; err_type is accepted and ignored as in the reference (:203-215);
    num_threads is the reference's CPU thread count and has no effect here."""
    if not (data_tree.N_fields == rand_tree.N_fields):
        raise ValueError("data and random trees must have same number of fields")
    estimator = _estimator(est_type)
    dd, dr, rr = _count_jobs([(data_tree, data_tree), (rand_tree, data_tree), (rand_tree, rand_tree)], radii)
    dd_array = np.diff(dd, axis=0)
    dr_array = np.diff(dr, axis=0)
    rr_array = np.diff(rr, axis=0)
    return autocorr_results(dd_array, dr_array, rr_array, data_tree, rand_tree, estimator, radii)


def _crosscorr(data_tree_1, data_tree_2, rand_tree_1, rand_tree_2, radii, est_type="landy-szalay",
               num_threads=4):
    """D1D2, D1R2, D2R1, R1R2 (clustering.py:237-291).  D1D2 and R1R2 query the
    larger tree with the other catalogue's points, as the reference does
    (:256-263, :275-284), so the [qf][tf] orientation follows the sizes."""
    if not (data_tree_1.N_fields == data_tree_2.N_fields == rand_tree_1.N_fields == rand_tree_2.N_fields):
        raise ValueError("data and random trees must have same number of fields")
    estimator = _estimator(est_type)
    if data_tree_1.size > data_tree_2.size:
        d1d2_pair = (data_tree_1, data_tree_2)
    else:
        d1d2_pair = (data_tree_2, data_tree_1)
    if rand_tree_1.size > rand_tree_2.size:
        r1r2_pair = (rand_tree_1, rand_tree_2)
    else:
        r1r2_pair = (rand_tree_2, rand_tree_1)
    d1d2, d1r2, d2r1, r1r2 = _count_jobs(
        [d1d2_pair, (rand_tree_2, data_tree_1), (rand_tree_1, data_tree_2), r1r2_pair], radii)
    return crosscorr_results(np.diff(d1d2, axis=0), np.diff(d1r2, axis=0), np.diff(d2r1, axis=0),
                             np.diff(r1r2, axis=0), data_tree_1, data_tree_2, rand_tree_1, rand_tree_2,
                             estimator, radii)


def autocorr_batch(pairs, radii, est_type="landy-szalay", precision=None):
    """Many independent auto-correlations in one GPU call: ``pairs`` is a
    list of (data_tree, rand_tree).  DD, DR and RR of every pair, at every
    radius, share one tile-pair enumeration and one persistent kernel launch.
    Returns one autocorr_results per pair, as _autocorr would (no reference
    counterpart: the reference runs each job, radius and pair type as its own
    pair_count call, clustering.py:218-230)."""
    estimator = _estimator(est_type)
    jobs = []
    for d, r in pairs:
        if not d.N_fields == r.N_fields:
            raise ValueError("data and random trees must have same number of fields")
        jobs += [(d, d), (r, d), (r, r)]
    out = _count_jobs(jobs, radii, precision)
    return _autocorr_results_of(pairs, out, estimator, radii)


class count_plan(object):
    """The batched count of ``pairs`` ((tree, query tree) pairs, as
    _count_jobs takes them) at ``radii``, made once and run many times: the
    library sizes the call on creation and records its whole device sequence
    into one HIP graph; run() replays it and returns the cumulative counts,
    one (len(radii), Nq, Nt) array per pair, equal to _count_jobs's.
    ``tuning`` is a dict of kd_tuning fields (blocks_per_cu, chunk,
    enum_blocks_per_cu; missing or 0: the library's defaults).  The plan keeps
    its trees alive.  No reference counterpart (the reference recounts
    everything per call, clustering.py:218-230)."""

    def __init__(self, pairs, radii, precision=None, tuning=None):
        radii = np.ascontiguousarray(radii, dtype=np.float64).ravel()
        if np.unique(radii).size > MAX_RADII_PER_CALL:
            raise ValueError("a plan takes at most {} distinct radii".format(MAX_RADII_PER_CALL))
        self.pairs = list(pairs)
        self.radii = radii
        self.tuning = dict(tuning or {})
        self._outs = [np.zeros((radii.size, q.N_fields, t.N_fields), dtype=np.int64) for t, q in self.pairs]
        jobs = (kd_job * len(self.pairs))()
        for i, (t, q) in enumerate(self.pairs):
            jobs[i].tree = t.ctree
            jobs[i].query = q.ctree
            jobs[i].num_fields_query = q.N_fields
            jobs[i].counts = self._outs[i].ctypes.data_as(POINTER(c_longlong))
        tu = kd_tuning(**{k: int(v) for k, v in self.tuning.items()})
        self._h = kdlib.pair_count_plan(jobs, len(self.pairs), _dp(radii), radii.size,
                                        _precision_code(precision), ctypes.byref(tu))
        if not self._h:
            raise RuntimeError("pair_count_plan failed: " + _err())
        self.items = int(kdlib.pair_count_plan_items(self._h))
        self.last_stats = None

    def run(self, record=True):
        """Replay the graph; returns copies of the counts."""
        global last_stats
        if not self._h:
            raise RuntimeError("the plan was closed")
        st = kd_stats()
        if kdlib.pair_count_plan_run(self._h, ctypes.byref(st)) != 0:
            raise RuntimeError("pair_count_plan_run failed: " + _err())
        self.last_stats = st.as_dict()
        if record:
            last_stats = self.last_stats
        return [o.copy() for o in self._outs]

    def close(self):
        if getattr(self, "_h", None):
            kdlib.pair_count_plan_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


class autocorr_plan(object):
    """autocorr_batch made once and run many times: DD, DR and RR of every
    (data_tree, rand_tree) pair at every radius in one graph launch per run
    (count_plan).  run() returns one autocorr_results per pair, equal to
    autocorr_batch's."""

    def __init__(self, pairs, radii, est_type="landy-szalay", precision=None, tuning=None):
        self.estimator = _estimator(est_type)
        self.pairs = list(pairs)
        self.radii = radii
        jobs = []
        for d, r in self.pairs:
            if not d.N_fields == r.N_fields:
                raise ValueError("data and random trees must have same number of fields")
            jobs += [(d, d), (r, d), (r, r)]
        self.plan = count_plan(jobs, radii, precision, tuning)

    def run(self):
        out = self.plan.run()
        return _autocorr_results_of(self.pairs, out, self.estimator, self.radii)

    def close(self):
        self.plan.close()


def _autocorr_results_of(pairs, out, estimator, radii):
    res = []
    for i, (d, r) in enumerate(pairs):
        dd, dr, rr = (np.diff(a, axis=0) for a in out[3 * i:3 * i + 3])
        res.append(autocorr_results(dd, dr, rr, d, r, estimator, radii))
    return res


class _Resampled(object):
    """field_sizes of the catalogue np.repeat(tree.points, w), as ``tree``
    would compute them (clustering.py:313-314, :338-349: with fields, the
    per-field point counts -- given, when the GPU summed them; without, the
    number of unique points)."""

    def __init__(self, t, w, field_sizes=None):
        self.N_fields = t.N_fields
        if t.fields is not None and field_sizes is not None:
            self.field_sizes = np.asarray(field_sizes, dtype=np.int64)
            return
        w = np.asarray(w, dtype=np.int64)
        if t.fields is None:
            # distinct points of the repeated catalogue: unique points with a copy
            u, inv = np.unique(t.points + 0.0, axis=0, return_inverse=True)
            self.field_sizes = np.array(int(np.count_nonzero(np.bincount(inv.ravel(), weights=w))))
        else:
            self.field_sizes = np.bincount(t.fields, weights=w, minlength=t.N_fields)[:t.N_fields].astype(np.int64)


def _set_weights(w, nsets):
    w = np.asarray(w)
    return w[None, :] if w.ndim == 1 else w


def autocorr_weighted(data_tree, rand_tree, radii, weights, est_type="landy-szalay", precision=None):
    """Auto-correlations of S resampled data catalogues in one GPU call:
    ``weights`` is (S, N_data) integer multiplicities of the data points
    (bootstrap draws), and set s gives the result of _autocorr on
    np.repeat(data, weights[s]) against the unweighted randoms.  DD is
    weighted on both sides (the data tree must have no duplicate points), DR
    on the query side, RR is counted once.  ``precision``: "fp64" or
    "fp32_guarded" (None: the module's form).  No reference counterpart: the
    reference would rebuild and recount every catalogue."""
    if not data_tree.N_fields == rand_tree.N_fields:
        raise ValueError("data and random trees must have same number of fields")
    estimator = _estimator(est_type)
    weights = np.asarray(weights)
    nsets = 1 if weights.ndim == 1 else weights.shape[0]
    # RR (unweighted) runs beside the weighted call, on its own stream
    ((dd, dr), fsz), (rr,) = _concurrently(
        lambda: _count_jobs([(rand_tree, rand_tree)], radii, precision, record=False),
        lambda: _count_weighted([(data_tree, data_tree, weights, weights),
                                 (rand_tree, data_tree, None, weights)], radii, nsets, sizes=True, prec=precision))
    rr = np.diff(rr, axis=0)
    ws = _set_weights(weights, nsets)
    dsz = fsz[0][1]
    return [autocorr_results(np.diff(dd[s], axis=0), np.diff(dr[s], axis=0), rr,
                             _Resampled(data_tree, ws[s], dsz[s]), rand_tree, estimator, radii)
            for s in range(nsets)]


def crosscorr_weighted(data_tree_1, data_tree_2, rand_tree_1, rand_tree_2, radii, weights_1, weights_2,
                       est_type="landy-szalay", precision=None):
    """Cross-correlations for S weight sets of both data catalogues in one GPU
    call (weights_k: (S, N_data_k) multiplicities).  D1D2 is weighted on both
    sides, D1R2 by weights_1, D2R1 by weights_2, R1R2 unweighted and counted
    once; the larger data tree is the tree side of D1D2 as in _crosscorr, and
    it must have no duplicate points.  ``precision``: "fp64" or
    "fp32_guarded" (None: the module's form).  Returns one crosscorr_results
    per set."""
    if not (data_tree_1.N_fields == data_tree_2.N_fields == rand_tree_1.N_fields == rand_tree_2.N_fields):
        raise ValueError("data and random trees must have same number of fields")
    estimator = _estimator(est_type)
    w1, w2 = np.asarray(weights_1), np.asarray(weights_2)
    nsets = 1 if w1.ndim == 1 else w1.shape[0]
    one_is_tree = data_tree_1.size > data_tree_2.size
    if one_is_tree:
        d1d2 = (data_tree_1, data_tree_2, w1, w2)
    else:
        d1d2 = (data_tree_2, data_tree_1, w2, w1)
    r1r2_pair = (rand_tree_1, rand_tree_2) if rand_tree_1.size > rand_tree_2.size else (rand_tree_2, rand_tree_1)
    # R1R2 (unweighted) runs beside the weighted call, on its own stream
    ((c12, c1r2, c2r1), fsz), (rr,) = _concurrently(
        lambda: _count_jobs([r1r2_pair], radii, precision, record=False),
        lambda: _count_weighted([d1d2, (rand_tree_2, data_tree_1, None, w1), (rand_tree_1, data_tree_2, None, w2)],
                                radii, nsets, sizes=True, prec=precision))
    sz1, sz2 = (fsz[0][0], fsz[0][1]) if one_is_tree else (fsz[0][1], fsz[0][0])
    rr = np.diff(rr, axis=0)
    ws1, ws2 = _set_weights(w1, nsets), _set_weights(w2, nsets)
    return [crosscorr_results(np.diff(c12[s], axis=0), np.diff(c1r2[s], axis=0), np.diff(c2r1[s], axis=0), rr,
                              _Resampled(data_tree_1, ws1[s], sz1[s]), _Resampled(data_tree_2, ws2[s], sz2[s]),
                              rand_tree_1, rand_tree_2, estimator, radii)
            for s in range(nsets)]


class tree:
    """A k-d tree of 3-D points built on the GPU (clustering.py:293-352).

    Attributes as in the reference: ``points``, ``fields``, ``N_fields``,
    ``field_sizes``, ``ctree`` and ``size`` (the number of unique points)."""

    def __init__(self, points, fields=None):
        points = validate_points(points)
        if fields is None:
            self.fields = None
            self.N_fields = 1
        else:
            fields = validate_fields(fields, points)
            # the ids must be exactly 0..N-1 (clustering.py:304-307); one
            # bincount gives both that check and the field sizes (:336-338)
            counts = np.bincount(fields) if fields.size and fields.min() >= 0 else None
            if counts is None or not np.all(counts > 0):
                raise ValueError("fields must be a 1-D array of integers from "
                                 "0 to the number of unique fields minus one")
            self.N_fields = counts.size
            if self.N_fields < 2:
                raise ValueError("At least two unique field IDs must be "
                                 "provided if using error fields")
            self.fields = fields
            self.field_sizes = counts
        self.points = points
        self.ctree = None
        self._make_tree()
        self.size = self.ctree.size
        if fields is None:
            self.field_sizes = np.array(self.size)

    def _calc_field_sizes(self):
        return np.bincount(self.fields, minlength=self.N_fields)[:self.N_fields]

    def _make_tree(self):
        ct = kdlib.tree_construct(_dp(self.points), _ip(self.fields), c_int(self.points.shape[0]),
                                  c_int(self.N_fields))
        if not ct.dev:
            raise RuntimeError("tree_construct failed: " + _err())
        self.ctree = ct

    def node_array(self):
        """The node array (memsize records, slot 0 unused) as a structured
        numpy array with the fields of src/kdtree.h:13-25."""
        ct = self.ctree
        if kdlib.tree_fetch_nodes(ctypes.byref(ct)) != 0:
            raise RuntimeError(_err())
        self.ctree = ct
        buf = (ctypes.c_char * (40 * ct.memsize)).from_address(ctypes.addressof(ct.node_data.contents))
        return np.frombuffer(bytes(buf), dtype=NODE_DTYPE)

    def heap_permutation(self):
        """(perm, inorder): heap index -> unique index and -> median position."""
        m = self.ctree.memsize
        perm = np.empty(m, dtype=np.int32)
        ino = np.empty(m, dtype=np.int32)
        if kdlib.tree_fetch_perm(self.ctree, _ip(perm), _ip(ino)) != 0:
            raise RuntimeError(_err())
        return perm, ino

    def save(self, path):
        """Write the tree to ``path`` as a numpy .npz archive (plain arrays,
        no pickled objects): the raw points and fields, the node array in the
        record layout of src/kdtree.h:13-25 (memsize 40-byte records, slot 0
        unused, padding zero) and the heap -> unique-index permutation.

        ``tree.load`` rebuilds the tree from the points and checks that the
        rebuilt node array and permutation equal the saved ones byte for
        byte: the GPU build of 10M points (about 0.25 s) is faster than
        reading its 671 MB node array back, so the file is an interchange
        and verification format, not a cache."""
        perm, _ = self.heap_permutation()
        arrays = {"format": np.array([TREE_FILE_VERSION, self.N_fields, self.points.shape[0], self.size],
                                     dtype=np.int64),
                  "points": self.points, "nodes": self.node_array().view(np.uint8).reshape(-1, 40), "perm": perm}
        if self.fields is not None:
            arrays["fields"] = self.fields
        np.savez(path, **arrays)

    @classmethod
    def load(cls, path, verify=True):
        """The tree written by ``save``: built again on the GPU from the
        saved points and fields; with ``verify`` the rebuilt node array and
        permutation must equal the saved ones (ValueError otherwise)."""
        with np.load(path, allow_pickle=False) as z:
            names = set(z.files)
            if not {"format", "points", "nodes", "perm"} <= names:
                raise ValueError("not a tree file: missing arrays")
            fmt = z["format"]
            if fmt.shape != (4,) or int(fmt[0]) != TREE_FILE_VERSION:
                raise ValueError("unsupported tree file version")
            points = z["points"]
            fields = z["fields"] if "fields" in names else None
            nodes = z["nodes"]
            perm = z["perm"]
        if points.ndim != 2 or points.shape != (int(fmt[2]), 3) or points.dtype != np.float64:
            raise ValueError("tree file: points must be ({:d}, 3) float64".format(int(fmt[2])))
        if nodes.dtype != np.uint8 or nodes.ndim != 2 or nodes.shape[1] != 40 or perm.shape != (nodes.shape[0],):
            raise ValueError("tree file: node array and permutation do not match")
        t = cls(points, fields)
        if t.N_fields != int(fmt[1]) or t.size != int(fmt[3]):
            raise ValueError("tree file: field count or unique size differs from the rebuilt tree")
        if verify:
            mine = t.node_array().view(np.uint8).reshape(-1, 40)
            perm2, _ = t.heap_permutation()
            if mine.shape != nodes.shape or not np.array_equal(mine, nodes) or not np.array_equal(perm2, perm):
                raise ValueError("tree file: the node array rebuilt from the saved points differs from the saved one")
        return t

    def __del__(self):
        ct = getattr(self, "ctree", None)
        if ct is not None and ct.dev:
            kdlib.destroy(ct)
            self.ctree = None


class query_set(object):
    """Raw query points (and fields) laid out for the pair kernels without a
    k-d build (``query_construct``): the query side of a job whose points are
    not a tree of their own, e.g. a sample of a catalogue's rows counted
    against the full tree.  Same attributes as ``tree`` for the query role:
    ``points``, ``fields``, ``N_fields``, ``ctree``."""

    def __init__(self, points, fields=None, N_fields=None):
        self.points = validate_points(points)
        self.fields = None if fields is None else validate_fields(fields, self.points)
        if N_fields is None:
            N_fields = 1 if self.fields is None else int(self.fields.max()) + 1
        self.N_fields = int(N_fields)
        self.ctree = None
        ct = kdlib.query_construct(_dp(self.points), _ip(self.fields), c_int(self.points.shape[0]),
                                   c_int(self.N_fields))
        if not ct.dev:
            raise RuntimeError("query_construct failed: " + _err())
        self.ctree = ct
        self.size = ct.size

    def __del__(self):
        ct = getattr(self, "ctree", None)
        if ct is not None and ct.dev:
            kdlib.destroy(ct)
            self.ctree = None


def trees(catalogues, max_workers=8):
    """Trees of several catalogues, built concurrently: each build is a
    library call on its own HIP stream, so many small builds overlap on the
    GPU.  ``catalogues`` holds point arrays or (points, fields) pairs; the
    trees come back in the same order (errors as from ``tree``)."""
    specs = [c if isinstance(c, tuple) else (c, None) for c in catalogues]
    if len(specs) <= 1 or max_workers <= 1:
        return [tree(p, f) for p, f in specs]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(max_workers, len(specs))) as ex:
        return list(ex.map(lambda pf: tree(*pf), specs))


NODE_DTYPE = np.dtype([("value", "<f8", (3,)), ("field", "u1"), ("weight", "u1"), ("pad0", "u1", (6,)),
                       ("bits", "<u4"), ("pad1", "<u4")])
TREE_FILE_VERSION = 1
