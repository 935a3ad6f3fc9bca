#!/usr/bin/env python3
"""Pair-count benchmark on one MI355X (3-D xi(r), 1M data / 10M randoms).

One step = DD + DR + RR cumulative counts at 20 radii, 16 jackknife fields,
for the trees already resident in HBM (one pair_count_jobs call).  Prints one
JSON line:

  value          pairs counted per second: the ordered (query, tree) pairs
                 within the largest radius over DD, DR and RR, as the
                 reference's counts hold them, divided by the step time
                 (whole job, one GPU); pair_evals_per_s beside it counts the
                 (query point, tree point) distances the pair kernel evaluated
  ms_per_step    DD+DR+RR wall time
  roofline       the pair kernels: 8 FLOP per evaluated pair / kernel time
                 (HIP events on the library's stream), against the MI355X
                 vector peak of the precision form; traffic from the rocprofv3
                 FETCH_SIZE / WRITE_SIZE record under profiles/
  compare        the other precision form on the same trees (same counts)
  tree_build     GPU k-d builds of both catalogues: wall time and the
                 algorithmic bytes of the build over that time
  cpu_baseline   the reference's own pair_count (src/kdquery.c compiled from
                 source into oracle/_ref) on all host cores, timed on a seeded
                 sample of the query points (every thread busy) and scaled per
                 pair type; the GPU counts the same sampled rows against the
                 full trees (counts_match_reference)

Usage: python bench.py [--steps K] [--warmup W] [--precision fp64|fp32_guarded]
       python bench.py --workload c3|c4|c5   (angular 2M/20M; 256 batched jobs
       in one graph launch; 200 bootstrap weight sets)

Both forms give identical counts.  FP64 is the default: on CDNA4 FP32 and FP64
vector instructions issue at the same rate unless FP32 is packed, so the
guarded FP32 form (FP32 distances on the matrix cores, a guard-band compare
per radius) runs slower than FP64 here.
"""
import argparse
import contextlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp32_guarded": 157.3, "fp64": 78.6}  # MI355X vector FP32 / FP64
FLOP_PER_PAIR = 8  # 3 sub + 3 mul + 2 add


def c2_workload(scale=1.0):
    """1M points (half uniform, half in 2000 Gaussian clusters of sigma 0.005,
    wrapped into the unit cube), 10M uniform randoms, fields 4*floor(4x) +
    floor(4y), 20 radii log-spaced 1e-3..1e-1 (seed 2)."""
    from tests import cases
    nd = int(1_000_000 * scale)
    nr = int(10_000_000 * scale)
    rng = np.random.default_rng(2)
    U = rng.random((nd // 2, 3))
    C = rng.random((max(1, 2000 * nd // 1_000_000), 3))
    per = (nd - nd // 2) // C.shape[0]
    K = np.repeat(C, per, 0) + rng.normal(0, 0.005, (C.shape[0] * per, 3))
    K -= np.floor(K)
    K[K == 1.0] = 0.0
    D = np.vstack([U, K])
    R = rng.random((nr, 3))
    fd = (4 * np.floor(4 * D[:, 0]) + np.floor(4 * D[:, 1])).astype(np.int32)
    fr = (4 * np.floor(4 * R[:, 0]) + np.floor(4 * R[:, 1])).astype(np.int32)
    radii = np.logspace(-3, -1, 20)
    del cases
    return D, fd, R, fr, radii


def c3_workload():
    """Angular w(theta) on a sky patch (seed 3): RA uniform in [0, 60) deg,
    DEC with density cos(DEC) on [-30, 30) deg; 2,000,000 data points of which
    600,000 sit in 5,000 clusters of 120 (Gaussian, sigma 0.05 deg, RA wrapped
    mod 60, DEC reflected at +-30), 20,000,000 uniform randoms; 15 angles
    log-spaced 0.001..1 deg.  Fields: the library's automatic division into 64
    (corruscant_amd.fields) is made in run_c3; returned here is the 8 x 8
    equal-area (RA, sin DEC) grid used with --grid-fields."""
    rng = np.random.default_rng(3)

    def patch(n):
        return rng.uniform(0, 60, n), np.degrees(np.arcsin(rng.uniform(-0.5, 0.5, n)))

    ura, udec = patch(1_400_000)
    cra, cdec = patch(5000)
    cra, cdec = np.repeat(cra, 120), np.repeat(cdec, 120)
    kdec = cdec + rng.normal(0, 0.05, cdec.size)
    kra = cra + rng.normal(0, 0.05, cra.size) / np.cos(np.radians(cdec))
    kra = np.mod(kra, 60.0)
    kdec = np.where(kdec > 30, 60 - kdec, np.where(kdec < -30, -60 - kdec, kdec))
    dra, ddec = np.concatenate([ura, kra]), np.concatenate([udec, kdec])
    rra, rdec = patch(20_000_000)

    def fields(ra, dec):
        fx = np.minimum((ra * (8 / 60.0)).astype(np.int32), 7)
        fy = np.minimum(((np.sin(np.radians(dec)) + 0.5) * 8).astype(np.int32), 7)
        return (8 * fx + fy).astype(np.int32)

    return dra, ddec, fields(dra, ddec), rra, rdec, fields(rra, rdec), np.logspace(-3, 0, 15)


def run_c3(args):
    from corruscant_amd import clustering
    from corruscant_amd.twopoint import angular
    dra, ddec, fd, rra, rdec, fr, deg = c3_workload()
    fields_s = None
    if not args.grid_fields:
        from corruscant_amd import fields
        t0 = time.perf_counter()
        tmpl = fields.template(64, ra=dra, dec=ddec)
        with contextlib.redirect_stdout(sys.stderr):  # the division's progress messages
            fd = tmpl.create_fields(min_ratio=0.8).astype(np.int32)
        fr = tmpl.assign_fields(ra=rra, dec=rdec).astype(np.int32)
        fields_s = time.perf_counter() - t0
        print("64 fields at Nside %d in %.2f s" % (tmpl.Nside, fields_s), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    D = angular.cartesian(dra, ddec)
    R = angular.cartesian(rra, rdec)
    conv_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    td, tr = clustering.trees([(D, fd), (R, fr)])
    build_s = time.perf_counter() - t0
    print("converted in %.2f s, trees built in %.2f s" % (conv_s, build_s), file=sys.stderr, flush=True)
    clustering.set_precision(args.precision)
    for _ in range(args.warmup):
        angular.autocorr(td, tr, deg)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        res = angular.autocorr(td, tr, deg)
        times.append(time.perf_counter() - t0)
        print("step %.3f s" % times[-1], file=sys.stderr, flush=True)
    st = clustering.last_stats
    sec = float(np.mean(times))
    achieved = st["evaluated_pairs"] * FLOP_PER_PAIR / (st["count_ms"] / 1e3) / 1e12
    return {
        "metric": "pairs counted/s (angular w(theta) DD+DR+RR pairs in the angular bins, 2M data / 20M randoms, "
                  "exact counts)",
        "value": int(res.dd.sum() + res.dr.sum() + res.rr.sum()) / sec, "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64 recheck" if args.precision == "fp32_guarded" else "f64",
        "data": "synthetic",
        "config": {"workload": "angular w(theta) on a 60 x 60 deg patch: 2,000,000 data (30%% in 5000 clusters "
                               "of sigma 0.05 deg) / 20,000,000 randoms, 64 fields (%s), 15 angles "
                               "0.001..1 deg as chord radii" % ("8 x 8 equal-area RA/sin DEC grid" if args.grid_fields
                                                                else "automatic HEALPix division"),
                   "precision": args.precision},
        "fields_s": fields_s, "coords_s": conv_s, "tree_build_s": build_s, "evaluated_pairs": st["evaluated_pairs"],
        "pair_evals_per_s": st["evaluated_pairs"] / sec,
        "useful_pairs_in_bins": int(res.dd.sum() + res.dr.sum() + res.rr.sum()),
        "work_items": st["work_items"], "enumerate_ms": st["enumerate_ms"],
        "rechecked_pairs": st["rechecked_pairs"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS[args.precision], "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS[args.precision], "traffic": None,
                     "kernel": ("k_pairs32 + k_pairs64 + k_pairs64_mixed" if args.precision == "fp32_guarded"
                       else "k_pairs64 + k_pairs64_mixed"),
                     "kernel_ms": st["count_ms"]},
    }


def c4_workload():
    """256 independent auto-correlation jobs (seed 4): data sizes log-uniform
    in [1e3, 2e5], uniform points in the unit cube, randoms 5x the data,
    fields floor(8x), 12 radii log-spaced 2e-3..2e-1."""
    rng = np.random.default_rng(4)
    jobs = []
    for _ in range(256):
        n = int(round(np.exp(rng.uniform(np.log(1e3), np.log(2e5)))))
        D = rng.random((n, 3))
        R = rng.random((5 * n, 3))
        jobs.append((D, np.minimum((D[:, 0] * 8).astype(np.int32), 7), R,
                     np.minimum((R[:, 0] * 8).astype(np.int32), 7)))
    return jobs, np.logspace(np.log10(2e-3), np.log10(2e-1), 12)


def c5_workload(nsets=200):
    """Bootstrap-weighted cross-correlation (seed 5): A = 1M points built as
    the 1M data catalogue above, B = 4M uniform points, R = 10M uniform
    randoms, 16 fields 4*floor(4x) + floor(4y), per-point multiplicities
    Poisson(1) as uint8 for each of nsets weight sets on A and on B, 20 radii
    log-spaced 1e-3..1e-1."""
    rng = np.random.default_rng(5)
    U = rng.random((500_000, 3))
    C = rng.random((2000, 3))
    K = np.repeat(C, 250, 0) + rng.normal(0, 0.005, (500_000, 3))
    K -= np.floor(K)
    K[K == 1.0] = 0.0
    A = np.vstack([U, K])
    B = rng.random((4_000_000, 3))
    R = rng.random((10_000_000, 3))
    f = lambda X: (4 * np.floor(4 * X[:, 0]) + np.floor(4 * X[:, 1])).astype(np.int32)
    wA = rng.poisson(1.0, (nsets, A.shape[0])).astype(np.uint8)
    wB = rng.poisson(1.0, (nsets, B.shape[0])).astype(np.uint8)
    return A, f(A), B, f(B), R, f(R), wA, wB, np.logspace(-3, -1, 20)


def run_c4(args):
    """The batched workload through one graph launch per step: the autotuner
    picks the tile size and the pair kernels' launch shape on the workload
    itself (graph replays, median GPU time), then a plan at that choice is
    replayed ``steps`` times.  A step is one replay plus the host's
    cumulative-count pass and the results objects of all 256 jobs."""
    from corruscant_amd import autotune, clustering
    jobs, radii = c4_workload()
    npts = sum(D.shape[0] + R.shape[0] for D, _, R, _ in jobs)
    t0 = time.perf_counter()
    clustering.set_tile_size(autotune.DEFAULT_TILE)
    built = clustering.trees([c for D, fd, R, fr in jobs for c in ((D, fd), (R, fr))],
                             max_workers=args.build_workers)
    build_s = time.perf_counter() - t0
    print("%d jobs, %d points, trees built in %.2f s" % (len(jobs), npts, build_s), file=sys.stderr, flush=True)
    del built
    if args.tune:
        tuned = autotune.tune_autocorr(jobs, radii, args.precision, tile_sizes=(32, 48, 64),
                                       blocks_per_cu=(4, 8, 12, 16), chunks=(2, 4, 16),
                                       enum_blocks_per_cu=(16, 32), reps=2, log=sys.stderr)
        pairs, tile, tuning, table = tuned.trees, tuned.tile_size, tuned.tuning, tuned.as_dict()["candidates"]
    else:
        pairs = autotune._build(jobs, autotune.DEFAULT_TILE)
        tile, tuning, table = autotune.DEFAULT_TILE, {}, []
    plan = clustering.autocorr_plan(pairs, radii, precision=args.precision, tuning=tuning)
    for _ in range(args.warmup):
        plan.run()
    times, gpu_ms = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        res = plan.run()
        times.append(time.perf_counter() - t0)
        gpu_ms.append(plan.plan.last_stats["count_ms"])
        print("step %.3f s (graph %.1f ms)" % (times[-1], gpu_ms[-1]), file=sys.stderr, flush=True)
    st = dict(plan.plan.last_stats)
    sec = float(np.mean(times))
    # the eager call on the same trees: same counts, its own enumeration and
    # pair-kernel times
    t0 = time.perf_counter()
    eager = clustering.autocorr_batch(pairs, radii, precision=args.precision)
    eager_s = time.perf_counter() - t0
    est = clustering.last_stats
    same = all(np.array_equal(a.dd, b.dd) and np.array_equal(a.dr, b.dr) and np.array_equal(a.rr, b.rr)
               for a, b in zip(res, eager))
    useful = int(sum(r.dd.sum() + r.dr.sum() + r.rr.sum() for r in res))
    achieved = est["evaluated_pairs"] * FLOP_PER_PAIR / (est["count_ms"] / 1e3) / 1e12
    plan.close()
    return {
        "metric": "pairs counted/s (256 batched auto-correlation jobs, DD+DR+RR pairs in the radius bins, "
                  "exact counts, one graph launch)",
        "value": useful / sec, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64 recheck" if args.precision == "fp32_guarded" else "f64",
        "data": "synthetic",
        "config": {"workload": "256 3-D auto-correlations, data log-uniform in [1e3, 2e5] points, randoms 5x, "
                               "8 fields, 12 radii 2e-3..2e-1, all jobs in one graph launch", "points": npts,
                   "precision": args.precision},
        "graph": {"gpu_ms_median": float(np.median(gpu_ms)), "work_items": st["work_items"],
                  "tile_size": tile, "tuning": tuning},
        "autotune": table,
        "eager": {"wall_s": eager_s, "enumerate_ms": est["enumerate_ms"], "pair_kernels_ms": est["count_ms"],
                  "counts_identical": bool(same)},
        "tree_build_s": build_s, "evaluated_pairs": st["evaluated_pairs"], "useful_pairs_in_bins": useful,
        "pair_evals_per_s": st["evaluated_pairs"] / sec,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS[args.precision], "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS[args.precision], "traffic": None,
                     "kernel": "k_pairs32 + k_pairs64 + k_pairs64_mixed" if args.precision == "fp32_guarded"
                     else "k_pairs64 + k_pairs64_mixed",
                     "kernel_ms": est["count_ms"]},
    }


def run_c5(args):
    from corruscant_amd import clustering
    A, fa, B, fb, R, fr, wA, wB, radii = c5_workload(args.sets)
    print("weights ready", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ta, tb, tr = clustering.trees([(A, fa), (B, fb), (R, fr)])
    build_s = time.perf_counter() - t0
    print("trees built in %.2f s" % build_s, file=sys.stderr, flush=True)

    def step():
        return clustering.crosscorr_weighted(ta, tb, tr, tr, radii, wA, wB, precision=args.precision)

    for _ in range(args.warmup):
        step()
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        res = step()
        times.append(time.perf_counter() - t0)
        print("step %.3f s" % times[-1], file=sys.stderr, flush=True)
    sec = float(np.mean(times))
    # the weighted call's statistics (the unweighted R1R2 call follows it)
    clustering._count_weighted([(tb, ta, wB, wA), (tr, ta, None, wA), (tr, tb, None, wB)], radii, args.sets,
                               prec=args.precision)
    st = clustering.last_stats
    weighted = int(sum(r.d1d2.sum() + r.d1r2.sum() + r.d2r1.sum() for r in res))
    return {
        "metric": "weighted pair updates/s (D1D2, D1R2, D2R1 for every weight set, plus R1R2)",
        "value": st["evaluated_pairs"] * args.sets / sec, "unit": "weighted pair updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64 recheck" if args.precision == "fp32_guarded" else "f64", "data": "synthetic",
        "config": {"workload": "bootstrap-weighted 3-D cross-correlation: A 1M clustered, B 4M uniform, "
                               "randoms 10M, 16 fields, 20 radii 1e-3..1e-1, Poisson(1) uint8 multiplicities",
                   "weight_sets": args.sets, "precision": args.precision},
        "fp32_pairs_per_set": st["fp32_pairs"], "rechecked_pairs": st["rechecked_pairs"],
        "tree_build_s": build_s,
        "weighted_call_ms": st["total_ms"], "weighted_kernel_ms": st["count_ms"],
        "enumerate_ms": st["enumerate_ms"], "evaluated_pairs_per_set": st["evaluated_pairs"],
        "sum_of_weighted_counts": weighted, "work_items": st["work_items"],
    }


def cpu_baseline(tree_d, tree_r, D, fd, R, fr, radii, budget_s, cores, precision):
    """The reference's pair_count on a seeded sample of query points, one call
    per radius and pair type as clustering._autocorr makes them, on the node
    arrays of this library's trees (bit-identical to the reference's
    tree_construct).

    Every thread gets work: the reference hands out chunks of
    max(1000, m / num_threads) queries (src/kdquery.c:181), so each pair
    type's sample holds at least 1000 x threads queries.  Times are scaled by
    N / sample per pair type.  Pair evaluations are the nodes the traversal
    visits, counted by the oracle's restatement of the same traversal.  The
    same sampled rows are also counted by the GPU library against the full
    trees (``query_set`` jobs), and the two sets of counts are compared."""
    from corruscant_amd import clustering
    from oracle import oracle as O
    if not O.ref_available():
        return None
    rng = np.random.default_rng(12345)
    nodes = {"D": tree_d.node_array(), "R": tree_r.node_array()}
    gpu_trees = {"D": tree_d, "R": tree_r}
    jobs = [("DD", "D", D, fd), ("DR", "R", D, fd), ("RR", "R", R, fr)]
    floor = 1000 * cores  # one chunk of 1000 queries per thread
    # calibrate: per-query cost of each type at the largest radius (which
    # dominates the traversal work), on one chunk per thread
    per_query = {}
    for name, t, Q, qf in jobs:
        idx = rng.choice(Q.shape[0], min(floor, Q.shape[0]), replace=False)
        rt = O.RefNodes(nodes[t], 16)
        t0 = time.perf_counter()
        rt.pair_count(Q[idx], qf[idx], 16, radii[-1], nthreads=max(1, min(cores, idx.size // 1000)))
        per_query[name] = (time.perf_counter() - t0) / idx.size
    full_est = sum(per_query[n] * Q.shape[0] for n, _, Q, _ in jobs) * radii.size / 3.0
    frac = min(1.0, budget_s / max(full_est, 1e-9))
    total_s = 0.0
    evals = 0
    useful = 0
    sample, stage_s = [], {}
    match = True
    for name, t, Q, qf in jobs:
        m = min(Q.shape[0], max(floor, int(round(Q.shape[0] * frac))))
        idx = np.sort(rng.choice(Q.shape[0], m, replace=False))
        rt = O.RefNodes(nodes[t], 16)
        thr = max(1, min(cores, m // 1000))
        t0 = time.perf_counter()
        cum = rt.cumulative(Q[idx], qf[idx], 16, radii, nthreads=thr)
        dt = time.perf_counter() - t0
        ot = O.OracleNodes(nodes[t], 16)
        v = 0
        for r in radii:
            vv, cnt = ot.traversal_visits(Q[idx], qf[idx], 16, r, nthreads=cores)
            v += vv
        assert np.array_equal(cnt, cum[-1]), "oracle traversal differs from the reference"
        # the GPU's counts of the same rows against the same (full) tree
        qs = clustering.query_set(Q[idx], qf[idx], 16)
        gpu = clustering._count_jobs([(gpu_trees[t], qs)], radii, precision, record=False)[0]
        del qs
        match = match and bool(np.array_equal(gpu, cum))
        scale = Q.shape[0] / m
        stage_s[name] = dt * scale
        total_s += dt * scale
        evals += v * scale
        useful += int(cum[-1].sum()) * scale
        sample.append("%s %d of %d queries (%d threads, %d chunks of >= 1000)" % (name, m, Q.shape[0], thr,
                                                                               max(1, m // max(1000, m // thr))))
    return {
        "value": useful / total_s,
        "unit": "pairs/s",
        "cores": cores,
        "kind": "reference",
        "sample": "; ".join(sample) + "; 20 radii, one pair_count call per radius; "
                  "times scaled by N/sample to DD+DR+RR = %.1f s; pair evals = tree nodes the traversal visits" % total_s,
        "dd_dr_rr_wall_s": total_s,
        "stage_s": stage_s,
        "pair_evals_per_s": evals / total_s,
        "sampled_rows_counts_match_gpu": match,
    }


def build_bytes(n, n_u):
    """Lower bound on the bytes a k-d build of n points (n_u unique) moves:
    read xyz + field, write the node array, per level of segments of 4 or more
    points 3 x (index + coordinate) for the variance and 2 x (read + write) of
    an index for the partition, and one read of each sort key per axis."""
    levels = 0
    s = n_u
    while s >= 4:
        levels += 1
        s = (s - 1) // 2
    pow2 = 1 << int(n_u).bit_length()
    return 28 * n + 40 * pow2 + levels * 52 * n_u + 3 * 24 * n_u


def build_line(td, tr, D, R, build_s):
    b = build_bytes(D.shape[0], td.size) + build_bytes(R.shape[0], tr.size)
    return {"wall_s": build_s, "algorithmic_bytes": b, "GBps": b / build_s / 1e9,
            "note": "both trees built concurrently (one stream each), wall time including the numpy "
                    "validation and the host-to-device copy of the points"}


def copy_calibration():
    """Three device-to-device copies of 1 GiB (2^28 float32): each reads and
    writes exactly 2^30 bytes of HBM."""
    import torch
    src = torch.full((1 << 28,), 1.0, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30, "dispatches": 3}))


def measured_traffic(name):
    """HBM bytes per launch of a kernel on a workload, as recorded in
    profiles/<name>.md by two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE)
    scaled by the same counters on a 1 GiB copy: the line
    "traffic_bytes_per_launch: <bytes>".  None when no such record exists."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name + ".md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp64", choices=sorted(PEAK_TFLOPS))
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the 1M/10M sizes")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip timing the other precision form")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5"],
                    help="c2: 1M/10M auto-correlation (the default line); c3: angular 2M/20M; "
                         "c4: 256 batched jobs; c5: bootstrap-weighted cross-correlation")
    ap.add_argument("--sets", type=int, default=200, help="weight sets of the c5 workload")
    ap.add_argument("--build-workers", type=int, default=8,
                    help="c4: host threads that build the 512 trees (one stream each)")
    ap.add_argument("--no-tune", dest="tune", action="store_false",
                    help="c4: skip the autotuner (tile size 64, default launch shape)")
    ap.add_argument("--grid-fields", action="store_true", help="c3: 8 x 8 RA/DEC grid instead of the automatic division")
    ap.add_argument("--build-only", action="store_true",
                    help="build the 10M-random tree of the default workload once (for build profiles) and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device three times: the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (got --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.build_only:
        from corruscant_amd import clustering
        _, _, R, fr, _ = c2_workload(args.scale)
        clustering.tree(R[:1000], fr[:1000])  # library and device set-up outside the timed build
        t0 = time.perf_counter()
        tr = clustering.tree(R, fr)
        build_s = time.perf_counter() - t0
        print(json.dumps({"build_only": {"points": R.shape[0], "unique": tr.size, "wall_s": build_s,
                                         "algorithmic_bytes": build_bytes(R.shape[0], tr.size)}}))
        return
    if args.workload == "c3":
        print(json.dumps(run_c3(args)))
        return
    if args.workload == "c4":
        print(json.dumps(run_c4(args)))
        return
    if args.workload == "c5":
        print(json.dumps(run_c5(args)))
        return

    from corruscant_amd import clustering

    D, fd, R, fr, radii = c2_workload(args.scale)
    print("workload: %d data, %d randoms, %d radii" % (D.shape[0], R.shape[0], radii.size), file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    td, tr = clustering.trees([(D, fd), (R, fr)])  # built concurrently, one stream each
    build_s = time.perf_counter() - t0
    print("trees built in %.3f s" % build_s, file=sys.stderr, flush=True)
    pairs = [(td, td), (tr, td), (tr, tr)]

    def step(prec):
        out = clustering._count_jobs(pairs, radii, prec)
        return out, clustering.last_stats

    for _ in range(args.warmup):
        step(args.precision)
    times, stats = [], None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out, stats = step(args.precision)
        times.append(time.perf_counter() - t0)
        print("step %.3f s" % times[-1], file=sys.stderr, flush=True)
    sec = float(np.mean(times))
    evaluated = stats["evaluated_pairs"]
    useful = int(sum(o[-1].sum() for o in out))
    kernel_s = stats["count_ms"] / 1e3
    achieved = evaluated * FLOP_PER_PAIR / kernel_s / 1e12
    peak = PEAK_TFLOPS[args.precision]
    line = {
        "metric": "pairs counted/s (DD+DR+RR ordered pairs within the largest radius, as the reference counts "
                  "them; 1M data / 10M randoms, 20 radii, 16 fields, exact counts)",
        "value": useful / sec,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 recheck" if args.precision == "fp32_guarded" else "f64",
        "data": "synthetic",
        "config": {
            "workload": "3-D xi(r) Landy-Szalay DD+DR+RR: 1,000,000 data (half in 2000 Gaussian clusters, "
                        "sigma 0.005) / 10,000,000 uniform randoms in the unit cube, 16 fields, 20 radii "
                        "log-spaced 1e-3..1e-1",
            "data_points": int(D.shape[0]),
            "random_points": int(R.shape[0]),
            "radii": int(radii.size),
            "fields": 16,
            "precision": args.precision,
        },
        "dd_dr_rr_wall_s": sec,
        "tree_build": build_line(td, tr, D, R, build_s),
        "evaluated_pairs": evaluated,
        "pair_evals_per_s": evaluated / sec,
        "useful_pairs": useful,
        "rechecked_pairs": stats["rechecked_pairs"],
        "recheck_fraction": stats["rechecked_pairs"] / max(evaluated, 1),
        "reach_checks": stats["reach_checks"],
        "work_items": stats["work_items"],
        "enumerate_ms": stats["enumerate_ms"],
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": (measured_traffic("traffic_k_pairs_%s_c2_1M_10M" % args.precision)
                        if args.scale == 1.0 else None),
            "kernel": "k_pairs32" if args.precision == "fp32_guarded" else "k_pairs64 + k_pairs64_mixed",
            "kernel_ms": stats["count_ms"],
        },
    }
    if not args.no_compare:
        other = "fp64" if args.precision == "fp32_guarded" else "fp32_guarded"
        t0 = time.perf_counter()
        out2, st2 = step(other)
        dt2 = time.perf_counter() - t0
        same = all(np.array_equal(a, b) for a, b in zip(out, out2))
        line["compare"] = {"precision": other, "ms_per_step": dt2 * 1e3, "kernel_ms": st2["count_ms"],
                           "pairs_counted_per_s": useful / dt2,
                           "pair_evals_per_s": st2["evaluated_pairs"] / dt2,
                           "rechecked_pairs": st2["rechecked_pairs"],
                           "recheck_fraction": st2["rechecked_pairs"] / max(st2["evaluated_pairs"], 1),
                           "counts_identical": bool(same)}
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))
        print("cpu baseline on %d cores ..." % cores, file=sys.stderr, flush=True)
        cb = cpu_baseline(td, tr, D, fd, R, fr, radii, args.cpu_seconds, cores, args.precision)
        line["cpu_baseline"] = cb
        if cb is not None:
            line["counts_match_reference"] = cb["sampled_rows_counts_match_gpu"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
