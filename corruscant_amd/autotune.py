"""Autotuner of the batched pair count.

Chooses, for a given workload, the tile size (points per tile, fixed when the
trees are built: the leaf size of the tile hierarchy) and the launch shape of
the pair kernels (workgroups per CU, the work-queue chunk a wave takes at
once, the enumeration's workgroups per CU), by timing graph replays of the
workload itself (clustering.count_plan): each candidate's time is the median
GPU time of ``reps`` replays after one warm-up replay.

The search is a coordinate descent: workgroups per CU, then the chunk, then
the enumeration's workgroups at the default tile size; then the tile size
with that shape; then, if another tile size won, the shape once more at that
size.  Counts never depend on the choice, and every candidate's counts are
checked against the first candidate's.

No reference counterpart: the reference's only knob is its thread count
(corruscant/clustering.py:218-230).
"""
from __future__ import print_function

import numpy as np

from . import clustering

TILE_SIZES = (16, 32, 48, 64)
BLOCKS_PER_CU = (4, 6, 8, 12, 16)
CHUNKS = (1, 2, 4, 16, 64)
ENUM_BLOCKS_PER_CU = (8, 16, 32)
DEFAULT_TILE = 64


class Tuned(object):
    """The autotuner's choice: tile_size, tuning (kd_tuning fields), ms (its
    median graph time), table (every candidate timed: tile size, tuning, ms,
    work items) and, for the chosen tile size, the built trees."""

    def __init__(self):
        self.tile_size = None
        self.tuning = None
        self.ms = None
        self.table = []
        self.trees = None

    def as_dict(self):
        return {"tile_size": self.tile_size, "tuning": dict(self.tuning), "graph_ms": self.ms,
                "candidates": [{"tile_size": t, "tuning": dict(u), "graph_ms": m, "work_items": n}
                               for t, u, m, n in self.table]}


def _build(catalogues, tile):
    """Trees of every (data, data_fields, randoms, random_fields) job at the
    given tile size; returns [(data_tree, rand_tree)]."""
    clustering.set_tile_size(tile)
    flat = clustering.trees([c for D, fd, R, fr in catalogues for c in ((D, fd), (R, fr))])
    return list(zip(flat[0::2], flat[1::2]))


def _jobs(pairs):
    return [j for d, r in pairs for j in ((d, d), (r, d), (r, r))]


def tune_autocorr(catalogues, radii, precision=None, tile_sizes=TILE_SIZES, blocks_per_cu=BLOCKS_PER_CU,
                  chunks=CHUNKS, enum_blocks_per_cu=ENUM_BLOCKS_PER_CU, reps=3, log=None):
    """Tune the batched auto-correlation of ``catalogues`` (a list of (data,
    data_fields, randoms, random_fields)) at ``radii``.  Leaves the library's
    tile size at the chosen value and returns a Tuned whose ``trees`` are the
    jobs' (data_tree, rand_tree) pairs built at it."""
    out = Tuned()
    ref = {}

    def say(msg):
        if log is not None:
            print(msg, file=log)
            log.flush()

    def timed(jobs, tile, tuning):
        plan = clustering.count_plan(jobs, radii, precision, tuning)
        try:
            first = plan.run(record=False)
            if "counts" not in ref:
                ref["counts"] = first
            elif not all(np.array_equal(a, b) for a, b in zip(first, ref["counts"])):
                raise RuntimeError("autotune: counts differ between launch shapes (tile %d, %r)" % (tile, tuning))
            ms = []
            for _ in range(reps):
                plan.run(record=False)
                ms.append(plan.last_stats["count_ms"])
            m = float(np.median(ms))
            out.table.append((tile, dict(tuning), m, plan.items))
            say("autotune: tile %2d %-52s %8.2f ms  %d items" % (tile, tuning, m, plan.items))
            return m
        finally:
            plan.close()

    def best_shape(jobs, tile, start):
        cur = dict(start)
        best = timed(jobs, tile, cur)
        for key, values in (("blocks_per_cu", blocks_per_cu), ("chunk", chunks),
                            ("enum_blocks_per_cu", enum_blocks_per_cu)):
            for v in values:
                if v == cur.get(key):
                    continue
                cand = dict(cur, **{key: v})
                m = timed(jobs, tile, cand)
                if m < best:
                    best, cur = m, cand
        return cur, best

    start = {"blocks_per_cu": 8, "chunk": 4, "enum_blocks_per_cu": 32}  # the library defaults
    tile0 = DEFAULT_TILE if DEFAULT_TILE in tile_sizes else tile_sizes[-1]
    pairs = _build(catalogues, tile0)
    shape, best = best_shape(_jobs(pairs), tile0, start)
    best_tile, best_pairs = tile0, pairs
    for tile in tile_sizes:
        if tile == tile0:
            continue
        cand_pairs = _build(catalogues, tile)
        ms = timed(_jobs(cand_pairs), tile, shape)
        if ms < best:
            best, best_tile, best_pairs = ms, tile, cand_pairs
        else:
            del cand_pairs
    if best_tile != tile0:
        shape, best = best_shape(_jobs(best_pairs), best_tile, shape)
    clustering.set_tile_size(best_tile)
    out.tile_size, out.tuning, out.ms, out.trees = best_tile, shape, best, best_pairs
    say("autotune: chose tile %d, %r: %.2f ms" % (best_tile, shape, best))
    return out
