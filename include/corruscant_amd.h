/*
 * corruscant_amd.h -- C ABI of the MI355X pair-counting library.
 *
 * libkdtree.so exports the tree and pair-count entry points, libcoords.so the
 * coordinate conversions; both are built from corruscant_amd/csrc for gfx950.
 * Every function that replaces a reference entry point keeps its name, its
 * argument order and its argument meaning; the citation next to each says
 * which reference interface it stands in for.
 *
 * Deliberate differences from the reference, all on error paths:
 *   - no function calls exit(); failures return an error value and leave a
 *     message for kd_last_error() (the reference prints and exits at
 *     src/kdbuild.c:501-511 and src/kdquery.c:176-179);
 *   - pair_count ignores num_threads (it is the CPU thread count of the
 *     reference, src/kdquery.c:181-234) instead of exiting when it exceeds N/20;
 *   - radii must be finite and >= 0, coordinates must not be NaN.
 */
#ifndef CORRUSCANT_AMD_H
#define CORRUSCANT_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Record layouts of src/kdtree.h:13-25 (x86-64: datum 32 bytes, node 40). */
typedef struct datum {
    double value[3];
    unsigned int field : 8;
    unsigned int weight : 8;
} datum_t;

typedef struct node {
    datum_t data;
    unsigned int has_lchild : 1;
    unsigned int has_rchild : 1;
    unsigned int dim : 8;
} node_t;

struct kd_device_tree;

/*
 * Tree handle, returned by value like src/kdtree.h:33-41 (64 bytes, same
 * offsets for node_data/size/memsize/num_fields).  The arrays live on the GPU;
 * node_data stays NULL until tree_fetch_nodes() copies the node array to the
 * host.  `dev` takes the slot of the reference's (dangling) `data` pointer.
 */
typedef struct kdtree {
    node_t *node_data;
    int size;       /* unique points (src/kdbuild.c:513) */
    int memsize;    /* heap slots, pow2ceil(size) (src/kdbuild.c:514) */
    int num_fields;
    struct kd_device_tree *dev;
    int *fields;    /* always NULL */
    int *args[3];   /* always NULL */
} kdtree_t;

/* ---- reference entry points ------------------------------------------ */

/* src/kdtree.h:43, src/kdbuild.c:451 -- builds the tree on the GPU.  xyz is
 * length x 3 row-major float64, fields int32[length] or NULL.  On error the
 * returned handle has dev == NULL and size == -1. */
kdtree_t tree_construct(double *xyz, int *fields, int length, int num_fields);

/* src/kdtree.h:45, src/kdquery.c:165 -- cumulative weighted count of
 * (query point, tree node) pairs with squared distance < r*r, exactly as the
 * reference's traversal counts them, as a malloc'd int64 matrix
 * [num_fields][tree.num_fields] indexed qf*num_fields + tf.  Free it with
 * kd_free().  NULL on error. */
long long *pair_count(kdtree_t tree, double *data, int *fields, int length, int num_fields,
                      double r, int num_threads);

/* src/kdbuild.c:303-306 -- frees the tree's device and host storage. */
void destroy(kdtree_t tree);

/* ---- additions --------------------------------------------------------- */

/* Message of the last failure on this thread ("" if none). */
const char *kd_last_error(void);

/* free() for the buffers this library returns (pair_count results). */
void kd_free(void *p);

/* Copies the node array (memsize records, slot 0 unused, src/kdbuild.c:516)
 * into tree->node_data (host, owned by the tree).  0 on success. */
int tree_fetch_nodes(kdtree_t *tree);

/* Build artefacts for parity checks: heap index -> unique index (memsize
 * ints, -1 for empty slots) and heap index -> in-order position (the segment
 * median).  Either pointer may be NULL.  0 on success. */
int tree_fetch_perm(kdtree_t tree, int *perm, int *inorder);

/* A query-only handle: raw points laid out for the pair kernels, without the
 * k-d build (used for arbitrary query sets).  size == length. */
kdtree_t query_construct(double *xyz, int *fields, int length, int num_fields);

/* One pair-count job: the raw points (and raw fields) of `query` against the
 * deduplicated, weighted nodes of `tree`.  Results are cumulative counts per
 * radius: counts[(ir * num_fields_query + qf) * num_fields_query ... ] laid out
 * as [nradii][num_fields_query][tree.num_fields], index qf*num_fields_query+tf
 * inside each radius block (src/kdquery.c:34, 137). */
typedef struct kd_job {
    kdtree_t tree;
    kdtree_t query;
    int num_fields_query;
    long long *counts; /* host, nradii * num_fields_query * tree.num_fields */
} kd_job_t;

/* Precision forms. */
#define KD_PRECISION_FP64 0      /* every distance in FP64 */
#define KD_PRECISION_FP32_GUARDED 1 /* FP32 on tile-centred coordinates, FP64 recheck in a guard band */

typedef struct kd_stats {
    long long evaluated_pairs; /* (query point, tree point) distances evaluated */
    long long rechecked_pairs; /* FP32 guard-band pairs recomputed in FP64 */
    long long reach_checks;    /* FP32 guarded form: work items left to the FP64 loop (radii below the FP32
                                  resolution of their tiles); 0 in the FP64 form */
    long long work_items;      /* (query tile, tree tile) items */
    double enumerate_ms;       /* GPU time of the item enumeration */
    double count_ms;           /* GPU time of the pair kernel */
    double total_ms;           /* wall time of the whole call */
    long long trimmed_items;   /* items whose end quads beyond the largest radius were cut off (FP32 guarded
                                  form; the FP64 loop skips such quads by its per-quad radius windows) */
    long long fp32_pairs;      /* of evaluated_pairs: those evaluated in FP32 on the matrix cores */
} kd_stats_t;

/* All radii of all jobs in one enumeration + one persistent kernel launch.
 * radii need not be sorted; at most 64 distinct values per call (the Python
 * front end splits longer lists into several calls, as the reference takes
 * any number of radii).  stats may be NULL.  0 on success. */
int pair_count_jobs(const kd_job_t *jobs, int njobs, const double *radii, int nradii,
                    int precision, kd_stats_t *stats);

/* pair_count_jobs for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, results and errors as
 * pair_count_jobs, but every device buffer of the call comes from a fixed
 * array of KD_NOALLOC_SCRATCH_BYTES in the library (one per GPU, held from
 * the library's first use on that GPU until the process ends); the call
 * makes no hipMalloc/hipFree and creates no events (stats->enumerate_ms and
 * count_ms are 0).  A call that needs more (its work items, 16 bytes each,
 * dominate: a few million pairs of 64-point tiles) fails with -1 and a
 * message in kd_last_error(); pair_count_jobs takes any size.  Calls on one
 * GPU share its array: this library serialises them within a process. */
#define KD_NOALLOC_SCRATCH_BYTES (256LL << 20)
int pair_count_jobs_noalloc(const kd_job_t *jobs, int njobs, const double *radii, int nradii,
                            int precision, kd_stats_t *stats);

/* Launch shape of the pair kernels; 0 in a field keeps the library's default.
 * Counts do not depend on it.  No counterpart in the reference (the
 * reference's only knob is num_threads, src/kdquery.c:120-171). */
typedef struct kd_tuning {
    int blocks_per_cu;      /* workgroups (4 waves each) of the pair kernels per CU, 1..64 (8) */
    int chunk;              /* items a wave takes from the work queue at once, 1..4096 (4) */
    int enum_blocks_per_cu; /* workgroups of the enumeration per CU, 1..64 (32) */
} kd_tuning_t;

/* A pair count made once and run many times: the batched jobs of
 * pair_count_jobs (same arguments) are sized on creation and their whole
 * device sequence -- enumeration, item scans, trim pass, pair kernels, copy
 * of the histograms to pinned host memory -- is recorded into one HIP graph.
 * pair_count_plan_run replays it (one graph launch, no allocation, no host
 * round trip inside) and writes each job's counts into the `counts` buffer
 * the job named at creation, exactly as pair_count_jobs would.  The trees and
 * the counts buffers must outlive the plan; runs of one plan must not
 * overlap (different plans may run concurrently).  stats->count_ms of a run
 * is the GPU time of the whole graph and enumerate_ms is 0.  tuning may be
 * NULL.  No counterpart in the reference, which recomputes everything per
 * call (corruscant/clustering.py:218-230).  Creation returns NULL on error
 * (kd_last_error()); run returns 0 on success, -1 on error (a replay whose
 * enumeration differs from the planned one, for a tree changed under it, is
 * refused). */
typedef struct kd_plan kd_plan_t;
kd_plan_t *pair_count_plan(const kd_job_t *jobs, int njobs, const double *radii, int nradii, int precision,
                           const kd_tuning_t *tuning);
int pair_count_plan_run(kd_plan_t *plan, kd_stats_t *stats);
/* Item count of the plan (the size of its work list; 16 bytes each). */
long long pair_count_plan_items(const kd_plan_t *plan);
void pair_count_plan_free(kd_plan_t *plan);

/* Weighted pair counts for S weight sets (bootstrap resampling): each set
 * gives every raw point an integer multiplicity (uint8), and the count of set
 * s is the reference's count for the catalogues np.repeat(points, w_s):
 * sum over pairs of wq_s(q) * wp_s(p) * node_weight(p) * [n < r*r].
 * No reference counterpart (the reference has no weights); it replaces one
 * full DD/DR/RR pass per weight set.  Tree-side weights need a tree without
 * duplicate points.  Weight arrays are nsets x N in the points' input order;
 * NULL means all ones.  counts: nsets x nradii x Nq x Nt, cumulative.  At most
 * 32 distinct radii.  precision: KD_PRECISION_FP64, or
 * KD_PRECISION_FP32_GUARDED (distances in FP32 on the matrix cores on
 * tile-centred coordinates, FP64 recheck of the pairs in a guard band around
 * each radius; the same counts).  stats->fp32_pairs and rechecked_pairs
 * report the FP32 form's share.  Tree-side multiplicities enter a
 * matrix-core contraction: as int8 operands when all of the call's are below
 * 128 (bootstrap draws are), else as bf16 (slower; the same counts). */
typedef struct kd_wjob {
    kdtree_t tree;
    kdtree_t query;
    int num_fields_query;
    const unsigned char *tree_weights;
    const unsigned char *query_weights;
    long long *counts;
    /* optional outputs (NULL to skip): per set and raw field id, the sum of
     * the multiplicities of the tree's (query's) input points -- the field
     * sizes of the resampled catalogues; nsets x tree.num_fields and
     * nsets x num_fields_query. */
    long long *tree_field_sizes;
    long long *query_field_sizes;
} kd_wjob_t;

int pair_count_weighted(const kd_wjob_t *jobs, int njobs, int nsets, const double *radii, int nradii,
                        int precision, kd_stats_t *stats);

/* Bootstrap resampling of the jackknife fields
 * (corruscant/clustering.py:416-446, results.bootstrap_pair_counts): for
 * trial t, h_t is the histogram of the field ids draws[t*nf .. t*nf+nf-1]
 * (each < nf), and for each of the nmats row-major nf x nf matrices A_m
 * (normalised count matrices, one per pair type and radius bin)
 *     out[m * ntrials + t] = sum_ij A_m[i][j] h_t[i] h_t[j],
 * the reference's tensordot of the counts with the outer product h_t h_t^T.
 * nf <= 256.  Host buffers; 0 on success, -1 with kd_last_error() set. */
int bootstrap_quadratic_forms(const double *mats, int nmats, int nf, const unsigned char *draws,
                              long long ntrials, double *out);

/* Device memory freed by the library (destroyed trees, per-call
 * temporaries) stays in a pool for reuse; this returns it to the driver.
 * 0 on success. */
int kd_trim_memory(void);

/* Library-wide tuning (leaf/tile sizes are fixed at construction time). */
int kd_set_tile_size(int points_per_tile);

/* Size of the pool in which a pair count's enumeration stages its work items
 * (16 bytes each) so that it walks the tile hierarchy once: 0 (the default)
 * sizes it per call, a negative value turns staging off, a positive one fixes
 * it (at most 2^31 items).  Counts do not depend on it; tiles that do not fit
 * are walked a second time.  No counterpart in the reference (a tuning knob
 * of this library).  0 on success, -1 on an invalid size. */
int kd_set_stage_pool(long long items);

#ifdef __cplusplus
}
#endif

#endif /* CORRUSCANT_AMD_H */
