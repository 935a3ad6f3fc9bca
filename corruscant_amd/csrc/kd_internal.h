// Host-side interfaces between the library's translation units.
#pragma once

#include <atomic>

#include "kd_common.h"

namespace kd {

// Radix sort and prefix sums (kd_scan.hip) with a reusable temporary buffer.
struct Sorter {
    DevBuf<unsigned char> tmp;
    DevBuf<unsigned> err;  // set by a sort whose look-back wait ran out (read by check)
    // stable sort of (kin, vin) by the 64-bit key into (kout, vout); the inputs are left as they were
    void pairs(const uint64_t *kin, uint64_t *kout, const int *vin, int *vout, int n, hipStream_t s);
    void check(hipStream_t s);  // waits for s; throws if a sort reported an error
    void incl_scan(const int *in, int *out, int n, hipStream_t s);
    void excl_scan_i64(const long long *in, long long *out, int n, hipStream_t s);
};

// kd_scan.hip: exclusive prefix sum with a caller-provided temporary of
// scan_parts(n) elements (no allocation: a graph can record it); in == out
// is allowed.  zero_async clears bytes (a multiple of 4) with a kernel.
long long scan_parts(long long n);
void excl_scan(const long long *in, long long *out, long long n, long long *part, hipStream_t s);
void zero_async(void *p, size_t bytes, hipStream_t s);

// Unique-point arrays the build leaves behind for the tree-side layout.
struct BuildExtras {
    DevBuf<double> ux, uy, uz;  // unique points, (x,y,z) order
    DevBuf<int> uf, uw;         // 8-bit field and weight
    DevBuf<int> heap_of;        // unique index -> heap index
    DevBuf<int> rep;            // unique index -> raw index of its representative
};

// kd_results.hip: out[m][t] = sum_ij mats[m][i][j] h_t[i] h_t[j], h_t the
// histogram of draws[t][0..nf-1] (host buffers)
void bootstrap_forms(const double *h_mats, int nmats, int nf, const unsigned char *h_draws, long long ntrials,
                     double *h_out, hipStream_t s);

// kd_build.hip
void build_tree(kd_device_tree *T, const double *d_xyz, const int *d_fields, int n, Sorter &so, BuildExtras &X,
                hipStream_t s);

// kd_layout.hip
// true when every field id is in [0, 255] (or there are no fields)
bool fields_in_byte(const int *d_fields, int n, hipStream_t s);
void layout_query(kd_device_tree *T, const double *d_xyz, const int *d_fields, int n, int cap, Sorter &so,
                  hipStream_t s);
void layout_tree(kd_device_tree *T, BuildExtras &X, int cap, Sorter &so, hipStream_t s);

// kd_pairs.hip
struct PairJob {
    const kd_device_tree *tree;
    const kd_device_tree *query;
    int nq_fields;  // row stride of the count matrix
    int nt_fields;
};
// counts_out: per job, nradii * nq_fields * nt_fields cumulative counts (host)
void count_pairs(const std::vector<PairJob> &jobs, const std::vector<double> &radii, int precision,
                 std::vector<long long *> &counts_out, kd_stats_t *stats, Sorter &so, hipStream_t s);
// The same call sized once and recorded into a graph (buffers from the
// calling context's stream s; the plan replays on a stream of its own).
struct PairPlan;
PairPlan *plan_pairs(const std::vector<PairJob> &jobs, const std::vector<double> &radii, int precision,
                     const std::vector<long long *> &counts_out, const kd_tuning_t *tuning, hipStream_t s);
void plan_run(PairPlan *plan, kd_stats_t *stats);
long long plan_items(const PairPlan *plan);
void plan_free(PairPlan *plan);

struct WPairJob {
    const kd_device_tree *tree;
    const kd_device_tree *query;
    int nq_fields;
    int nt_fields;
    const unsigned char *tree_weights;   // host, nsets x tree raw points (input order) or null
    const unsigned char *query_weights;  // host, nsets x query raw points or null
    long long *tree_field_sizes;         // host out, nsets x nt_fields, or null
    long long *query_field_sizes;        // host out, nsets x nq_fields, or null
};
// counts_out: per job, nsets * nradii * nq_fields * nt_fields cumulative counts (host)
void count_pairs_weighted(const std::vector<WPairJob> &jobs, int nsets, const std::vector<double> &radii,
                          int precision, std::vector<long long *> &counts_out, kd_stats_t *stats, Sorter &so,
                          hipStream_t s);

static constexpr int kMaxEdges = 64;
static constexpr int kDefaultTile = 64;

// Items of the enumeration's staging pool (kd_pairs.hip): 0 sizes it from the
// call, a negative value turns staging off, a positive one fixes the size.
extern std::atomic<long long> g_stage_items;

}  // namespace kd
