// C ABI of libkdtree.so (declared in include/corruscant_amd.h).
//
// Entry points keep the reference's names and argument meaning
// (src/kdtree.h:43-45, src/kdbuild.c:303-306); the work happens on the GPU.
// Every call leases a context (a HIP stream and the sort/scan scratch) from a
// pool and returns it when done, so calls from several host threads run on
// separate streams and overlap on the device (building many small trees, for
// instance); each call synchronises its own stream before it returns.  The
// functions are therefore safe to call concurrently (the reference's
// pair_count is not reentrant: src/kdquery.c:25-28); trees are read-only after
// their build and may be shared by concurrent calls.
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "kd_internal.h"

namespace {

struct Ctx {
    hipStream_t s = nullptr;
    kd::Sorter so;
};

std::once_flag g_init_once;
std::string g_init_err;
hipMemPool_t g_pool = nullptr;
std::mutex g_ctx_mu;            // the context pool
std::vector<Ctx *> g_free_ctx;
std::mutex g_host_mu;           // host copies of node arrays
std::atomic<int> g_tile{kd::kDefaultTile};
thread_local std::string g_err;
thread_local Ctx *t_ctx = nullptr;  // the context of the call running on this thread
thread_local kd::Arena *t_arena = nullptr;  // the scratch arena of a call that allocates nothing

void init_device()
{
    std::call_once(g_init_once, [] {
        try {
            int n = 0;
            if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
                throw kd::Error("no HIP device: this library runs its kernels on an MI355X (gfx950) GPU");
            int dev = 0;
            KD_HIP(hipGetDevice(&dev));
            KD_HIP(hipDeviceGetDefaultMemPool(&g_pool, dev));
            // keep freed blocks in the pool: trees and per-call temporaries
            // come and go, and returning memory to the driver costs a device sync
            uint64_t keep = ~0ull;
            KD_HIP(hipMemPoolSetAttribute(g_pool, hipMemPoolAttrReleaseThreshold, &keep));
        } catch (const std::exception &e) {
            g_init_err = e.what();
        }
    });
    if (!g_init_err.empty()) throw kd::Error(g_init_err);
}

// A context for the duration of one call (nested calls on the same thread
// reuse the outer one).
struct Lease {
    Ctx *c = nullptr;
    bool own = false;
    Lease()
    {
        if (t_ctx) {
            c = t_ctx;
            return;
        }
        init_device();
        {
            std::lock_guard<std::mutex> lk(g_ctx_mu);
            if (!g_free_ctx.empty()) {
                c = g_free_ctx.back();
                g_free_ctx.pop_back();
            }
        }
        if (!c) {
            std::unique_ptr<Ctx> fresh(new Ctx());
            KD_HIP(hipStreamCreateWithFlags(&fresh->s, hipStreamNonBlocking));
            c = fresh.release();
        }
        own = true;
        t_ctx = c;
    }
    ~Lease()
    {
        if (!own) return;
        t_ctx = nullptr;
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        g_free_ctx.push_back(c);
    }
    Lease(const Lease &) = delete;
    Lease &operator=(const Lease &) = delete;
    hipStream_t s() const { return c->s; }
    kd::Sorter &so() const { return c->so; }
};

void set_error(const char *m) { g_err = m ? m : "unknown error"; }

kdtree_t bad_tree()
{
    kdtree_t t;
    memset(&t, 0, sizeof(t));
    t.size = -1;
    return t;
}

struct HostCopy {
    kd::DevBuf<double> xyz;
    kd::DevBuf<int> fields;
    HostCopy(const double *h_xyz, const int *h_fields, int n, hipStream_t s)
    {
        xyz.alloc(3 * (size_t)n);
        KD_HIP(hipMemcpyAsync(xyz.p, h_xyz, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, s));
        if (h_fields) {
            fields.alloc(n);
            KD_HIP(hipMemcpyAsync(fields.p, h_fields, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s));
        }
        KD_HIP(hipStreamSynchronize(s));
    }
};

// The jobs of a batched call in the library's form, checked.
void to_pair_jobs(const kd_job_t *jobs, int njobs, const double *radii, int nradii, const std::string &who,
                  std::vector<kd::PairJob> &pj, std::vector<long long *> &outs)
{
    KD_REQUIRE(jobs && njobs >= 1, who + ": no jobs");
    KD_REQUIRE(radii && nradii >= 1, who + ": no radii");
    pj.resize(njobs);
    outs.resize(njobs);
    for (int i = 0; i < njobs; i++) {
        KD_REQUIRE(jobs[i].tree.dev && jobs[i].query.dev, who + ": job with an invalid tree handle");
        KD_REQUIRE(jobs[i].counts, who + ": job without an output buffer");
        KD_REQUIRE(jobs[i].num_fields_query >= 1 && jobs[i].tree.num_fields >= 1, who + ": field counts must be >= 1");
        pj[i].tree = jobs[i].tree.dev;
        pj[i].query = jobs[i].query.dev;
        pj[i].nq_fields = jobs[i].num_fields_query;
        pj[i].nt_fields = jobs[i].tree.num_fields;
        outs[i] = jobs[i].counts;
    }
}

}  // namespace

// Opaque handle of a plan.
struct kd_plan {
    kd::PairPlan *p = nullptr;
};

namespace kd {

void *dev_alloc(size_t bytes)
{
    if (t_arena) {
        Arena &a = *t_arena;
        const size_t off = (a.used + 255) & ~(size_t)255;
        if (off > a.size || bytes > a.size - off)
            throw Error(std::string(a.who) + ": the call needs more device memory than the fixed scratch of " +
                        std::to_string(a.size >> 20) + " MiB (" + std::to_string((off + bytes) >> 20) +
                        " MiB requested so far); use the allocating entry point for this size");
        a.used = off + bytes;
        return a.base + off;
    }
    KD_REQUIRE(t_ctx, "internal error: device allocation outside a library call");
    void *p = nullptr;
    KD_HIP(hipMallocAsync(&p, bytes, t_ctx->s));
    return p;
}

void dev_free(void *p)
{
    if (t_arena && (unsigned char *)p >= t_arena->base && (unsigned char *)p < t_arena->base + t_arena->size)
        return;  // arena blocks are not freed one by one
    // stream-ordered on the calling context's stream; no call leaves work in
    // flight on a buffer it frees or that another stream still reads
    if (p) (void)hipFreeAsync(p, t_ctx ? t_ctx->s : nullptr);
}

bool arena_active() { return t_arena != nullptr; }

}  // namespace kd

extern "C" {

const char *kd_last_error(void) { return g_err.c_str(); }

void kd_free(void *p) { free(p); }

int kd_set_tile_size(int points_per_tile)
{
    if (points_per_tile < 1 || points_per_tile > 64) {
        set_error("kd_set_tile_size: tile size must be in [1, 64]");
        return -1;
    }
    g_tile = points_per_tile;
    return 0;
}

int kd_set_stage_pool(long long items)
{
    if (items > (1LL << 31)) {
        set_error("kd_set_stage_pool: at most 2^31 items");
        return -1;
    }
    kd::g_stage_items = items;
    return 0;
}

kdtree_t tree_construct(double *xyz, int *fields, int length, int num_fields)
{
    g_err.clear();
    try {
        KD_REQUIRE(xyz != nullptr, "tree_construct: xyz is NULL");
        KD_REQUIRE(length >= 1, "tree_construct: length must be at least 1");
        Lease L;
        hipStream_t s = L.s();
        const int tile = g_tile;
        std::unique_ptr<kd_device_tree> T(new kd_device_tree());
        HostCopy in(xyz, fields, length, s);
        kd::BuildExtras X;
        kd::build_tree(T.get(), in.xyz.p, in.fields.p, length, L.so(), X, s);
        T->num_fields = num_fields;
        T->symmetric_ok = T->n_u == T->n_raw && kd::fields_in_byte(in.fields.p, length, s);
        kd::layout_query(T.get(), in.xyz.p, in.fields.p, length, tile, L.so(), s);
        kd::layout_tree(T.get(), X, tile, L.so(), s);
        kdtree_t t;
        memset(&t, 0, sizeof(t));
        t.size = T->n_u;
        t.memsize = T->memsize;
        t.num_fields = num_fields;
        t.dev = T.release();
        return t;
    } catch (const std::exception &e) {
        set_error(e.what());
        return bad_tree();
    }
}

kdtree_t query_construct(double *xyz, int *fields, int length, int num_fields)
{
    g_err.clear();
    try {
        KD_REQUIRE(xyz != nullptr, "query_construct: xyz is NULL");
        KD_REQUIRE(length >= 1, "query_construct: length must be at least 1");
        Lease L;
        hipStream_t s = L.s();
        std::unique_ptr<kd_device_tree> T(new kd_device_tree());
        HostCopy in(xyz, fields, length, s);
        T->query_only = true;
        T->n_raw = length;
        T->n_u = length;
        T->num_fields = num_fields;
        kd::layout_query(T.get(), in.xyz.p, in.fields.p, length, g_tile, L.so(), s);
        T->cmax = T->q.t.cmax;
        kdtree_t t;
        memset(&t, 0, sizeof(t));
        t.size = length;
        t.memsize = 0;
        t.num_fields = num_fields;
        t.dev = T.release();
        return t;
    } catch (const std::exception &e) {
        set_error(e.what());
        return bad_tree();
    }
}

void destroy(kdtree_t tree)
{
    if (!tree.dev) return;
    try {
        Lease L;  // the device buffers are released on this call's stream
        free(tree.dev->host_nodes);
        delete tree.dev;
    } catch (const std::exception &e) {
        set_error(e.what());
    }
}

int kd_trim_memory(void)
{
    g_err.clear();
    try {
        init_device();
        KD_HIP(hipDeviceSynchronize());
        KD_HIP(hipMemPoolTrimTo(g_pool, 0));
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int tree_fetch_nodes(kdtree_t *tree)
{
    g_err.clear();
    try {
        KD_REQUIRE(tree && tree->dev && !tree->dev->query_only, "tree_fetch_nodes: not a built tree");
        std::lock_guard<std::mutex> lk(g_host_mu);
        kd_device_tree *T = tree->dev;
        if (!T->host_nodes) {
            T->host_nodes = (node_t *)malloc((size_t)T->memsize * sizeof(node_t));
            KD_REQUIRE(T->host_nodes, "tree_fetch_nodes: out of host memory");
            KD_HIP(hipMemcpy(T->host_nodes, T->nodes.p, T->nodes.bytes(), hipMemcpyDeviceToHost));
        }
        tree->node_data = T->host_nodes;
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int tree_fetch_perm(kdtree_t tree, int *perm, int *inorder)
{
    g_err.clear();
    try {
        KD_REQUIRE(tree.dev && !tree.dev->query_only, "tree_fetch_perm: not a built tree");
        kd_device_tree *T = tree.dev;
        if (perm) KD_HIP(hipMemcpy(perm, T->perm.p, T->perm.bytes(), hipMemcpyDeviceToHost));
        if (inorder) KD_HIP(hipMemcpy(inorder, T->inorder.p, T->inorder.bytes(), hipMemcpyDeviceToHost));
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int pair_count_jobs(const kd_job_t *jobs, int njobs, const double *radii, int nradii, int precision,
                    kd_stats_t *stats)
{
    g_err.clear();
    try {
        std::vector<kd::PairJob> pj;
        std::vector<long long *> outs;
        to_pair_jobs(jobs, njobs, radii, nradii, "pair_count_jobs", pj, outs);
        std::vector<double> r(radii, radii + nradii);
        Lease L;
        kd::count_pairs(pj, r, precision, outs, stats, L.so(), L.s());
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

kd_plan_t *pair_count_plan(const kd_job_t *jobs, int njobs, const double *radii, int nradii, int precision,
                           const kd_tuning_t *tuning)
{
    g_err.clear();
    try {
        std::vector<kd::PairJob> pj;
        std::vector<long long *> outs;
        to_pair_jobs(jobs, njobs, radii, nradii, "pair_count_plan", pj, outs);
        std::vector<double> r(radii, radii + nradii);
        Lease L;
        std::unique_ptr<kd_plan> h(new kd_plan());
        h->p = kd::plan_pairs(pj, r, precision, outs, tuning, L.s());
        return h.release();
    } catch (const std::exception &e) {
        set_error(e.what());
        return nullptr;
    }
}

int pair_count_plan_run(kd_plan_t *plan, kd_stats_t *stats)
{
    g_err.clear();
    try {
        KD_REQUIRE(plan && plan->p, "pair_count_plan_run: no plan");
        kd::plan_run(plan->p, stats);
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

long long pair_count_plan_items(const kd_plan_t *plan)
{
    return plan && plan->p ? kd::plan_items(plan->p) : -1;
}

void pair_count_plan_free(kd_plan_t *plan)
{
    if (!plan) return;
    try {
        Lease L;  // the plan's device buffers return to the pool on this stream
        kd::plan_free(plan->p);
    } catch (const std::exception &e) {
        set_error(e.what());
    }
    delete plan;
}

}  // extern "C"

// ---- pair_count_jobs without device allocations ------------------------------
//
// Every device buffer of the call (the job table, the enumeration's counts,
// offsets and work items, the histograms, counters and the scan scratch) is
// carved from `scratch`, a fixed array in the library's code object (each GPU
// holds its own copy from the library's first use on it until the process
// ends).  Host code never passes `scratch` itself to the device: its name is a
// host-side stand-in; the device address comes from hipGetSymbolAddress, once
// per device, with that device current.  Calls on one device are serialised
// by a mutex (they would otherwise race on the one array) and every call
// synchronises its stream before it releases the array.  Everything the
// kernels accumulate into is cleared by the call itself (fill kernels over
// the histograms, counters and error word), so nothing depends on what the
// previous call left in the array.
static __device__ __align__(256) unsigned char scratch[KD_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr size_t kScratchBytes = (size_t)KD_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu[kMaxDevices];  // one call at a time per device
unsigned char *g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device (under the mutex)

unsigned char *scratch_on(int dev)
{
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw kd::Error(std::string("pair_count_jobs_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                            hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    return g_scratch_ptr[dev];
}

struct ArenaScope {
    explicit ArenaScope(kd::Arena *a) { t_arena = a; }
    ~ArenaScope() { t_arena = nullptr; }
};

}  // namespace

extern "C" {

int pair_count_jobs_noalloc(const kd_job_t *jobs, int njobs, const double *radii, int nradii, int precision,
                            kd_stats_t *stats)
{
    g_err.clear();
    try {
        KD_REQUIRE(!t_arena, "pair_count_jobs_noalloc: nested call");
        std::vector<kd::PairJob> pj;
        std::vector<long long *> outs;
        to_pair_jobs(jobs, njobs, radii, nradii, "pair_count_jobs_noalloc", pj, outs);
        std::vector<double> r(radii, radii + nradii);
        Lease L;
        int dev = 0;
        KD_HIP(hipGetDevice(&dev));
        KD_REQUIRE(dev >= 0 && dev < kMaxDevices, "pair_count_jobs_noalloc: device index out of range");
        std::lock_guard<std::mutex> lk(g_scratch_mu[dev]);
        kd::Arena arena;
        arena.base = scratch_on(dev);
        arena.size = kScratchBytes;
        arena.who = "pair_count_jobs_noalloc";
        ArenaScope scope(&arena);
        try {
            kd::Sorter so;  // its scan scratch comes from the arena too
            kd::count_pairs(pj, r, precision, outs, stats, so, L.s());
        } catch (...) {
            // kernels of the call may still be running on the array
            (void)hipStreamSynchronize(L.s());
            throw;
        }
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int pair_count_weighted(const kd_wjob_t *jobs, int njobs, int nsets, const double *radii, int nradii,
                        int precision, kd_stats_t *stats)
{
    g_err.clear();
    try {
        KD_REQUIRE(jobs && njobs >= 1, "pair_count_weighted: no jobs");
        KD_REQUIRE(radii && nradii >= 1, "pair_count_weighted: no radii");
        KD_REQUIRE(nsets >= 1, "pair_count_weighted: nsets must be >= 1");
        std::vector<kd::WPairJob> pj(njobs);
        std::vector<long long *> outs(njobs);
        for (int i = 0; i < njobs; i++) {
            KD_REQUIRE(jobs[i].tree.dev && jobs[i].query.dev, "pair_count_weighted: job with an invalid tree handle");
            KD_REQUIRE(jobs[i].counts, "pair_count_weighted: job without an output buffer");
            KD_REQUIRE(jobs[i].num_fields_query >= 1 && jobs[i].tree.num_fields >= 1,
                       "pair_count_weighted: field counts must be >= 1");
            pj[i].tree = jobs[i].tree.dev;
            pj[i].query = jobs[i].query.dev;
            pj[i].nq_fields = jobs[i].num_fields_query;
            pj[i].nt_fields = jobs[i].tree.num_fields;
            pj[i].tree_weights = jobs[i].tree_weights;
            pj[i].query_weights = jobs[i].query_weights;
            pj[i].tree_field_sizes = jobs[i].tree_field_sizes;
            pj[i].query_field_sizes = jobs[i].query_field_sizes;
            outs[i] = jobs[i].counts;
        }
        std::vector<double> r(radii, radii + nradii);
        Lease L;
        kd::count_pairs_weighted(pj, nsets, r, precision, outs, stats, L.so(), L.s());
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int bootstrap_quadratic_forms(const double *mats, int nmats, int nf, const unsigned char *draws,
                              long long ntrials, double *out)
{
    g_err.clear();
    try {
        Lease L;
        kd::bootstrap_forms(mats, nmats, nf, draws, ntrials, out, L.s());
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

long long *pair_count(kdtree_t tree, double *data, int *fields, int length, int num_fields, double r,
                      int num_threads)
{
    (void)num_threads;  // CPU thread count of the reference; the GPU path has no use for it
    g_err.clear();
    kdtree_t q = query_construct(data, fields, length, num_fields);
    if (!q.dev) return nullptr;
    long long *out = nullptr;
    try {
        KD_REQUIRE(tree.dev, "pair_count: invalid tree handle");
        KD_REQUIRE(num_fields >= 1, "pair_count: num_fields must be >= 1");
        size_t cnt = (size_t)num_fields * (size_t)tree.num_fields;
        out = (long long *)calloc(cnt ? cnt : 1, sizeof(long long));
        KD_REQUIRE(out, "pair_count: out of host memory");
        kd_job_t job;
        job.tree = tree;
        job.query = q;
        job.num_fields_query = num_fields;
        job.counts = out;
        if (pair_count_jobs(&job, 1, &r, 1, KD_PRECISION_FP64, nullptr) != 0) throw kd::Error(g_err);
    } catch (const std::exception &e) {
        std::string m = e.what();
        free(out);
        out = nullptr;
        destroy(q);
        set_error(m.c_str());
        return nullptr;
    }
    destroy(q);
    return out;
}

}  // extern "C"
