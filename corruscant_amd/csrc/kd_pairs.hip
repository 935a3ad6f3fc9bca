// Pair counting: tile-pair enumeration and the persistent leaf-pair kernel.
//
// What is counted (src/kdquery.c:30-113, called once per radius by
// corruscant/clustering.py:218-230): for every query point q (raw, with its
// own field) and every tree node p (deduplicated, 8-bit field, weight w),
// cumulative over the sorted radii r_e,
//     C_e[qf][tf] += w   iff   n < r_e*r_e  and the reference's traversal
//                              reaches p and counts it at radius r_e,
// with n = fma(dz,dz, fma(dy,dy, dx*dx)), d = p - q in FP64 (src/kdquery.c:39-47).
//
// The traversal's pruning (src/kdquery.c:96-109) can only drop a pair whose
// distance is within a few ulps of r_e along one axis: if an ancestor with
// split value s on axis k prunes p, then |p_k - q_k| > r_e - 4u(cmax + r_e),
// so the pair's computed n exceeds lo_e = (r_e - 4u(cmax + r_e))^2 (1 - 16u).
// Pairs with n < lo_e are therefore counted at r_e exactly when n < r_e^2, and
// only pairs in the band lo_e <= n < r_e^2 replay the reference's pruning
// tests by walking p's ancestors in the node array (reach()).  Counting at r_e
// is monotone in r_e, so each pair contributes to one histogram bin, the
// first radius that counts it, and the cumulative counts are prefix sums.
//
// Work decomposition: one wave per (query tile, tree tile) item.  Lanes hold
// up to 64 query points; the tree tile's points are read as wave-uniform
// values.  Items list only the candidate radii [elo, ehi) whose thresholds
// fall between the two tiles' minimum and maximum squared box distances.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "kd_internal.h"

namespace kd {

static inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

struct EdgeTable {
    int ne;
    double r[kMaxEdges];
    double rsq[kMaxEdges];
    double lo[kMaxEdges];
};

struct JobDesc {
    // query side (raw points)
    const double *qx, *qy, *qz;
    const int *qf;
    const int *qstart;
    const Box *qbox;
    const float *qcx, *qcy, *qcz;
    const double *qcentre;
    int nqt;
    // tree side (unique points)
    const TPoint *tp;
    const TPoint32 *tp32;
    const double *tcentre;
    const int *tstart;
    const Box *hbox;
    int P;
    int ntt;
    const double *split;
    const uint32_t *nbits;
    int nq_fields;
    int nslots;
    long long hist_base;  // first histogram row of this job
};

struct Item {
    int job;
    int q;
    int t;
    unsigned short elo, ehi;
};
static_assert(sizeof(Item) == 16, "Item is 16 bytes");

static constexpr double kBoxMargin = 1e-12;

__device__ __forceinline__ double box_d2min(const Box &a, const Box &b)
{
    double s = 0.0;
    for (int d = 0; d < 3; d++) {
        double g = fmax(0.0, fmax(b.lo[d] - a.hi[d], a.lo[d] - b.hi[d]));
        s += g * g;
    }
    return s;
}

__device__ __forceinline__ double box_d2max(const Box &a, const Box &b)
{
    double s = 0.0;
    for (int d = 0; d < 3; d++) {
        double g = fmax(fabs(a.hi[d] - b.lo[d]), fabs(b.hi[d] - a.lo[d]));
        s += g * g;
    }
    return s;
}

// One thread per (job, query tile): walk the tree-tile hierarchy, emit the
// tile pairs whose pairs can be counted at some radius.  mode 0 counts,
// mode 1 writes at offs[g].
__global__ void k_enum(const JobDesc *__restrict__ jobs, const long long *__restrict__ qt_prefix, int njobs,
                       long long total, EdgeTable E, int mode, int *__restrict__ counts,
                       const long long *__restrict__ offs, Item *__restrict__ out)
{
    long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (g >= total) return;
    int lo = 0, hi = njobs;  // last job with prefix <= g
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (qt_prefix[mid] <= g) lo = mid;
        else hi = mid;
    }
    int job = lo;
    const JobDesc &J = jobs[job];
    int qt = (int)(g - qt_prefix[job]);
    Box qb = J.qbox[qt];
    const int ne = E.ne;
    const double rsq_last = E.rsq[ne - 1];
    int stack[48];
    int sp = 0;
    stack[sp++] = 1;
    int cnt = 0;
    long long w = mode ? offs[g] : 0;
    while (sp) {
        int node = stack[--sp];
        Box hb = J.hbox[node];
        double d2 = box_d2min(qb, hb);
        if (!(d2 * (1.0 - kBoxMargin) < rsq_last)) continue;
        if (node >= J.P) {
            int t = node - J.P;
            double dmin = d2 * (1.0 - kBoxMargin);
            double dmax = box_d2max(qb, hb) * (1.0 + kBoxMargin);
            int elo = 0;
            while (elo < ne && !(E.rsq[elo] > dmin)) elo++;
            int ehi = elo;
            while (ehi < ne && !(E.lo[ehi] > dmax)) ehi++;
            if (elo < ne) {
                if (mode) {
                    Item it;
                    it.job = job;
                    it.q = qt;
                    it.t = t;
                    it.elo = (unsigned short)elo;
                    it.ehi = (unsigned short)ehi;
                    out[w + cnt] = it;
                }
                cnt++;
            }
        } else {
            stack[sp++] = 2 * node + 1;
            stack[sp++] = 2 * node;
        }
    }
    if (!mode) counts[g] = cnt;
}

__global__ void k_widen(const int *__restrict__ c, long long *__restrict__ o, long long n)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) o[i] = c[i];
}

// Replays the reference's pruning for tree node h and query q at radius r:
// a two-child node counts itself only if both of its own tests pass, a left
// child is reached iff !(q_k - r > s), a right child iff !(q_k + r < s),
// one-child nodes always reach their left child (src/kdquery.c:86-110).
__device__ bool reach(double qx, double qy, double qz, int h, double r, const double *__restrict__ split,
                      const uint32_t *__restrict__ nbits)
{
    uint32_t b = nbits[h];
    if ((b & 3u) == 3u) {
        int k = (b >> 2) & 0xFF;
        double qk = k == 0 ? qx : (k == 1 ? qy : qz);
        double s = split[h];
        if ((qk + r) < s || (qk - r) > s) return false;
    }
    int c = h;
    while (c > 1) {
        int a = c >> 1;
        uint32_t ab = nbits[a];
        if ((ab & 3u) == 3u) {
            int k = (ab >> 2) & 0xFF;
            double qk = k == 0 ? qx : (k == 1 ? qy : qz);
            double s = split[a];
            if ((c & 1) == 0) {
                if ((qk - r) > s) return false;
            } else {
                if ((qk + r) < s) return false;
            }
        }
        c = a;
    }
    return true;
}

__device__ __forceinline__ unsigned wave_sum_u32(unsigned v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

static constexpr int kRows = 8;  // per-wave direct-mapped histogram row cache

struct WaveLds {
    // layout inside the dynamic LDS block of one wave:
    //   unsigned hist[(ne+1) * 64]      lane-private bins of the current item
    //   long long rows[kRows][ne+1]     row cache of (job, slot) histograms
    //   long long tags[kRows]
    unsigned *hist;
    long long *rows;
    long long *tags;
};

__device__ __forceinline__ void row_writeback(WaveLds &L, int r, int nb, long long *__restrict__ hist_g, int lane)
{
    long long tag = L.tags[r];
    if (tag >= 0) {
        for (int b = lane; b < nb; b += 64) {
            long long v = L.rows[r * nb + b];
            if (v) atomicAdd((unsigned long long *)&hist_g[tag * nb + b], (unsigned long long)v);
            L.rows[r * nb + b] = 0;
        }
    }
}

// Adds the current item's lane bins into the row cache, per query-field group.
__device__ void flush_bins(WaveLds &L, const JobDesc &J, int tf, int elo, int k, int nb, bool valid, int qf,
                           long long *__restrict__ hist_g, unsigned *__restrict__ err, int lane)
{
    unsigned long long remaining = __ballot(valid);
    while (remaining) {
        int leader = __ffsll((long long)remaining) - 1;
        int g = __shfl(qf, leader, 64);
        bool in = valid && qf == g;
        remaining &= ~__ballot(in);
        long long slot = (long long)g * J.nq_fields + tf;
        if (g < 0 || slot < 0 || slot >= J.nslots) {
            if (lane == 0) atomicOr(err, 1u);
            continue;
        }
        long long tag = J.hist_base + slot;
        int r = (int)(tag % kRows);
        if (L.tags[r] != tag) {
            row_writeback(L, r, nb, hist_g, lane);
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) L.tags[r] = tag;
            __builtin_amdgcn_wave_barrier();
        }
        for (int b = 0; b <= k; b++) {
            unsigned v = in ? L.hist[b * 64 + lane] : 0u;
            unsigned s = wave_sum_u32(v);
            if (lane == 0 && s) L.rows[r * nb + elo + b] += (long long)s;
        }
        __builtin_amdgcn_wave_barrier();
    }
    for (int b = 0; b <= k; b++) L.hist[b * 64 + lane] = 0u;
    __builtin_amdgcn_wave_barrier();
}

// FP64 form: every distance in FP64 with the reference's operation order.
__global__ __launch_bounds__(256) void k_pairs_fp64(const JobDesc *__restrict__ jobs, const Item *__restrict__ items,
                                                   long long nitems, unsigned long long *__restrict__ next,
                                                   int chunk, EdgeTable E, long long *__restrict__ hist_g,
                                                   unsigned long long *__restrict__ stats, unsigned *__restrict__ err)
{
    extern __shared__ unsigned char lds_raw[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int ne = E.ne, nb = ne + 1;
    const size_t per_wave = (size_t)nb * 64 * 4 + (size_t)kRows * nb * 8 + kRows * 8;
    unsigned char *base = lds_raw + wid * ((per_wave + 15) & ~(size_t)15);
    WaveLds L;
    L.hist = (unsigned *)base;
    L.rows = (long long *)(base + (size_t)nb * 64 * 4);
    L.tags = (long long *)(base + (size_t)nb * 64 * 4 + (size_t)kRows * nb * 8);
    for (int b = lane; b < nb * 64; b += 64) L.hist[b] = 0u;
    for (int b = lane; b < kRows * nb; b += 64) L.rows[b] = 0;
    if (lane < kRows) L.tags[lane] = -1;
    __builtin_amdgcn_wave_barrier();

    unsigned long long evaluated = 0, reached = 0;
    int cur_job = -1, cur_q = -1;
    double qx = 0, qy = 0, qz = 0;
    int qf = 0;
    bool valid = false;
    int qcount = 0;

    while (true) {
        unsigned long long start = 0;
        if (lane == 0) start = atomicAdd(next, (unsigned long long)chunk);
        start = __shfl(start, 0, 64);
        if (start >= (unsigned long long)nitems) break;
        long long end = min((long long)start + chunk, nitems);
        for (long long i = (long long)start; i < end; i++) {
            Item it = items[i];
            const JobDesc &J = jobs[it.job];
            if (it.job != cur_job || it.q != cur_q) {
                cur_job = it.job;
                cur_q = it.q;
                int b0 = J.qstart[it.q];
                qcount = J.qstart[it.q + 1] - b0;
                valid = lane < qcount;
                if (valid) {
                    qx = J.qx[b0 + lane];
                    qy = J.qy[b0 + lane];
                    qz = J.qz[b0 + lane];
                    qf = J.qf[b0 + lane];
                }
            }
            const int elo = it.elo, k = it.ehi - it.elo;
            const int t0 = J.tstart[it.t], t1 = J.tstart[it.t + 1];
            int cur_tf = -1;
            if (k > 0) evaluated += (unsigned long long)qcount * (unsigned long long)(t1 - t0);
            for (int j = t0; j < t1; j++) {
                const TPoint p = J.tp[j];
                const int tf = (p.meta >> 8) & 0xFF;
                const unsigned w = p.meta & 0xFF;
                if (tf != cur_tf) {
                    if (cur_tf >= 0) flush_bins(L, J, cur_tf, elo, k, nb, valid, qf, hist_g, err, lane);
                    cur_tf = tf;
                }
                int b = k;
                if (k > 0) {
                    const double dx = p.x - qx, dy = p.y - qy, dz = p.z - qz;
                    const double n = fma(dz, dz, fma(dy, dy, dx * dx));
                    bool band = false;
                    for (int e = it.ehi - 1; e >= elo; e--) {
                        const bool safe = n < E.lo[e];
                        b = safe ? e - elo : b;
                        band = band || (!safe && n < E.rsq[e]);
                    }
                    band = band && valid;
                    if (__ballot(band)) {
                        if (band) {
                            for (int e = elo; e < elo + b; e++) {
                                if (n < E.rsq[e] && reach(qx, qy, qz, p.heap, E.r[e], J.split, J.nbits)) {
                                    b = e - elo;
                                    break;
                                }
                            }
                        }
                        reached += __popcll(__ballot(band));
                    }
                }
                if (valid) L.hist[b * 64 + lane] += w;
            }
            if (cur_tf >= 0) flush_bins(L, J, cur_tf, elo, k, nb, valid, qf, hist_g, err, lane);
        }
    }
    for (int r = 0; r < kRows; r++) row_writeback(L, r, nb, hist_g, lane);
    if (lane == 0) {
        atomicAdd(&stats[0], evaluated);
        atomicAdd(&stats[2], reached);
    }
}

void count_pairs(const std::vector<PairJob> &jobs, const std::vector<double> &radii, int precision,
                 std::vector<long long *> &counts_out, kd_stats_t *stats, Sorter &so, hipStream_t s)
{
    auto wall0 = std::chrono::steady_clock::now();
    const int nr = (int)radii.size();
    KD_REQUIRE(nr >= 1, "pair count: at least one radius is required");
    for (double r : radii)
        KD_REQUIRE(std::isfinite(r) && r >= 0.0, "pair count: radii must be finite and >= 0");
    std::vector<double> edges(radii);
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    const int ne = (int)edges.size();
    KD_REQUIRE(ne <= kMaxEdges, "pair count: at most 64 distinct radii per call");
    KD_REQUIRE(precision == KD_PRECISION_FP64 || precision == KD_PRECISION_FP32_GUARDED,
               "pair count: unknown precision form");

    double cmax = 0.0;
    for (auto &j : jobs) cmax = std::max({cmax, j.tree->cmax, j.query->cmax});
    EdgeTable E;
    memset(&E, 0, sizeof(E));
    E.ne = ne;
    const double u = 1.1102230246251565e-16;
    for (int e = 0; e < ne; e++) {
        double r = edges[e];
        E.r[e] = r;
        E.rsq[e] = r * r;
        double t = r - 4.0 * u * (cmax + r);
        E.lo[e] = t > 0.0 ? t * t * (1.0 - 16.0 * u) : 0.0;
    }

    // job table
    const int nj = (int)jobs.size();
    std::vector<JobDesc> hj(nj);
    std::vector<long long> qt_prefix(nj + 1, 0);
    long long rows = 0;
    for (int i = 0; i < nj; i++) {
        const kd_device_tree *T = jobs[i].tree, *Q = jobs[i].query;
        KD_REQUIRE(!T->query_only, "pair count: the tree side must be a built tree");
        JobDesc &J = hj[i];
        J.qx = Q->q.x.p;
        J.qy = Q->q.y.p;
        J.qz = Q->q.z.p;
        J.qf = Q->q.f.p;
        J.qstart = Q->q.t.start.p;
        J.qbox = Q->q.t.box.p;
        J.qcx = Q->q.cx.p;
        J.qcy = Q->q.cy.p;
        J.qcz = Q->q.cz.p;
        J.qcentre = Q->q.centre.p;
        J.nqt = Q->q.t.ntiles;
        J.tp = T->tr.pts.p;
        J.tp32 = T->tr.pts32.p;
        J.tcentre = T->tr.centre.p;
        J.tstart = T->tr.t.start.p;
        J.hbox = T->tr.t.hbox.p;
        J.P = T->tr.t.P;
        J.ntt = T->tr.t.ntiles;
        J.split = T->split.p;
        J.nbits = T->nbits.p;
        J.nq_fields = jobs[i].nq_fields;
        J.nslots = jobs[i].nq_fields * jobs[i].nt_fields;
        J.hist_base = rows;
        rows += J.nslots;
        qt_prefix[i + 1] = qt_prefix[i] + J.nqt;
    }
    const long long total_qt = qt_prefix[nj];
    const int nb = ne + 1;
    DevBuf<JobDesc> djobs(nj);
    DevBuf<long long> dprefix(nj + 1);
    KD_HIP(hipMemcpyAsync(djobs.p, hj.data(), sizeof(JobDesc) * nj, hipMemcpyHostToDevice, s));
    KD_HIP(hipMemcpyAsync(dprefix.p, qt_prefix.data(), sizeof(long long) * (nj + 1), hipMemcpyHostToDevice, s));

    hipEvent_t ev0, ev1, ev2;
    KD_HIP(hipEventCreate(&ev0));
    KD_HIP(hipEventCreate(&ev1));
    KD_HIP(hipEventCreate(&ev2));
    KD_HIP(hipEventRecord(ev0, s));
    DevBuf<int> cnt(total_qt);
    DevBuf<long long> cnt64(total_qt), offs(total_qt);
    hipLaunchKernelGGL(k_enum, dim3(nblk(total_qt, 128)), dim3(128), 0, s, djobs.p, dprefix.p, nj, total_qt, E, 0,
                       cnt.p, (const long long *)nullptr, (Item *)nullptr);
    hipLaunchKernelGGL(k_widen, dim3(nblk(total_qt)), dim3(256), 0, s, cnt.p, cnt64.p, total_qt);
    so.excl_scan_i64(cnt64.p, offs.p, (int)total_qt, s);
    long long last_off = 0;
    int last_cnt = 0;
    KD_HIP(hipMemcpyAsync(&last_off, offs.p + (total_qt - 1), sizeof(long long), hipMemcpyDeviceToHost, s));
    KD_HIP(hipMemcpyAsync(&last_cnt, cnt.p + (total_qt - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    const long long nitems = last_off + last_cnt;
    DevBuf<Item> items(std::max<long long>(nitems, 1));
    hipLaunchKernelGGL(k_enum, dim3(nblk(total_qt, 128)), dim3(128), 0, s, djobs.p, dprefix.p, nj, total_qt, E, 1,
                       cnt.p, offs.p, items.p);
    KD_HIP(hipEventRecord(ev1, s));

    DevBuf<long long> hist(std::max<long long>(rows * nb, 1));
    DevBuf<unsigned long long> dnext(1), dstats(4);
    DevBuf<unsigned> derr(1);
    KD_HIP(hipMemsetAsync(hist.p, 0, hist.bytes(), s));
    KD_HIP(hipMemsetAsync(dnext.p, 0, dnext.bytes(), s));
    KD_HIP(hipMemsetAsync(dstats.p, 0, dstats.bytes(), s));
    KD_HIP(hipMemsetAsync(derr.p, 0, derr.bytes(), s));
    int dev = 0, ncu = 0;
    KD_HIP(hipGetDevice(&dev));
    KD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t per_wave = ((size_t)nb * 64 * 4 + (size_t)kRows * nb * 8 + kRows * 8 + 15) & ~(size_t)15;
    const int wpb = (int)std::max<size_t>(1, std::min<size_t>(4, 65536 / per_wave));
    const size_t lds = per_wave * wpb;
    const int nblocks = std::max(1, ncu * 16 / wpb);
    const int nwaves = nblocks * wpb;
    int chunk = (int)std::max<long long>(1, std::min<long long>(256, nitems / ((long long)nwaves * 8) + 1));
    if (nitems > 0) {
        (void)precision;
        hipLaunchKernelGGL(k_pairs_fp64, dim3(nblocks), dim3(64 * wpb), lds, s, djobs.p, items.p, nitems, dnext.p, chunk, E,
                           hist.p, dstats.p, derr.p);
        KD_HIP(hipGetLastError());
    }
    KD_HIP(hipEventRecord(ev2, s));
    std::vector<long long> hh(rows * nb);
    unsigned long long hstats[4];
    unsigned herr = 0;
    KD_HIP(hipMemcpyAsync(hh.data(), hist.p, sizeof(long long) * rows * nb, hipMemcpyDeviceToHost, s));
    KD_HIP(hipMemcpyAsync(hstats, dstats.p, sizeof(hstats), hipMemcpyDeviceToHost, s));
    KD_HIP(hipMemcpyAsync(&herr, derr.p, sizeof(herr), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    float ms_enum = 0, ms_pairs = 0;
    KD_HIP(hipEventElapsedTime(&ms_enum, ev0, ev1));
    KD_HIP(hipEventElapsedTime(&ms_pairs, ev1, ev2));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    (void)hipEventDestroy(ev2);
    KD_REQUIRE(!herr, "pair count: a field id gives a count-matrix index outside [0, Nq*Nt) "
                      "(index = query_field * num_fields + tree_field)");

    // cumulative counts per original radius
    std::vector<int> emap(nr);
    for (int i = 0; i < nr; i++)
        emap[i] = (int)(std::lower_bound(edges.begin(), edges.end(), radii[i]) - edges.begin());
    for (int jb = 0; jb < nj; jb++) {
        const JobDesc &J = hj[jb];
        long long *out = counts_out[jb];
        std::vector<long long> cum((size_t)J.nslots * ne);
        for (int sl = 0; sl < J.nslots; sl++) {
            long long acc = 0;
            const long long *row = &hh[(J.hist_base + sl) * nb];
            for (int e = 0; e < ne; e++) {
                acc += row[e];
                cum[(size_t)sl * ne + e] = acc;
            }
        }
        for (int i = 0; i < nr; i++)
            for (int sl = 0; sl < J.nslots; sl++) out[(size_t)i * J.nslots + sl] = cum[(size_t)sl * ne + emap[i]];
    }
    if (stats) {
        stats->evaluated_pairs = (long long)hstats[0];
        stats->rechecked_pairs = (long long)hstats[1];
        stats->reach_checks = (long long)hstats[2];
        stats->work_items = nitems;
        stats->enumerate_ms = ms_enum;
        stats->count_ms = ms_pairs;
        stats->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    }
}

}  // namespace kd
