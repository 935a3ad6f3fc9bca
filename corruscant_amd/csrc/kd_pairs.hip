// Pair counting: tile-pair enumeration and the persistent leaf-pair kernels.
//
// What is counted (src/kdquery.c:30-113, called once per radius by
// corruscant/clustering.py:218-230): for every query point q (raw, with its
// own field) and every tree node p (deduplicated, 8-bit field, weight w),
// cumulative over the radii r_e,
//     C_e[qf * Nq + tf] += w   iff   n < fl(r_e * r_e),
// n = fma(dz,dz, fma(dy,dy, dx*dx)), d = p - q in FP64 (src/kdquery.c:39-47).
//
// The reference's traversal also prunes on fl(q_k -/+ r) against split values
// (src/kdquery.c:96-109), but that never removes a pair with n < r*r: a
// failed test means q_k - r > s (or q_k + r < s) exactly, every point of the
// pruned subtree lies at or beyond s on axis k, so |p_k - q_k| > r, the
// rounded difference is >= r and the FMA chain gives n >= fl(r*r).  The
// count is therefore the plain predicate above, evaluated here per pair.
//
// Work decomposition.  Items are (query tile, tree tile) pairs listing the
// candidate radii [elo, ehi): radii below elo cannot count any pair of the two
// tiles, radii from ehi on count all of them (added in bulk).  One wave takes
// one item: lanes hold up to 64 query points, tree points are wave-uniform
// scalar loads, and for each candidate radius one compare per lane gives a
// 64-bit ballot whose popcount (scalar unit) accumulates the count.
//
// FP32 guarded form: coordinates relative to the tree tile's centre in FP32,
// thresholds lo_e < r_e^2 < hi_e that bracket the FP32 rounding error of the
// tile pair (derived below); pairs with lo_e <= n32 < hi_e are recomputed with
// the FP64 formula above.  Counts are identical to the FP64 form.
//
// Auto-correlation jobs (query set == tree set, no duplicates, fields < 256)
// evaluate each unordered tile pair once and add the mirrored counts.

#include <algorithm>
#include <map>
#include <utility>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "kd_internal.h"
#include "kd_item.h"

namespace kd {

static inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

struct EdgeTable {
    int ne;
    double r[kMaxEdges];
    double rsq[kMaxEdges];
};

// Thresholds of the FP32 form for tiles of one size class (fp32_thresholds).
static constexpr int kMfClasses = 16;
struct MfClass {
    unsigned lo32[kMaxEdges];
    unsigned band_w;
    float eps32;  // added to every D so that it is positive
    int e32;
    int pad;
};

struct JobDesc {
    // query side (raw points)
    const double *qx, *qy, *qz;
    const int *qf;
    const int *qstart;
    const Box *qbox;
    const int *qfmin, *qfmax;  // per query tile: smallest and largest field id
    int nqt;
    // tree side (unique points)
    const TPoint *tp;
    const double *tquad;         // the coordinates in padded quads of 4 points
    const int *tquad_start;      // ntt + 1
    const unsigned *tquad_meta;  // per quad: field | points << 8
    const BoxF *tquad_box;       // per quad: box of its points (FP32, rounded outwards)
    const double *tcentre;
    const float *tgrp;     // FP32 form: operand groups of 16 points (64 floats each)
    const int *tgstart;    // ntt + 1: first group of each tile
    const float *tquad32;  // FP32 form of the weighted kernels: the quads' operands, 16 bytes per point
    const unsigned *tmeta;
    const Box *tbox;
    const int *tstart;
    const Box *hbox;
    const unsigned *hfield;    // 2P: field of a subtree's points (kHMixed / kHEmpty)
    const long long *hweight;  // 2P: total weight of a subtree
    const long long *hfw;      // (P/64) x nfw: per-field weights of subtrees of >= 128 tile slots
    int nfw;                   // 0: not kept
    int P;
    int ntt;
    int symmetric;  // count each unordered tile pair once
    int relax;      // runs may merge across lower window ends (merge 3): the job's loop has quad windows
    int nq_fields;
    int nslots;
    long long hist_base;  // first histogram row of this job
    // FP32 form (matrix cores), per size class k of the tree tile (tile
    // meta bits 16..19): a pair is inside radius e when the bits of its value
    // D (> 0) are below lo32[k][e]; [lo32[k][e], lo32[k][e] + band_w[k]) is
    // the guard band, rechecked in FP64.  Radii below e32[k] run in FP64.
    const MfClass *mfc;  // kMfClasses entries (per call, device)
};

// Items (kd_item.h): tree tiles t .. t + tiles() - 1 (contiguous points)
// against query tile q.  A subtree item stands for every tree tile below
// hierarchy node t: every pair with the query tile lies inside radius ehi and
// beyond radius ehi - 1 (elo == ehi), so the item is a bulk count of (query
// points) x (subtree weight of each tree field) and nothing is compared.
//
// Every kernel that indexes tile or quad arrays with an item's fields checks
// them first (tile_item_ok / subtree_item_ok) and, on a bad item, sets
// kErrItem and skips it instead of reading out of bounds.
__device__ __forceinline__ bool tile_item_ok(const Item &it, int ntt)
{
    const int n = it.tiles();
    return it.t >= 0 && n >= 1 && it.t <= ntt - n;
}
__device__ __forceinline__ bool subtree_item_ok(const Item &it, int P) { return it.t >= 1 && it.t < 2 * P; }

static constexpr double kBoxMargin = 1e-12;
// Radii per pass of the pair loop.  Seven keeps the thresholds, counters and
// a prefetched quad of the K = KC loop within the SGPR budget of eight waves
// per SIMD; at eight the loop spilled (17 lane moves and 2 scratch loads per
// four points) and a second pass over the tile is cheaper than the spills.
static constexpr int KC = 7;

__device__ __forceinline__ double box_d2min(const Box &a, const Box &b)
{
    double s = 0.0;
    for (int d = 0; d < 3; d++) {
        double g = fmax(0.0, fmax(b.lo[d] - a.hi[d], a.lo[d] - b.hi[d]));
        s += g * g;
    }
    return s;
}

__device__ __forceinline__ double box_d2min(const Box &a, const BoxF &f)
{
    Box b;
    for (int d = 0; d < 3; d++) {
        b.lo[d] = f.lo[d];
        b.hi[d] = f.hi[d];
    }
    return box_d2min(a, b);
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

template <typename T>
__device__ __forceinline__ T sload(const T *p, int j)
{
    // wave-uniform read of read-only data through the scalar cache: the
    // constant address space lets the compiler use s_load for it
    static_assert(sizeof(T) % 4 == 0, "word-sized records only");
    typedef const unsigned __attribute__((address_space(4))) *cwords;
    const cwords w = (cwords)(const void *)(p + j);
    union {
        T v;
        unsigned u[sizeof(T) / 4];
    } r;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); k++) r.u[k] = w[k];
    return r.v;
}

__device__ __forceinline__ int find_job(const long long *__restrict__ prefix, int njobs, long long g)
{
    int lo = 0, hi = njobs;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (prefix[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One wave per (job, query tile) enumerates the tile pairs some radius can
// count, lanes = nodes of the tree-tile box hierarchy (an implicit complete
// binary tree, leaves = tree tiles).  An LDS stack holds the pending nodes in
// tile order, leftmost on top.  Each step pops the leftmost 64 nodes and
// tests their boxes; leaves that pass get their radius range [elo, ehi).
// Results left of the first surviving inner node are final and are emitted
// in order; everything from that node on goes back on the stack (children of
// inner nodes, and passing leaves, which are tested again later), so tiles
// come out in increasing order.  The stack holds fewer than 64 nodes per
// hierarchy level (plus 64).
// With `merge`, runs of adjacent tree tiles (inside 64-tile windows) with the
// same radius range and the same single-field, unit-weight kind become one
// item (merge 2, the FP32 form: and the same size class; merge 3, the FP64
// form: the same upper end ehi, the lower end elo of the run the smallest of
// its tiles', as long as every window fits one pass of KC radii, since that
// loop narrows the window per quad anyway).  mode 0 counts the items of each
// query tile, mode 1 writes them.
__device__ __forceinline__ unsigned long long lanes_below(int lane) { return (1ull << lane) - 1ull; }

static constexpr int kXcds = 8;  // XCDs of an MI355X

// A static assignment of n work units (query tiles) to the waves of the
// grid that keeps each XCD's L2 on one region: the units are cut into
// ns = min(kXcds, gridDim.x) contiguous slices, workgroup b (dispatched to
// XCD b % kXcds, round-robin) works on slice b % ns, and its waves take
// adjacent units.  Every unit goes to exactly one wave for any grid size.
struct XcdSlice {
    long long first, stride, end;
};
__device__ __forceinline__ XcdSlice xcd_slice(long long n, int wid, int waves_per_block)
{
    const int ns = (int)gridDim.x < kXcds ? (int)gridDim.x : kXcds;
    const int x = (int)blockIdx.x % ns;
    const long long blocks = ((long long)gridDim.x - x + ns - 1) / ns;  // workgroups on slice x
    return {n * x / ns + (long long)((int)blockIdx.x / ns) * waves_per_block + wid, blocks * waves_per_block,
            n * (x + 1) / ns};
}

// Staging of the items in the counting pass: each wave appends the items of
// its query tiles to chunks of a pool; the writing pass then copies a staged
// tile's items instead of walking the hierarchy again.  A tile's items may
// run from one chunk into a second; a tile that would need a third chunk, or
// that finds the pool used up, is walked again.
static constexpr long long kStageChunk = 2048;  // items per chunk (32 KiB)
static constexpr long long kStageMin = 256;     // a chunk with less room left is abandoned

// mode 0: count the items of every query tile (counts[g]).  mode 2: count and
// stage (counts[g]; pos[g] = the pool index of the tile's first item, -1 if
// not staged; when the items continue in a second chunk, the first n1[g] of
// them are at pos[g] and the rest at pos2[g], else n1[g] = -1; out is the
// pool).  mode 1: write the items of tile g at offs[g], copying them from src
// when staged (pos may be null); a tile whose items would end past item_cap
// (out's size) writes nothing and sets kErrPlan.
// Latency-bound (box loads per popped node): six waves per SIMD (80 VGPRs;
// at seven the staging state spills).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_enum(
    const JobDesc *__restrict__ jobs, const long long *__restrict__ qt_prefix, int njobs, long long total, EdgeTable E,
    int mode, int merge, int subtree, int cap, int *__restrict__ counts, const long long *__restrict__ offs,
    Item *__restrict__ out, unsigned *__restrict__ err, long long pool_cap, unsigned long long *__restrict__ pool_next,
    long long *__restrict__ pos, long long *__restrict__ pos2, int *__restrict__ n1s, const Item *__restrict__ src,
    long long item_cap)
{
    extern __shared__ int lds_stack[];
    __shared__ double s_rsq[kMaxEdges];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int e = threadIdx.x; e < E.ne; e += blockDim.x) s_rsq[e] = E.rsq[e];
    __syncthreads();
    int *stk = lds_stack + (size_t)wid * cap;
    const int ne = E.ne;
    const double rsq_last = s_rsq[ne - 1];
    // query tiles strided over all waves: XCD slices (as in k_wm) took the
    // 1M/10M enumeration from 6.6 to 8.7 ms, their jobs' tiles cost too
    // unevenly for contiguous slices
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    long long cur = 0, cend = 0;  // mode 2: the wave's free range of its chunk
    for (long long g = (long long)blockIdx.x * (blockDim.x >> 6) + wid; g < total; g += nw) {
        long long w = 0;
        if (mode == 1) {
            w = offs[g];
            if (w + counts[g] > item_cap) {
                if (lane == 0) atomicOr(err, kErrPlan);
                continue;
            }
            const long long pg = pos ? pos[g] : -1;
            if (pg >= 0) {  // staged: copy
                const int n = counts[g], n1 = n1s[g];
                const long long p2 = pos2[g];
                for (int k = lane; k < n; k += 64) out[w + k] = src[(n1 < 0 || k < n1) ? pg + k : p2 + (k - n1)];
                continue;
            }
        } else if (mode == 2) {
            if (cend - cur < kStageMin) {  // a fresh chunk
                unsigned long long b = 0;
                if (lane == 0) b = atomicAdd(pool_next, (unsigned long long)kStageChunk);
                b = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(b >> 32)) << 32) |
                    (unsigned)__builtin_amdgcn_readfirstlane((unsigned)b);
                const bool fits = (long long)b + kStageChunk <= pool_cap;
                cur = fits ? (long long)b : 0;
                cend = fits ? (long long)b + kStageChunk : 0;
            }
            w = cur;
        }
        bool spill = mode == 2 && cend - cur < kStageMin;  // mode 2: the tile is not staged
        const long long p0 = w;  // mode 2: where the tile's first item goes
        long long p1 = 0;        // mode 2: the second chunk, when the tile reaches one
        int n1 = -1;             // mode 2: the tile's items in the first chunk, -1 if one chunk holds all
        const int job = find_job(qt_prefix, njobs, g);
        const JobDesc &J = jobs[job];
        const int qt = (int)(g - qt_prefix[job]);
        const Box qb = sload(J.qbox, qt);
        const int levels = 31 - __clz(J.P);
        int cnt = 0;
        // the open (not yet written) item: tiles o_t .. o_t + o_n - 1
        bool open = false;
        int o_t = 0, o_n = 0, o_jb = 0, o_elo = 0, o_ehi = 0;
        unsigned o_meta = 0;
        int size = 1;
        if (lane == 0) stk[0] = 1;  // the root
        __builtin_amdgcn_wave_barrier();
        while (size > 0) {
            const int m = size < 64 ? size : 64;
            const bool valid = lane < m;
            const int node = valid ? stk[size - 1 - lane] : 1;
            size -= m;
            __builtin_amdgcn_wave_barrier();
            bool cand = false, kids = false, whole = false;
            int elo = 0, ehi = 0;
            bool skip = !valid;
            if (valid && J.symmetric) {  // subtrees whose tiles all precede qt
                const int depth = 31 - __clz(node);
                const long long last = (((long long)node + 1) << (levels - depth)) - J.P - 1;
                skip = last < qt;
            }
            if (!skip) {
                const Box hb = J.hbox[node];
                const double d2 = box_d2min(qb, hb);
                if (d2 * (1.0 - kBoxMargin) < rsq_last) {
                    // a subtree of at least 128 tile slots (smaller ones end
                    // up as merged runs of tiles anyway) with one field, or
                    // with its per-field weights kept, entirely after qt in a
                    // symmetric job, whose farthest pair lies inside the
                    // largest radius may be taken whole; a tile always
                    const double dmax = box_d2max(qb, hb) * (1.0 + kBoxMargin);
                    bool rank = node >= J.P;
                    const int below = levels - (31 - __clz(node));  // log2 of the subtree's tile slots
                    const unsigned hf =
                        !rank && subtree && J.hfield && below >= 7 && dmax < rsq_last ? J.hfield[node] : kHEmpty;
                    if (hf < kHEmpty || (hf == kHMixed && J.nfw > 0)) {
                        const long long first = ((long long)node << below) - J.P;
                        rank = !J.symmetric || first > qt;
                    }
                    if (rank) {
                        const double dmin = d2 * (1.0 - kBoxMargin);
                        // first radius with r^2 > dmin, first with r^2 > dmax
                        int a = 0, b = ne;
                        while (a < b) {
                            const int mid = (a + b) >> 1;
                            if (s_rsq[mid] > dmin) b = mid;
                            else a = mid + 1;
                        }
                        elo = a;
                        b = ne;
                        while (a < b) {
                            const int mid = (a + b) >> 1;
                            if (s_rsq[mid] > dmax) b = mid;
                            else a = mid + 1;
                        }
                        ehi = a;
                        if (node >= J.P) {
                            cand = elo < ne;
                        } else if (elo == ehi) {  // no radius between the nearest and farthest pair
                            cand = true;
                            whole = true;
                        } else {
                            kids = true;
                        }
                    } else {
                        kids = true;
                    }
                }
            }
            // results left of the first inner node are final
            const unsigned long long kmask = __ballot(kids);
            const int k = kmask ? __ffsll((long long)kmask) - 1 : 64;
            const bool emit = cand && lane < k;
            const bool back = cand && lane >= k;
            const unsigned long long b1 = __ballot(kids || back), b2 = __ballot(kids);
            const int T = __popcll(b1) + __popcll(b2);
            if (size + T > cap) {  // not reached with cap >= 64 * (levels + 3)
                if (lane == 0) atomicOr(err, 2u);
                break;
            }
            const int pre = __popcll(b1 & lanes_below(lane)) + __popcll(b2 & lanes_below(lane));
            if (back) stk[size + T - 1 - pre] = node;
            if (kids) {
                stk[size + T - 1 - pre] = 2 * node;
                stk[size + T - 2 - pre] = 2 * node + 1;
            }
            size += T;
            __builtin_amdgcn_wave_barrier();

            // emit in lane order, merging runs of adjacent tiles (at most 64
            // items per step, and one more at the end of the tile)
            const unsigned long long C = __ballot(emit);
            if (mode == 2 && !spill && w + cnt + 66 > cend) {  // continue in a second chunk, once
                if (n1 >= 0) {
                    spill = true;
                } else {
                    unsigned long long b = 0;
                    if (lane == 0) b = atomicAdd(pool_next, (unsigned long long)kStageChunk);
                    b = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(b >> 32)) << 32) |
                        (unsigned)__builtin_amdgcn_readfirstlane((unsigned)b);
                    if ((long long)b + kStageChunk <= pool_cap) {
                        n1 = cnt;
                        p1 = (long long)b;
                        w = p1 - cnt;  // item i of the tile at w + i from here on
                        cend = p1 + kStageChunk;
                    } else {
                        spill = true;
                    }
                }
            }
            const bool wr = mode == 1 || (mode == 2 && !spill);
            if (C) {
                const int t = whole ? node : node - J.P;
                const int jb = job | ((J.symmetric && (whole || t != qt)) ? (int)0x80000000 : 0) |
                               (whole ? kSubtreeItem : 0);
                // a subtree item has meta 0 and never joins a run
                const unsigned meta = emit && !whole ? J.tmeta[t] : 0u;
                // compare with the previous candidate of this step, or with the open item
                const unsigned long long below = C & lanes_below(lane);
                const int pl = below ? 63 - __clzll((long long)below) : lane;
                const int p_t = __shfl(t, pl, 64), p_jb = __shfl(jb, pl, 64), p_elo = __shfl(elo, pl, 64),
                          p_ehi = __shfl(ehi, pl, 64);
                const unsigned p_meta = __shfl(meta, pl, 64);
                const bool have = below ? true : open;
                const int q_t = below ? p_t : o_t + o_n - 1, q_jb = below ? p_jb : o_jb,
                          q_elo = below ? p_elo : o_elo, q_ehi = below ? p_ehi : o_ehi;
                const unsigned q_meta = below ? p_meta : o_meta;
                // merge 2 (FP32 form) and 4 (FP32 form of the weighted
                // kernels): tiles of a run share the size class too
                const unsigned mmask = (merge == 2 || merge == 4) ? ~0u : 0xFFFFu;
                // merge 3 joins differing elo only while every tile's window
                // fits one pass of the pair loop (KC radii), so a run never
                // takes more passes than its tiles would
                // (merge 2: the FP32 form sends items with ehi == ne to the
                // FP64 loop, which may take them relaxed too; merge 4 relaxes
                // as merge 3 within a size class)
                const bool relax = (merge == 3 || merge == 4 || (merge == 2 && ehi == ne)) && J.relax;
                const bool join = emit && merge && have && t == q_t + 1 && (t & 63) != 0 && jb == q_jb &&
                                  (elo == q_elo || (relax && ehi - elo <= KC && ehi - q_elo <= KC)) &&
                                  ehi == q_ehi && (meta & 3u) == 3u && ((meta ^ q_meta) & mmask) == 0;
                const unsigned long long H = __ballot(emit && !join);
                // merge 3: the smallest elo of each run, a suffix minimum over
                // the lanes from each head up to the next (non-candidate lanes
                // neutral); lane 0 then holds the part before the first head
                int run_elo = elo;
                if (relax) {
                    const unsigned long long after = H & ~lanes_below(lane + 1);
                    const int nh = after ? __ffsll((long long)after) - 1 : 64;
                    run_elo = emit ? elo : 64;
#pragma unroll
                    for (int off = 1; off < 64; off <<= 1) {
                        const int o = __shfl_down(run_elo, off, 64);
                        if (lane + off < nh) run_elo = min(run_elo, o);
                    }
                }
                const int pre_elo = __builtin_amdgcn_readfirstlane(run_elo);  // (lane 0's)
                if (!H) {  // every candidate extends the open item
                    o_n += __popcll(C);
                    if (relax) o_elo = min(o_elo, pre_elo);
                } else {
                    const int h0 = __ffsll((long long)H) - 1;
                    if (open) {  // close the open item; the candidates before h0 extend it
                        if (wr && lane == 0) {
                            Item it;
                            it.q = qt;
                            it.t = o_t;
                            it.job = o_jb;
                            it.elo = (unsigned char)(relax && h0 > 0 ? min(o_elo, pre_elo) : o_elo);
                            it.ehi = (unsigned char)o_ehi;
                            it.set_tiles(o_n + __popcll(C & lanes_below(h0)));
                            out[w + cnt] = it;
                        }
                        cnt++;
                    }
                    // every head but the last starts a complete item
                    const int hl = 63 - __clzll((long long)H);
                    const bool head = (H >> lane) & 1ull;
                    if (wr && head && lane != hl) {
                        const unsigned long long above = H & ~lanes_below(lane + 1);
                        const int nh = __ffsll((long long)above) - 1;
                        Item it;
                        it.q = qt;
                        it.t = t;
                        it.job = jb;
                        it.elo = (unsigned char)run_elo;
                        it.ehi = (unsigned char)ehi;
                        it.set_tiles(__popcll(C & lanes_below(nh) & ~lanes_below(lane)));
                        out[w + cnt + __popcll(H & lanes_below(lane))] = it;
                    }
                    cnt += __popcll(H) - 1;
                    // the last head stays open
                    open = true;
                    o_n = __popcll(C & ~lanes_below(hl));
                    o_t = __builtin_amdgcn_readlane(t, hl);
                    o_jb = __builtin_amdgcn_readlane(jb, hl);
                    o_elo = __builtin_amdgcn_readlane(run_elo, hl);
                    o_ehi = __builtin_amdgcn_readlane(ehi, hl);
                    o_meta = __builtin_amdgcn_readlane(meta, hl);
                }
            }
        }
        if (open) {
            if ((mode == 1 || (mode == 2 && !spill)) && lane == 0) {
                Item it;
                it.q = qt;
                it.t = o_t;
                it.job = o_jb;
                it.elo = (unsigned char)o_elo;
                it.ehi = (unsigned char)o_ehi;
                it.set_tiles(o_n);
                out[w + cnt] = it;
            }
            cnt++;
        }
        if (mode != 1 && lane == 0) counts[g] = cnt;
        if (mode == 2) {
            if (lane == 0) {
                pos[g] = spill ? -1 : p0;
                pos2[g] = p1;
                n1s[g] = n1;
            }
            cur = spill ? cend : w + cnt;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ void k_widen(const int *__restrict__ c, long long *__restrict__ o, long long n)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i <= n) o[i] = i < n ? c[i] : 0;
}

// The work items of all jobs: counts per query tile, their exclusive scan
// (total_qt + 1 entries, the last one the total), then the items.  The items
// of query tile g (numbered over all jobs) are offs[g] .. offs[g + 1] - 1.
struct ItemList {
    DevBuf<Item> items;
    DevBuf<long long> offs;
    long long nitems = 0;
    long long nsplit = 0;  // split lists: items 0..nsplit-1 belong to query tiles of one field
};

// Split item lists: the items of query tiles whose points share one field
// first (in query-tile order), then those of tiles with several fields.
__global__ void k_split_counts(const JobDesc *__restrict__ jobs, const long long *__restrict__ qt_prefix, int njobs,
                               long long total, const int *__restrict__ c, long long *__restrict__ one,
                               long long *__restrict__ several, unsigned char *__restrict__ mixed)
{
    long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (g > total) return;
    if (g == total) {
        one[g] = 0;
        several[g] = 0;
        return;
    }
    const int job = find_job(qt_prefix, njobs, g);
    const JobDesc &J = jobs[job];
    const int qt = (int)(g - qt_prefix[job]);
    const bool m = J.qfmin[qt] != J.qfmax[qt];
    one[g] = m ? 0 : c[g];
    several[g] = m ? c[g] : 0;
    mixed[g] = m ? 1 : 0;
}

// Items of unit-weight runs of one tree field whose radius window reaches
// past the largest radius (so they carry no bulk count): the quads at the
// start of the first tile and at the end of the last tile whose boxes fail
// the enumeration's box test for the largest radius against the query tile's
// box are cut off (every pair they hold lies beyond every radius).  At least
// one quad of each end tile stays, so a cut never reaches the partial last
// quad of the first tile.  The cuts (at most 15 each) are stored in the item
// (Item::set_cuts).  A grid-stride loop over the items; stats[2] counts the
// items cut (one atomic per wave at the end, not per item).
__global__ void k_trim_items(const JobDesc *__restrict__ jobs, Item *__restrict__ items, long long n,
                             double rsq_last, int ne, unsigned long long *__restrict__ stats,
                             unsigned *__restrict__ err)
{
    unsigned cut_items = 0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const Item it = items[i];
        if (it.job & kSubtreeItem) continue;
        const JobDesc &J = jobs[it.job & kJobMask];
        if (!tile_item_ok(it, J.ntt)) {
            atomicOr(err, kErrItem);
            continue;
        }
        if (it.ehi < ne || (J.tmeta[it.t] & 3u) != 3u) continue;
        const int nt = it.tiles();
        const Box qb = J.qbox[it.q];
        const int a0 = J.tquad_start[it.t], a1 = J.tquad_start[it.t + 1];
        const int b0 = J.tquad_start[it.t + nt - 1], b1 = J.tquad_start[it.t + nt];
        int lo = 0;
        while (lo < Item::kMaxCut && lo < a1 - a0 - 1 &&
               !(box_d2min(qb, J.tquad_box[a0 + lo]) * (1.0 - kBoxMargin) < rsq_last))
            lo++;
        const int keep = nt == 1 ? a0 + lo + 1 : b0 + 1;  // the last kept quad is at or after this one, minus one
        int hi = 0;
        while (hi < Item::kMaxCut && b1 - 1 - hi >= keep &&
               !(box_d2min(qb, J.tquad_box[b1 - 1 - hi]) * (1.0 - kBoxMargin) < rsq_last))
            hi++;
        if (lo | hi) {
            Item w = it;
            w.set_cuts(lo, hi);
            items[i] = w;
            cut_items++;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cut_items += __shfl_xor(cut_items, off, 64);
    if ((threadIdx.x & 63) == 0 && cut_items) atomicAdd(&stats[2], (unsigned long long)cut_items);
}

__global__ void k_split_offs(const long long *__restrict__ one, const long long *__restrict__ several,
                             const unsigned char *__restrict__ mixed, long long total, long long *__restrict__ offs)
{
    long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (g > total) return;
    offs[g] = g == total ? one[total] + several[total] : (mixed[g] ? one[total] + several[g] : one[g]);
}

// subtree: whole subtrees may become single bulk items, for the jobs whose
// tree-side hierarchy fields are set (the unweighted kernel and k_wq read
// them; tree-weighted jobs leave them unset and get tiles only)
// split: order the items as described at k_split_counts (nsplit set; offs[g]
// is then where query tile g's items start, not a boundary between tiles)
std::atomic<long long> g_stage_items{0};

static void enumerate_items(const JobDesc *djobs, const long long *dprefix, int nj, long long total_qt, int maxP,
                            const EdgeTable &E, int merge, int subtree, int split, Sorter &so, hipStream_t s,
                            ItemList &out)
{
    KD_REQUIRE(total_qt < 0x7fffffffLL, "pair count: too many query tiles for one call");
    const int levels = 31 - __builtin_clz((unsigned)std::max(maxP, 1));
    const int cap = 64 * (levels + 3);
    const size_t lds = (size_t)4 * cap * sizeof(int);
    int dev = 0, ncu = 0;
    KD_HIP(hipGetDevice(&dev));
    KD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int nblocks = (int)std::max<long long>(1, std::min<long long>((total_qt + 3) / 4, (long long)ncu * 32));
    DevBuf<int> cnt(total_qt);
    DevBuf<long long> cnt64(total_qt + 1);
    DevBuf<unsigned> derr(1);
    out.offs.alloc(total_qt + 1);
    KD_HIP(hipMemsetAsync(derr.p, 0, sizeof(unsigned), s));
    // the staging pool: room for 512 items per query tile and for two chunks
    // per wave, at least 2^20 and at most 4 GiB of items (without it, or past
    // it, tiles are walked twice)
    DevBuf<Item> pool;
    DevBuf<long long> pos, pos2;
    DevBuf<int> n1s;
    DevBuf<unsigned long long> pnext;
    long long pool_cap = 0;
    const long long stage_set = g_stage_items.load();
    // (not in a call that allocates nothing: tiles are then walked twice)
    if (stage_set >= 0 && !arena_active()) {
        const long long want =
            stage_set > 0 ? stage_set
                          : std::min((4LL << 30) / (long long)sizeof(Item),
                                     std::max({1LL << 20, 512LL * total_qt, 2LL * kStageChunk * nblocks * 4}));
        void *pp = nullptr;
        if (hipMallocAsync(&pp, (size_t)want * sizeof(Item), s) == hipSuccess && pp) {
            pool.p = (Item *)pp;
            pool.n = (size_t)want;
            pool_cap = want;
            pos.alloc(total_qt);
            pos2.alloc(total_qt);
            n1s.alloc(total_qt);
            pnext.alloc(1);
            KD_HIP(hipMemsetAsync(pnext.p, 0, sizeof(unsigned long long), s));
        } else {
            (void)hipGetLastError();
        }
    }
    hipLaunchKernelGGL(k_enum, dim3(nblocks), dim3(256), lds, s, djobs, dprefix, nj, total_qt, E, pool_cap ? 2 : 0,
                       merge, subtree, cap, cnt.p, (const long long *)nullptr, pool.p, derr.p, pool_cap, pnext.p,
                       pos.p, pos2.p, n1s.p, (const Item *)nullptr, 0LL);
    if (split) {
        DevBuf<long long> sev(total_qt + 1), o1(total_qt + 1), o2(total_qt + 1);
        DevBuf<unsigned char> mixed(total_qt);
        hipLaunchKernelGGL(k_split_counts, dim3(nblk(total_qt + 1)), dim3(256), 0, s, djobs, dprefix, nj, total_qt,
                           cnt.p, cnt64.p, sev.p, mixed.p);
        so.excl_scan_i64(cnt64.p, o1.p, (int)(total_qt + 1), s);
        so.excl_scan_i64(sev.p, o2.p, (int)(total_qt + 1), s);
        hipLaunchKernelGGL(k_split_offs, dim3(nblk(total_qt + 1)), dim3(256), 0, s, o1.p, o2.p, mixed.p, total_qt,
                           out.offs.p);
        KD_HIP(hipMemcpyAsync(&out.nsplit, o1.p + total_qt, sizeof(long long), hipMemcpyDeviceToHost, s));
    } else {
        hipLaunchKernelGGL(k_widen, dim3(nblk(total_qt + 1)), dim3(256), 0, s, cnt.p, cnt64.p, total_qt);
        so.excl_scan_i64(cnt64.p, out.offs.p, (int)(total_qt + 1), s);
    }
    unsigned herr = 0;
    KD_HIP(hipMemcpyAsync(&out.nitems, out.offs.p + total_qt, sizeof(long long), hipMemcpyDeviceToHost, s));
    KD_HIP(hipMemcpyAsync(&herr, derr.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    KD_REQUIRE(!herr, "pair count: tile-pair enumeration stack overflow");
    KD_REQUIRE(out.nitems < 0x7fffffffLL, "pair count: too many tile pairs for one call");
    if (!split) out.nsplit = out.nitems;
    out.items.alloc(std::max<long long>(out.nitems, 1));
    hipLaunchKernelGGL(k_enum, dim3(nblocks), dim3(256), lds, s, djobs, dprefix, nj, total_qt, E, 1, merge, subtree,
                       cap, cnt.p, out.offs.p, out.items.p, derr.p, pool_cap, pnext.p,
                       pool_cap ? pos.p : (long long *)nullptr, pos2.p, n1s.p, (const Item *)pool.p, out.nitems);
    KD_HIP(hipGetLastError());
}

// Global histogram: per (job, slot) a row of 2*ne+1 int64: direct cumulative
// counts for radii [0, ne) followed by ne+1 bulk entries, where bulk[e] adds
// to every radius >= e.  Each wave keeps a small fully associative cache of
// rows in LDS (least recently used row replaced) and writes evicted rows back
// with 64-bit atomics.  A query tile's working set is its own field against
// the few tree fields around it, plus the mirrored slots of auto-correlation
// jobs, which a direct-mapped cache would map onto one row.

static constexpr int kRows = 8;
static constexpr int kNextStride = 32;  // one 256-byte line per segment counter

struct WaveLds {
    long long *rows;  // kRows * rowlen
    long long *tags;  // kRows
    unsigned *stamp;  // kRows: last use (0 = never)
};

// Per-wave statistics of the pair kernels (LDS, after the row cache):
// evaluated pairs, rechecked pairs, FP64 items of the FP32 form, pairs
// evaluated on the matrix cores.  Kept out of registers: as 64-bit values
// live across the item loop they were spilled to scratch and updated there.
static constexpr int kStatEval = 0, kStatRecheck = 1, kStatFp64Items = 2, kStatMf = 3, kStats = 4;

__host__ __device__ constexpr size_t wave_lds_bytes(int rowlen)
{
    return ((size_t)kRows * rowlen * 8 + kRows * 8 + kRows * 4 + kStats * 8 + 15) & ~(size_t)15;
}

// lane 0 adds v (wave-uniform) to a per-wave statistic
__device__ __forceinline__ void stat_add(unsigned long long *stat, int k, unsigned long long v, int lane)
{
    if (lane == 0 && v) atomicAdd(&stat[k], v);
}

__device__ void row_writeback(WaveLds &L, int r, int rowlen, long long *__restrict__ hist_g, int lane)
{
    long long tag = L.tags[r];
    if (tag >= 0) {
        for (int b = lane; b < rowlen; b += 64) {
            long long v = L.rows[r * rowlen + b];
            if (v) atomicAdd((unsigned long long *)&hist_g[tag * rowlen + b], (unsigned long long)v);
            L.rows[r * rowlen + b] = 0;
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// Row of `tag` in the wave's cache: lanes 0..kRows-1 compare the tags in
// parallel; on a miss the least recently used row is written back and taken.
__device__ __forceinline__ int row_get(WaveLds &L, long long tag, int rowlen, long long *__restrict__ hist_g,
                                       int lane, unsigned &clock)
{
    const long long mine = lane < kRows ? L.tags[lane] : -2;
    const unsigned long long hit = __ballot(mine == tag);
    int r;
    if (hit) {
        r = __ffsll((long long)hit) - 1;
    } else {
        // (stamp, row) minimum over lanes 0..kRows-1 (shuffles within groups of kRows)
        unsigned long long key = lane < kRows ? ((unsigned long long)L.stamp[lane] << 8) | (unsigned)lane : ~0ull;
#pragma unroll
        for (int off = 1; off < kRows; off <<= 1) {
            const unsigned long long o = __shfl_xor(key, off, 64);
            key = o < key ? o : key;
        }
        r = (int)(__builtin_amdgcn_readfirstlane((unsigned)key) & 0xFF);
        row_writeback(L, r, rowlen, hist_g, lane);
        if (lane == 0) L.tags[r] = tag;
    }
    if (lane == 0) L.stamp[r] = ++clock;
    __builtin_amdgcn_wave_barrier();
    return r;
}

// The wave's row-cache state: the LRU clock and the last row it flushed to
// (consecutive items of a query tile mostly hit the same row, which then
// skips the tag lookup).  Every miss goes through row_get, which updates
// the pair, so the cached row is never one that was evicted since.
struct RowState {
    unsigned clock;
    long long last_tag;
    int last_row;
};

// Adds kc direct counts (radii e0..e0+kc-1) and, once per item, the bulk
// count for radii >= ehi, for (query field g, tree field tf).  The counters
// move from SGPRs into lanes 0..kc-1 (lane selects) and go to the LDS row with
// one 64-bit LDS atomic per lane.  A mirrored item of a symmetric job (a tile
// pair counted once for both orders) goes to the job's mirror plane, which
// the host adds to the direct plane both as is and transposed.
template <int N = KC>
__device__ __forceinline__ void flush(WaveLds &L, const JobDesc &J, int ne, int g, int tf, int e0, int kc,
                                      const unsigned *c, long long bulk, int ehi, bool mirror,
                                      long long *__restrict__ hist_g, unsigned *__restrict__ err, int lane,
                                      RowState &rs)
{
    const int rowlen = 2 * ne + 1;
    const long long slot = (long long)g * J.nq_fields + tf, tslot = (long long)tf * J.nq_fields + g;
    if (g < 0 || slot < 0 || slot >= J.nslots || (mirror && (tslot < 0 || tslot >= J.nslots))) {
        if (lane == 0) atomicOr(err, 1u);
        return;
    }
    unsigned v = 0;
#pragma unroll
    for (int e = 0; e < N; e++) v = lane == e ? c[e] : v;
    const long long tag = J.hist_base + (mirror ? J.nslots : 0) + slot;
    int r;
    if (tag == rs.last_tag) {
        r = rs.last_row;
        if (lane == 0) L.stamp[r] = ++rs.clock;
    } else {
        r = row_get(L, tag, rowlen, hist_g, lane, rs.clock);
        rs.last_tag = tag;
        rs.last_row = r;
    }
    if (lane < kc && v) atomicAdd((unsigned long long *)&L.rows[r * rowlen + e0 + lane], (unsigned long long)v);
    if (lane == 63 && bulk) atomicAdd((unsigned long long *)&L.rows[r * rowlen + ne + ehi], (unsigned long long)bulk);
    __builtin_amdgcn_wave_barrier();
}

// ---- the pair kernel ---------------------------------------------------------


__device__ __forceinline__ unsigned wave_sum_u32(unsigned v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return __builtin_amdgcn_readfirstlane(v);
}

// ---- the FP32 form on the matrix cores ----------------------------------------
//
// With a = q - c and b = p - c relative to the tree tile's centre c, the
// squared distance is a bilinear form plus the two norms,
//     n = |b|^2 + (|a|^2 - 2 a.b)  =  C + sum_k A_k B_k,
// A = (-2a, |a|^2 + eps) per query point, B = (b, 1) and C = |b|^2 per tree
// point (the tile's operand groups, kd_layout.hip, hold b and |b|^2).
// One v_mfma_f32_16x16x4_f32 gives D for 16 query x 16 tree points (lane l
// holds A[l%16][l/16], B[l/16][l%16] and D[4(l/16)+r][l%16] in register r),
// so a group of 16 tree points against the 64-point query tile is four MFMAs
// and sixteen D registers; the VALU only compares and the SALU counts.  The
// query lanes are re-centred on each tree tile of a run (mf_centre).
//
// Guard band (per job and tile size class, set on the host,
// fp32_thresholds): every D lies within G_e + eps of the reference's
// n64 = fma(dz,dz,fma(dy,dy,dx*dx)) for pairs near radius e, and eps keeps
// D > 0, so the bits of D order like D: bits(D) < lo32[e] means n64 < r_e^2,
// bits(D) >= lo32[e] + band_w means n64 >= r_e^2, and the pairs in between
// are recomputed with the FP64 formula.  One subtract-with-borrow per pair
// and radius gives both the count (the borrow) and the band test (the
// running minimum of the difference).
typedef float floatx4 __attribute__((ext_vector_type(4)));
static constexpr int KM = 4;                   // radii per pass of the FP32 loop
static constexpr size_t kMfLds = 4 * 64 * 4;   // per wave: -2a (3 x 64) and |a|^2 + eps (64)

// Re-centres the query lanes on tree tile t's centre and returns the A
// operands: A[b] (lane l: query 16b + l%16, component k = l/16: -2 a_k for
// k < 3, |a|^2 + eps for k = 3).  Invalid lanes get |a|^2 = +inf, so every D
// of theirs is +inf.
__device__ __forceinline__ void mf_centre(const JobDesc &J, int t, float eps, double qx, double qy, double qz,
                                          bool valid, float *mfl, int lane, float A[4])
{
    const double c0 = sload(J.tcentre, 3 * t), c1 = sload(J.tcentre, 3 * t + 1), c2 = sload(J.tcentre, 3 * t + 2);
    const float ax = (float)(qx - c0), ay = (float)(qy - c1), az = (float)(qz - c2);
    const float an =
        valid ? (float)((double)ax * ax + (double)ay * ay + (double)az * az + (double)eps) : INFINITY;
    mfl[lane] = -2.f * ax;
    mfl[64 + lane] = -2.f * ay;
    mfl[128 + lane] = -2.f * az;
    mfl[192 + lane] = an;
    __builtin_amdgcn_wave_barrier();
    const int k = lane >> 4, i = lane & 15;
#pragma unroll
    for (int b = 0; b < 4; b++) A[b] = mfl[64 * k + 16 * b + i];
    __builtin_amdgcn_wave_barrier();
}

// Pairs of D below the threshold (bits) lo, counted on the SALU from the
// borrow of one subtract; the difference goes into the band minimum (a
// borrow wraps it past every band).  Written out because the compiler turns
// the borrow into a select and a compare before the ballot.  (Measured on
// the 1M / 10M workload: this form beat the VOP3 form with the borrow in an
// SGPR pair of the compiler's choosing, and moving part of the count onto
// the VALU with add-with-carry, although both balance the issue counts.)
__device__ __forceinline__ unsigned below_count(float d, unsigned lo, unsigned &bmin)
{
    unsigned x, cnt;
    asm("v_subrev_co_u32_e32 %0, vcc, %2, %3\n\ts_bcnt1_i32_b64 %1, vcc"
        : "=v"(x), "=s"(cnt)
        : "s"(lo), "v"(__float_as_uint(d))
        : "vcc", "scc");
    bmin = x < bmin ? x : bmin;
    return cnt;
}

// Groups g0..g1-1 of one tree tile (its first group gs, first point pt0)
// against the re-centred query tile (first query point qb0): counts of the
// K radii e0..e0+K-1 into c, FP64 rechecks of the guard-band pairs.
template <int K>
__device__ __forceinline__ void mf_groups(const JobDesc &J, const MfClass &M, const EdgeTable &E, int e0, int g0,
                                          int g1, int gs, int pt0, int qb0, const float A[4], unsigned *c,
                                          unsigned long long *stat, int lane)
{
    const unsigned W = M.band_w;
    unsigned lo[K];
#pragma unroll
    for (int e = 0; e < K; e++) lo[e] = M.lo32[e0 + e];
    const int j = lane & 15;
    // the operands of group g + 1 are loaded while group g is counted
    typedef const float __attribute__((address_space(1))) *gfloat;
    const gfloat grp = (gfloat)J.tgrp;
    float v_next = grp[64 * (size_t)g0 + lane], bn_next = grp[64 * (size_t)g0 + 48 + j];
    for (int g = g0; g < g1; g++) {
        const float v = v_next, bn = bn_next;
        const int gn = g + 1 < g1 ? g + 1 : g;
        v_next = grp[64 * (size_t)gn + lane];
        bn_next = grp[64 * (size_t)gn + 48 + j];
        const float B = lane < 48 ? v : 1.0f;
        const floatx4 C = {bn, bn, bn, bn};
        floatx4 D[4];
#pragma unroll
        for (int b = 0; b < 4; b++) D[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[b], B, C, 0, 0, 0);
        // the compares below are inline assembly, before which the compiler
        // inserts no wait states for the MFMAs' writes (an early read sees the
        // register's old value): 19 wait states here, the eight-pass 16x16x4
        // form's need, with D as operands so that none of its readers moves
        // above them
        asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 2" : "+v"(D[0]), "+v"(D[1]), "+v"(D[2]), "+v"(D[3]));
        unsigned bmin = 0xFFFFFFFFu;
#pragma unroll
        for (int e = 0; e < K; e++) {
            unsigned s = 0;
#pragma unroll
            for (int w = 0; w < 16; w++) s += below_count(D[w >> 2][w & 3], lo[e], bmin);
            c[e] += s;
        }
        if (__ballot(bmin < W)) {
            // a pair of this group lies in a guard band (rare): recompute
            // the band pairs of each radius with the reference's FP64 formula
            for (int e = 0; e < K; e++) {
                const unsigned loe = M.lo32[e0 + e];
                const double rsq = E.rsq[e0 + e];
                unsigned add = 0;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    for (int r = 0; r < 4; r++) {
                        const float d = r == 0 ? D[b][0] : (r == 1 ? D[b][1] : (r == 2 ? D[b][2] : D[b][3]));
                        const unsigned x = __float_as_uint(d) - loe;
                        const unsigned long long m = __ballot(x < W);
                        if (!m) continue;
                        bool in = false;
                        if (x < W) {
                            const int qi = qb0 + 16 * b + 4 * (lane >> 4) + r;
                            const TPoint p = J.tp[pt0 + 16 * (g - gs) + j];
                            const double dx = p.x - J.qx[qi], dy = p.y - J.qy[qi], dz = p.z - J.qz[qi];
                            in = fma(dz, dz, fma(dy, dy, dx * dx)) < rsq;
                        }
                        add += (unsigned)__popcll(__ballot(in));
                        stat_add(stat, kStatRecheck, (unsigned long long)__popcll(m), lane);
                    }
                }
#pragma unroll
                for (int k = 0; k < K; k++)
                    if (k == e) c[k] += add;
            }
        }
    }
}

// FP64 loops over the padded quads of a tile range (padding points sit at
// +inf and never count).  Issue balance: a ballot count costs one VALU
// compare and two SALU ops (popcount, add), a per-lane count two VALU ops and
// no SALU; the last lane_radii(K) radii of a pass may be counted per lane to
// move issue from the SALU to the VALU.  With the quad windows the loop is
// VALU-bound (on the 1M / 10M workload: 88% of VALU issue, 60% of SALU), so
// every radius is counted by ballot (measured 1% faster than one radius per
// lane at K >= 5).
__host__ __device__ constexpr int lane_radii(int) { return 0; }

// A per-lane value v in 0..7 as three lane masks of its bits; the sum of v
// over the lanes of a mask is then three popcounts on the SALU.  (Ballots of
// `sel && bit` made the compiler move each mask through a VGPR and back:
// two VALU instructions per ballot.)
struct LaneBits3 {
    unsigned long long b0, b1, b2;
};
__device__ __forceinline__ LaneBits3 lane_bits3(unsigned v)
{
    return {__ballot((v & 1u) != 0u), __ballot((v & 2u) != 0u), __ballot((v & 4u) != 0u)};
}
__device__ __forceinline__ unsigned masked_sum3(unsigned long long m, const LaneBits3 &B)
{
    return (unsigned)__popcll(m & B.b0) + 2u * (unsigned)__popcll(m & B.b1) + 4u * (unsigned)__popcll(m & B.b2);
}

// The quads in mask m (bit j: quad j of the batch at p) against the query
// lanes, counting radii A .. K-1 of the pass (the quads' windows start at A).
template <int K, int A>
__device__ __forceinline__ void quads_from(const double *__restrict__ p, unsigned long long m, const double *thr64,
                                           double qx, double qy, double qz, unsigned *c, unsigned (*cv)[4])
{
    constexpr int KV = lane_radii(K);
    constexpr int KS = K - KV;
    while (m) {
        const int j = __ffsll((long long)m) - 1;
        m &= m - 1;
        const double *q = p + 12 * j;
        double px[4], py[4], pz[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            px[u] = sload(q, u);
            py[u] = sload(q, 4 + u);
            pz[u] = sload(q, 8 + u);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double dx = px[u] - qx, dy = py[u] - qy, dz = pz[u] - qz;
            const double n = fma(dz, dz, fma(dy, dy, dx * dx));
#pragma unroll
            for (int e = A; e < KS; e++) c[e] += (unsigned)__popcll(__ballot(n < thr64[e]));
#pragma unroll
            for (int e = (A > KS ? A : KS); e < K; e++) cv[e - KS][u] += (unsigned)(n < thr64[e]);
        }
    }
}

template <int K, int A>
__device__ __forceinline__ void quads_by_entry(const double *__restrict__ p, unsigned long long live, int lo,
                                               const double *thr64, double qx, double qy, double qz, unsigned *c,
                                               unsigned (*cv)[4])
{
    if constexpr (A < K) {
        const unsigned long long m = live & __ballot(lo == A);
        if (m) quads_from<K, A>(p, m, thr64, qx, qy, qz, c, cv);
        quads_by_entry<K, A + 1>(p, live, lo, thr64, qx, qy, qz, c, cv);
    }
}

// The radius window of quad gi (none when gi >= g1: lo = hi = K, np = 0)
// against the query tile's box qb over the pass's K radii thr[0..K-1], with
// the enumeration's margins: lo = radii whose r^2 is at most the smallest
// possible squared distance, hi = those at most the largest.  first: the
// window does not start below the pass (thr_prev: r^2 of the radius just
// below it, negative for an item's first pass).
template <int K>
__device__ __forceinline__ void quad_window(const BoxF *__restrict__ qboxes, const unsigned *__restrict__ qmeta,
                                            int gi, int g1, const double *thr, double thr_prev, const Box &qb,
                                            int &lo, int &hi, unsigned &np, bool &first)
{
    lo = K;
    hi = K;
    np = 0;
    first = false;
    if (gi < g1) {
        const BoxF f = qboxes[gi];
        np = (qmeta[gi] >> 8) & 7u;
        double dmin = 0.0, dmax = 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double flo = f.lo[d], fhi = f.hi[d];
            const double gap = fmax(0.0, fmax(flo - qb.hi[d], qb.lo[d] - fhi));
            const double ext = fmax(fabs(qb.hi[d] - flo), fabs(fhi - qb.lo[d]));
            dmin += gap * gap;
            dmax += ext * ext;
        }
        dmin *= 1.0 - kBoxMargin;
        dmax *= 1.0 + kBoxMargin;
        lo = 0;
        hi = 0;
#pragma unroll
        for (int e = 0; e < K; e++) {
            lo += thr[e] <= dmin ? 1 : 0;
            hi += thr[e] <= dmax ? 1 : 0;
        }
        first = !(dmin < thr_prev);
    }
}

// pure_loop64 with a radius window per quad.  The quads are taken 64 at a
// time, one per lane, and each lane tests its quad's box against the query
// tile's box (with the enumeration's margins) over the pass's K radii:
//   lo = radii with r^2 <= the smallest possible squared distance (no pair of
//        the quad counts for them),
//   hi = radii with r^2 <= the largest possible one (every pair counts from
//        hi on).
// A quad with lo == hi needs no distance: its points count in bulk for the
// radii from hi on (none at all when hi == K: a quad beyond the pass, or
// beyond every radius).  The others are evaluated for radii lo .. K-1, in
// groups of equal lo (one compare sequence per entry radius).  thr_prev is
// r^2 of the radius just below the pass (negative for the item's first pass):
// a quad whose window starts below it was evaluated, and counted in
// `evaluated`, in an earlier pass.
template <int K>
__device__ __forceinline__ void pure_loop64w(const double *__restrict__ quad, const BoxF *__restrict__ qboxes,
                                             const unsigned *__restrict__ qmeta, int g0, int g1, const double *thr64,
                                             double thr_prev, const Box &qb, double qx, double qy, double qz, int nv,
                                             unsigned *c, unsigned long long *stat, int lane)
{
    constexpr int KV = lane_radii(K);
    constexpr int KS = K - KV;
    unsigned cv[KV > 0 ? KV : 1][4];
#pragma unroll
    for (int e = 0; e < (KV > 0 ? KV : 1); e++)
#pragma unroll
        for (int u = 0; u < 4; u++) cv[e][u] = 0u;
    unsigned full[K];
#pragma unroll
    for (int e = 0; e < K; e++) full[e] = 0u;
    unsigned first_pts = 0;
    for (int gb = g0; gb < g1; gb += 64) {
        int lo, hi;
        unsigned np;
        bool first;
        quad_window<K>(qboxes, qmeta, gb + lane, g1, thr64, thr_prev, qb, lo, hi, np, first);
        const unsigned long long live = __ballot(lo < hi);
        const LaneBits3 nb = lane_bits3(np);
        const unsigned long long dead = ~live;
        if (dead & __ballot(hi < K)) {  // quads entirely inside some radius of the pass (rare)
#pragma unroll
            for (int e = 0; e < K; e++) full[e] += masked_sum3(dead & __ballot(hi <= e), nb);
        }
        first_pts += masked_sum3(live & __ballot(first), nb);
        quads_by_entry<K, 0>(quad + 12LL * gb, live, lo, thr64, qx, qy, qz, c, cv);
    }
    stat_add(stat, kStatEval, (unsigned long long)nv * first_pts, lane);
#pragma unroll
    for (int e = 0; e < K; e++) c[e] += (unsigned)nv * full[e];
#pragma unroll
    for (int e = 0; e < KV; e++) c[KS + e] += wave_sum_u32(cv[e][0] + cv[e][1] + cv[e][2] + cv[e][3]);
}

// c += (n < thr) as one compare into VCC and one add-with-carry.  Written out
// because the compiler turns a sum of several compare results into bit
// packing and a popcount (up to four VALU ops per compare instead of two).
// The compare is ordered: a NaN distance (a lane parked at +inf against a
// +inf padding point) never counts.
__device__ __forceinline__ void count_below(unsigned &c, double n, double thr)
{
    asm("v_cmp_gt_f64_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, 0, %0, vcc" : "+v"(c) : "s"(thr), "v"(n) : "vcc");
}

// Per-lane counts over the padded quads g0..g1-1: cnt[e] += number of tree
// points with squared distance below thr[e] (a compare and an add-with-carry
// per point and radius), for query tiles whose lanes belong to several query
// fields: one pass over the tree points serves every field group, whose
// counts are then summed over its lanes.
template <int K>
__device__ __forceinline__ void lane_loop64(const double *__restrict__ quad, int g0, int g1, const double *thr64,
                                            double qx, double qy, double qz, unsigned *cnt)
{
    const double *p = quad + 12LL * g0;
    for (int g = g0; g < g1; g++, p += 12) {
        double px[4], py[4], pz[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            px[u] = sload(p, u);
            py[u] = sload(p, 4 + u);
            pz[u] = sload(p, 8 + u);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double dx = px[u] - qx, dy = py[u] - qy, dz = pz[u] - qz;
            const double n = fma(dz, dz, fma(dy, dy, dx * dx));
#pragma unroll
            for (int e = 0; e < K; e++) count_below(cnt[e], n, thr64[e]);
        }
    }
}

// Per-lane counts of the quads in mask m (bit j: quad j of the batch at p)
// for radii A .. K-1 (the quads' windows start at A).
template <int K, int A>
__device__ __forceinline__ void lane_quads_from(const double *__restrict__ p, unsigned long long m, const double *thr,
                                                double qx, double qy, double qz, unsigned *cnt)
{
    while (m) {
        const int j = __ffsll((long long)m) - 1;
        m &= m - 1;
        const double *q = p + 12 * j;
        double px[4], py[4], pz[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            px[u] = sload(q, u);
            py[u] = sload(q, 4 + u);
            pz[u] = sload(q, 8 + u);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double dx = px[u] - qx, dy = py[u] - qy, dz = pz[u] - qz;
            const double n = fma(dz, dz, fma(dy, dy, dx * dx));
#pragma unroll
            for (int e = A; e < K; e++) count_below(cnt[e], n, thr[e]);
        }
    }
}

template <int K, int A>
__device__ __forceinline__ void lane_quads_by_entry(const double *__restrict__ p, unsigned long long live, int lo,
                                                    const double *thr, double qx, double qy, double qz, unsigned *cnt)
{
    if constexpr (A < K) {
        const unsigned long long m = live & __ballot(lo == A);
        if (m) lane_quads_from<K, A>(p, m, thr, qx, qy, qz, cnt);
        lane_quads_by_entry<K, A + 1>(p, live, lo, thr, qx, qy, qz, cnt);
    }
}

// lane_loop64 with the radius windows of pure_loop64w: per-lane counts
// cnt[e] += points of quads g0..g1-1 under radius e, quads beyond the pass
// skipped, quads entirely inside radius e added in bulk; returns the points
// of the quads evaluated for the first time in this pass.
template <int K>
__device__ __forceinline__ unsigned lane_loop64w(const double *__restrict__ quad, const BoxF *__restrict__ qboxes,
                                                 const unsigned *__restrict__ qmeta, int g0, int g1, const double *thr,
                                                 double thr_prev, const Box &qb, double qx, double qy, double qz,
                                                 unsigned *cnt, int lane)
{
    unsigned first_pts = 0;
    for (int gb = g0; gb < g1; gb += 64) {
        int lo, hi;
        unsigned np;
        bool first;
        quad_window<K>(qboxes, qmeta, gb + lane, g1, thr, thr_prev, qb, lo, hi, np, first);
        const unsigned long long live = __ballot(lo < hi);
        const LaneBits3 nb = lane_bits3(np);
        const unsigned long long dead = ~live;
        if (dead & __ballot(hi < K)) {
#pragma unroll
            for (int e = 0; e < K; e++) cnt[e] += masked_sum3(dead & __ballot(hi <= e), nb);
        }
        first_pts += masked_sum3(live & __ballot(first), nb);
        lane_quads_by_entry<K, 0>(quad + 12LL * gb, live, lo, thr, qx, qy, qz, cnt);
    }
    return first_pts;
}

// A run of one tree field (quads g0..g1-1; np points when the bulk count is
// due, else 0) against a query tile whose lanes hold several query fields:
// per-lane counts in one pass over the points, then each field group's sum
// over its lanes.
__device__ __forceinline__ unsigned group_run(WaveLds &L, const JobDesc &J, int ne, const double *qd, int g0, int g1,
                                              const double *thr64, double thr_prev, const Box &qb, double lx,
                                              double ly, double lz, bool valid, int qf, unsigned long long vm, int tf,
                                              int e0, int kc, int np, int ehi, bool mirror,
                                              long long *__restrict__ hist_g, unsigned *__restrict__ err, int lane,
                                              RowState &rs)
{
    unsigned cnt[KC], c[KC];
#pragma unroll
    for (int e = 0; e < KC; e++) cnt[e] = 0u;
    unsigned fresh = 0;  // points of the quads evaluated for the first time
#define KD_GW(k) fresh = lane_loop64w<k>(qd, J.tquad_box, J.tquad_meta, g0, g1, thr64, thr_prev, qb, lx, ly, lz, cnt, lane)
    switch (kc) {
    case 1: KD_GW(1); break;
    case 2: KD_GW(2); break;
    case 3: KD_GW(3); break;
    case 4: KD_GW(4); break;
    case 5: KD_GW(5); break;
    case 6: KD_GW(6); break;
    case 7: KD_GW(7); break;
    default: break;
    }
#undef KD_GW
    unsigned long long remaining = vm;
    while (remaining) {  // each query field group: its lanes' counts
        const int leader = __ffsll((long long)remaining) - 1;
        const int gq = __builtin_amdgcn_readlane(qf, leader);
        const bool ing = valid && qf == gq;
        const unsigned long long gm = __ballot(ing);
        remaining &= ~gm;
#pragma unroll
        for (int e = 0; e < KC; e++) c[e] = e < kc ? wave_sum_u32(ing ? cnt[e] : 0u) : 0u;
        flush(L, J, ne, gq, tf, e0, kc, c, (long long)__popcll(gm) * np, ehi, mirror, hist_g, err, lane, rs);
    }
    return fresh;
}

// The pair loop over one item list.  SINGLE: every query tile of the list
// holds one query field (the FP64 kernel's first list); otherwise a tile may
// hold several, and the FP64 kernel counts unit-weight tree runs per lane for
// all field groups at once.
template <bool MF, bool SINGLE>
__device__ __forceinline__ void pairs_body(const JobDesc *__restrict__ jobs, const Item *__restrict__ items,
                                           long long nitems, unsigned long long *__restrict__ next, int chunk,
                                           const EdgeTable &E, long long *__restrict__ hist_g,
                                           unsigned long long *__restrict__ stats, unsigned *__restrict__ err,
                                           int skip_mf)
{
    extern __shared__ unsigned char lds_raw[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int ne = E.ne;
    const int rowlen = 2 * ne + 1;
    // per wave: the row cache, then (FP32 form) the re-centred query operands
    unsigned char *base = lds_raw + wid * (wave_lds_bytes(rowlen) + (MF ? kMfLds : 0));
    float *mfl = (float *)(base + wave_lds_bytes(rowlen));
    WaveLds L;
    L.rows = (long long *)base;
    L.tags = (long long *)(base + (size_t)kRows * rowlen * 8);
    L.stamp = (unsigned *)(base + (size_t)kRows * rowlen * 8 + kRows * 8);
    unsigned long long *stat = (unsigned long long *)(base + (size_t)kRows * rowlen * 8 + kRows * 8 + kRows * 4);
    for (int b = lane; b < kRows * rowlen; b += 64) L.rows[b] = 0;
    if (lane < kRows) {
        L.tags[lane] = -1;
        L.stamp[lane] = 0u;
    }
    if (lane < kStats) stat[lane] = 0ull;
    RowState rs = {0u, -1LL, 0};
    __builtin_amdgcn_wave_barrier();

    int cur_job = -1, cur_q = -1;
    double qx = 0, qy = 0, qz = 0;
    int qf = 0;
    bool valid = false;

    // The items (in query-tile order) are cut into kXcds contiguous segments.
    // Workgroups are dispatched to the XCDs round-robin, so block b works
    // first on segment b % kXcds, which keeps each XCD's L2 on one region of
    // the trees, then helps with the other segments until all are drained.
    int seg = blockIdx.x & (kXcds - 1), visited = 0;
    while (true) {
        const long long s0 = nitems * seg / kXcds, s1 = nitems * (seg + 1) / kXcds;
        unsigned long long off = 0;
        if (lane == 0) off = atomicAdd(&next[seg * kNextStride], (unsigned long long)chunk);
        off = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(off >> 32)) << 32) |
              (unsigned)__builtin_amdgcn_readfirstlane((unsigned)off);
        if (off >= (unsigned long long)(s1 - s0)) {
            if (++visited == kXcds) break;
            seg = (seg + 1) & (kXcds - 1);
            continue;
        }
        const long long start = s0 + (long long)off;
        const long long end = min(start + chunk, s1);
        for (long long i = start; i < end; i++) {
            const Item it = sload(items, (int)i);
            const int job = it.job & kJobMask;
            const bool mirror = it.job < 0;
            const JobDesc &J = jobs[job];
            if (job != cur_job || it.q != cur_q) {
                cur_job = job;
                cur_q = it.q;
                const int b0 = J.qstart[it.q];
                const int qcount = J.qstart[it.q + 1] - b0;
                valid = lane < qcount;
                if (valid) {
                    qx = J.qx[b0 + lane];
                    qy = J.qy[b0 + lane];
                    qz = J.qz[b0 + lane];
                    qf = J.qf[b0 + lane];
                }
            }
            if (it.job & kSubtreeItem) {
                // bulk only: (query points of each field group) x (subtree
                // weight of each tree field); one field, or the kept per-field
                // weights of a subtree with several
                if (MF) continue;  // the FP64 kernel takes these
                if (!subtree_item_ok(it, J.P)) {
                    if (lane == 0) atomicOr(err, kErrItem);
                    continue;
                }
                const unsigned hf = sload(J.hfield, it.t);
                const int nf = hf < kHEmpty ? 1 : J.nfw;
                unsigned c0[KC];
#pragma unroll
                for (int e = 0; e < KC; e++) c0[e] = 0u;
                for (int k = 0; k < nf; k++) {
                    const int tf = hf < kHEmpty ? (int)hf : k;
                    const long long wsum = hf < kHEmpty ? sload(J.hweight, it.t) : sload(J.hfw, it.t * J.nfw + k);
                    if (!wsum) continue;
                    unsigned long long remaining = __ballot(valid);
                    while (remaining) {
                        const int leader = __ffsll((long long)remaining) - 1;
                        const int g = __builtin_amdgcn_readlane(qf, leader);
                        const unsigned long long gm = __ballot(valid && qf == g);
                        remaining &= ~gm;
                        flush(L, J, ne, g, tf, it.ehi, 0, c0, (long long)__popcll(gm) * wsum, it.ehi, mirror, hist_g,
                              err, lane, rs);
                    }
                }
                continue;
            }
            if (!tile_item_ok(it, J.ntt)) {
                if (lane == 0) atomicOr(err, kErrItem);
                continue;
            }
            const int t = it.t;
            const int t0 = __builtin_amdgcn_readfirstlane(J.tstart[t]);
            const int nt = it.tiles();
            const int t1 = __builtin_amdgcn_readfirstlane(J.tstart[t + nt]);
            if (t0 >= t1) continue;
            const int elo = it.elo, ehi = it.ehi;
            const unsigned tmeta = sload(J.tmeta, t);
            const bool fast = (tmeta & 3u) == 3u;
            const unsigned long long vm = __ballot(valid);
            // items of the matrix-core loop: one tree field, unit weights and
            // every candidate radius above the FP32 resolution of the tiles;
            // the FP32 kernel takes these and only these, and the FP64 kernel
            // skips them when it runs after it on the same list (skip_mf)
            const int mcls = (int)((tmeta >> 16) & 0xFu);  // the size class of the run's tiles
            // (items whose window reaches past the largest radius go to the
            // FP64 loop, whose quad windows skip the quads beyond it)
            const bool mf_item = (MF || skip_mf) && fast && vm && ehi > elo && ehi < ne && elo >= J.mfc[mcls].e32;
            if (MF != mf_item) {
                if (MF || skip_mf) continue;
            } else if (MF) {
                // one tree field, unit weights, every candidate radius above
                // the FP32 resolution of the tiles: the matrix-core loop
                const MfClass &M = J.mfc[mcls];
                const int tf = (int)((tmeta >> 8) & 0xFFu);
                const int g_lead = __builtin_amdgcn_readlane(qf, __ffsll((long long)vm) - 1);
                const int cut_lo = it.cut_lo(), cut_hi = it.cut_hi();
                const long long nv = __popcll(vm);
                const int qb0 = __builtin_amdgcn_readfirstlane(J.qstart[it.q]);
                for (int e0 = elo; e0 < ehi; e0 += KM) {
                    const int kc = ehi - e0 < KM ? ehi - e0 : KM;
                    unsigned c[KM];
#pragma unroll
                    for (int e = 0; e < KM; e++) c[e] = 0u;
                    for (int tt = t; tt < t + nt; tt++) {
                        float A[4];
                        mf_centre(J, tt, M.eps32, qx, qy, qz, valid, mfl, lane, A);
                        const int gs = __builtin_amdgcn_readfirstlane(J.tgstart[tt]);
                        const int pt0 = __builtin_amdgcn_readfirstlane(J.tstart[tt]);
                        const int np = __builtin_amdgcn_readfirstlane(J.tstart[tt + 1]) - pt0;
                        // groups left after the trim pass's cuts (whole quads of
                        // four points off the first tile's start and the last's end)
                        int g0 = gs, g1 = __builtin_amdgcn_readfirstlane(J.tgstart[tt + 1]);
                        if (tt == t) g0 += cut_lo >> 2;
                        if (tt == t + nt - 1 && cut_hi) g1 = gs + (4 * (((np + 3) >> 2) - cut_hi) + 15) / 16;
                        if (g0 >= g1) continue;
                        if (e0 == elo) {
                            const unsigned long long ev = nv * (unsigned long long)(min(np, 16 * (g1 - gs)) - 16 * (g0 - gs));
                            stat_add(stat, kStatEval, ev, lane);
                            stat_add(stat, kStatMf, ev, lane);
                        }
                        switch (kc) {
                        case 1: mf_groups<1>(J, M, E, e0, g0, g1, gs, pt0, qb0, A, c, stat, lane); break;
                        case 2: mf_groups<2>(J, M, E, e0, g0, g1, gs, pt0, qb0, A, c, stat, lane); break;
                        case 3: mf_groups<3>(J, M, E, e0, g0, g1, gs, pt0, qb0, A, c, stat, lane); break;
                        default: mf_groups<4>(J, M, E, e0, g0, g1, gs, pt0, qb0, A, c, stat, lane); break;
                        }
                    }
                    // the bulk count (radii from ehi on) once, with the first pass
                    flush<KM>(L, J, ne, g_lead, tf, e0, kc, c, e0 == elo && ehi < ne ? nv * (long long)(t1 - t0) : 0,
                          ehi, mirror, hist_g, err, lane, rs);
                }
                continue;
            }
            if (skip_mf && ehi > elo) stat_add(stat, kStatFp64Items, 1ull, lane);
            if ((tmeta & 2u) && vm) {
                // unit weights: the tree points in runs of one field (a run
                // of pure tiles, or the field runs of one mixed tile), each
                // over whole quads
                const int g_lead = __builtin_amdgcn_readlane(qf, __ffsll((long long)vm) - 1);
                const double lx = valid ? qx : INFINITY, ly = valid ? qy : INFINITY, lz = valid ? qz : INFINITY;
                const int cut_lo = it.cut_lo(), cut_hi = it.cut_hi();
                const int qg0 = __builtin_amdgcn_readfirstlane(J.tquad_start[t]) + cut_lo;
                const int qg1 = __builtin_amdgcn_readfirstlane(J.tquad_start[t + nt]) - cut_hi;
                if (qg0 > qg1) {  // cuts beyond the run's quads
                    if (lane == 0) atomicOr(err, kErrItem);
                    continue;
                }
                // points left after the cuts: whole quads off the first tile; off
                // the last, its last (partial) quad and whole ones before it
                int npts = t1 - t0 - 4 * cut_lo;
                if (cut_hi) {
                    const int nl = t1 - __builtin_amdgcn_readfirstlane(J.tstart[t + nt - 1]);
                    npts -= nl - 4 * (((nl + 3) >> 2) - cut_hi);
                }
                const double *qd = J.tquad;
                const Box qb = sload(J.qbox, it.q);
                const int nv = __popcll(vm);
                for (int e0 = elo;; e0 += KC) {
                    const int kc = ehi - e0 >= KC ? KC : (ehi - e0 > 0 ? ehi - e0 : 0);
                    const bool first = e0 == elo;
                    double thr64[KC];
#pragma unroll
                    for (int e = 0; e < KC; e++) thr64[e] = E.rsq[min(e0 + e, kMaxEdges - 1)];
                    const double thr_prev = first ? -1.0 : E.rsq[e0 - 1];
                    for (int g = qg0; g < qg1;) {
                        int tf = (int)((tmeta >> 8) & 0xFFu), ge = qg1, np = npts;
                        if (!fast) {  // the run ends at the next quad of another field
                            const unsigned m0 = sload(J.tquad_meta, g);
                            tf = (int)(m0 & 0xFFu);
                            np = (int)(m0 >> 8);
                            for (ge = g + 1; ge < qg1; ge++) {
                                const unsigned m = sload(J.tquad_meta, ge);
                                if ((m & 0xFFu) != (unsigned)tf) break;
                                np += (int)(m >> 8);
                            }
                        }
                        unsigned c[KC];
#pragma unroll
                        for (int e = 0; e < KC; e++) c[e] = 0u;
                        if (SINGLE) {
                            const BoxF *qbx = J.tquad_box;
                            const unsigned *qm = J.tquad_meta;
#define KD_W(k) pure_loop64w<k>(qd, qbx, qm, g, ge, thr64, thr_prev, qb, lx, ly, lz, nv, c, stat, lane)
                            switch (kc) {
                            case 1: KD_W(1); break;
                            case 2: KD_W(2); break;
                            case 3: KD_W(3); break;
                            case 4: KD_W(4); break;
                            case 5: KD_W(5); break;
                            case 6: KD_W(6); break;
                            case 7: KD_W(7); break;
                            default: break;
                            }
#undef KD_W
                            flush(L, J, ne, g_lead, tf, e0, kc, c,
                                  first && ehi < ne ? (long long)__popcll(vm) * np : 0, ehi, mirror, hist_g, err,
                                  lane, rs);
                        } else {
                            const unsigned fresh =
                                group_run(L, J, ne, qd, g, ge, thr64, thr_prev, qb, lx, ly, lz, valid, qf, vm, tf, e0,
                                          kc, first ? np : 0, ehi, mirror, hist_g, err, lane, rs);
                            stat_add(stat, kStatEval, (unsigned long long)nv * fresh, lane);
                        }
                        g = ge;
                    }
                    if (e0 + KC >= ehi) break;
                }
                continue;
            }

            // general path: weights above one (duplicates) in the tile; FP64
            // in both forms
            unsigned long long remaining = __ballot(valid);
            while (remaining) {
                const int leader = __ffsll((long long)remaining) - 1;
                const int g = __builtin_amdgcn_readlane(qf, leader);
                const bool ing = valid && qf == g;
                const unsigned long long gm = __ballot(ing);
                remaining &= ~gm;
                const int ng = __popcll(gm);
                // query coordinates of this group; lanes outside it are parked at +inf
                const double gx = ing ? qx : INFINITY, gy = ing ? qy : INFINITY, gz = ing ? qz : INFINITY;
                for (int e0 = elo;; e0 += KC) {
                    const int kc = ehi - e0 >= KC ? KC : (ehi - e0 > 0 ? ehi - e0 : 0);
                    const bool first = e0 == elo;
                    double thr64[KC];
#pragma unroll
                    for (int e = 0; e < KC; e++) thr64[e] = E.rsq[min(e0 + e, kMaxEdges - 1)];
                    unsigned c[KC];
#pragma unroll
                    for (int e = 0; e < KC; e++) c[e] = 0u;
                    long long wseg = 0;
                    int cur_tf = -1;
                    for (int j = t0; j < t1; j++) {
                        const TPoint p = sload(J.tp, j);
                        const int tf = (p.meta >> 8) & 0xFF;
                        const unsigned w = p.meta & 0xFF;
                        if (tf != cur_tf) {
                            if (cur_tf >= 0)
                                flush(L, J, ne, g, cur_tf, e0, kc, c, first ? (long long)ng * wseg : 0, ehi, mirror,
                                      hist_g, err, lane, rs);
#pragma unroll
                            for (int e = 0; e < KC; e++) c[e] = 0u;
                            wseg = 0;
                            cur_tf = tf;
                        }
                        wseg += w;
                        if (kc == 0) continue;
                        const double dx = p.x - gx, dy = p.y - gy, dz = p.z - gz;
                        const double n = fma(dz, dz, fma(dy, dy, dx * dx));
#pragma unroll
                        for (int e = 0; e < KC; e++)
                            if (e < kc) c[e] += (unsigned)__popcll(__ballot(n < thr64[e])) * w;
                    }
                    if (cur_tf >= 0)
                        flush(L, J, ne, g, cur_tf, e0, kc, c, first ? (long long)ng * wseg : 0, ehi, mirror, hist_g,
                              err, lane, rs);
                    if (first && kc > 0)
                        stat_add(stat, kStatEval, (unsigned long long)ng * (unsigned long long)(t1 - t0), lane);
                    if (e0 + KC >= ehi) break;
                }
            }
        }
    }
    for (int r = 0; r < kRows; r++) row_writeback(L, r, rowlen, hist_g, lane);
    if (lane == 0) {
        atomicAdd(&stats[0], stat[kStatEval]);
        atomicAdd(&stats[1], stat[kStatRecheck]);
        atomicAdd(&stats[3], stat[kStatFp64Items]);
        if (MF) atomicAdd(&stats[4], stat[kStatMf]);
    }
}

// FP64, the query tiles of one field.  Eight waves per SIMD (64 VGPRs): the
// loops are issue-bound and need co-resident waves to pair VALU with SALU
// issue; the few spills this costs sit outside the loops.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_pairs64(
    const JobDesc *__restrict__ jobs, const Item *__restrict__ items, long long nitems,
    unsigned long long *__restrict__ next, int chunk, EdgeTable E, long long *__restrict__ hist_g,
    unsigned long long *__restrict__ stats, unsigned *__restrict__ err, int skip_mf)
{
    pairs_body<false, true>(jobs, items, nitems, next, chunk, E, hist_g, stats, err, skip_mf);
}

// FP64, the query tiles that hold several fields.  Its per-lane loop needs no
// SALU issue beside it; at least seven waves per SIMD (at most 72 VGPRs) fit
// it without spills.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_pairs64_mixed(
    const JobDesc *__restrict__ jobs, const Item *__restrict__ items, long long nitems,
    unsigned long long *__restrict__ next, int chunk, EdgeTable E, long long *__restrict__ hist_g,
    unsigned long long *__restrict__ stats, unsigned *__restrict__ err)
{
    pairs_body<false, false>(jobs, items, nitems, next, chunk, E, hist_g, stats, err, 0);
}

// FP32 guarded, the query tiles of one field: the matrix-core loop for runs
// of one tree field with unit weights whose radii are above the tiles' FP32
// resolution, the FP64 loops for the rest.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_pairs32(
    const JobDesc *__restrict__ jobs, const Item *__restrict__ items, long long nitems,
    unsigned long long *__restrict__ next, int chunk, EdgeTable E, long long *__restrict__ hist_g,
    unsigned long long *__restrict__ stats, unsigned *__restrict__ err)
{
    pairs_body<true, true>(jobs, items, nitems, next, chunk, E, hist_g, stats, err, 0);
}

// Thresholds of the FP32 form (the guard band of mf_groups) for tree tiles
// whose points lie within h of the tile centre.  For a pair near radius r:
// |a| + |b| <= r + 2h, and |D - eps - n64| is at most
//   6 u32 (sum |A_k B_k| + |C|)        MFMA sums (the probe saw 3 u32)
//   + u32 (|a'|^2 + eps) + u32 |b'|^2  rounding of the norms
//   + 2.01 u32 r (r + 2h)              a, b rounded to FP32
//   + 7 u64 r^2                        the reference's own FP64 rounding
// <= G = u32 (9 (r + 2h)^2 + 8 eps + 2.01 r (r + 2h)) + 7 u64 r^2, and
// eps = 40 u32 h^2 keeps every D positive.  Radii whose band is wider than
// 2e-5 r^2 (small against the tiles) run in FP64.
static void fp32_thresholds(MfClass &M, const EdgeTable &E, double h)
{
    const double u32 = std::ldexp(1.0, -24), u64 = std::ldexp(1.0, -53);
    const double eps = 40.0 * u32 * h * h;
    float e32f = (float)eps;
    if ((double)e32f < eps) e32f = std::nextafter(e32f, INFINITY);
    memset(&M, 0, sizeof(M));
    M.eps32 = e32f;
    M.e32 = E.ne;
    M.band_w = 1;
    for (int e = E.ne - 1; e >= 0; e--) {
        const double r = E.r[e], rsq = E.rsq[e];
        const double G = u32 * (9.0 * (r + 2 * h) * (r + 2 * h) + 8.0 * (double)e32f + 2.01 * r * (r + 2 * h)) +
                         7.0 * u64 * rsq;
        if (!(G <= 2e-5 * rsq)) break;
        const double xl = rsq - G + (double)e32f, xh = rsq + G + (double)e32f;
        float lo = (float)xl, hi = (float)xh;
        if ((double)lo > xl) lo = std::nextafter(lo, 0.0f);
        if ((double)hi < xh) hi = std::nextafter(hi, INFINITY);
        unsigned lb, hb;
        memcpy(&lb, &lo, 4);
        memcpy(&hb, &hi, 4);
        M.lo32[e] = lb;
        M.band_w = std::max(M.band_w, hb - lb + 1);
        M.e32 = e;
    }
}

// ---- one pair-count call ----------------------------------------------------
//
// A call is prepared on the host (radii, job table, histogram rows), then runs
// as one device sequence over buffers sized in advance:
//   clears -> counting pass of the enumeration -> scans (item offsets)
//   -> writing pass -> trim pass -> pair kernels -> copy-out
// The item list's length is only known after the counting pass.  The eager
// call (count_pairs) reads it back between the two passes; a plan (PairPlan)
// runs the whole sequence once to size the list, then records it into a graph
// whose replays allocate nothing and never wait on the host.  The enumeration
// is deterministic, so every replay writes the same number of items; the
// writing pass still refuses to write past the list and a check kernel flags
// a replay whose totals differ (kErrPlan).

struct CallSetup {
    std::vector<double> radii, edges;  // as given; distinct and sorted
    EdgeTable E;
    int ne = 0, rowlen = 0, merge = 1, maxP = 1;
    bool fp32 = false;
    std::vector<JobDesc> hj;
    std::vector<MfClass> hmfc;
    std::vector<long long> qt_prefix;
    long long rows = 0, total_qt = 0;
};

static void setup_call(const std::vector<PairJob> &jobs, const std::vector<double> &radii, int precision,
                       CallSetup &c)
{
    const int nr = (int)radii.size();
    KD_REQUIRE(nr >= 1, "pair count: at least one radius is required");
    for (double r : radii)
        KD_REQUIRE(std::isfinite(r) && r >= 0.0, "pair count: radii must be finite and >= 0");
    KD_REQUIRE(precision == KD_PRECISION_FP64 || precision == KD_PRECISION_FP32_GUARDED,
               "pair count: unknown precision form");
    c.radii = radii;
    c.edges = radii;
    std::sort(c.edges.begin(), c.edges.end());
    c.edges.erase(std::unique(c.edges.begin(), c.edges.end()), c.edges.end());
    c.ne = (int)c.edges.size();
    KD_REQUIRE(c.ne <= kMaxEdges, "pair count: at most 64 distinct radii per call");
    memset(&c.E, 0, sizeof(c.E));
    c.E.ne = c.ne;
    for (int e = 0; e < c.ne; e++) {
        c.E.r[e] = c.edges[e];
        c.E.rsq[e] = c.edges[e] * c.edges[e];
    }
    c.rowlen = 2 * c.ne + 1;
    c.fp32 = precision == KD_PRECISION_FP32_GUARDED;
    c.merge = c.fp32 ? 2 : 3;

    const int nj = (int)jobs.size();
    c.hj.assign(nj, JobDesc{});
    c.hmfc.assign((size_t)nj * kMfClasses, MfClass{});
    c.qt_prefix.assign(nj + 1, 0);
    c.rows = 0;
    c.maxP = 1;
    for (int i = 0; i < nj; i++) {
        const kd_device_tree *T = jobs[i].tree, *Q = jobs[i].query;
        KD_REQUIRE(!T->query_only, "pair count: the tree side must be a built tree");
        JobDesc &J = c.hj[i];
        memset(&J, 0, sizeof(J));
        J.qx = Q->q.x.p;
        J.qy = Q->q.y.p;
        J.qz = Q->q.z.p;
        J.qf = Q->q.f.p;
        J.qstart = Q->q.t.start.p;
        J.qbox = Q->q.t.box.p;
        J.qfmin = Q->q.fmin.p;
        J.qfmax = Q->q.fmax.p;
        J.nqt = Q->q.t.ntiles;
        J.tp = T->tr.pts.p;
        J.tquad = T->tr.quad.p;
        J.tquad_start = T->tr.quad_start.p;
        J.tquad_meta = T->tr.quad_meta.p;
        J.tquad_box = T->tr.quad_box.p;
        J.tcentre = T->tr.centre.p;
        J.tgrp = T->tr.grp.p;
        J.tgstart = T->tr.gstart.p;
        for (int k = 0; k < kMfClasses; k++)
            fp32_thresholds(c.hmfc[(size_t)i * kMfClasses + k], c.E,
                            T->tr.t.hmax * std::ldexp(1.0, k - (kMfClasses - 1)));
        J.tmeta = T->tr.meta.p;
        J.tbox = T->tr.t.box.p;
        J.tstart = T->tr.t.start.p;
        J.hbox = T->tr.t.hbox.p;
        J.hfield = T->tr.hfield.p;
        J.hweight = T->tr.hweight.p;
        J.hfw = T->tr.hfw.p;
        J.nfw = T->tr.nfw;
        J.P = T->tr.t.P;
        J.ntt = T->tr.t.ntiles;
        J.symmetric = (T == Q && T->symmetric_ok && jobs[i].nq_fields == jobs[i].nt_fields) ? 1 : 0;
        J.relax = 1;
        J.nq_fields = jobs[i].nq_fields;
        J.nslots = jobs[i].nq_fields * jobs[i].nt_fields;
        J.hist_base = c.rows;
        c.rows += J.symmetric ? 2LL * J.nslots : J.nslots;  // symmetric jobs: direct and mirror planes
        c.qt_prefix[i + 1] = c.qt_prefix[i] + J.nqt;
        c.maxP = std::max(c.maxP, J.P);
    }
    c.total_qt = c.qt_prefix[nj];
    KD_REQUIRE(c.total_qt < 0x7fffffffLL, "pair count: too many query tiles for one call");
}

// Cumulative counts per job and original radius from the histogram rows hh.
static void cumulative_counts(const CallSetup &c, const long long *hh, const std::vector<long long *> &outs)
{
    const int ne = c.ne, rowlen = c.rowlen, nr = (int)c.radii.size();
    std::vector<int> emap(nr);
    for (int i = 0; i < nr; i++)
        emap[i] = (int)(std::lower_bound(c.edges.begin(), c.edges.end(), c.radii[i]) - c.edges.begin());
    std::vector<long long> cum;
    for (size_t jb = 0; jb < c.hj.size(); jb++) {
        const JobDesc &J = c.hj[jb];
        long long *out = outs[jb];
        cum.assign((size_t)J.nslots * ne, 0);
        for (int sl = 0; sl < J.nslots; sl++) {
            // direct plane, plus for a symmetric job the mirror plane at this
            // slot and at the transposed one
            const int a = sl / J.nq_fields, b = sl % J.nq_fields;
            const long long *rows3[3] = {hh + (J.hist_base + sl) * rowlen,
                                         J.symmetric ? hh + (J.hist_base + J.nslots + sl) * rowlen : nullptr,
                                         J.symmetric ? hh + (J.hist_base + J.nslots + (long long)b * J.nq_fields + a) * rowlen
                                                     : nullptr};
            long long bulk = 0;
            for (int e = 0; e < ne; e++) {
                long long d = 0;
                for (int k = 0; k < 3; k++)
                    if (rows3[k]) {
                        bulk += rows3[k][ne + e];
                        d += rows3[k][e];
                    }
                cum[(size_t)sl * ne + e] = d + bulk;
            }
        }
        for (int i = 0; i < nr; i++)
            for (int sl = 0; sl < J.nslots; sl++) out[(size_t)i * J.nslots + sl] = cum[(size_t)sl * ne + emap[i]];
    }
}

__global__ void k_plan_check(const long long *__restrict__ offs, const long long *__restrict__ one, long long total,
                             long long nitems, long long nsplit, unsigned *__restrict__ err)
{
    if (threadIdx.x == 0 && (offs[total] != nitems || one[total] != nsplit)) atomicOr(err, kErrPlan);
}

// The device side of one call.  prepare() allocates everything but the item
// list, size_items() the item list; the enqueue_* steps allocate nothing.
struct PairCall {
    const CallSetup &c;
    kd_tuning_t tune;
    int ncu = 0, enum_blocks = 0, levels = 0, cap = 0;
    size_t enum_lds = 0;
    DevBuf<JobDesc> djobs;
    DevBuf<long long> dprefix;
    DevBuf<MfClass> dmfc;
    // enumeration: per query tile counts, split lists, offsets, staging pool
    DevBuf<int> cnt;
    DevBuf<long long> cnt64, sev, o1, o2, offs, part;
    DevBuf<unsigned char> mixed;
    DevBuf<Item> pool;
    long long pool_cap = 0;
    DevBuf<long long> pos, pos2;
    DevBuf<int> n1s;
    DevBuf<Item> items;
    long long nitems = 0, nsplit = 0;
    // pair kernels
    DevBuf<long long> hist;
    DevBuf<unsigned long long> dnext, dstats, pnext;
    DevBuf<unsigned> derr;
    DevBuf<long long> dtot;  // nitems, nsplit as enumerated

    PairCall(const CallSetup &setup, const kd_tuning_t *t) : c(setup)
    {
        memset(&tune, 0, sizeof(tune));
        if (t) tune = *t;
        KD_REQUIRE(tune.blocks_per_cu >= 0 && tune.blocks_per_cu <= 64 && tune.chunk >= 0 && tune.chunk <= 4096 &&
                       tune.enum_blocks_per_cu >= 0 && tune.enum_blocks_per_cu <= 64,
                   "pair count: tuning values out of range (blocks per CU 0..64, chunk 0..4096)");
    }

    void prepare(hipStream_t s)
    {
        const int nj = (int)c.hj.size();
        int dev = 0;
        KD_HIP(hipGetDevice(&dev));
        KD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        djobs.alloc(nj);
        dprefix.alloc(nj + 1);
        dmfc.alloc(c.hmfc.size());
        std::vector<JobDesc> hj(c.hj);
        for (int i = 0; i < nj; i++) hj[i].mfc = dmfc.p + (size_t)i * kMfClasses;
        KD_HIP(hipMemcpyAsync(dmfc.p, c.hmfc.data(), sizeof(MfClass) * c.hmfc.size(), hipMemcpyHostToDevice, s));
        KD_HIP(hipMemcpyAsync(djobs.p, hj.data(), sizeof(JobDesc) * nj, hipMemcpyHostToDevice, s));
        KD_HIP(hipMemcpyAsync(dprefix.p, c.qt_prefix.data(), sizeof(long long) * (nj + 1), hipMemcpyHostToDevice,
                              s));
        const long long T = c.total_qt;
        levels = 31 - __builtin_clz((unsigned)std::max(c.maxP, 1));
        cap = 64 * (levels + 3);
        enum_lds = (size_t)4 * cap * sizeof(int);
        const int epc = tune.enum_blocks_per_cu ? tune.enum_blocks_per_cu : 32;
        enum_blocks = (int)std::max<long long>(1, std::min<long long>((T + 3) / 4, (long long)ncu * epc));
        cnt.alloc(std::max<long long>(T, 1));
        cnt64.alloc(T + 1);
        sev.alloc(T + 1);
        o1.alloc(T + 1);
        o2.alloc(T + 1);
        offs.alloc(T + 1);
        part.alloc(scan_parts(T + 1));
        mixed.alloc(std::max<long long>(T, 1));
        hist.alloc(std::max<long long>(c.rows * c.rowlen, 1));
        dnext.alloc(3 * kXcds * kNextStride);
        dstats.alloc(5);
        pnext.alloc(1);
        derr.alloc(1);
        dtot.alloc(2);
        // the staging pool: room for 512 items per query tile and for two
        // chunks per wave, at least 2^20 and at most 4 GiB of items (without
        // it, or past it, tiles are walked twice); not in a call that
        // allocates nothing
        const long long stage_set = g_stage_items.load();
        if (stage_set >= 0 && !arena_active()) {
            const long long want =
                stage_set > 0 ? stage_set
                              : std::min((4LL << 30) / (long long)sizeof(Item),
                                         std::max({1LL << 20, 512LL * T, 2LL * kStageChunk * enum_blocks * 4}));
            void *pp = nullptr;
            if (hipMallocAsync(&pp, (size_t)want * sizeof(Item), s) == hipSuccess && pp) {
                pool.p = (Item *)pp;
                pool.n = (size_t)want;
                pool_cap = want;
                pos.alloc(std::max<long long>(T, 1));
                pos2.alloc(std::max<long long>(T, 1));
                n1s.alloc(std::max<long long>(T, 1));
            } else {
                (void)hipGetLastError();
            }
        }
    }

    // clears, counting pass, item offsets (query tiles of one field first)
    void enqueue_count(hipStream_t s)
    {
        const long long T = c.total_qt;
        const int nj = (int)c.hj.size();
        zero_async(derr.p, derr.bytes(), s);
        zero_async(dstats.p, dstats.bytes(), s);
        zero_async(dnext.p, dnext.bytes(), s);
        zero_async(pnext.p, pnext.bytes(), s);
        hipLaunchKernelGGL(k_enum, dim3(enum_blocks), dim3(256), enum_lds, s, djobs.p, dprefix.p, nj, T, c.E,
                           pool_cap ? 2 : 0, c.merge, 1, cap, cnt.p, (const long long *)nullptr, pool.p, derr.p,
                           pool_cap, pnext.p, pos.p, pos2.p, n1s.p, (const Item *)nullptr, 0LL);
        hipLaunchKernelGGL(k_split_counts, dim3(nblk(T + 1)), dim3(256), 0, s, djobs.p, dprefix.p, nj, T, cnt.p,
                           cnt64.p, sev.p, mixed.p);
        excl_scan(cnt64.p, o1.p, T + 1, part.p, s);
        excl_scan(sev.p, o2.p, T + 1, part.p, s);
        hipLaunchKernelGGL(k_split_offs, dim3(nblk(T + 1)), dim3(256), 0, s, o1.p, o2.p, mixed.p, T, offs.p);
        KD_HIP(hipGetLastError());
    }

    // after enqueue_count in an eager call: the list's length, then the list
    void size_items(hipStream_t s)
    {
        long long h[2] = {0, 0};
        unsigned herr = 0;
        KD_HIP(hipMemcpyAsync(&h[0], offs.p + c.total_qt, sizeof(long long), hipMemcpyDeviceToHost, s));
        KD_HIP(hipMemcpyAsync(&h[1], o1.p + c.total_qt, sizeof(long long), hipMemcpyDeviceToHost, s));
        KD_HIP(hipMemcpyAsync(&herr, derr.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        KD_HIP(hipStreamSynchronize(s));
        KD_REQUIRE(!(herr & kErrStack), "pair count: tile-pair enumeration stack overflow");
        KD_REQUIRE(h[0] < 0x7fffffffLL, "pair count: too many tile pairs for one call");
        nitems = h[0];
        nsplit = h[1];
        items.alloc(std::max<long long>(nitems, 1));
    }

    // the planned totals against this run's (graph replays)
    void enqueue_check(hipStream_t s)
    {
        hipLaunchKernelGGL(k_plan_check, dim3(1), dim3(64), 0, s, (const long long *)offs.p,
                           (const long long *)o1.p, c.total_qt, nitems, nsplit, derr.p);
    }

    void enqueue_write(hipStream_t s)
    {
        const int nj = (int)c.hj.size();
        hipLaunchKernelGGL(k_enum, dim3(enum_blocks), dim3(256), enum_lds, s, djobs.p, dprefix.p, nj, c.total_qt,
                           c.E, 1, c.merge, 1, cap, cnt.p, (const long long *)offs.p, items.p, derr.p, pool_cap,
                           pnext.p, pool_cap ? pos.p : (long long *)nullptr, pos2.p, n1s.p, (const Item *)pool.p,
                           nitems);
        // the trim pass serves the matrix-core loop of the FP32 form (and the
        // per-lane loop of query tiles with several fields); the FP64 loop's
        // quad windows skip every quad it would cut
        if (nitems > 0 && c.fp32)
            hipLaunchKernelGGL(k_trim_items, dim3(std::min(nblk(nitems), 4096)), dim3(256), 0, s, djobs.p,
                               items.p, nitems, c.E.rsq[c.ne - 1], c.ne, dstats.p, derr.p);
        KD_HIP(hipGetLastError());
    }

    // query tiles of one field: in the FP32 form k_pairs32 takes the items
    // of its matrix-core loop and k_pairs64 the rest of the same list, in the
    // FP64 form k_pairs64 takes all; query tiles with several fields:
    // k_pairs64_mixed in both forms
    void enqueue_pairs(hipStream_t s)
    {
        zero_async(hist.p, hist.bytes(), s);
        const size_t per_wave = wave_lds_bytes(c.rowlen);
        const int wpb = 4;
        const size_t lds = per_wave * wpb, lds32 = (per_wave + kMfLds) * wpb;
        const int nblocks = std::max(1, ncu * (tune.blocks_per_cu ? tune.blocks_per_cu : 8));
        const int nwaves = nblocks * wpb;
        // items per take: 4 unless tuned (on the 1M/10M workload 4 items
        // 156.4 ms, 16 157.1, 64 162.2: small takes balance the end of the
        // pass; the batched workload's tuner also picks 4), fewer for a list
        // too short to give every wave 64 items
        auto chunk_of = [&](long long n) {
            if (tune.chunk) return tune.chunk;
            return (int)std::max<long long>(1, std::min<long long>(4, n / ((long long)nwaves * 64)));
        };
        const long long nmixed = nitems - nsplit;
        if (nsplit > 0 && c.fp32)
            hipLaunchKernelGGL(k_pairs32, dim3(nblocks), dim3(64 * wpb), lds32, s, djobs.p, items.p, nsplit,
                               dnext.p + 2 * kXcds * kNextStride, chunk_of(nsplit), c.E, hist.p, dstats.p, derr.p);
        if (nsplit > 0)
            hipLaunchKernelGGL(k_pairs64, dim3(nblocks), dim3(64 * wpb), lds, s, djobs.p, items.p, nsplit, dnext.p,
                               chunk_of(nsplit), c.E, hist.p, dstats.p, derr.p, c.fp32 ? 1 : 0);
        if (nmixed > 0) {
            const int nb2 = (int)std::max<long long>(1, std::min<long long>(nblocks, (nmixed + 63) / 64));
            hipLaunchKernelGGL(k_pairs64_mixed, dim3(nb2), dim3(64 * wpb), lds, s, djobs.p, items.p + nsplit,
                               nmixed, dnext.p + kXcds * kNextStride, chunk_of(nmixed), c.E, hist.p, dstats.p,
                               derr.p);
        }
        KD_HIP(hipGetLastError());
    }

    void enqueue_copy_out(hipStream_t s, long long *h_hist, unsigned long long *h_stats, unsigned *h_err)
    {
        KD_HIP(hipMemcpyAsync(h_hist, hist.p, sizeof(long long) * c.rows * c.rowlen, hipMemcpyDeviceToHost, s));
        KD_HIP(hipMemcpyAsync(h_stats, dstats.p, sizeof(unsigned long long) * 5, hipMemcpyDeviceToHost, s));
        KD_HIP(hipMemcpyAsync(h_err, derr.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    }
};

static void check_error_word(unsigned herr)
{
    KD_REQUIRE(!(herr & kErrItem), "pair count: internal error: a work item lies outside its job's tiles");
    KD_REQUIRE(!(herr & kErrStack), "pair count: tile-pair enumeration stack overflow");
    KD_REQUIRE(!(herr & kErrPlan), "pair count plan: the enumeration gave other work items than when the plan was "
                                   "made (were its trees changed or freed?)");
    KD_REQUIRE(!herr, "pair count: a field id gives a count-matrix index outside [0, Nq*Nt) "
                      "(index = query_field * num_fields + tree_field)");
}

static void fill_stats(kd_stats_t *stats, const unsigned long long *hs, long long nitems, float ms_enum,
                       float ms_pairs, std::chrono::steady_clock::time_point wall0)
{
    if (!stats) return;
    stats->evaluated_pairs = (long long)hs[0];
    stats->rechecked_pairs = (long long)hs[1];
    stats->reach_checks = (long long)hs[3];
    stats->trimmed_items = (long long)hs[2];
    stats->fp32_pairs = (long long)hs[4];
    stats->work_items = nitems;
    stats->enumerate_ms = ms_enum;
    stats->count_ms = ms_pairs;
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
}

void count_pairs(const std::vector<PairJob> &jobs, const std::vector<double> &radii, int precision,
                 std::vector<long long *> &counts_out, kd_stats_t *stats, Sorter &, hipStream_t s)
{
    auto wall0 = std::chrono::steady_clock::now();
    CallSetup c;
    setup_call(jobs, radii, precision, c);
    PairCall pc(c, nullptr);
    pc.prepare(s);
    // GPU time of the enumeration and of the pair kernels (HIP events on this
    // call's stream); a call that allocates nothing creates no events
    const bool timed = !arena_active();
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    if (timed)
        for (auto &e : ev) KD_HIP(hipEventCreate(&e));
    auto mark = [&](int k) {
        if (timed) KD_HIP(hipEventRecord(ev[k], s));
    };
    mark(0);
    pc.enqueue_count(s);
    pc.size_items(s);
    pc.enqueue_write(s);
    mark(1);
    mark(2);
    pc.enqueue_pairs(s);
    mark(3);
    std::vector<long long> hh(c.rows * c.rowlen);
    unsigned long long hstats[5];
    unsigned herr = 0;
    pc.enqueue_copy_out(s, hh.data(), hstats, &herr);
    KD_HIP(hipStreamSynchronize(s));
    float ms_enum = 0, ms_pairs = 0;
    if (timed) {
        KD_HIP(hipEventElapsedTime(&ms_enum, ev[0], ev[1]));
        KD_HIP(hipEventElapsedTime(&ms_pairs, ev[2], ev[3]));
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    check_error_word(herr);
    cumulative_counts(c, hh.data(), counts_out);
    fill_stats(stats, hstats, pc.nitems, ms_enum, ms_pairs, wall0);
}

// ---- plans: one call recorded into a graph ----------------------------------

struct PairPlan {
    CallSetup c;
    std::unique_ptr<PairCall> pc;
    std::vector<long long *> outs;
    hipStream_t s = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    long long *h_hist = nullptr;  // pinned host memory, the graph's copy targets
    unsigned long long *h_stats = nullptr;
    unsigned *h_err = nullptr;
    int runs = 0;

    ~PairPlan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (s) (void)hipStreamSynchronize(s);
        if (h_hist) (void)hipHostFree(h_hist);
        if (h_stats) (void)hipHostFree(h_stats);
        if (h_err) (void)hipHostFree(h_err);
        if (s) (void)hipStreamDestroy(s);
    }
};

PairPlan *plan_pairs(const std::vector<PairJob> &jobs, const std::vector<double> &radii, int precision,
                     const std::vector<long long *> &counts_out, const kd_tuning_t *tuning, hipStream_t ls)
{
    std::unique_ptr<PairPlan> P(new PairPlan());
    setup_call(jobs, radii, precision, P->c);
    P->outs = counts_out;
    P->pc.reset(new PairCall(P->c, tuning));
    PairCall &pc = *P->pc;
    // buffers come from the calling context's stream; the sizing run too
    pc.prepare(ls);
    pc.enqueue_count(ls);
    pc.size_items(ls);
    KD_HIP(hipHostMalloc((void **)&P->h_hist, sizeof(long long) * std::max<long long>(P->c.rows * P->c.rowlen, 1),
                         hipHostMallocDefault));
    KD_HIP(hipHostMalloc((void **)&P->h_stats, sizeof(unsigned long long) * 5, hipHostMallocDefault));
    KD_HIP(hipHostMalloc((void **)&P->h_err, sizeof(unsigned), hipHostMallocDefault));
    KD_HIP(hipStreamSynchronize(ls));
    KD_HIP(hipStreamCreateWithFlags(&P->s, hipStreamNonBlocking));
    for (auto &e : P->ev) KD_HIP(hipEventCreate(&e));
    KD_HIP(hipStreamBeginCapture(P->s, hipStreamCaptureModeThreadLocal));
    try {
        pc.enqueue_count(P->s);
        pc.enqueue_check(P->s);
        pc.enqueue_write(P->s);
        pc.enqueue_pairs(P->s);
        pc.enqueue_copy_out(P->s, P->h_hist, P->h_stats, P->h_err);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(P->s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    KD_HIP(hipStreamEndCapture(P->s, &P->graph));
    KD_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    return P.release();
}

void plan_run(PairPlan *P, kd_stats_t *stats)
{
    auto wall0 = std::chrono::steady_clock::now();
    KD_HIP(hipEventRecord(P->ev[0], P->s));
    KD_HIP(hipGraphLaunch(P->exec, P->s));
    KD_HIP(hipEventRecord(P->ev[1], P->s));
    KD_HIP(hipStreamSynchronize(P->s));
    float ms = 0;
    KD_HIP(hipEventElapsedTime(&ms, P->ev[0], P->ev[1]));
    P->runs++;
    check_error_word(*P->h_err);
    cumulative_counts(P->c, P->h_hist, P->outs);
    fill_stats(stats, P->h_stats, P->pc->nitems, 0.0f, ms, wall0);
}

long long plan_items(const PairPlan *P) { return P->pc->nitems; }

void plan_free(PairPlan *P) { delete P; }


// ============================================================================
// Weighted pair counts (bootstrap weight sets).
//
// For S weight sets given as per-point multiplicities (uint8) on either side,
//     C_e^s[qf * Nq + tf] = sum_q sum_p wq_s(q) * wp_s(p) * w(p) * [n < r_e^2],
// w(p) the node's own duplicate weight -- the reference's count for the
// catalogues np.repeat(points, w_s) (repeated tree points merge into the node
// weight, repeated query points count one by one).  Tree-side weights need a
// tree without duplicate points.
//
// Tree side unweighted (k_wq): one wave takes all items of one query tile,
// histograms each lane's first counting radius in LDS (per lane, per tree
// field, per bin), then weights the per-lane cumulative counts by the query
// multiplicities of every set, lanes = sets, in one pass per query tile.
// Tree side weighted (k_wm): for one query tile and one tree item the
// counts of all sets under radius e are a contraction
//     G_e[q][s] = sum_p [n(q,p) < r_e^2] * wt_s(p)      (64 x P) x (P x S)
//     C_e^s    += sum_q wq_s(q) * G_e[q][s]
// whose first step runs on the matrix cores: the 0/1 indicator matrix and
// the tree multiplicities as int8 with int32 sums while every multiplicity
// is below 128, else as bf16 with FP32 sums (exact: at most 64 * 255 < 2^24
// per pass), then the query weighting in 32-bit integers on the vector ALUs.
// ============================================================================

static constexpr int kWMaxEdges = 32;  // radii per weighted call

// ---- the weighted kernels' FP32 form ------------------------------------------
//
// The weighted kernels count per lane (one query point per lane), and for
// that the 16-block v_mfma_f32_4x4x1f32 has the right layout: with A =
// component k of point l % 4 of a tree quad (one 16-byte load of quad32 per
// lane) and B = -2 a_k of query point l, register r of lane l accumulates the
// value of (tree point r, query point l).  Four MFMAs, one per term of
//     n = (|a|^2 + eps) + sum_k b_k (-2 a_k) + |b|^2 * 1
// (a = q - c, b = p - c relative to the tree tile's centre c; |a|^2 + eps
// starts the accumulator), give the lane's own query against the quad's four
// points, which the FP64 loops compute with six FP64 instructions per point.
// (Operand and result layout checked with an in-bounds probe on the MI355X;
// the 4x4x1 form runs at the full FP32 matrix rate, 256 multiply-adds per 8
// cycles.)  Per radius, a subtract-with-borrow of the threshold bits puts
// "below" into VCC and an add-with-carry moves it into the lane's count (or,
// for a slot mask, shifts it in); the difference goes into the band minimum.
// Pairs in the guard band are recomputed with the reference's FP64 formula,
// per lane.
//
// Thresholds per job and tree tile size class (fp32_thresholds4): for a pair
// near radius r, |a| <= r + h and |b| <= h (h the class's bound on a tile
// point's distance from the tile centre), and D is the accumulator start
// C = fl(|a'|^2 + eps) plus the four terms, added in four MFMA steps with at
// most two roundings each, so
//   |D - eps - n64| <= 8 u32 ((r + 2h)^2 + eps)          the MFMA steps
//                      + u32 (|a'|^2 + eps) + u32 |b'|^2  rounding of the norms
//                      + 2.01 u32 r (r + 2h)              a, b rounded to FP32
//                      + 7 u64 r^2                        the reference's own rounding
//                   <= G = u32 (11 (r + 2h)^2 + 9 eps + 2.01 r (r + 2h)) + 7 u64 r^2
// (the 11 covers second-order terms); eps = 48 u32 h^2 (at least 2^-100, a
// normal float) keeps every D positive, so D's bits order like D:
// bits(D) < lo32[e] = fl32_down(r_e^2 - G + eps) means n64 < r_e^2,
// bits(D) >= fl32_up(r_e^2 + G + eps) means n64 >= r_e^2, the rest is
// rechecked.  The band decides how many pairs are rechecked, never a count:
// radii with G <= kBandMax4 r^2 take the FP32 loops (G / r^2 falls as r
// grows), radii with r^2 > 1e36 (whose FP32 values could overflow) and the
// rest FP64.  A pass uses the widest band of its radii.
struct MfClass4 {
    unsigned lo32[kWMaxEdges];
    unsigned band_w[kWMaxEdges];  // in units of the last place of lo32
    float eps32;
    int e32, e32hi;  // radii e32 .. e32hi-1 take the FP32 loops
    int pad;
};
static constexpr double kBandMax4 = 1e-3;

static void fp32_thresholds4(MfClass4 &M, const EdgeTable &E, double h)
{
    const double u32 = std::ldexp(1.0, -24), u64 = std::ldexp(1.0, -53);
    const double eps = std::max(48.0 * u32 * h * h, std::ldexp(1.0, -100));
    float e32f = (float)eps;
    if ((double)e32f < eps) e32f = std::nextafter(e32f, INFINITY);
    memset(&M, 0, sizeof(M));
    M.eps32 = e32f;
    int e = E.ne - 1;
    while (e >= 0 && !(E.rsq[e] <= 1e36)) e--;
    M.e32hi = e + 1;
    M.e32 = e + 1;
    for (; e >= 0; e--) {
        const double r = E.r[e], rsq = E.rsq[e];
        const double G = u32 * (11.0 * (r + 2 * h) * (r + 2 * h) + 9.0 * (double)e32f + 2.01 * r * (r + 2 * h)) +
                         7.0 * u64 * rsq;
        if (!(rsq > 0.0 && G <= kBandMax4 * rsq)) break;
        const double xl = rsq - G + (double)e32f, xh = rsq + G + (double)e32f;
        float lo = (float)xl, hi = (float)xh;
        if ((double)lo > xl) lo = std::nextafter(lo, 0.0f);
        if ((double)hi < xh) hi = std::nextafter(hi, INFINITY);
        unsigned lb, hb;
        memcpy(&lb, &lo, 4);
        memcpy(&hb, &hi, 4);
        M.lo32[e] = lb;
        M.band_w[e] = hb - lb + 1;
        M.e32 = e;
    }
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// The query lanes against tree tile t: B operands -2a (a = q - c in FP32,
// rounded from the FP64 difference) and the accumulator start |a|^2 + eps
// (the FP32 components' squares summed in FP64, rounded once); invalid lanes
// (and lanes parked at +inf) start at +inf, so every D of theirs is +inf.
struct MfQuery {
    float bx, by, bz, c;
};
__device__ __forceinline__ MfQuery mf4_query(const JobDesc &J, int t, float eps, double qx, double qy, double qz,
                                             bool valid)
{
    const double c0 = sload(J.tcentre, 3 * t), c1 = sload(J.tcentre, 3 * t + 1), c2 = sload(J.tcentre, 3 * t + 2);
    const float ax = (float)(qx - c0), ay = (float)(qy - c1), az = (float)(qz - c2);
    MfQuery m;
    m.bx = -2.f * ax;
    m.by = -2.f * ay;
    m.bz = -2.f * az;
    const double an = (double)ax * ax + (double)ay * ay + (double)az * az + (double)eps;
    m.c = valid && an < 1e38 ? (float)an : INFINITY;
    return m;
}

// D[r] = the FP32 value of (tree point r of the quad whose operands are a,
// the lane's query)
__device__ __forceinline__ floatx4 mf4_dist(const floatx4 &a, const MfQuery &Q)
{
    floatx4 D = {Q.c, Q.c, Q.c, Q.c};
    D = __builtin_amdgcn_mfma_f32_4x4x1f32(a[0], Q.bx, D, 0, 0, 0);
    D = __builtin_amdgcn_mfma_f32_4x4x1f32(a[1], Q.by, D, 0, 0, 0);
    D = __builtin_amdgcn_mfma_f32_4x4x1f32(a[2], Q.bz, D, 0, 0, 0);
    D = __builtin_amdgcn_mfma_f32_4x4x1f32(a[3], 1.0f, D, 0, 0, 0);
    // The compares that read D are inline assembly, before which the compiler
    // inserts no wait states for the MFMA's write (a VALU read issued too
    // soon sees the register's old value): 10 wait states here, more than the
    // two-pass 4x4x1 form needs.  D as an operand keeps them between the MFMA
    // and its readers.
    asm volatile("s_nop 7\n\ts_nop 1" : "+v"(D));
    return D;
}

// c += [bits(d) < lo] as a subtract-with-borrow and an add-with-carry (the
// difference into the band minimum)
__device__ __forceinline__ void lane_below(unsigned &c, float d, unsigned lo, unsigned &bmin)
{
    unsigned x;
    asm("v_subrev_co_u32_e32 %0, vcc, %3, %2\n\tv_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "=&v"(x), "+v"(c)
        : "v"(__float_as_uint(d)), "s"(lo)
        : "vcc");
    bmin = x < bmin ? x : bmin;
}

// Guard-band pairs of one quad (index g) for radii a .. K-1 recomputed per
// lane with the reference's FP64 formula and added to the lane's counts;
// seen counts the band pairs.
template <int K>
__device__ __forceinline__ void mf4_lane_recheck(const JobDesc &J, int g, int a, const floatx4 &D,
                                                 const unsigned *lo32, unsigned W, const double *thr, double qx,
                                                 double qy, double qz, unsigned *cnt, unsigned &seen)
{
    const double *pq = J.tquad + 12LL * g;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const unsigned bits = __float_as_uint(D[r]);
        bool any = false;
#pragma unroll
        for (int e = 0; e < K; e++) any = any || (e >= a && bits - lo32[e] < W);
        if (!__ballot(any)) continue;
        const double dx = sload(pq, r) - qx, dy = sload(pq, 4 + r) - qy, dz = sload(pq, 8 + r) - qz;
        const double n = fma(dz, dz, fma(dy, dy, dx * dx));
#pragma unroll
        for (int e = 0; e < K; e++) {
            const bool band = e >= a && bits - lo32[e] < W;
            seen += band ? 1u : 0u;
            cnt[e] += (band && n < thr[e]) ? 1u : 0u;
        }
    }
}

// Per-lane counts of the quads in mask m (nonzero; bit j: quad gb + j) for
// radii A .. K-1 (the quads' windows start at A), the query lanes re-centred
// as Q; the next quad's operands load while one is counted.
template <int K, int A>
__device__ __forceinline__ void mf4_lane_quads_from(const JobDesc &J, int gb, unsigned long long m,
                                                    const unsigned *lo32, unsigned W, const double *thr,
                                                    const MfQuery &Q, double qx, double qy, double qz, unsigned *cnt,
                                                    unsigned &seen, int lane)
{
    typedef const floatx4 __attribute__((address_space(1))) *g4;
    const g4 ops = (g4)J.tquad32 + 4LL * gb + (lane & 3);  // point lane % 4 of quad gb + j: ops[4 j]
    int j = __ffsll((long long)m) - 1;
    m &= m - 1;
    floatx4 nxt = ops[4 * j];
    while (true) {
        const floatx4 a = nxt;
        const int jc = j;
        const bool more = m != 0;
        if (more) {
            j = __ffsll((long long)m) - 1;
            m &= m - 1;
            nxt = ops[4 * j];
        }
        const floatx4 D = mf4_dist(a, Q);
        unsigned bmin = 0xFFFFFFFFu;
#pragma unroll
        for (int e = A; e < K; e++)
#pragma unroll
            for (int r = 0; r < 4; r++) lane_below(cnt[e], D[r], lo32[e], bmin);
        if (__ballot(bmin < W)) mf4_lane_recheck<K>(J, gb + jc, A, D, lo32, W, thr, qx, qy, qz, cnt, seen);
        if (!more) break;
    }
}

template <int K, int A>
__device__ __forceinline__ void mf4_lane_by_entry(const JobDesc &J, int gb, unsigned long long live, int lo,
                                                  const unsigned *lo32, unsigned W, const double *thr,
                                                  const MfQuery &Q, double qx, double qy, double qz, unsigned *cnt,
                                                  unsigned &seen, int lane)
{
    if constexpr (A < K) {
        const unsigned long long m = live & __ballot(lo == A);
        if (m) mf4_lane_quads_from<K, A>(J, gb, m, lo32, W, thr, Q, qx, qy, qz, cnt, seen, lane);
        mf4_lane_by_entry<K, A + 1>(J, gb, live, lo, lo32, W, thr, Q, qx, qy, qz, cnt, seen, lane);
    }
}

// Bits a .. b-1 of a 64-bit mask (0 <= a <= b <= 64).
__device__ __forceinline__ unsigned long long span_mask(int a, int b)
{
    return b - a >= 64 ? ~0ull : (((1ull << (b - a)) - 1ull) << a);
}

// lane_loop64w with the distances from the matrix cores: per-lane counts
// cnt[e] += points of quads g0..g1-1 (a run whose first quad lies in tree
// tile t) under radius e0 + e, for radii in the FP32 range of the run's size
// class M.  The window tests and bulk counts are lane_loop64w's; the live
// quads of each tile are counted after re-centring the query lanes on it.
// Returns the points of the quads evaluated for the first time in this pass.
template <int K>
__device__ __forceinline__ unsigned mf4_lane_loop(const JobDesc &J, const MfClass4 &M, int t, int g0, int g1, int e0,
                                                  const double *thr, double thr_prev, const Box &qb, double qx,
                                                  double qy, double qz, bool valid, unsigned *cnt, unsigned &seen,
                                                  int lane)
{
    unsigned lo32[K];
    unsigned W = 1;
#pragma unroll
    for (int e = 0; e < K; e++) {
        lo32[e] = __builtin_amdgcn_readfirstlane(sload(M.lo32, e0 + e));
        W = max(W, (unsigned)__builtin_amdgcn_readfirstlane(sload(M.band_w, e0 + e)));
    }
    unsigned first_pts = 0;
    // quad sg lies in tile tt: [tquad_start[tt], tq_end); g1 <= tquad_start[t
    // + tiles of the run], so tt stays inside the run
    int tt = t;
    int tq_end = __builtin_amdgcn_readfirstlane(J.tquad_start[tt + 1]);
    int centred = -1;
    MfQuery Q = {0.f, 0.f, 0.f, INFINITY};
    for (int gb = g0; gb < g1; gb += 64) {
        int lo, hi;
        unsigned np;
        bool first;
        quad_window<K>(J.tquad_box, J.tquad_meta, gb + lane, g1, thr, thr_prev, qb, lo, hi, np, first);
        const unsigned long long live = __ballot(lo < hi);
        const LaneBits3 nb = lane_bits3(np);
        const unsigned long long dead = ~live;
        if (dead & __ballot(hi < K)) {
#pragma unroll
            for (int e = 0; e < K; e++) cnt[e] += masked_sum3(dead & __ballot(hi <= e), nb);
        }
        first_pts += masked_sum3(live & __ballot(first), nb);
        const int bend = min(gb + 64, g1);
        for (int sg = gb; sg < bend;) {
            while (tq_end <= sg) {
                tt++;
                tq_end = __builtin_amdgcn_readfirstlane(J.tquad_start[tt + 1]);
            }
            const int se = min(bend, tq_end);
            const unsigned long long lv = live & span_mask(sg - gb, se - gb);
            if (lv) {
                if (tt != centred) {
                    Q = mf4_query(J, tt, M.eps32, qx, qy, qz, valid);
                    centred = tt;
                }
                mf4_lane_by_entry<K, 0>(J, gb, lv, lo, lo32, W, thr, Q, qx, qy, qz, cnt, seen, lane);
            }
            sg = se;
        }
    }
    return first_pts;
}

struct WJobDesc {
    JobDesc b;
    const unsigned char *wq;  // [nq][spad] point-major in tile order (sets zero padded), or null (all 1)
    int nq;
    int spad;                 // row length of wq: the sets rounded up to 32
    long long whist_base;  // first int64 of this job's [S][nslots][2ne+1] block
    const MfClass4 *mfc4;  // FP32 form: kMfClasses classes of the tree's tiles; null: FP64
};

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}


// per-lane cumulative counts of one tree-field slot, weighted by every set
// The per-lane histograms of k_wq: 16-bit counts, lanes l and l + 32 in one
// dword, so that each half-wave touches 32 distinct banks.
__device__ __forceinline__ int hix(int row, int lane) { return row * 64 + (((lane & 31) << 1) | (lane >> 5)); }

__device__ void wq_flush_slot(const WJobDesc &W, unsigned short *H, int slot, int tf, int ne, int S, bool valid,
                              int qf, int b0, long long *__restrict__ hist_w, unsigned *__restrict__ err, int lane)
{
    const JobDesc &J = W.b;
    const int nb = ne + 1, rowlen = 2 * ne + 1;
    // bins to cumulative counts (each lane's total stays below 2^16)
    unsigned run = 0;
    for (int b = 0; b < ne; b++) {
        run += H[hix(slot * nb + b, lane)];
        H[hix(slot * nb + b, lane)] = (unsigned short)run;
    }
    __builtin_amdgcn_wave_barrier();
    unsigned long long remaining = __ballot(valid);
    while (remaining) {
        const int leader = __ffsll((long long)remaining) - 1;
        const int gq = __builtin_amdgcn_readlane(qf, leader);
        const unsigned long long gm = __ballot(valid && qf == gq);
        remaining &= ~gm;
        const long long sl = (long long)gq * J.nq_fields + tf;
        if (gq < 0 || sl < 0 || sl >= J.nslots) {
            if (lane == 0) atomicOr(err, 1u);
            continue;
        }
        // lanes = weight sets; eight radii at a time keep the accumulators
        // in 16 VGPRs
        for (int s0 = 0; s0 < S; s0 += 64) {
            const int sset = s0 + lane;
            const bool sv = sset < S;
            for (int e0 = 0; e0 < ne; e0 += 8) {
                unsigned long long acc[8];
#pragma unroll
                for (int e = 0; e < 8; e++) acc[e] = 0;
                unsigned long long bits = gm;
                while (bits) {
                    const int q = __ffsll((long long)bits) - 1;
                    bits &= bits - 1;
                    const unsigned wgt = sv ? (W.wq ? (unsigned)W.wq[(size_t)(b0 + q) * W.spad + sset] : 1u) : 0u;
#pragma unroll
                    for (int e = 0; e < 8; e++)
                        if (e0 + e < ne) acc[e] += (unsigned long long)wgt * H[hix(slot * nb + e0 + e, q)];
                }
                if (sv) {
                    long long *row = hist_w + W.whist_base + ((long long)sset * J.nslots + sl) * rowlen;
#pragma unroll
                    for (int e = 0; e < 8; e++)
                        if (e0 + e < ne && acc[e]) atomicAdd((unsigned long long *)&row[e0 + e], acc[e]);
                }
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void wq_zero(unsigned short *H, int rows, int lane)
{
    unsigned *h = (unsigned *)H;  // rows x 64 shorts = rows x 32 words
    for (int b = lane; b < rows * 32; b += 64) h[b] = 0u;
    __builtin_amdgcn_wave_barrier();
}

// Both tree-field slots (fields tf0, tf1; -1 when unused) to the histogram,
// then the histograms cleared.
__device__ __forceinline__ void wq_flush_all(const WJobDesc &W, unsigned short *H, int tf0, int tf1, int ne, int S,
                                             bool valid, int qf, int b0, long long *__restrict__ hist_w,
                                             unsigned *__restrict__ err, int lane)
{
    if (tf0 >= 0) wq_flush_slot(W, H, 0, tf0, ne, S, valid, qf, b0, hist_w, err, lane);
    if (tf1 >= 0) wq_flush_slot(W, H, 1, tf1, ne, S, valid, qf, b0, hist_w, err, lane);
    wq_zero(H, 2 * (ne + 1), lane);
}

// The slot for tree field tf: the one holding it, else a free one, else -1
// (both hold other fields).
__device__ __forceinline__ int wq_pick(int tf, int tf0, int tf1)
{
    return tf0 == tf ? 0 : (tf1 == tf ? 1 : (tf0 < 0 ? 0 : (tf1 < 0 ? 1 : -1)));
}

// A whole subtree under radius ehi with weight w in tree field tf: for each
// weight set, w times the set's multiplicities summed over each query field
// group, into the bulk entry of radius ehi (it adds to every radius >= ehi).
__device__ void wq_bulk(const WJobDesc &W, int tf, int ehi, long long w, int ne, int S, bool valid, int qf, int b0,
                        long long *__restrict__ hist_w, unsigned *__restrict__ err, int lane)
{
    const JobDesc &J = W.b;
    const int rowlen = 2 * ne + 1;
    unsigned long long remaining = __ballot(valid);
    while (remaining) {
        const int leader = __ffsll((long long)remaining) - 1;
        const int gq = __builtin_amdgcn_readlane(qf, leader);
        const unsigned long long gm = __ballot(valid && qf == gq);
        remaining &= ~gm;
        const long long sl = (long long)gq * J.nq_fields + tf;
        if (gq < 0 || sl < 0 || sl >= J.nslots) {
            if (lane == 0) atomicOr(err, 1u);
            continue;
        }
        for (int s0 = 0; s0 < S; s0 += 64) {
            const int sset = s0 + lane;
            const bool sv = sset < S;
            unsigned long long acc = 0;
            unsigned long long bits = gm;
            while (bits) {
                const int q = __ffsll((long long)bits) - 1;
                bits &= bits - 1;
                acc += sv ? (W.wq ? (unsigned)W.wq[(size_t)(b0 + q) * W.spad + sset] : 1u) : 0u;
            }
            if (sv && acc) {
                long long *row = hist_w + W.whist_base + ((long long)sset * J.nslots + sl) * rowlen;
                atomicAdd((unsigned long long *)&row[ne + ehi], acc * (unsigned long long)w);
            }
        }
    }
}

// Query-weighted jobs: one wave per query tile (a work queue over the query
// tiles of all jobs).  Each lane keeps a histogram of its own pair counts per
// radius bin for two tree fields (16-bit counts in LDS); a flush multiplies
// the histograms by the query multiplicities of every weight set.  `load`
// bounds every lane's total since the last flush, which is taken before the
// total could pass 2^16 - 1.  At least seven waves per SIMD (at most 72
// VGPRs, no spills); 5.4 KB of LDS per wave at 20 radii, so 28 waves per CU
// take 151 KB.
template <bool F32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_wq(
    const WJobDesc *__restrict__ jobs, const long long *__restrict__ qt_prefix, int njobs, long long total_qt,
    const Item *__restrict__ items, const long long *__restrict__ offs, EdgeTable E, int S,
    unsigned long long *__restrict__ next, long long *__restrict__ hist_w, unsigned long long *__restrict__ stats,
    unsigned *__restrict__ err)
{
    extern __shared__ unsigned lds_u[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int ne = E.ne, nb = ne + 1;
    unsigned short *H = (unsigned short *)lds_u + (size_t)wid * 2 * nb * 64;
    constexpr unsigned kLoadMax = 0xFFFFu;
    unsigned long long evaluated = 0, fp32_evaluated = 0;
    unsigned seen = 0;  // FP32 form: guard-band pairs rechecked in FP64 (this lane)
    // The query tiles are cut into kXcds contiguous segments, claimed one
    // tile at a time: workgroups are dispatched to the XCDs round-robin, so
    // workgroup b drains segment b % kXcds first (each XCD's L2 then holds
    // the tree region of one region of query tiles) and then helps with the
    // others.  next: one counter per segment, kNextStride apart.
    int seg = blockIdx.x & (kXcds - 1), visited = 0;
    while (true) {
        const long long s0 = total_qt * seg / kXcds, s1 = total_qt * (seg + 1) / kXcds;
        unsigned long long off = 0;
        if (lane == 0) off = atomicAdd(&next[seg * kNextStride], 1ull);
        off = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(off >> 32)) << 32) |
              (unsigned)__builtin_amdgcn_readfirstlane((unsigned)off);
        if (off >= (unsigned long long)(s1 - s0)) {
            if (++visited == kXcds) break;
            seg = (seg + 1) & (kXcds - 1);
            continue;
        }
        const long long g = s0 + (long long)off;
        const int job = find_job(qt_prefix, njobs, g);
        const WJobDesc &W = jobs[job];
        const JobDesc &J = W.b;
        if (J.nqt < 0) continue;  // a tree-weighted job: k_wm counts it
        const int qt = (int)(g - qt_prefix[job]);
        const int b0 = J.qstart[qt];
        const int qcount = J.qstart[qt + 1] - b0;
        const bool valid = lane < qcount;
        double qx = 0, qy = 0, qz = 0;
        int qf = 0;
        if (valid) {
            qx = J.qx[b0 + lane];
            qy = J.qy[b0 + lane];
            qz = J.qz[b0 + lane];
            qf = J.qf[b0 + lane];
        }
        wq_zero(H, 2 * nb, lane);
        int tf0 = -1, tf1 = -1;
        unsigned load = 0u;
        const long long i0 = offs[g], i1 = offs[g + 1];
        for (long long i = i0; i < i1; i++) {
            const Item it = sload(items, (int)i);
            if (it.job & kSubtreeItem) {
                // a whole subtree under radius ehi, per tree field
                if (!subtree_item_ok(it, J.P)) {
                    if (lane == 0) atomicOr(err, kErrItem);
                    continue;
                }
                if (it.ehi >= ne) continue;
                const unsigned hf = sload(J.hfield, it.t);
                const int nf = hf < kHEmpty ? 1 : J.nfw;
                for (int k = 0; k < nf; k++) {
                    const int tf = hf < kHEmpty ? (int)hf : k;
                    const long long w = hf < kHEmpty ? sload(J.hweight, it.t) : sload(J.hfw, it.t * J.nfw + k);
                    if (w) wq_bulk(W, tf, it.ehi, w, ne, S, valid, qf, b0, hist_w, err, lane);
                }
                continue;
            }
            if (!tile_item_ok(it, J.ntt)) {
                if (lane == 0) atomicOr(err, kErrItem);
                continue;
            }
            const int t0 = __builtin_amdgcn_readfirstlane(J.tstart[it.t]);
            const int t1 = __builtin_amdgcn_readfirstlane(J.tstart[it.t + it.tiles()]);
            const int elo = it.elo, ehi = it.ehi;
            const unsigned tmeta = sload(J.tmeta, it.t);
            if ((tmeta & 3u) != 3u && ehi > elo) evaluated += (unsigned long long)qcount * (unsigned long long)(t1 - t0);
            if ((tmeta & 3u) == 3u) {
                // one tree field and unit weights: per-lane counts of each
                // direct radius, turned into the lane's bin histogram
                const unsigned npts = (unsigned)(t1 - t0);
                const int tf = (int)((tmeta >> 8) & 0xFFu);
                int slot = wq_pick(tf, tf0, tf1);
                if (slot < 0 || load + npts > kLoadMax) {
                    wq_flush_all(W, H, tf0, tf1, ne, S, valid, qf, b0, hist_w, err, lane);
                    tf0 = -1;
                    tf1 = -1;
                    load = 0u;
                    slot = 0;
                }
                tf0 = slot == 0 ? tf : tf0;
                tf1 = slot == 1 ? tf : tf1;
                load += npts;
                // the run's quads with a radius window each (lane_loop64w):
                // the quads beyond every radius cost one box test
                const int nt = it.tiles();
                const int g0 = __builtin_amdgcn_readfirstlane(J.tquad_start[it.t]);
                const int g1 = __builtin_amdgcn_readfirstlane(J.tquad_start[it.t + nt]);
                const Box qb = sload(J.qbox, qt);
                // FP32 form: the radii [e32, e32hi) of the run's size class
                // (its tiles share one) in passes of the matrix-core loop
                const MfClass4 *M = F32 ? W.mfc4 + ((tmeta >> 16) & 0xFu) : nullptr;
                const int f32lo = F32 ? sload(&M->e32, 0) : ne, f32hi = F32 ? sload(&M->e32hi, 0) : ne;
                unsigned prev = 0u;  // count under the radius before the chunk
                int kc = 0;
                for (int e0 = elo; e0 < ehi; e0 += kc) {
                    const int top = min(ehi, e0 < f32lo ? f32lo : (e0 < f32hi ? f32hi : ne));
                    kc = top - e0 < 8 ? top - e0 : 8;
                    const bool mf = e0 >= f32lo && e0 < f32hi;
                    double thr[8];
#pragma unroll
                    for (int e = 0; e < 8; e++) thr[e] = E.rsq[min(e0 + e, kMaxEdges - 1)];
                    const double thr_prev = e0 == elo ? -1.0 : E.rsq[e0 - 1];
                    unsigned cnt[8];
#pragma unroll
                    for (int e = 0; e < 8; e++) cnt[e] = 0u;
                    unsigned fresh = 0;
                    if (F32 && mf) {
#define KD_MW(k) fresh = mf4_lane_loop<k>(J, *M, it.t, g0, g1, e0, thr, thr_prev, qb, qx, qy, qz, valid, cnt, seen, lane)
                        switch (kc) {
                        case 1: KD_MW(1); break;
                        case 2: KD_MW(2); break;
                        case 3: KD_MW(3); break;
                        case 4: KD_MW(4); break;
                        case 5: KD_MW(5); break;
                        case 6: KD_MW(6); break;
                        case 7: KD_MW(7); break;
                        default: KD_MW(8); break;
                        }
#undef KD_MW
                        fp32_evaluated += (unsigned long long)qcount * fresh;
                    } else {
#define KD_LW(k) fresh = lane_loop64w<k>(J.tquad, J.tquad_box, J.tquad_meta, g0, g1, thr, thr_prev, qb, qx, qy, qz, cnt, lane)
                        switch (kc) {
                        case 1: KD_LW(1); break;
                        case 2: KD_LW(2); break;
                        case 3: KD_LW(3); break;
                        case 4: KD_LW(4); break;
                        case 5: KD_LW(5); break;
                        case 6: KD_LW(6); break;
                        case 7: KD_LW(7); break;
                        default: KD_LW(8); break;
                        }
#undef KD_LW
                    }
                    evaluated += (unsigned long long)qcount * fresh;
                    if (valid) {
#pragma unroll
                        for (int e = 0; e < 8; e++) {
                            if (e < kc) {
                                H[hix(slot * nb + e0 + e, lane)] += (unsigned short)(cnt[e] - prev);
                                prev = cnt[e];
                            }
                        }
                    }
                }
                // pairs beyond the last direct radius fall in bin ehi (every pair of
                // the item is under radius ehi); bins >= ne are not counted
                if (valid && ehi < ne) H[hix(slot * nb + ehi, lane)] += (unsigned short)(npts - prev);
                __builtin_amdgcn_wave_barrier();
                continue;
            }
            for (int j = t0; j < t1; j++) {
                const TPoint p = sload(J.tp, j);
                const unsigned w = p.meta & 0xFF;
                const int tf = (int)((p.meta >> 8) & 0xFF);
                int slot = wq_pick(tf, tf0, tf1);
                if (slot < 0 || load + w > kLoadMax) {
                    wq_flush_all(W, H, tf0, tf1, ne, S, valid, qf, b0, hist_w, err, lane);
                    tf0 = -1;
                    tf1 = -1;
                    load = 0u;
                    slot = 0;
                }
                tf0 = slot == 0 ? tf : tf0;
                tf1 = slot == 1 ? tf : tf1;
                load += w;
                const double dx = p.x - qx, dy = p.y - qy, dz = p.z - qz;
                const double n = fma(dz, dz, fma(dy, dy, dx * dx));
                int bin = elo;
                for (int e = elo; e < ehi; e++) bin += n < E.rsq[e] ? 0 : 1;
                if (valid && bin < ne) H[hix(slot * nb + bin, lane)] += (unsigned short)w;
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (tf0 >= 0) wq_flush_slot(W, H, 0, tf0, ne, S, valid, qf, b0, hist_w, err, lane);
        if (tf1 >= 0) wq_flush_slot(W, H, 1, tf1, ne, S, valid, qf, b0, hist_w, err, lane);
    }
    const unsigned long long rechecked = wave_sum_u64(seen);
    if (lane == 0) {
        atomicAdd(&stats[0], evaluated);
        if (rechecked) atomicAdd(&stats[1], rechecked);
        if (fp32_evaluated) atomicAdd(&stats[3], fp32_evaluated);
    }
}

// Rows of a point-major [n][spad] array in tile order: out[j] = in[order[j]],
// 16 bytes per thread.
__global__ void k_gather_rows(const unsigned char *__restrict__ in, const int *__restrict__ order, int n, int spad,
                              unsigned char *__restrict__ out)
{
    const int per = spad / 16;
    long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (k >= (long long)n * per) return;
    const int j = (int)(k / per), c = (int)(k - (long long)j * per);
    *(uint4 *)&out[(size_t)j * spad + 16 * c] = *(const uint4 *)&in[(size_t)order[j] * spad + 16 * c];
}

// ---- tree-weighted jobs on the matrix cores --------------------------------

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;
typedef int intx4 __attribute__((ext_vector_type(4)));
typedef int intx16 __attribute__((ext_vector_type(16)));

struct WMJob {
    JobDesc b;            // query side, tree tiles (tstart, tmeta)
    const int *kstart;    // ntt + 1: first 16-slot K-step of each tree tile
    const double *px, *py, *pz;  // nks * 16 slots, tile order, +inf padding
    const unsigned *pmeta;       // nks * 16 slots: point meta (field in bits 8..15)
    const uint4 *wb;      // [nks][Spad][16] bf16 tree multiplicities (two uint4 per set; bf16 form)
    const uint4 *wb8;     // [nks][Spad][16] u8 tree multiplicities, all < 128 (one uint4 per set; int8 form)
    const unsigned *wt;   // [ntt][Spad] tile sums of the tree multiplicities
    const unsigned *wq;   // [nqt][Spad][64] u8 query multiplicities (16 words per set)
    long long whist_base;
    int on;               // 1: a tree-weighted job
    // FP32 form (null: FP64): the slots' FP32 operands, the tile of each
    // K-step, and the guard-band thresholds of the tree's tile size classes
    const floatx4 *p32;
    const int *kstile;
    const MfClass4 *mfc4;
};

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v)
{
    union {
        uint4 u;
        bf16x8 b;
    } c;
    c.u = v;
    return c.b;
}

struct Quad4 {
    double v[4];
};

// m = 2m + (n < thr): one compare into VCC and one add-with-carry of m to
// itself (ordered compare: NaN never sets the bit).
__device__ __forceinline__ void shift_in_below(unsigned &m, double n, double thr)
{
    asm("v_cmp_gt_f64_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc" : "+v"(m) : "s"(thr), "v"(n) : "vcc");
}

// Slots k4 .. k0 (k4 >= k0, steps of 4, highest first) into the masks m:
// slot k ends at bit k - k0 of each radius' mask.
template <int KN>
__device__ __forceinline__ void wm_mask_run(const WMJob &W, int base, int k4, int k0, const double *thr, double gx,
                                            double gy, double gz, unsigned *m)
{
    for (; k4 >= k0; k4 -= 4) {
        const Quad4 X = sload((const Quad4 *)(W.px + base), k4 >> 2), Y = sload((const Quad4 *)(W.py + base), k4 >> 2),
                    Z = sload((const Quad4 *)(W.pz + base), k4 >> 2);
#pragma unroll
        for (int u = 3; u >= 0; u--) {
            const double dx = X.v[u] - gx, dy = Y.v[u] - gy, dz = Z.v[u] - gz;
            const double n = fma(dz, dz, fma(dy, dy, dx * dx));
#pragma unroll
            for (int e = 0; e < KN; e++) shift_in_below(m[e], n, thr[e]);
        }
    }
}

// Per-lane slot masks of one pass (<= 4 K-steps, 64 tree slots) for KN
// radii: bit k = tree slot k is under the radius (slots at +inf never are).
// Radius KN-1 is the bulk indicator (infinite radius) when `bulk`.
template <int KN>
__device__ __forceinline__ void wm_masks(const WMJob &W, int base, int nslots, const EdgeTable &E, int e0, bool bulk,
                                         double gx, double gy, double gz, unsigned *__restrict__ ml, int lane)
{
    double thr[KN];
#pragma unroll
    for (int e = 0; e < KN; e++) thr[e] = (bulk && e == KN - 1) ? INFINITY : E.rsq[min(e0 + e, kMaxEdges - 1)];
    unsigned lo[KN], hi[KN];
#pragma unroll
    for (int e = 0; e < KN; e++) lo[e] = hi[e] = 0u;
    // slots 32.. into hi, slots 0..31 into lo (nslots is a multiple of 16)
    wm_mask_run<KN>(W, base, nslots - 4, 32, thr, gx, gy, gz, hi);
    wm_mask_run<KN>(W, base, (nslots < 32 ? nslots : 32) - 4, 0, thr, gx, gy, gz, lo);
#pragma unroll
    for (int e = 0; e < KN; e++) {
        ml[(2 * e) * 64 + lane] = lo[e];
        ml[(2 * e + 1) * 64 + lane] = hi[e];
    }
}

// m = 2m + [bits(d) < lo]: a subtract-with-borrow and an add-with-carry of m
// to itself (the difference into the band minimum).
__device__ __forceinline__ void shift_in_below32(unsigned &m, float d, unsigned lo, unsigned &bmin)
{
    unsigned x;
    asm("v_subrev_co_u32_e32 %0, vcc, %3, %2\n\tv_addc_co_u32_e32 %1, vcc, %1, %1, vcc"
        : "=&v"(x), "+v"(m)
        : "v"(__float_as_uint(d)), "s"(lo)
        : "vcc");
    bmin = x < bmin ? x : bmin;
}

// wm_mask_run with the distances from the matrix cores (the FP32 form): the
// slots of each 4-slot group (one tile's) against the lanes' query points
// re-centred on that tile (Q, for tile `centred`), shifted in highest slot
// first; guard-band pairs recomputed with the reference's FP64 formula and
// their bits set before the next group shifts them up.
template <int KN>
__device__ __forceinline__ void wm_mask_run32(const WMJob &W, const MfClass4 &M, int base, int k4, int k0,
                                              const unsigned *lo32, unsigned Wb, const double *thr, double gx,
                                              double gy, double gz, unsigned *m, int &centred, MfQuery &Q, int lane)
{
    typedef const floatx4 __attribute__((address_space(1))) *g4;
    const g4 ops = (g4)W.p32;
    const bool valid = gx != INFINITY;
    for (; k4 >= k0; k4 -= 4) {
        const int tt = sload(W.kstile, (base + k4) >> 4);
        if (tt != centred) {
            Q = mf4_query(W.b, tt, M.eps32, gx, gy, gz, valid);
            centred = tt;
        }
        const floatx4 a = ops[base + k4 + (lane & 3)];
        const floatx4 D = mf4_dist(a, Q);
        unsigned bmin = 0xFFFFFFFFu;
#pragma unroll
        for (int e = 0; e < KN; e++)
#pragma unroll
            for (int r = 3; r >= 0; r--) shift_in_below32(m[e], D[r], lo32[e], bmin);
        if (__ballot(bmin < Wb)) {  // a pair of the group in a guard band (rare)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const unsigned bits = __float_as_uint(D[r]);
                bool any = false;
#pragma unroll
                for (int e = 0; e < KN; e++) any = any || bits - lo32[e] < Wb;
                if (!__ballot(any)) continue;
                const int sl = base + k4 + r;
                const double dx = sload(W.px, sl) - gx, dy = sload(W.py, sl) - gy, dz = sload(W.pz, sl) - gz;
                const double n = fma(dz, dz, fma(dy, dy, dx * dx));
#pragma unroll
                for (int e = 0; e < KN; e++)
                    if (bits - lo32[e] < Wb && n < thr[e]) m[e] |= 1u << r;
            }
        }
    }
}

// wm_masks in the FP32 form, for kn direct radii e0 .. e0+KN-1 of the FP32
// range of the run's size class M (no bulk indicator).
template <int KN>
__device__ __forceinline__ void wm_masks32(const WMJob &W, const MfClass4 &M, int base, int nslots,
                                           const EdgeTable &E, int e0, double gx, double gy, double gz,
                                           unsigned *__restrict__ ml, int lane)
{
    double thr[KN];
    unsigned lo32[KN];
    unsigned Wb = 1;
#pragma unroll
    for (int e = 0; e < KN; e++) {
        thr[e] = E.rsq[min(e0 + e, kMaxEdges - 1)];
        lo32[e] = __builtin_amdgcn_readfirstlane(sload(M.lo32, e0 + e));
        Wb = max(Wb, (unsigned)__builtin_amdgcn_readfirstlane(sload(M.band_w, e0 + e)));
    }
    unsigned lo[KN], hi[KN];
#pragma unroll
    for (int e = 0; e < KN; e++) lo[e] = hi[e] = 0u;
    int centred = -1;
    MfQuery Q = {0.f, 0.f, 0.f, INFINITY};
    // slots 32.. into hi, slots 0..31 into lo (nslots is a multiple of 16)
    wm_mask_run32<KN>(W, M, base, nslots - 4, 32, lo32, Wb, thr, gx, gy, gz, hi, centred, Q, lane);
    wm_mask_run32<KN>(W, M, base, (nslots < 32 ? nslots : 32) - 4, 0, lo32, Wb, thr, gx, gy, gz, lo, centred, Q,
                      lane);
#pragma unroll
    for (int e = 0; e < KN; e++) {
        ml[(2 * e) * 64 + lane] = lo[e];
        ml[(2 * e + 1) * 64 + lane] = hi[e];
    }
}

// For the kn radii of a pass: the indicator rows as MFMA A operands (bit
// patterns expanded to 1.0/0 bf16 through a 256-entry LDS table), then per
// 32-set column block -- its tree and query multiplicities loaded once for all
// radii -- and per radius and 32-query row block the v_mfma_f32_32x32x16_bf16
// products over the pass's four 16-slot K-steps (exact: at most 64 x 255 <
// 2^24 in FP32) and the query-weighted column sums, added to the histogram
// row.  The form for tree multiplicities up to 255 (wm_contract_i8 takes them
// below 128).  Operand layouts (lane l, element i; verified on the device):
// A[l%32][8(l/32)+i], B[8(l/32)+i][l%32]; D lane l value v -> row
// (v/4)*8 + (l/32)*4 + v%4, column l%32.
__device__ void wm_contract(const WMJob &W, int qt, int kk0, int nks, int kn, int e0, int bulk_out,
                            unsigned long long slotmask, unsigned *__restrict__ ml, const uint4 *__restrict__ table,
                            int S, int Spad, long long *__restrict__ row, int lane)
{
    const unsigned smlo = (unsigned)slotmask, smhi = (unsigned)(slotmask >> 32);
    const bool upper = lane >= 32;
    const int hsel = upper ? 1 : 0;
    const int s_col = lane & 31;
    const int nc = Spad / 32;
    // the indicator pattern words of each radius (row block 0 = query lanes
    // 0..31, row block 1 = 32..63; lane l holds row l%32 and slots
    // 8*(l/32)..+8 of each K-step: pattern byte 2ks + l/32) replace the masks
    // in LDS; each lane reads and writes only its own two words.  Slots past
    // the pass's nks K-steps have zero mask bits, so all four K-steps run
    // without branches.
    unsigned live = 0u;
    for (int e = 0; e < kn; e++) {
        const unsigned olo = ml[(2 * e) * 64 + lane] & smlo, ohi = ml[(2 * e + 1) * 64 + lane] & smhi;
        if (__ballot(olo | ohi)) live |= 1u << e;
        const unsigned slo = __shfl_xor(olo, 32, 64), shi = __shfl_xor(ohi, 32, 64);
        unsigned pat[2];
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            const bool own = (rb == 0) != upper;
            const unsigned mlo = own ? olo : slo, mhi = own ? ohi : shi;
            unsigned pk = 0u;
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                const int bytei = 2 * ks + hsel;
                const unsigned word = bytei < 4 ? mlo : mhi;
                pk |= ((word >> (8 * (bytei & 3))) & 0xffu) << (8 * ks);
            }
            pat[rb] = pk;
        }
        ml[(2 * e) * 64 + lane] = pat[0];
        ml[(2 * e + 1) * 64 + lane] = pat[1];
    }
    __builtin_amdgcn_wave_barrier();
    if (!live) return;
    for (int c = 0; c < nc; c++) {
        const int sidx = 32 * c + s_col;
        // the tree and query multiplicities of this column block, once for
        // all radii (K-steps past nks read step kk0: their A rows are zero)
        uint4 Bv[4];
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int kks = kk0 + (ks < nks ? ks : 0);
            Bv[ks] = W.wb[((size_t)kks * Spad + sidx) * 2 + hsel];
        }
        const unsigned *wqs = W.wq + ((size_t)qt * Spad + sidx) * 16;
        unsigned w[2][4];
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int g4 = 0; g4 < 4; g4++) w[rb][g4] = wqs[8 * rb + 2 * g4 + hsel];
        const unsigned live_s = __builtin_amdgcn_readfirstlane(live);
        for (int e = 0; e < kn; e++) {
            if (!((live_s >> e) & 1u)) continue;
            const int out = (bulk_out >= 0 && e == kn - 1) ? bulk_out : e0 + e;
            // both row blocks' A operands read from LDS together, then two
            // independent MFMA chains
            const unsigned pat0 = ml[(2 * e) * 64 + lane], pat1 = ml[(2 * e + 1) * 64 + lane];
            bf16x8 A0[4], A1[4];
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                A0[ks] = as_bf16x8(table[(pat0 >> (8 * ks)) & 0xffu]);
                A1[ks] = as_bf16x8(table[(pat1 >> (8 * ks)) & 0xffu]);
            }
            floatx16 C0 = {}, C1 = {};
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                C0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A0[ks], as_bf16x8(Bv[ks]), C0, 0, 0, 0);
                C1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1[ks], as_bf16x8(Bv[ks]), C1, 0, 0, 0);
            }
            // The entries are exact integers below 2^24 and the weights
            // bytes: 24-bit multiply-adds (full rate; a 32-bit multiply is
            // quarter rate)
            unsigned acc = 0u;
#pragma unroll
            for (int g4 = 0; g4 < 4; g4++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    acc += __umul24((unsigned)C0[4 * g4 + r], (w[0][g4] >> (8 * r)) & 0xffu) +
                           __umul24((unsigned)C1[4 * g4 + r], (w[1][g4] >> (8 * r)) & 0xffu);
            // lanes l and l+32 hold the two halves of column l%32
            const auto sw = __builtin_amdgcn_permlane32_swap(acc, acc, false, false);
            acc = sw[0] + sw[1];
            if (!upper && sidx < S && acc)
                atomicAdd((unsigned long long *)&row[(long long)out * Spad + sidx], (unsigned long long)acc);
        }
    }
}

// wm_contract for tree multiplicities below 128, as signed int8 operands of
// v_mfma_i32_16x16x64_i8: one MFMA per 16-query x 16-set output block covers
// the pass's 64 slots, so each radius and 32-set column block takes eight
// independent MFMAs (four row blocks x two column blocks); 0/1 bytes from a
// table of 8-byte entries, int32 products (exact: at most 64 x 127).  Against
// the bf16 form: half the matrix-core work, half the operand bytes, no
// float-to-integer conversions.  Layouts (lane l, element i; verified on the
// device): A[l%16][16(l/16)+i], B[16(l/16)+i][l%16]; D lane l value v -> row
// 4(l/16)+v, column l%16.
__device__ void wm_contract_i8(const WMJob &W, int qt, int kk0, int nks, int kn, int e0, int bulk_out,
                               unsigned long long slotmask, unsigned *__restrict__ ml, const uint4 *__restrict__ table,
                               int S, int Spad, long long *__restrict__ row, int lane)
{
    const int j = lane >> 4, l16 = lane & 15;  // lane's 16-slot K-step of the pass, its row/column in a block
    const unsigned smlo = (unsigned)slotmask, smhi = (unsigned)(slotmask >> 32);
    const unsigned sm = ((j < 2 ? smlo : smhi) >> (16 * (j & 1))) & 0xffffu;
    const int nc = Spad / 32;
    // pattern words of each radius: lane l holds, for row block rb (queries
    // 16rb..16rb+15), the mask bits of query 16rb + l%16 for slots
    // 16(l/16)..+16, rb 0/1 in the first word, 2/3 in the second (read from the
    // other lanes' masks in LDS before this lane's words are overwritten)
    unsigned live = 0u;
    for (int e = 0; e < kn; e++) {
        const unsigned olo = ml[(2 * e) * 64 + lane] & smlo, ohi = ml[(2 * e + 1) * 64 + lane] & smhi;
        if (__ballot(olo | ohi)) live |= 1u << e;
        unsigned p[4];
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
            p[rb] = (ml[(2 * e + (j >> 1)) * 64 + 16 * rb + l16] >> (16 * (j & 1))) & sm;
        __builtin_amdgcn_wave_barrier();
        ml[(2 * e) * 64 + lane] = p[0] | (p[1] << 16);
        ml[(2 * e + 1) * 64 + lane] = p[2] | (p[3] << 16);
    }
    __builtin_amdgcn_wave_barrier();
    if (!live) return;
    const uint2 *table8 = (const uint2 *)table;
    const int kks = kk0 + (j < nks ? j : 0);  // K-steps past nks: zero A rows
    for (int c = 0; c < nc; c++) {
        intx4 B[2];
        unsigned w[4][2];
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
            const int sidx = 32 * c + 16 * cb + l16;
            const uint4 b = W.wb8[(size_t)kks * Spad + sidx];
            B[cb] = intx4{(int)b.x, (int)b.y, (int)b.z, (int)b.w};
            const unsigned *wqs = W.wq + ((size_t)qt * Spad + sidx) * 16;
#pragma unroll
            for (int rb = 0; rb < 4; rb++) w[rb][cb] = wqs[4 * rb + j];
        }
        const unsigned live_s = __builtin_amdgcn_readfirstlane(live);
        for (int e = 0; e < kn; e++) {
            if (!((live_s >> e) & 1u)) continue;
            const int out = (bulk_out >= 0 && e == kn - 1) ? bulk_out : e0 + e;
            const unsigned pw[2] = {ml[(2 * e) * 64 + lane], ml[(2 * e + 1) * 64 + lane]};
            intx4 A[4];
#pragma unroll
            for (int rb = 0; rb < 4; rb++) {
                const unsigned p = pw[rb >> 1] >> (16 * (rb & 1));
                const uint2 a = table8[p & 0xffu], b = table8[(p >> 8) & 0xffu];
                A[rb] = intx4{(int)a.x, (int)a.y, (int)b.x, (int)b.y};
            }
            unsigned acc[2] = {0u, 0u};
#pragma unroll
            for (int rb = 0; rb < 4; rb++)
#pragma unroll
                for (int cb = 0; cb < 2; cb++) {
                    const intx4 D = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[rb], B[cb], intx4{0, 0, 0, 0}, 0, 0, 0);
#pragma unroll
                    for (int v = 0; v < 4; v++) acc[cb] += __umul24((unsigned)D[v], (w[rb][cb] >> (8 * v)) & 0xffu);
                }
            // sums over the four lane groups (rows 4(l/16)..+4 of each block)
#pragma unroll
            for (int cb = 0; cb < 2; cb++) {
                const auto s32 = __builtin_amdgcn_permlane32_swap(acc[cb], acc[cb], false, false);
                const unsigned h = s32[0] + s32[1];
                const auto s16 = __builtin_amdgcn_permlane16_swap(h, h, false, false);
                acc[cb] = s16[0] + s16[1];
            }
            const int so = 32 * c + lane;
            const unsigned v = lane < 16 ? acc[0] : acc[1];
            if (lane < 32 && so < S && v)
                atomicAdd((unsigned long long *)&row[(long long)out * Spad + so], (unsigned long long)v);
        }
    }
}

// All passes of one (item, query group, tree field) for kn <= 8 radii from e0
// (the last one the bulk indicator when bulk_out >= 0).
template <bool I8>
__device__ __forceinline__ void wm_passes(int kn, const WMJob &W, int qt, int ks0, int ks1, const EdgeTable &E, int e0,
                                          int bulk_out, unsigned long long slotmask, double gx, double gy, double gz,
                                          unsigned *__restrict__ ml, const uint4 *__restrict__ table, int S, int Spad,
                                          long long *__restrict__ row, int lane, const MfClass4 *M32 = nullptr)
{
    const bool bulk = bulk_out >= 0;
    for (int kk0 = ks0; kk0 < ks1; kk0 += 4) {
        const int nks = ks1 - kk0 < 4 ? ks1 - kk0 : 4;
        const int base = kk0 * 16, nslots = nks * 16;
        if (M32) {  // FP32 form: direct radii only
            switch (kn) {
            case 1: wm_masks32<1>(W, *M32, base, nslots, E, e0, gx, gy, gz, ml, lane); break;
            case 2: wm_masks32<2>(W, *M32, base, nslots, E, e0, gx, gy, gz, ml, lane); break;
            case 3: wm_masks32<3>(W, *M32, base, nslots, E, e0, gx, gy, gz, ml, lane); break;
            case 4: wm_masks32<4>(W, *M32, base, nslots, E, e0, gx, gy, gz, ml, lane); break;
            case 5: wm_masks32<5>(W, *M32, base, nslots, E, e0, gx, gy, gz, ml, lane); break;
            case 6: wm_masks32<6>(W, *M32, base, nslots, E, e0, gx, gy, gz, ml, lane); break;
            case 7: wm_masks32<7>(W, *M32, base, nslots, E, e0, gx, gy, gz, ml, lane); break;
            default: wm_masks32<8>(W, *M32, base, nslots, E, e0, gx, gy, gz, ml, lane); break;
            }
            __builtin_amdgcn_wave_barrier();
            if (I8)
                wm_contract_i8(W, qt, kk0, nks, kn, e0, bulk_out, slotmask, ml, table, S, Spad, row, lane);
            else
                wm_contract(W, qt, kk0, nks, kn, e0, bulk_out, slotmask, ml, table, S, Spad, row, lane);
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        switch (kn) {
        case 1: wm_masks<1>(W, base, nslots, E, e0, bulk, gx, gy, gz, ml, lane); break;
        case 2: wm_masks<2>(W, base, nslots, E, e0, bulk, gx, gy, gz, ml, lane); break;
        case 3: wm_masks<3>(W, base, nslots, E, e0, bulk, gx, gy, gz, ml, lane); break;
        case 4: wm_masks<4>(W, base, nslots, E, e0, bulk, gx, gy, gz, ml, lane); break;
        case 5: wm_masks<5>(W, base, nslots, E, e0, bulk, gx, gy, gz, ml, lane); break;
        case 6: wm_masks<6>(W, base, nslots, E, e0, bulk, gx, gy, gz, ml, lane); break;
        case 7: wm_masks<7>(W, base, nslots, E, e0, bulk, gx, gy, gz, ml, lane); break;
        default: wm_masks<8>(W, base, nslots, E, e0, bulk, gx, gy, gz, ml, lane); break;
        }
        __builtin_amdgcn_wave_barrier();
        if (I8)
                wm_contract_i8(W, qt, kk0, nks, kn, e0, bulk_out, slotmask, ml, table, S, Spad, row, lane);
            else
                wm_contract(W, qt, kk0, nks, kn, e0, bulk_out, slotmask, ml, table, S, Spad, row, lane);
        __builtin_amdgcn_wave_barrier();
    }
}

// One wave per query tile of the tree-weighted jobs.  Pure items (one tree
// field): the bulk count for radii >= ehi is (sum of the group's query
// weights) x (tile sums of the tree weights); the direct radii go through
// wm_pass.  Impure tiles: one pass per tree field with a slot mask, the bulk
// as an indicator with an infinite radius.
// Six waves per SIMD for the FP64 form with int8 operands (80 VGPRs, some
// of the item loop's state spilled: 117 -> 113 ms on the bootstrap workload;
// five waves 118, seven 123, eight 139 ms), four for the others, whose
// larger register state spills more.
template <bool F32, bool I8>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((F32 || !I8) ? 4 : 6))) void k_wm(const WMJob *__restrict__ jobs, const long long *__restrict__ qt_prefix,
                                           int njobs, const long long *__restrict__ granges, int nranges,
                                           long long total_tiles, const Item *__restrict__ items,
                                           const long long *__restrict__ offs, EdgeTable E, int S, int Spad,
                                           long long *__restrict__ hist_w, unsigned long long *__restrict__ stats,
                                           unsigned *__restrict__ err)
{
    // pattern byte b -> eight operand elements (bit j -> element j): 0/1
    // bytes (I8, uint2 entries) or bf16 0/1.0 (uint4 entries)
    __shared__ uint4 table[256];
    __shared__ unsigned s_masks[4][16 * 64];  // per wave: 8 radii x {lo, hi} x 64 lanes
    for (int b = threadIdx.x; b < 256; b += blockDim.x) {
        if (I8) {
            unsigned w[2] = {0u, 0u};
            for (int j = 0; j < 8; j++) w[j >> 2] |= (unsigned)((b >> j) & 1) << (8 * (j & 3));
            ((uint2 *)table)[b] = make_uint2(w[0], w[1]);
        } else {
            unsigned w[4];
            for (int h = 0; h < 4; h++) {
                const unsigned lo = ((b >> (2 * h)) & 1) ? 0x3F80u : 0u,
                               hi = ((b >> (2 * h + 1)) & 1) ? 0x3F80u : 0u;
                w[h] = lo | (hi << 16);
            }
            table[b] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    unsigned *ml = s_masks[wid];
    const int ne = E.ne, rowlen = 2 * ne + 1;
    unsigned long long evaluated = 0, fp32_evaluated = 0;
    // The query tiles of the tree-weighted jobs (ranges [granges[2r],
    // granges[2r+1]) of the global tile index g, numbered one after the
    // other), one wave per tile, in one contiguous slice per XCD, so that
    // each XCD's L2 holds the tree region of one region of query tiles.
    const XcdSlice sl = xcd_slice(total_tiles, wid, blockDim.x >> 6);
    for (long long v = sl.first; v < sl.end; v += sl.stride) {
        long long g = -1, before = 0;
        for (int r = 0; r < nranges; r++) {
            const long long len = granges[2 * r + 1] - granges[2 * r];
            if (v < before + len) {
                g = granges[2 * r] + (v - before);
                break;
            }
            before += len;
        }
        if (g < 0) break;
        const int job = __builtin_amdgcn_readfirstlane(find_job(qt_prefix, njobs, g));
        const WMJob &W = jobs[job];
        if (!W.on) {
            if (lane == 0) atomicOr(err, kErrItem);
            continue;
        }
        const JobDesc &J = W.b;
        const int qt = (int)(g - qt_prefix[job]);
        const int b0 = J.qstart[qt];
        const int qcount = J.qstart[qt + 1] - b0;
        const bool valid = lane < qcount;
        double qx = 0, qy = 0, qz = 0;
        int qf = 0;
        if (valid) {
            qx = J.qx[b0 + lane];
            qy = J.qy[b0 + lane];
            qz = J.qz[b0 + lane];
            qf = J.qf[b0 + lane];
        }
        const long long i1 = offs[g + 1];
        for (long long i = offs[g]; i < i1; i++) {
            const Item it = sload(items, (int)i);
            // the enumeration gives tree-weighted jobs tile items only
            if ((it.job & kSubtreeItem) || !tile_item_ok(it, J.ntt)) {
                if (lane == 0) atomicOr(err, kErrItem);
                continue;
            }
            const int t = it.t, nt = it.tiles(), elo = it.elo, ehi = it.ehi;
            const int ks0 = __builtin_amdgcn_readfirstlane(W.kstart[t]);
            const int ks1 = __builtin_amdgcn_readfirstlane(W.kstart[t + nt]);
            const int p0 = __builtin_amdgcn_readfirstlane(J.tstart[t]);
            const int p1 = __builtin_amdgcn_readfirstlane(J.tstart[t + nt]);
            const unsigned tmeta = sload(J.tmeta, t);
            const bool pure = (tmeta & 1u) != 0u;
            unsigned long long remaining = __ballot(valid);
            while (remaining) {
                const int leader = __ffsll((long long)remaining) - 1;
                const int gq = __builtin_amdgcn_readlane(qf, leader);
                const bool ing = valid && qf == gq;
                const unsigned long long gm = __ballot(ing);
                remaining &= ~gm;
                const double gx = ing ? qx : INFINITY, gy = ing ? qy : INFINITY, gz = ing ? qz : INFINITY;
                if (ehi > elo) evaluated += (unsigned long long)__popcll(gm) * (unsigned long long)(p1 - p0);
                if (pure) {
                    const int tf = (int)((tmeta >> 8) & 0xFFu);
                    const long long sl = (long long)gq * J.nq_fields + tf;
                    if (gq < 0 || sl < 0 || sl >= J.nslots) {
                        if (lane == 0) atomicOr(err, 1u);
                        continue;
                    }
                    long long *row = hist_w + W.whist_base + sl * rowlen * Spad;
                    if (ehi < ne) {  // bulk: every pair of the item counts for radii >= ehi
                        for (int s0 = 0; s0 < S; s0 += 64) {
                            const int sidx = s0 + lane;
                            if (sidx < S) {
                                const unsigned *wqs = W.wq + ((size_t)qt * Spad + sidx) * 16;
                                unsigned qsum = 0u;
                                for (int k = 0; k < 16; k++) {
                                    const unsigned bits = (unsigned)(gm >> (4 * k)) & 0xfu;
                                    const unsigned w4 = wqs[k];
                                    qsum += (bits & 1u ? (w4 & 0xffu) : 0u) + (bits & 2u ? ((w4 >> 8) & 0xffu) : 0u) +
                                            (bits & 4u ? ((w4 >> 16) & 0xffu) : 0u) + (bits & 8u ? (w4 >> 24) : 0u);
                                }
                                unsigned long long tsum = 0;
                                for (int tt = t; tt < t + nt; tt++) tsum += W.wt[(size_t)tt * Spad + sidx];
                                const unsigned long long v = tsum * qsum;
                                if (v) atomicAdd((unsigned long long *)&row[(long long)(ne + ehi) * Spad + sidx], v);
                            }
                        }
                    }
                    // FP32 form: the radii [e32, e32hi) of the run's size
                    // class (its tiles share one) in passes of the FP32 masks
                    const MfClass4 *M = F32 ? W.mfc4 + ((tmeta >> 16) & 0xFu) : nullptr;
                    const int f32lo = F32 ? sload(&M->e32, 0) : ne, f32hi = F32 ? sload(&M->e32hi, 0) : ne;
                    int kn = 0;
                    for (int e0 = elo; e0 < ehi; e0 += kn) {
                        const int top = min(ehi, e0 < f32lo ? f32lo : (e0 < f32hi ? f32hi : ne));
                        kn = top - e0 < 8 ? top - e0 : 8;
                        const bool mf = F32 && e0 >= f32lo && e0 < f32hi;
                        if (mf && e0 == max(elo, f32lo))  // the item's first FP32 pass
                            fp32_evaluated += (unsigned long long)__popcll(gm) * (unsigned long long)(p1 - p0);
                        wm_passes<I8>(kn, W, qt, ks0, ks1, E, e0, -1, ~0ull, gx, gy, gz, ml, table, S, Spad, row, lane,
                                  mf ? M : nullptr);
                    }
                } else {
                    // impure tile (items of impure tiles are single tiles): one pass per tree field
                    unsigned long long done = 0ull;
                    const int nslot = (ks1 - ks0) * 16;
                    while (true) {
                        int tf = -1;
                        unsigned long long smask = 0ull;
                        for (int k = 0; k < nslot && k < 64; k++) {
                            if ((done >> k) & 1ull) continue;
                            if (k >= p1 - p0) break;
                            const int f = (int)((sload(W.pmeta, ks0 * 16 + k) >> 8) & 0xFFu);
                            if (tf < 0) tf = f;
                            if (f == tf) smask |= 1ull << k;
                        }
                        if (tf < 0) break;
                        done |= smask;
                        const long long sl = (long long)gq * J.nq_fields + tf;
                        if (gq < 0 || sl < 0 || sl >= J.nslots) {
                            if (lane == 0) atomicOr(err, 1u);
                            continue;
                        }
                        long long *row = hist_w + W.whist_base + sl * rowlen * Spad;
                        // direct radii, then the bulk as an infinite radius
                        int e0 = elo;
                        while (true) {
                            const int nd = ehi - e0 < 7 ? ehi - e0 : 7;
                            const bool with_bulk = e0 + nd >= ehi && ehi < ne;
                            const int kn = nd + (with_bulk ? 1 : 0);
                            if (kn == 0) break;
                            wm_passes<I8>(kn, W, qt, ks0, ks1, E, e0, with_bulk ? ne + ehi : -1, smask, gx, gy, gz, ml,
                                      table, S, Spad, row, lane);
                            e0 += nd;
                            if (e0 >= ehi) break;
                        }
                    }
                }
            }
        }
    }
    if (lane == 0) {
        atomicAdd(&stats[0], evaluated);
        if (fp32_evaluated) atomicAdd(&stats[3], fp32_evaluated);
    }
}

// Weighted field sizes: out[s][f] += sum of w[s][order[j]] over the points j
// (tile order) with field f; one block per (set, chunk), LDS partial sums.
// Field of each point in input order (the query side keeps them in tile order).
__global__ void k_fields_input_order(const int *__restrict__ order, const int *__restrict__ f, int n,
                                     int *__restrict__ f_in)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) f_in[order[j]] = f[j];
}

// Per set (blockIdx.y) and field, the sum of the multiplicities: the set's row
// and the fields both read in input order, an LDS histogram per block.
__global__ void k_field_sizes(const unsigned char *__restrict__ w, const int *__restrict__ f_in, int n, int nf,
                              unsigned long long *__restrict__ out)
{
    __shared__ unsigned part[256];
    const int sset = blockIdx.y;
    for (int b = threadIdx.x; b < nf; b += blockDim.x) part[b] = 0u;
    __syncthreads();
    const int chunk = (n + gridDim.x - 1) / gridDim.x;
    const int j0 = blockIdx.x * chunk, j1 = min(n, j0 + chunk);
    for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
        const int fj = f_in[j];
        const unsigned wj = w ? w[(size_t)sset * n + j] : 1u;
        if (fj >= 0 && fj < nf && wj) atomicAdd(&part[fj], wj);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nf; b += blockDim.x)
        if (part[b]) atomicAdd(&out[(size_t)sset * nf + b], (unsigned long long)part[b]);
}

// d_w: the device copy of the weights ([S][n_raw], input order) or null (all 1)
static void field_sizes(const kd_device_tree *X, const unsigned char *d_w, int S, int nf, long long *h_out,
                        hipStream_t s)
{
    KD_REQUIRE(nf >= 1 && nf <= 256, "weighted field sizes: at most 256 fields");
    const int n = X->n_raw;
    DevBuf<unsigned long long> out((size_t)S * nf);
    KD_HIP(hipMemsetAsync(out.p, 0, out.bytes(), s));
    const int chunks = std::max(1, std::min(256, n / 65536 + 1));
    DevBuf<int> f_in(n);
    hipLaunchKernelGGL(k_fields_input_order, dim3(nblk(n)), dim3(256), 0, s, X->q.order.p, X->q.f.p, n, f_in.p);
    hipLaunchKernelGGL(k_field_sizes, dim3(chunks, S), dim3(256), 0, s, d_w, f_in.p, n, nf, out.p);
    KD_HIP(hipGetLastError());
    KD_HIP(hipMemcpyAsync(h_out, out.p, out.bytes(), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
}

// ---- layouts for k_wm -------------------------------------------------------

__global__ void k_tile_ksteps(const int *__restrict__ tstart, int ntiles, int *__restrict__ nk)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntiles) nk[t] = (tstart[t + 1] - tstart[t] + 15) >> 4;
}

__device__ __forceinline__ int tile_of_kstep(const int *__restrict__ kstart, int ntiles, int ks)
{
    int lo = 0, hi = ntiles;  // tile t with kstart[t] <= ks < kstart[t+1]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (kstart[mid] <= ks) lo = mid;
        else hi = mid;
    }
    return lo;
}

// slot k of K-step ks: point tstart[t] + 16 (ks - kstart[t]) + k, or padding.
// With centre (the FP32 form): the slot's FP32 operands in p32 ({bx, by, bz,
// |b|^2} relative to the tile's centre, as the quads' quad32; padding
// {0, 0, 0, +inf}) and the tile of each K-step in kstile.
__global__ void k_fill_slots(const TPoint *__restrict__ pts, const int *__restrict__ tstart,
                             const int *__restrict__ kstart, int ntiles, int nks, double *__restrict__ px,
                             double *__restrict__ py, double *__restrict__ pz, unsigned *__restrict__ pmeta,
                             int *__restrict__ pidx, const double *__restrict__ centre, float4 *__restrict__ p32,
                             int *__restrict__ kstile)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= 16LL * nks) return;
    const int ks = (int)(i >> 4), k = (int)(i & 15);
    const int t = tile_of_kstep(kstart, ntiles, ks);
    const int j = tstart[t] + 16 * (ks - kstart[t]) + k;
    if (centre && k == 0) kstile[ks] = t;
    if (j < tstart[t + 1]) {
        px[i] = pts[j].x;
        py[i] = pts[j].y;
        pz[i] = pts[j].z;
        pmeta[i] = pts[j].meta;
        pidx[i] = j;
        if (centre) {
            const float bx = (float)(pts[j].x - centre[3 * t]), by = (float)(pts[j].y - centre[3 * t + 1]),
                        bz = (float)(pts[j].z - centre[3 * t + 2]);
            p32[i] = make_float4(bx, by, bz, (float)((double)bx * bx + (double)by * by + (double)bz * bz));
        }
    } else {
        px[i] = py[i] = pz[i] = INFINITY;
        pmeta[i] = 0u;
        pidx[i] = -1;
        if (centre) p32[i] = make_float4(0.f, 0.f, 0.f, INFINITY);
    }
}

// bf16 tree multiplicities [ks][Spad][16] (0 for padding slots and sets >= S)
// Multiplicities [S][n] u8 (input order) -> [n][Spad] u8, zero padded: one
// 64-set x 64-point tile per block through LDS, so that both the reads (runs
// of a set's row) and the writes (runs of a point's row) are contiguous.
__global__ __launch_bounds__(256) void k_transpose_w(const unsigned char *__restrict__ w, int n, int S, int Spad,
                                                     unsigned char *__restrict__ out)
{
    __shared__ unsigned char tile[64][64 + 4];
    const int t = threadIdx.x;
    const long long p0 = (long long)blockIdx.x * 64;
    const int s0 = blockIdx.y * 64;
    {
        const int r = t >> 2, c0 = (t & 3) * 16;  // set row r, points c0..c0+15
        const int sset = s0 + r;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const long long p = p0 + c0 + k;
            tile[r][c0 + k] = (sset < S && p < n) ? w[(size_t)sset * n + p] : (unsigned char)0;
        }
    }
    __syncthreads();
    const int c = t >> 2, r0 = (t & 3) * 16;  // point c, sets r0..r0+15
    const long long p = p0 + c;
    if (p < n && s0 + r0 < Spad) {
        unsigned v[4];
#pragma unroll
        for (int q = 0; q < 4; q++)
            v[q] = (unsigned)tile[r0 + 4 * q][c] | ((unsigned)tile[r0 + 4 * q + 1][c] << 8) |
                   ((unsigned)tile[r0 + 4 * q + 2][c] << 16) | ((unsigned)tile[r0 + 4 * q + 3][c] << 24);
        *(uint4 *)&out[(size_t)p * Spad + s0 + r0] = make_uint4(v[0], v[1], v[2], v[3]);
    }
}

// Bit 0 of *big set when any byte of w[0, n) is >= 128 (a multiplicity that
// the signed int8 operand cannot hold).
__global__ void k_any_big(const unsigned char *__restrict__ w, long long n, unsigned *__restrict__ big)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    unsigned any = 0u;
    const long long n16 = n / 16;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n16; i += stride) {
        const uint4 b = ((const uint4 *)w)[i];
        any |= b.x | b.y | b.z | b.w;
    }
    for (long long i = 16 * n16 + blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) any |= w[i];
    if (any & 0x80808080u) atomicOr(big, 1u);
}

// Tree multiplicities [K-step][Spad][16] as the MFMA B operand, from the
// point-major copy: one thread per slot reads its point's Spad bytes; bytes
// (I8, every multiplicity < 128) or bf16 (exact for bytes).
template <bool I8>
__global__ void k_fill_wb(const unsigned char *__restrict__ wT, const int *__restrict__ raw,
                          const int *__restrict__ pidx, int nks, int Spad, void *__restrict__ out_v)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= 16LL * nks) return;
    const int ks = (int)(i >> 4), k = (int)(i & 15);
    const int j = pidx[i];
    const unsigned char *row = j >= 0 ? wT + (size_t)raw[j] * Spad : nullptr;
    for (int s16 = 0; s16 < Spad; s16 += 16) {
        uint4 b = make_uint4(0u, 0u, 0u, 0u);
        if (row) b = *(const uint4 *)(row + s16);
        const unsigned words[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const unsigned v = (words[q >> 2] >> (8 * (q & 3))) & 0xFFu;
            const size_t o = ((size_t)ks * Spad + s16 + q) * 16 + k;
            if (I8)
                ((unsigned char *)out_v)[o] = (unsigned char)v;
            else
                ((unsigned short *)out_v)[o] = (unsigned short)(__float_as_uint((float)v) >> 16);  // exact
        }
    }
}

// Per tree tile and set, the sum of the tree multiplicities: one thread per
// (tile, four sets), reading a u32 of each point's row.
__global__ void k_tile_wsum(const unsigned char *__restrict__ wT, const int *__restrict__ raw,
                            const int *__restrict__ tstart, int ntiles, int Spad, unsigned *__restrict__ out)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int ng = Spad / 4;
    if (i >= (long long)ntiles * ng) return;
    const int t = (int)(i / ng), g4 = (int)(i - (long long)t * ng);
    unsigned sum[4] = {0u, 0u, 0u, 0u};
    for (int j = tstart[t]; j < tstart[t + 1]; j++) {
        const unsigned v = *(const unsigned *)(wT + (size_t)raw[j] * Spad + 4 * g4);
#pragma unroll
        for (int q = 0; q < 4; q++) sum[q] += (v >> (8 * q)) & 0xFFu;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) out[(size_t)t * Spad + 4 * g4 + q] = sum[q];
}

// query multiplicities [qt][Spad][64] u8 (1 for every point without weights, 0 for padding)
__global__ void k_fill_wq64(const unsigned char *__restrict__ wtile, const int *__restrict__ qstart, int nqt, int nq,
                            int S, int Spad, unsigned char *__restrict__ out)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= (long long)nqt * Spad * 64) return;
    const int q = (int)(i & 63);
    const long long r = i >> 6;
    const int qt = (int)(r / Spad), sset = (int)(r - (long long)qt * Spad);
    const int j = qstart[qt] + q;
    unsigned char v = 0;
    if (j < qstart[qt + 1] && sset < S) v = wtile ? wtile[(size_t)j * Spad + sset] : (unsigned char)1;
    out[i] = v;
}

// Cumulative counts of one weighted job in its output layout [set][radius][slot]:
// one thread per (set, slot) walks the row (direct counts, then the bulk
// entries that add to every radius from theirs on) in either histogram
// layout ([set][slot][element], or [slot][element][set] with sets padded to
// spad when tree weights are set).
__global__ void k_weighted_cum(const long long *__restrict__ hist, int S, int nslots, int ne, int rowlen, int spad,
                               int tree_layout, const int *__restrict__ emap, int nr, long long *__restrict__ out)
{
    const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (k >= (long long)S * nslots) return;
    const int ss = (int)(k / nslots), sl = (int)(k - (long long)ss * nslots);
    const long long base = tree_layout ? (long long)sl * rowlen * spad + ss : ((long long)ss * nslots + sl) * rowlen;
    const long long es = tree_layout ? spad : 1;
    long long cum[kMaxEdges];
    long long b = 0;
    for (int e = 0; e < ne; e++) {
        b += hist[base + (long long)(ne + e) * es];
        cum[e] = hist[base + (long long)e * es] + b;
    }
    for (int i = 0; i < nr; i++) out[((long long)ss * nr + i) * nslots + sl] = cum[emap[i]];
}

void count_pairs_weighted(const std::vector<WPairJob> &jobs, int S, const std::vector<double> &radii, int precision,
                          std::vector<long long *> &counts_out, kd_stats_t *stats, Sorter &so, hipStream_t s)
{
    auto wall0 = std::chrono::steady_clock::now();
    const int nr = (int)radii.size();
    KD_REQUIRE(nr >= 1, "weighted pair count: at least one radius is required");
    KD_REQUIRE(S >= 1, "weighted pair count: at least one weight set is required");
    KD_REQUIRE(precision == KD_PRECISION_FP64 || precision == KD_PRECISION_FP32_GUARDED,
               "weighted pair count: unknown precision form");
    const bool fp32 = precision == KD_PRECISION_FP32_GUARDED;
    for (double r : radii)
        KD_REQUIRE(std::isfinite(r) && r >= 0.0, "weighted pair count: radii must be finite and >= 0");
    std::vector<double> edges(radii);
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    const int ne = (int)edges.size();
    KD_REQUIRE(ne <= kWMaxEdges, "weighted pair count: at most 32 distinct radii per call");
    EdgeTable E;
    memset(&E, 0, sizeof(E));
    E.ne = ne;
    for (int e = 0; e < ne; e++) {
        E.r[e] = edges[e];
        E.rsq[e] = edges[e] * edges[e];
    }
    const int rowlen = 2 * ne + 1;
    const int nj = (int)jobs.size();
    for (auto &jb : jobs)
        KD_REQUIRE((long long)jb.nq_fields * jb.nt_fields * rowlen < (1LL << 19),
                   "weighted pair count: too many field pairs x radii per job");
    // weight sets padded to whole 32-set column blocks.  Histogram blocks:
    // tree-weighted jobs [slot][2ne+1][Spad] (a wave's 32 sets contiguous),
    // the others [S][slot][2ne+1] (a wave's 64 sets on separate lines, which
    // spreads the atomics of many waves flushing the same slot)
    const int Spad = (S + 31) / 32 * 32;
    std::vector<WJobDesc> hw(nj);
    std::vector<JobDesc> hj(nj);
    std::vector<long long> qt_prefix(nj + 1, 0);
    std::vector<DevBuf<unsigned char>> wq_tile(nj);
    // each distinct weight array goes to the device once per call (the same
    // catalogue's weights often serve two jobs: as the query side of D1D2 and
    // D1R2, or the tree side of D1D2 and the query side of D2R1), and its
    // query-order copy once per (array, query catalogue)
    std::map<std::pair<const void *, size_t>, DevBuf<unsigned char>> uploads;
    std::map<std::pair<const void *, const void *>, const unsigned char *> gathered;
    auto upload = [&](const unsigned char *h, size_t nb) -> const unsigned char * {
        DevBuf<unsigned char> &d = uploads[std::make_pair((const void *)h, nb)];
        if (!d.p) {
            d.alloc(nb);
            KD_HIP(hipMemcpyAsync(d.p, h, nb, hipMemcpyHostToDevice, s));
        }
        return d.p;
    };
    // every weight array goes up first, so that the copies run while the GPU
    // is still busy with earlier work (the kernels below wait for them in
    // stream order anyway)
    for (int i = 0; i < nj; i++) {
        if (jobs[i].query_weights) upload(jobs[i].query_weights, (size_t)S * jobs[i].query->n_raw);
        if (jobs[i].tree_weights) upload(jobs[i].tree_weights, (size_t)S * jobs[i].tree->n_raw);
    }
    std::vector<char> tw(nj, 0);
    long long hist_len = 0;
    bool any_tw = false, any_qonly = false;
    // FP32 form: the guard-band thresholds of every job and tree tile size class
    std::vector<MfClass4> hmfc(fp32 ? (size_t)nj * kMfClasses : 0);
    DevBuf<MfClass4> dmfc(std::max<size_t>(hmfc.size(), 1));
    for (int i = 0; fp32 && i < nj; i++)
        for (int k = 0; k < kMfClasses; k++)
            fp32_thresholds4(hmfc[(size_t)i * kMfClasses + k], E,
                             jobs[i].tree->tr.t.hmax * std::ldexp(1.0, k - (kMfClasses - 1)));
    if (fp32)
        KD_HIP(hipMemcpyAsync(dmfc.p, hmfc.data(), sizeof(MfClass4) * hmfc.size(), hipMemcpyHostToDevice, s));
    for (int i = 0; i < nj; i++) {
        const kd_device_tree *T = jobs[i].tree, *Q = jobs[i].query;
        KD_REQUIRE(!T->query_only, "weighted pair count: the tree side must be a built tree");
        JobDesc &J = hw[i].b;
        memset(&J, 0, sizeof(J));
        J.qx = Q->q.x.p;
        J.qy = Q->q.y.p;
        J.qz = Q->q.z.p;
        J.qf = Q->q.f.p;
        J.qstart = Q->q.t.start.p;
        J.qbox = Q->q.t.box.p;
        J.qfmin = Q->q.fmin.p;
        J.qfmax = Q->q.fmax.p;
        J.nqt = Q->q.t.ntiles;
        J.tp = T->tr.pts.p;
        J.tquad = T->tr.quad.p;
        J.tquad_start = T->tr.quad_start.p;
        J.tquad_meta = T->tr.quad_meta.p;
        J.tquad_box = T->tr.quad_box.p;
        J.tmeta = T->tr.meta.p;
        J.tbox = T->tr.t.box.p;
        J.tstart = T->tr.t.start.p;
        J.hbox = T->tr.t.hbox.p;
        J.P = T->tr.t.P;
        J.ntt = T->tr.t.ntiles;
        J.tcentre = T->tr.centre.p;
        J.tquad32 = T->tr.quad32.p;
        if (!jobs[i].tree_weights) {  // k_wq takes whole subtrees too; k_wm takes tiles only
            J.hfield = T->tr.hfield.p;
            J.hweight = T->tr.hweight.p;
            J.hfw = T->tr.hfw.p;
            J.nfw = T->tr.nfw;
        }
        J.symmetric = 0;
        J.relax = jobs[i].tree_weights ? 0 : 1;  // k_wq narrows windows per quad, k_wm does not
        J.nq_fields = jobs[i].nq_fields;
        J.nslots = jobs[i].nq_fields * jobs[i].nt_fields;
        hj[i] = J;
        hw[i].nq = Q->n_raw;
        hw[i].spad = Spad;
        hw[i].mfc4 = fp32 ? dmfc.p + (size_t)i * kMfClasses : nullptr;
        hw[i].whist_base = hist_len;
        hist_len += (long long)(jobs[i].tree_weights ? Spad : S) * J.nslots * rowlen;
        qt_prefix[i + 1] = qt_prefix[i] + J.nqt;
        hw[i].wq = nullptr;
        if (jobs[i].query_weights) {
            const size_t nb = (size_t)S * Q->n_raw;
            const unsigned char *&g = gathered[std::make_pair((const void *)jobs[i].query_weights, (const void *)Q)];
            if (!g) {
                const unsigned char *raw = upload(jobs[i].query_weights, nb);
                const size_t nrow = (size_t)Q->n_raw * Spad;
                DevBuf<unsigned char> wT(nrow);
                hipLaunchKernelGGL(k_transpose_w, dim3((Q->n_raw + 63) / 64, (Spad + 63) / 64), dim3(256), 0, s, raw,
                                   Q->n_raw, S, Spad, wT.p);
                wq_tile[i].alloc(nrow);
                hipLaunchKernelGGL(k_gather_rows, dim3(nblk((long long)Q->n_raw * (Spad / 16))), dim3(256), 0, s,
                                   wT.p, Q->q.order.p, Q->n_raw, Spad, wq_tile[i].p);
                g = wq_tile[i].p;
            }
            hw[i].wq = g;
        }
        if (jobs[i].tree_weights) {
            KD_REQUIRE(T->n_u == T->n_raw, "weighted pair count: tree-side weights need a tree without duplicate points");
            tw[i] = 1;
            any_tw = true;
        } else {
            any_qonly = true;
        }
    }

    // layouts of the tree-weighted jobs for the matrix-core kernel
    struct WMBufs {
        DevBuf<int> kstart, kstile;
        DevBuf<double> px, py, pz;
        DevBuf<float4> p32;
        DevBuf<unsigned> pmeta, wt;
        DevBuf<unsigned short> wb;
        DevBuf<unsigned char> wq, wb8;
    };
    std::vector<WMBufs> wmb(nj);
    std::vector<WMJob> hm(nj);
    // int8 matrix-core operands when every tree multiplicity of the
    // tree-weighted jobs is below 128, else bf16
    bool wm_i8 = true;
    if (any_tw) {
        DevBuf<unsigned> dbig(1);
        KD_HIP(hipMemsetAsync(dbig.p, 0, sizeof(unsigned), s));
        for (int i = 0; i < nj; i++) {
            if (!tw[i]) continue;
            const long long nb = (long long)S * jobs[i].tree->n_raw;
            hipLaunchKernelGGL(k_any_big, dim3((unsigned)std::min<long long>(nblk(nb / 16 + 1), 2048)), dim3(256), 0,
                               s, upload(jobs[i].tree_weights, (size_t)nb), nb, dbig.p);
        }
        unsigned hbig = 0u;
        KD_HIP(hipMemcpyAsync(&hbig, dbig.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        KD_HIP(hipStreamSynchronize(s));
        wm_i8 = hbig == 0u;
    }
    for (int i = 0; i < nj; i++) {
        memset(&hm[i], 0, sizeof(WMJob));
        hm[i].b = hw[i].b;
        if (!tw[i]) continue;
        const kd_device_tree *T = jobs[i].tree, *Q = jobs[i].query;
        WMBufs &B = wmb[i];
        const int ntt = T->tr.t.ntiles, nqt = Q->q.t.ntiles;
        DevBuf<int> nk(ntt);
        B.kstart.alloc(ntt + 1);
        hipLaunchKernelGGL(k_tile_ksteps, dim3(nblk(ntt)), dim3(256), 0, s, T->tr.t.start.p, ntt, nk.p);
        KD_HIP(hipMemsetAsync(B.kstart.p, 0, sizeof(int), s));
        so.incl_scan(nk.p, B.kstart.p + 1, ntt, s);
        int nks = 0;
        KD_HIP(hipMemcpyAsync(&nks, B.kstart.p + ntt, sizeof(int), hipMemcpyDeviceToHost, s));
        KD_HIP(hipStreamSynchronize(s));
        const long long nslot = 16LL * nks;
        B.px.alloc(nslot);
        B.py.alloc(nslot);
        B.pz.alloc(nslot);
        B.pmeta.alloc(nslot);
        DevBuf<int> pidx(nslot);
        if (fp32) {
            B.p32.alloc(nslot);
            B.kstile.alloc(nks);
        }
        hipLaunchKernelGGL(k_fill_slots, dim3(nblk(nslot)), dim3(256), 0, s, T->tr.pts.p, T->tr.t.start.p,
                           B.kstart.p, ntt, nks, B.px.p, B.py.p, B.pz.p, B.pmeta.p, pidx.p,
                           fp32 ? (const double *)T->tr.centre.p : nullptr, B.p32.p, B.kstile.p);
        const size_t nb = (size_t)S * T->n_raw;
        const unsigned char *raw = upload(jobs[i].tree_weights, nb);
        DevBuf<unsigned char> wT((size_t)T->n_raw * Spad);  // point-major, zero-padded sets
        hipLaunchKernelGGL(k_transpose_w, dim3((T->n_raw + 63) / 64, Spad / 64 + (Spad % 64 ? 1 : 0)), dim3(256), 0, s,
                           raw, T->n_raw, S, Spad, wT.p);
        if (wm_i8) {
            B.wb8.alloc((size_t)nslot * Spad);
            hipLaunchKernelGGL(k_fill_wb<true>, dim3(nblk(nslot)), dim3(256), 0, s, wT.p, T->tr.raw.p, pidx.p, nks,
                               Spad, (void *)B.wb8.p);
        } else {
            B.wb.alloc((size_t)nslot * Spad);
            hipLaunchKernelGGL(k_fill_wb<false>, dim3(nblk(nslot)), dim3(256), 0, s, wT.p, T->tr.raw.p, pidx.p, nks,
                               Spad, (void *)B.wb.p);
        }
        B.wt.alloc((size_t)ntt * Spad);
        hipLaunchKernelGGL(k_tile_wsum, dim3(nblk((long long)ntt * (Spad / 4))), dim3(256), 0, s, wT.p, T->tr.raw.p,
                           T->tr.t.start.p, ntt, Spad, B.wt.p);
        B.wq.alloc((size_t)nqt * Spad * 64);
        hipLaunchKernelGGL(k_fill_wq64, dim3(nblk((long long)nqt * Spad * 64)), dim3(256), 0, s, hw[i].wq,
                           Q->q.t.start.p, nqt, Q->n_raw, S, Spad, B.wq.p);
        KD_HIP(hipGetLastError());
        KD_HIP(hipStreamSynchronize(s));
        WMJob &M = hm[i];
        M.kstart = B.kstart.p;
        M.px = B.px.p;
        M.py = B.py.p;
        M.pz = B.pz.p;
        M.pmeta = B.pmeta.p;
        M.wb = (const uint4 *)B.wb.p;
        M.wb8 = (const uint4 *)B.wb8.p;
        M.wt = B.wt.p;
        M.wq = (const unsigned *)B.wq.p;
        M.whist_base = hw[i].whist_base;
        M.on = 1;
        M.p32 = (const floatx4 *)B.p32.p;
        M.kstile = B.kstile.p;
        M.mfc4 = hw[i].mfc4;
    }
    for (int i = 0; i < nj; i++) {
        if (jobs[i].tree_field_sizes)
            field_sizes(jobs[i].tree,
                        jobs[i].tree_weights ? upload(jobs[i].tree_weights, (size_t)S * jobs[i].tree->n_raw) : nullptr,
                        S, jobs[i].nt_fields, jobs[i].tree_field_sizes, s);
        if (jobs[i].query_field_sizes)
            field_sizes(jobs[i].query,
                        jobs[i].query_weights ? upload(jobs[i].query_weights, (size_t)S * jobs[i].query->n_raw)
                                              : nullptr,
                        S, jobs[i].nq_fields, jobs[i].query_field_sizes, s);
    }
    const long long total_qt = qt_prefix[nj];
    DevBuf<JobDesc> djobs(nj);
    DevBuf<WJobDesc> dw(nj);
    DevBuf<long long> dprefix(nj + 1);
    KD_HIP(hipMemcpyAsync(djobs.p, hj.data(), sizeof(JobDesc) * nj, hipMemcpyHostToDevice, s));
    KD_HIP(hipMemcpyAsync(dw.p, hw.data(), sizeof(WJobDesc) * nj, hipMemcpyHostToDevice, s));
    KD_HIP(hipMemcpyAsync(dprefix.p, qt_prefix.data(), sizeof(long long) * (nj + 1), hipMemcpyHostToDevice, s));

    // enumeration (runs of pure tiles merge into one item)
    hipEvent_t ev0, ev1, ev2;
    KD_HIP(hipEventCreate(&ev0));
    KD_HIP(hipEventCreate(&ev1));
    KD_HIP(hipEventCreate(&ev2));
    KD_HIP(hipEventRecord(ev0, s));
    int maxP = 1;
    for (int i = 0; i < nj; i++) maxP = std::max(maxP, hj[i].P);
    ItemList il;
    // (FP32 form: runs of tiles of one size class, whose thresholds the run uses)
    enumerate_items(djobs.p, dprefix.p, nj, total_qt, maxP, E, fp32 ? 4 : 3, 1, 0, so, s, il);
    const long long nitems = il.nitems;
    DevBuf<long long> hist(std::max<long long>(hist_len, 1));
    DevBuf<unsigned long long> dnext(kXcds * kNextStride), dstats(4);
    DevBuf<unsigned> derr(1);
    KD_HIP(hipMemsetAsync(dstats.p, 0, dstats.bytes(), s));
    KD_HIP(hipMemsetAsync(derr.p, 0, derr.bytes(), s));
    // no trim pass: k_wq's quad windows skip the quads beyond every radius,
    // and k_wm works on whole tiles
    KD_HIP(hipEventRecord(ev1, s));

    KD_HIP(hipMemsetAsync(hist.p, 0, hist.bytes(), s));
    KD_HIP(hipMemsetAsync(dnext.p, 0, dnext.bytes(), s));
    int dev = 0, ncu = 0;
    KD_HIP(hipGetDevice(&dev));
    KD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (nitems > 0 && any_qonly) {
        // the wave-per-query-tile kernel covers the jobs without tree weights
        std::vector<WJobDesc> hq(hw);
        // query tiles of tree-weighted jobs are skipped inside the kernel
        for (int i = 0; i < nj; i++)
            if (tw[i]) hq[i].b.nqt = -1;
        DevBuf<WJobDesc> dq(nj);
        KD_HIP(hipMemcpyAsync(dq.p, hq.data(), sizeof(WJobDesc) * nj, hipMemcpyHostToDevice, s));
        const size_t lds = (size_t)4 * 2 * (ne + 1) * 64 * sizeof(unsigned short);
        const void *kq = fp32 ? (const void *)k_wq<true> : (const void *)k_wq<false>;
        KD_HIP(hipFuncSetAttribute(kq, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        if (fp32)
            hipLaunchKernelGGL(k_wq<true>, dim3(ncu * 7), dim3(256), lds, s, dq.p, dprefix.p, nj, total_qt,
                               il.items.p, il.offs.p, E, S, dnext.p, hist.p, dstats.p, derr.p);
        else
            hipLaunchKernelGGL(k_wq<false>, dim3(ncu * 7), dim3(256), lds, s, dq.p, dprefix.p, nj, total_qt,
                               il.items.p, il.offs.p, E, S, dnext.p, hist.p, dstats.p, derr.p);
        KD_HIP(hipGetLastError());
        KD_HIP(hipStreamSynchronize(s));
    }
    if (nitems > 0 && any_tw) {
        // the matrix-core kernel covers the jobs with tree weights (the others are off there)
        DevBuf<WMJob> dm(nj);
        KD_HIP(hipMemcpyAsync(dm.p, hm.data(), sizeof(WMJob) * nj, hipMemcpyHostToDevice, s));
        // the query-tile ranges of the tree-weighted jobs
        std::vector<long long> granges;
        long long tiles = 0;
        for (int i = 0; i < nj; i++)
            if (tw[i] && qt_prefix[i + 1] > qt_prefix[i]) {
                if (!granges.empty() && granges.back() == qt_prefix[i])
                    granges.back() = qt_prefix[i + 1];
                else {
                    granges.push_back(qt_prefix[i]);
                    granges.push_back(qt_prefix[i + 1]);
                }
                tiles += qt_prefix[i + 1] - qt_prefix[i];
            }
        DevBuf<long long> dgr(std::max<size_t>(granges.size(), 1));
        if (!granges.empty())
            KD_HIP(hipMemcpyAsync(dgr.p, granges.data(), sizeof(long long) * granges.size(), hipMemcpyHostToDevice,
                                  s));
        // one wave per query tile, at most 16 workgroups of 4 waves per CU,
        // a multiple of kXcds workgroups
        const int nblocks = (int)std::max<long long>(
                                1, std::min<long long>((tiles + 4 * kXcds - 1) / (4 * kXcds), (long long)ncu * 16 / kXcds)) *
                            kXcds;
        auto km = fp32 ? (wm_i8 ? k_wm<true, true> : k_wm<true, false>)
                       : (wm_i8 ? k_wm<false, true> : k_wm<false, false>);
        hipLaunchKernelGGL(km, dim3(nblocks), dim3(256), 0, s, dm.p, dprefix.p, nj, dgr.p, (int)(granges.size() / 2),
                           tiles, il.items.p, il.offs.p, E, S, Spad, hist.p, dstats.p, derr.p);
        KD_HIP(hipGetLastError());
        KD_HIP(hipStreamSynchronize(s));  // before dgr goes
    }
    KD_HIP(hipEventRecord(ev2, s));
    // cumulative counts per original radius, made on the device and copied
    // straight into the caller's arrays
    std::vector<int> emap(nr);
    for (int i = 0; i < nr; i++)
        emap[i] = (int)(std::lower_bound(edges.begin(), edges.end(), radii[i]) - edges.begin());
    DevBuf<int> demap(nr);
    KD_HIP(hipMemcpyAsync(demap.p, emap.data(), sizeof(int) * nr, hipMemcpyHostToDevice, s));
    std::vector<DevBuf<long long>> dout(nj);
    for (int jb = 0; jb < nj; jb++) {
        const long long n = (long long)S * nr * hj[jb].nslots;
        dout[jb].alloc(std::max<long long>(n, 1));
        hipLaunchKernelGGL(k_weighted_cum, dim3(nblk((long long)S * hj[jb].nslots)), dim3(256), 0, s,
                           hist.p + hw[jb].whist_base, S, hj[jb].nslots, ne, rowlen, Spad, (int)tw[jb], demap.p, nr,
                           dout[jb].p);
        KD_HIP(hipMemcpyAsync(counts_out[jb], dout[jb].p, sizeof(long long) * n, hipMemcpyDeviceToHost, s));
    }
    KD_HIP(hipGetLastError());
    unsigned long long hstats[4];
    unsigned herr = 0;
    KD_HIP(hipMemcpyAsync(hstats, dstats.p, sizeof(hstats), hipMemcpyDeviceToHost, s));
    KD_HIP(hipMemcpyAsync(&herr, derr.p, sizeof(herr), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    float ms_enum = 0, ms_pairs = 0;
    KD_HIP(hipEventElapsedTime(&ms_enum, ev0, ev1));
    KD_HIP(hipEventElapsedTime(&ms_pairs, ev1, ev2));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    (void)hipEventDestroy(ev2);
    KD_REQUIRE(!(herr & kErrItem), "weighted pair count: internal error: a work item lies outside its job's tiles");
    KD_REQUIRE(!herr, "weighted pair count: a field id gives a count-matrix index outside [0, Nq*Nt)");
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->evaluated_pairs = (long long)hstats[0];
        stats->rechecked_pairs = (long long)hstats[1];
        stats->trimmed_items = (long long)hstats[2];
        stats->fp32_pairs = (long long)hstats[3];
        stats->work_items = nitems;
        stats->enumerate_ms = ms_enum;
        stats->count_ms = ms_pairs;
        stats->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    }
}

}  // namespace kd
