// GPU build of corruscant's balanced k-d tree, node array bit-identical to
// the reference's tree_construct (src/kdbuild.c:451-537).
//
// Steps (all on the device):
//   1. lexicographic (x,y,z) order of the raw points: three stable LSD radix
//      passes over order-preserving 64-bit keys, with -0.0 keyed as +0.0 so
//      that the IEEE == of src/kdbuild.c:82-96 decides equality;
//   2. duplicate runs -> one unique point each, carrying the lowest original
//      index's value and field, weight = run length (src/kdbuild.c:484-511);
//   3. the three cyclic-superkey argsorts (src/kdbuild.c:221-243) and their
//      rank arrays;
//   4. a level-synchronous median build.  The segment [l, r] of every heap
//      node depends only on the number of unique points, so each level is one
//      pass: split dimension by maximum variance (src/kdbuild.c:308-351),
//      node records, then a stable partition of the other two argsort arrays
//      (src/kdbuild.c:251-291) done as a segmented prefix sum of
//      rank-compare flags.
//
// Split-dimension decisions reproduce the reference's sequential FP64 sums.
// Segments of up to kSeqMax points replay the sequential arithmetic exactly
// (one thread per segment and dimension).  Larger segments use a parallel sum
// whose error bound is compared with the bound on the reference's sequential
// rounding: the parallel decision is kept only when it is certain; otherwise
// the segment falls back to the exact sequential replay.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "kd_internal.h"

namespace kd {

static constexpr int kSeqMax = 1024;
static constexpr int kChunk = 16384;

__device__ __forceinline__ uint64_t ordkey(double v)
{
    if (v == 0.0) v = 0.0;  // one key for -0.0 and +0.0
    uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__global__ void k_scan_input(const double *__restrict__ xyz, size_t n3, unsigned *__restrict__ nanflag,
                             unsigned long long *__restrict__ amax)
{
    unsigned long long m = 0;
    unsigned bad = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        double v = xyz[i];
        if (v != v) bad = 1;
        unsigned long long b = (unsigned long long)__double_as_longlong(fabs(v));
        m = b > m ? b : m;
    }
    if (bad) atomicOr(nanflag, 1u);
    // wave max then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_down(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(amax, m);
}

__global__ void k_fill_int(int *a, int n, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_iota(int *a, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// keys[j] = ordkey(v[stride * idx[j] + off])
__global__ void k_gather_key(const double *__restrict__ v, int stride, int off, const int *__restrict__ idx,
                             uint64_t *__restrict__ keys, int n)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int i = idx ? idx[j] : j;
    keys[j] = ordkey(v[(size_t)stride * i + off]);
}

// *tie = 1 when two neighbours of the sorted keys are equal
__global__ void k_any_tie(const uint64_t *__restrict__ keys, int n, unsigned *__restrict__ tie)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= 1 && j < n && keys[j] == keys[j - 1]) *tie = 1u;
}

__global__ void k_run_flags(const double *__restrict__ xyz, const int *__restrict__ order, int n, int *__restrict__ flag)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int f = 1;
    if (j > 0) {
        const double *a = xyz + 3 * (size_t)order[j];
        const double *b = xyz + 3 * (size_t)order[j - 1];
        f = !(a[0] == b[0] && a[1] == b[1] && a[2] == b[2]);
    }
    flag[j] = f;
}

__global__ void k_unique(const double *__restrict__ xyz, const int *__restrict__ fields, const int *__restrict__ order,
                         const int *__restrict__ flag, const int *__restrict__ incl, int n, double *__restrict__ ux,
                         double *__restrict__ uy, double *__restrict__ uz, int *__restrict__ uf,
                         int *__restrict__ runpos, int *__restrict__ urep)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || !flag[j]) return;
    int u = incl[j] - 1;
    int rep = order[j];  // lowest original index of the run (the passes are stable)
    ux[u] = xyz[3 * (size_t)rep];
    uy[u] = xyz[3 * (size_t)rep + 1];
    uz[u] = xyz[3 * (size_t)rep + 2];
    uf[u] = fields ? (fields[rep] & 0xFF) : 0;
    runpos[u] = j;
    urep[u] = rep;
}

__global__ void k_weights(const int *__restrict__ runpos, int nu, int n, int *__restrict__ uw, unsigned *__restrict__ err)
{
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    int len = (u + 1 < nu ? runpos[u + 1] : n) - runpos[u];
    if (len > 255) atomicOr(err, 1u);
    uw[u] = len & 0xFF;
}

__global__ void k_rank(const int *__restrict__ args, int n, int *__restrict__ rank)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rank[args[i]] = i;
}

// segments of the heap nodes of one level, from their parents
// Segments [l, r] of the level's nodes lo..hi-1 from their parents'; also
// clears the level's accumulators of the large-segment sums (nacc doubles)
// and its per-node flags.
__global__ void k_level_segs(int lo, int hi, int nu, int *__restrict__ seg_l, int *__restrict__ seg_r,
                             double *__restrict__ acc, int nacc, unsigned char *__restrict__ flagged)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (int k = tid; k < nacc; k += gridDim.x * blockDim.x) acc[k] = 0.0;
    int o = lo + tid;
    if (o >= hi) return;
    flagged[o - lo] = 0;
    if (o == 1) {
        seg_l[1] = 0;
        seg_r[1] = nu - 1;
        return;
    }
    int p = o >> 1;
    int l = seg_l[p], r = seg_r[p];
    int L = -1, R = -1;
    if (l >= 0) {
        int s = r - l + 1, med = (l + r + 1) / 2;
        if ((o & 1) == 0) {
            if (s >= 2) { L = l; R = med - 1; }
        } else {
            if (s >= 3) { L = med + 1; R = r; }
        }
    }
    seg_l[o] = L;
    seg_r[o] = R;
}

struct UPtrs {
    const double *v[3];
    const int *a[3];
    const double *c[3];  // c[d][i] = v[d][a[d][i]]: the coordinates in args order
};

// Exact replay of src/kdbuild.c:318-338 for one (node, dim): sequential sum
// in args order, division by the size, sequential fma variance.
__global__ void k_var_seq(int lo, int hi, const int *__restrict__ seg_l, const int *__restrict__ seg_r, int smin,
                          int smax, const unsigned char *__restrict__ only_flagged, UPtrs P, double *__restrict__ var3)
{
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    int o = lo + tid / 3, d = tid % 3;
    if (o >= hi) return;
    int l = seg_l[o], r = seg_r[o];
    if (l < 0) return;
    int s = r - l + 1;
    if (s < smin || s > smax) return;
    if (only_flagged && !only_flagged[o - lo]) return;
    const double *__restrict__ c = P.c[d];
    double sum = 0.0;
    int i = l;
    for (; i + 8 <= r + 1; i += 8) {
        double t[8];
#pragma unroll
        for (int k = 0; k < 8; k++) t[k] = c[i + k];
#pragma unroll
        for (int k = 0; k < 8; k++) sum += t[k];
    }
    for (; i <= r; i++) sum += c[i];
    double avg = sum / (double)s;
    double var = 0.0;
    i = l;
    for (; i + 8 <= r + 1; i += 8) {
        double t[8];
#pragma unroll
        for (int k = 0; k < 8; k++) t[k] = c[i + k];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            double e = t[k] - avg;
            var = fma(e, e, var);
        }
    }
    for (; i <= r; i++) {
        double e = c[i] - avg;
        var = fma(e, e, var);
    }
    var3[(size_t)(o - lo) * 3 + d] = var;
}

// strict-> argmax with dim 0 as the default (src/kdbuild.c:340-348)
__global__ void k_pick_tdim(int lo, int hi, const int *__restrict__ seg_l, const int *__restrict__ seg_r, int smin,
                            int smax, const unsigned char *__restrict__ only_flagged, const double *__restrict__ var3,
                            unsigned char *__restrict__ tdim)
{
    int o = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= hi) return;
    int l = seg_l[o], r = seg_r[o];
    if (l < 0) return;
    int s = r - l + 1;
    if (s < 4) {
        tdim[o] = 0;
        return;
    }
    if (s < smin || s > smax) return;
    if (only_flagged && !only_flagged[o - lo]) return;
    const double *w = var3 + (size_t)(o - lo) * 3;
    int dim = 0;
    double best = w[0];
    for (int j = 1; j < 3; j++)
        if (w[j] > best) {
            best = w[j];
            dim = j;
        }
    tdim[o] = (unsigned char)dim;
}

// ---- certified parallel variance for large segments ----------------------

struct BigNode {
    int o, l, r;
};

// acc layout per big node: [d][0] sum, [d][1] sum |v|, [d][2] sum (v-m)^2
__global__ void k_big_sum(const BigNode *__restrict__ nodes, const int *__restrict__ chunk_node,
                          const int *__restrict__ chunk_off, UPtrs P, double *__restrict__ acc, int pass,
                          const double *__restrict__ mean)
{
    int c = blockIdx.x;
    int bn = chunk_node[c];
    BigNode nd = nodes[bn];
    int b0 = nd.l + chunk_off[c] * kChunk;
    int b1 = min(nd.r + 1, b0 + kChunk);
    __shared__ double red[3][2][4];
    for (int d = 0; d < 3; d++) {
        const double *__restrict__ c = P.c[d];
        double s0 = 0.0, s1 = 0.0;
        double m = pass ? mean[bn * 3 + d] : 0.0;
        for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
            double x = c[i];
            if (pass) {
                double e = x - m;
                s0 += e * e;
            } else {
                s0 += x;
                s1 += fabs(x);
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            s0 += __shfl_down(s0, off, 64);
            s1 += __shfl_down(s1, off, 64);
        }
        int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            red[d][0][w] = s0;
            red[d][1][w] = s1;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int d = threadIdx.x;
        int nw = blockDim.x >> 6;
        double s0 = 0.0, s1 = 0.0;
        for (int w = 0; w < nw; w++) {
            s0 += red[d][0][w];
            s1 += red[d][1][w];
        }
        if (pass) {
            atomicAdd(&acc[bn * 9 + d * 3 + 2], s0);
        } else {
            atomicAdd(&acc[bn * 9 + d * 3 + 0], s0);
            atomicAdd(&acc[bn * 9 + d * 3 + 1], s1);
        }
    }
}

__global__ void k_big_mean(const BigNode *__restrict__ nodes, int nbig, const double *__restrict__ acc,
                           double *__restrict__ mean)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbig * 3) return;
    int bn = t / 3, d = t % 3;
    int s = nodes[bn].r - nodes[bn].l + 1;
    mean[t] = acc[bn * 9 + d * 3] / (double)s;
}

// Decide the split dimension of a large segment when the bounds allow it.
//   u = 2^-53, s points, A = sum |v|, S, V = parallel sum and sum of squares
//   (any order), m = S/s.  With G = (s+16)u/(1-(s+16)u):
//     |S_any - S_exact| <= G*A, so the exact mean and the reference's rounded
//     mean both lie within delta = 2*(G+u)*A/s of m;
//     V(mean) in [V/(1+G) - s*delta^2, V/(1-G)]  (V(x) = V(mean) + s(x-mean)^2);
//     the reference's sequential value lies in
//     [V(mean)(1-u)^2(1-G), (V(mean) + s*delta^2)(1+u)^2(1+G)].
//   The parallel argmax is kept when its lower bound exceeds every other
//   dimension's upper bound; otherwise the node is flagged for the replay.
__global__ void k_big_decide(const BigNode *__restrict__ nodes, int nbig, const double *__restrict__ acc,
                             unsigned char *__restrict__ tdim, unsigned char *__restrict__ flagged, int lo)
{
    int bn = blockIdx.x * blockDim.x + threadIdx.x;
    if (bn >= nbig) return;
    const double u = 1.1102230246251565e-16;
    BigNode nd = nodes[bn];
    double s = (double)(nd.r - nd.l + 1);
    double g = (s + 16.0) * u;
    double G = g / (1.0 - g);
    double lo_b[3], hi_b[3], mid[3];
    for (int d = 0; d < 3; d++) {
        double A = acc[bn * 9 + d * 3 + 1] * (1.0 + 2.0 * G);
        double V = acc[bn * 9 + d * 3 + 2];
        double delta = 2.0 * (G + u) * A / s;
        double sd2 = s * delta * delta;
        double vlo = V / (1.0 + G) - sd2;
        double vhi = V / (1.0 - G);
        double rl = vlo * (1.0 - u) * (1.0 - u) * (1.0 - G);
        double rh = (vhi + sd2) * (1.0 + u) * (1.0 + u) * (1.0 + G);
        lo_b[d] = rl - 8.0 * u * fabs(rl);
        hi_b[d] = rh + 8.0 * u * fabs(rh);
        mid[d] = V;
    }
    int best = 0;
    for (int d = 1; d < 3; d++)
        if (mid[d] > mid[best]) best = d;
    bool ok = true;
    for (int d = 0; d < 3; d++)
        if (d != best && !(lo_b[best] > hi_b[d])) ok = false;
    if (ok) {
        tdim[nd.o] = (unsigned char)best;
        flagged[nd.o - lo] = 0;
    } else {
        flagged[nd.o - lo] = 1;
    }
}

// ---- node records, partition ---------------------------------------------

__global__ void k_write_nodes(int lo, int hi, const int *__restrict__ seg_l, const int *__restrict__ seg_r,
                              const unsigned char *__restrict__ tdim, UPtrs P, const int *__restrict__ uf,
                              const int *__restrict__ uw, unsigned long long *__restrict__ nodes,
                              double *__restrict__ split, uint32_t *__restrict__ nbits, int *__restrict__ perm,
                              int *__restrict__ inorder, int *__restrict__ heap_of)
{
    int o = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= hi) return;
    int l = seg_l[o], r = seg_r[o];
    if (l < 0) return;
    int s = r - l + 1, med = (l + r + 1) / 2;
    int t = tdim[o];
    int u = P.a[t][med];
    double x = P.v[0][u], y = P.v[1][u], z = P.v[2][u];
    uint32_t bits = (uint32_t)(s >= 2) | ((uint32_t)(s >= 3) << 1) | ((uint32_t)t << 2);
    unsigned long long *rec = nodes + (size_t)o * 5;
    rec[0] = (unsigned long long)__double_as_longlong(x);
    rec[1] = (unsigned long long)__double_as_longlong(y);
    rec[2] = (unsigned long long)__double_as_longlong(z);
    rec[3] = (unsigned long long)(uf[u] & 0xFF) | ((unsigned long long)(uw[u] & 0xFF) << 8);
    rec[4] = bits;
    split[o] = t == 0 ? x : (t == 1 ? y : z);
    nbits[o] = bits;
    perm[o] = u;
    inorder[o] = med;
    heap_of[u] = o;
}

struct RPtrs {
    const int *rank[3];  // rank[0] == nullptr means identity
};

__device__ __forceinline__ int rank_of(const RPtrs &R, int t, int u) { return R.rank[t] ? R.rank[t][u] : u; }

// The ranks that travel with the argsort arrays of the global levels: an
// element of array d is compared with the median of a split dimension t != d
// by its rank in order t, which is its index for t = 0 (the unique points
// are in (x,y,z) order) and otherwise one of these four arrays, moved with
// the indices by every partition so the flags read them contiguously:
//   r[0], r[1]: ranks in orders 1 and 2 of the elements of array 0;
//   r[2]: rank in order 2 of the elements of array 1;
//   r[3]: rank in order 1 of the elements of array 2.
struct CarriedRanks {
    int *r[4];
};

__device__ __forceinline__ int carried_rank(const CarriedRanks &C, const UPtrs &P, int d, int t, int i)
{
    if (t == 0) return P.a[d][i];
    if (d == 0) return C.r[t == 1 ? 0 : 1][i];
    return C.r[d == 1 ? 2 : 3][i];
}

__global__ void k_carry_ranks(const int *__restrict__ args1, const int *__restrict__ args2,
                              const int *__restrict__ rank1, const int *__restrict__ rank2, int nu, CarriedRanks C)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu) return;
    C.r[0][i] = rank1[i];  // array 0 is the identity
    C.r[1][i] = rank2[i];
    C.r[2][i] = rank2[args1[i]];
    C.r[3][i] = rank1[args2[i]];
}

// The stable partition of one level (src/kdbuild.c:251-291) for all three
// argsort arrays at once.  Element i of array d (in a segment of at least four
// points whose split dimension t is not d) is below or above the median by
// its rank in order t; flags[d * nu + i] holds 1 (below), 1 << 31 (the
// median) or 0 (above).  One exclusive 32-bit scan over the three arrays
// together then gives each element its place: the scatter uses differences
// of the scan inside one segment of one array (modulo 2^32, where the carry
// of everything before the segment cancels): the low 31 bits count the
// elements below the median before it, the top bit whether the median is
// among them, and the rest of the segment before it lies above.
__global__ void k_part_flags3(int nu, const int *__restrict__ owner, const int *__restrict__ seg_l,
                              const int *__restrict__ seg_r, const unsigned char *__restrict__ tdim, UPtrs P,
                              RPtrs R, CarriedRanks C, unsigned *__restrict__ flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu) return;
    const int o = owner[i];
    unsigned f[3] = {0u, 0u, 0u};
    if (o) {
        const int l = seg_l[o], r = seg_r[o];
        const int t = tdim[o];
        if (r - l + 1 >= 4) {
            const int med = (l + r + 1) / 2;
            const int rm = rank_of(R, t, P.a[t][med]);  // one address per segment
#pragma unroll
            for (int d = 0; d < 3; d++) {
                if (d == t) continue;
                const int rx = carried_rank(C, P, d, t, i);
                f[d] = rx < rm ? 1u : (rx > rm ? 0u : 0x80000000u);
            }
        }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) flags[(size_t)d * nu + i] = f[d];
}

// Scatter of the three arrays (indices with their coordinates and carried
// ranks) to their places in the children's segments, and the owner of every
// position for the next level (the median's position leaves the build).
__global__ void k_part_scatter3(int nu, int *__restrict__ owner, const int *__restrict__ seg_l,
                                const int *__restrict__ seg_r, const unsigned char *__restrict__ tdim, UPtrs P,
                                CarriedRanks C, const unsigned *__restrict__ flags, const unsigned *__restrict__ scan,
                                int *__restrict__ out0, int *__restrict__ out1, int *__restrict__ out2,
                                double *__restrict__ cout0, double *__restrict__ cout1, double *__restrict__ cout2,
                                CarriedRanks Cout)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu) return;
    const int o = owner[i];
    int l = 0, r = -1, t = 0, med = 0;
    if (o) {
        l = seg_l[o];
        r = seg_r[o];
        t = tdim[o];
        med = (l + r + 1) / 2;
    }
    int *outs[3] = {out0, out1, out2};
    double *couts[3] = {cout0, cout1, cout2};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int x = P.a[d][i];
        const double cx = P.c[d][i];  // the coordinate travels with its index
        int dst = i;
        if (o && r - l + 1 >= 4 && t != d) {
            const size_t off = (size_t)d * nu;
            const unsigned diff = scan[off + i] - scan[off + l], f = flags[off + i];
            const int below = (int)(diff & 0x7fffffffu), median = (int)(diff >> 31);
            if (f == 1u) dst = l + below;                                   // below the median: [l, med-1]
            else if (f == 0u) dst = med + 1 + ((i - l) - below - median);   // above: [med+1, r]
            else dst = med;  // the median itself; this slot is never read again
        }
        outs[d][dst] = x;
        couts[d][dst] = cx;
        if (d == 0) {
            Cout.r[0][dst] = C.r[0][i];
            Cout.r[1][dst] = C.r[1][i];
        } else {
            Cout.r[d + 1][dst] = C.r[d + 1][i];
        }
    }
    if (o) owner[i] = i < med ? 2 * o : (i > med ? 2 * o + 1 : 0);
}

__global__ void k_gather_vals(const double *__restrict__ v, const int *__restrict__ a, int n, double *__restrict__ out)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = v[a[j]];
}

// The bottom of the tree in one launch: once no segment of a level holds
// more than kSubMax points, each workgroup takes one segment of that level and
// builds its whole subtree in LDS, level by level, doing what the level
// kernels do for it:
//   - the variances of every segment of four points or more, one thread per
//     (segment, dimension), by the exact sequential replay of k_var_seq, and
//     the split dimension by the strict argmax of k_pick_tdim (dimension 0
//     below four points);
//   - the node records of k_write_nodes;
//   - the stable partition of the argsort arrays with their coordinates by
//     rank against the median in the split dimension's order, as
//     k_part_flags3 / k_part_scatter3 do it: per-element flags, one block-wide
//     exclusive scan over the subtree (below / above counts in the low / high
//     16 bits), and positions from the differences inside each segment.
// Each element finds its segment of the level by descending from the subtree
// root (a median met on the way has left the build).  The arrays of the
// subtree's range belong to it alone, and nothing reads them afterwards.
static constexpr int kSubMax = 512;
static constexpr int kSubThreads = 256;

// Segment [l, r] (relative to the subtree root's start) of node j of level
// k below a root of n0 points; l < 0 when the node is empty.
__device__ __forceinline__ void sub_segment(int n0, int k, int j, int &l, int &r)
{
    l = 0;
    r = n0 - 1;
    for (int b = k - 1; b >= 0 && l >= 0; b--) {
        const int sz = r - l + 1, med = (l + r + 1) / 2;
        if ((j >> b) & 1) {
            if (sz >= 3) l = med + 1;
            else l = -1;
        } else {
            if (sz >= 2) r = med - 1;
            else l = -1;
        }
    }
}

// Block-wide exclusive sum of v (kSubThreads threads); total in *tot.
__device__ __forceinline__ unsigned sub_excl(unsigned v, unsigned *wsum, unsigned &tot)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    unsigned inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned u = __shfl_up(inc, off, 64);
        if (lane >= off) inc += u;
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    unsigned before = 0;
    tot = 0;
#pragma unroll
    for (int w = 0; w < kSubThreads / 64; w++) {
        if (w < wid) before += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    return before + inc - v;
}

__global__ __launch_bounds__(kSubThreads) void k_subtrees(int lo, int hi, const int *__restrict__ seg_l,
                                                          const int *__restrict__ seg_r, UPtrs P, RPtrs R,
                                                          const int *__restrict__ uf, const int *__restrict__ uw,
                                                          unsigned long long *__restrict__ nodes,
                                                          double *__restrict__ split, uint32_t *__restrict__ nbits,
                                                          int *__restrict__ perm, int *__restrict__ inorder,
                                                          int *__restrict__ heap_of, int memsize,
                                                          unsigned *__restrict__ err)
{
    __shared__ int sa[3][kSubMax], ta[kSubMax];
    __shared__ double sc[3][kSubMax], tc[kSubMax];
    __shared__ double svar[kSubMax][3];
    __shared__ unsigned char st[kSubMax];  // split dimension of each segment of the level
    __shared__ short sseg[kSubMax];        // segment of each element at the level (-1: placed)
    __shared__ unsigned sscan[kSubMax];
    __shared__ unsigned wsum[kSubThreads / 64];
    const int o0 = lo + blockIdx.x;
    if (o0 >= hi) return;
    const int l0 = seg_l[o0], r0 = seg_r[o0];
    if (l0 < 0) return;
    const int n0 = r0 - l0 + 1;
    if (n0 > kSubMax) {  // excluded by the caller's level choice
        if (threadIdx.x == 0) atomicOr(err, 1u);
        return;
    }
    for (int d = 0; d < 3; d++)
        for (int i = threadIdx.x; i < n0; i += kSubThreads) {
            sa[d][i] = P.a[d][l0 + i];
            sc[d][i] = P.c[d][l0 + i];
        }
    __syncthreads();
    for (int k = 0; ((long long)o0 << k) < memsize; k++) {
        const int first = o0 << k;
        const int nseg = min(1 << k, kSubMax);
        // variances (segments of four points or more)
        for (int x = threadIdx.x; x < 3 * nseg; x += kSubThreads) {
            const int j = x / 3, d = x - 3 * j;
            int l, r;
            sub_segment(n0, k, j, l, r);
            if (l < 0 || r - l + 1 < 4) continue;
            const int sz = r - l + 1;
            double sum = 0.0;
            for (int i = l; i <= r; i++) sum += sc[d][i];
            const double avg = sum / (double)sz;
            double var = 0.0;
            for (int i = l; i <= r; i++) {
                const double e = sc[d][i] - avg;
                var = fma(e, e, var);
            }
            svar[j][d] = var;
        }
        __syncthreads();
        // split dimensions and node records
        int live = 0;
        for (int j = threadIdx.x; j < nseg; j += kSubThreads) {
            int l, r;
            sub_segment(n0, k, j, l, r);
            if (l < 0) continue;
            live = 1;
            const int o = first + j;
            const int sz = r - l + 1, med = (l + r + 1) / 2;
            int t = 0;
            if (sz >= 4) {
                double best = svar[j][0];
                for (int d = 1; d < 3; d++)
                    if (svar[j][d] > best) {
                        best = svar[j][d];
                        t = d;
                    }
            }
            st[j] = (unsigned char)t;
            const int u = sa[t][med];
            const double x = P.v[0][u], y = P.v[1][u], z = P.v[2][u];
            const uint32_t bits = (uint32_t)(sz >= 2) | ((uint32_t)(sz >= 3) << 1) | ((uint32_t)t << 2);
            unsigned long long *rec = nodes + (size_t)o * 5;
            rec[0] = (unsigned long long)__double_as_longlong(x);
            rec[1] = (unsigned long long)__double_as_longlong(y);
            rec[2] = (unsigned long long)__double_as_longlong(z);
            rec[3] = (unsigned long long)(uf[u] & 0xFF) | ((unsigned long long)(uw[u] & 0xFF) << 8);
            rec[4] = bits;
            split[o] = t == 0 ? x : (t == 1 ? y : z);
            nbits[o] = bits;
            perm[o] = u;
            inorder[o] = l0 + med;
            heap_of[u] = o;
        }
        if (!__syncthreads_or(live)) break;
        // the segment of every element at this level
        for (int i = threadIdx.x; i < n0; i += kSubThreads) {
            int l = 0, r = n0 - 1, j = 0, b = 0;
            for (; b < k; b++) {
                const int med = (l + r + 1) / 2;
                if (i == med) break;
                if (i < med) {
                    r = med - 1;
                    j = 2 * j;
                } else {
                    l = med + 1;
                    j = 2 * j + 1;
                }
            }
            sseg[i] = (short)(b < k ? -1 : j);
        }
        __syncthreads();
        // stable partition of each array whose order is not the split one
        for (int d = 0; d < 3; d++) {
            // two elements per thread: flags packed as below | above << 16
            unsigned f2[2];
            int segl[2], segmed[2], move[2];
            unsigned mine = 0;
            for (int q = 0; q < 2; q++) {
                const int i = 2 * threadIdx.x + q;
                f2[q] = 0;
                move[q] = 0;
                if (i < n0 && sseg[i] >= 0) {
                    const int j = sseg[i];
                    int l, r;
                    sub_segment(n0, k, j, l, r);
                    const int sz = r - l + 1, med = (l + r + 1) / 2;
                    const int t = st[j];
                    segl[q] = l;
                    segmed[q] = med;
                    if (sz >= 4 && t != d) {
                        move[q] = 1;
                        const int rm = rank_of(R, t, sa[t][med]);
                        const int rx = rank_of(R, t, sa[d][i]);
                        f2[q] = rx < rm ? 1u : (rx > rm ? 1u << 16 : 0u);
                    }
                }
                mine += f2[q];
            }
            unsigned tot;
            const unsigned ex = sub_excl(mine, wsum, tot);
            if (2 * threadIdx.x < n0) sscan[2 * threadIdx.x] = ex;
            if (2 * threadIdx.x + 1 < n0) sscan[2 * threadIdx.x + 1] = ex + f2[0];
            __syncthreads();
            for (int q = 0; q < 2; q++) {
                const int i = 2 * threadIdx.x + q;
                if (i >= n0) continue;
                int dst = i;
                if (move[q]) {
                    const unsigned base = sscan[segl[q]], cur = sscan[i];
                    if (f2[q] & 0xFFFFu) dst = segl[q] + (int)((cur & 0xFFFFu) - (base & 0xFFFFu));
                    else if (f2[q] >> 16) dst = segmed[q] + 1 + (int)((cur >> 16) - (base >> 16));
                    else dst = segmed[q];
                }
                ta[dst] = sa[d][i];
                tc[dst] = sc[d][i];
            }
            __syncthreads();
            for (int i = threadIdx.x; i < n0; i += kSubThreads) {
                sa[d][i] = ta[i];
                sc[d][i] = tc[i];
            }
            __syncthreads();
        }
    }
}

// The owner of every position for the next level, on levels without a
// partition (every segment below four points).
__global__ void k_owner_update(int nu, int *__restrict__ owner, const int *__restrict__ seg_l,
                               const int *__restrict__ seg_r)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu) return;
    int o = owner[i];
    if (!o) return;
    int med = (seg_l[o] + seg_r[o] + 1) / 2;
    owner[i] = i < med ? 2 * o : (i > med ? 2 * o + 1 : 0);
}

static inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

void Sorter::pairs(const uint64_t *kin, uint64_t *kout, const int *vin, int *vout, int n, hipStream_t s)
{
    size_t need = 0;
    KD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, vin, vout, n, 0, 64, s));
    if (tmp.n < need) tmp.alloc(need);
    KD_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, need, kin, kout, vin, vout, n, 0, 64, s));
}

// True when the sorted keys hold two equal neighbours.
static bool any_tie(const uint64_t *keys, int n, hipStream_t s)
{
    DevBuf<unsigned> tie(1);
    unsigned h = 0u;
    KD_HIP(hipMemsetAsync(tie.p, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_any_tie, dim3(nblk(n)), dim3(256), 0, s, keys, n, tie.p);
    KD_HIP(hipMemcpyAsync(&h, tie.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    return h != 0u;
}

// Point sets from this size on try the single sort by the leading key first:
// below it the sorts are cheap and the tie check's host round trip costs
// more than the sorts it saves (the 512 small trees of the batched workload
// built 9% slower with it).
static constexpr int kTieCheckMin = 1 << 20;

// Stable argsort of the points (stored with the given stride) by the
// lexicographic key (v[offs3[0]], v[offs3[1]], v[offs3[2]]): for large sets
// one radix sort by the first key, and only if two points share it the
// stable LSD passes from the last key (ties keep the original index order
// either way); small sets take the LSD passes directly.
static void lex_argsort(Sorter &so, const double *v, int stride, const int *offs3, int n, DevBuf<int> &out,
                        hipStream_t s)
{
    DevBuf<uint64_t> k0(n), k1(n);
    DevBuf<int> a(n), b(n);
    hipLaunchKernelGGL(k_iota, dim3(nblk(n)), dim3(256), 0, s, a.p, n);
    if (n >= kTieCheckMin) {
        hipLaunchKernelGGL(k_gather_key, dim3(nblk(n)), dim3(256), 0, s, v, stride, offs3[0], a.p, k0.p, n);
        so.pairs(k0.p, k1.p, a.p, b.p, n, s);
        if (!any_tie(k1.p, n, s)) {
            out = std::move(b);
            return;
        }
    }
    for (int pass = 2; pass >= 0; pass--) {
        hipLaunchKernelGGL(k_gather_key, dim3(nblk(n)), dim3(256), 0, s, v, stride, offs3[pass], a.p, k0.p, n);
        so.pairs(k0.p, k1.p, a.p, b.p, n, s);
        std::swap(a, b);
    }
    out = std::move(a);
}

// Segment sizes per level depend only on n (max child size is floor(s/2)).
static int build_levels(int nu)
{
    int L = 0;
    for (int s = nu; s > 0; s /= 2) L++;
    return L;
}

void build_tree(kd_device_tree *T, const double *d_xyz, const int *d_fields, int n, Sorter &so, BuildExtras &X,
                hipStream_t s)
{
    KD_REQUIRE(n >= 1, "tree_construct: length must be at least 1");
    // NaN check and |coordinate| maximum
    DevBuf<unsigned> flag(2);
    DevBuf<unsigned long long> amax(1);
    KD_HIP(hipMemsetAsync(flag.p, 0, flag.bytes(), s));
    KD_HIP(hipMemsetAsync(amax.p, 0, amax.bytes(), s));
    hipLaunchKernelGGL(k_scan_input, dim3(std::min(nblk(3ll * n), 4096)), dim3(256), 0, s, d_xyz, 3 * (size_t)n,
                       flag.p, amax.p);

    // 1. lexicographic order of raw points (ties keep the original index order)
    DevBuf<int> order;
    const int offs_xyz[3] = {0, 1, 2};
    lex_argsort(so, d_xyz, 3, offs_xyz, n, order, s);

    // 2. duplicate runs
    DevBuf<int> rflag(n), incl(n);
    hipLaunchKernelGGL(k_run_flags, dim3(nblk(n)), dim3(256), 0, s, d_xyz, order.p, n, rflag.p);
    so.incl_scan(rflag.p, incl.p, n, s);
    int nu = 0;
    unsigned hflag[2];
    unsigned long long hmax = 0;
    KD_HIP(hipMemcpyAsync(&nu, incl.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    KD_HIP(hipMemcpyAsync(hflag, flag.p, sizeof(hflag), hipMemcpyDeviceToHost, s));
    KD_HIP(hipMemcpyAsync(&hmax, amax.p, sizeof(hmax), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    KD_REQUIRE(!hflag[0], "tree_construct: NaN coordinate (unsupported by the reference's comparisons)");
    double cmax;
    memcpy(&cmax, &hmax, sizeof(cmax));

    X.ux.alloc(nu);
    X.uy.alloc(nu);
    X.uz.alloc(nu);
    X.uf.alloc(nu);
    X.uw.alloc(nu);
    X.rep.alloc(nu);
    DevBuf<int> runpos(nu);
    hipLaunchKernelGGL(k_unique, dim3(nblk(n)), dim3(256), 0, s, d_xyz, d_fields, order.p, rflag.p, incl.p, n, X.ux.p,
                       X.uy.p, X.uz.p, X.uf.p, runpos.p, X.rep.p);
    KD_HIP(hipMemsetAsync(flag.p, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_weights, dim3(nblk(nu)), dim3(256), 0, s, runpos.p, nu, n, X.uw.p, flag.p);
    KD_HIP(hipMemcpyAsync(hflag, flag.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    if (hflag[0]) throw Error("Invalid weights: a point occurs 256 or more times (8-bit node weight)");
    order.release();
    rflag.release();
    incl.release();
    runpos.release();

    int memsize = 1;
    {
        int x = nu;
        while (x > 1) { x >>= 1; memsize <<= 1; }
        memsize <<= 1;
    }
    T->n_raw = n;
    T->n_u = nu;
    T->memsize = memsize;
    T->cmax = cmax;

    // 3. cyclic superkey argsorts of the unique points (args[0] is the identity)
    DevBuf<int> args[3], nxt[3];
    args[0].alloc(nu);
    hipLaunchKernelGGL(k_iota, dim3(nblk(nu)), dim3(256), 0, s, args[0].p, nu);
    // Superkey (d, d+1, d+2).  When no two unique points share coordinate d
    // (the usual case for real-valued data), one radix sort by it gives the
    // order (tried for large sets).  Otherwise stable LSD passes, least
    // significant first; the unique points are distinct, so any starting
    // order already sorted by the least significant key gives the same
    // result: the (x,y,z) order (args[0], sorted by x) starts (y,z,x), which
    // then skips its x pass, and the (y,z,x) order (sorted by y) starts
    // (z,x,y), which skips its y pass.
    {
        const double *U[3] = {X.ux.p, X.uy.p, X.uz.p};
        DevBuf<uint64_t> k0(nu), k1(nu);
        for (int d = 1; d < 3; d++) {
            DevBuf<int> a(nu), b(nu);
            KD_HIP(hipMemcpyAsync(a.p, args[d - 1].p, sizeof(int) * (size_t)nu, hipMemcpyDeviceToDevice, s));
            if (nu >= kTieCheckMin) {
                hipLaunchKernelGGL(k_gather_key, dim3(nblk(nu)), dim3(256), 0, s, U[d], 1, 0, a.p, k0.p, nu);
                so.pairs(k0.p, k1.p, a.p, b.p, nu, s);
                if (!any_tie(k1.p, nu, s)) {
                    args[d] = std::move(b);
                    continue;
                }
            }
            for (int pass = 1; pass >= 0; pass--) {
                int dim = (d + pass) % 3;
                hipLaunchKernelGGL(k_gather_key, dim3(nblk(nu)), dim3(256), 0, s, U[dim], 1, 0, a.p, k0.p, nu);
                so.pairs(k0.p, k1.p, a.p, b.p, nu, s);
                std::swap(a, b);
            }
            args[d] = std::move(a);
        }
    }
    DevBuf<int> rank1(nu), rank2(nu);
    hipLaunchKernelGGL(k_rank, dim3(nblk(nu)), dim3(256), 0, s, args[1].p, nu, rank1.p);
    hipLaunchKernelGGL(k_rank, dim3(nblk(nu)), dim3(256), 0, s, args[2].p, nu, rank2.p);
    for (int d = 0; d < 3; d++) nxt[d].alloc(nu);
    // the coordinates of each dimension in its args order, moved by every
    // partition with the indices, so the variance sums read them contiguously
    DevBuf<double> cv[3], cvn[3];
    {
        const double *U[3] = {X.ux.p, X.uy.p, X.uz.p};
        for (int d = 0; d < 3; d++) {
            cv[d].alloc(nu);
            cvn[d].alloc(nu);
            hipLaunchKernelGGL(k_gather_vals, dim3(nblk(nu)), dim3(256), 0, s, U[d], args[d].p, nu, cv[d].p);
        }
    }

    // 4. level-synchronous build
    T->nodes.alloc((size_t)memsize * 40);
    T->split.alloc(memsize);
    T->nbits.alloc(memsize);
    T->perm.alloc(memsize);
    T->inorder.alloc(memsize);
    X.heap_of.alloc(nu);
    KD_HIP(hipMemsetAsync(T->nodes.p, 0, T->nodes.bytes(), s));
    KD_HIP(hipMemsetAsync(T->split.p, 0, T->split.bytes(), s));
    KD_HIP(hipMemsetAsync(T->nbits.p, 0, T->nbits.bytes(), s));
    KD_HIP(hipMemsetAsync(T->perm.p, 0xFF, T->perm.bytes(), s));
    KD_HIP(hipMemsetAsync(T->inorder.p, 0xFF, T->inorder.bytes(), s));
    DevBuf<int> seg_l(memsize), seg_r(memsize), owner(nu);
    DevBuf<unsigned char> tdim(memsize);
    KD_HIP(hipMemsetAsync(seg_l.p, 0xFF, seg_l.bytes(), s));
    KD_HIP(hipMemsetAsync(seg_r.p, 0xFF, seg_r.bytes(), s));
    KD_HIP(hipMemsetAsync(tdim.p, 0, tdim.bytes(), s));
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(nu)), dim3(256), 0, s, owner.p, nu, 1);
    KD_REQUIRE(3LL * nu < 0x7fffffffLL, "tree_construct: too many unique points");
    DevBuf<unsigned> flags(3 * (size_t)nu), scan(3 * (size_t)nu);
    // ranks carried with the argsort arrays (double-buffered like them)
    DevBuf<int> crk[4], crn[4];
    for (int k = 0; k < 4; k++) {
        crk[k].alloc(nu);
        crn[k].alloc(nu);
    }
    hipLaunchKernelGGL(k_carry_ranks, dim3(nblk(nu)), dim3(256), 0, s, args[1].p, args[2].p, rank1.p, rank2.p, nu,
                       CarriedRanks{{crk[0].p, crk[1].p, crk[2].p, crk[3].p}});
    int nlev = build_levels(nu);
    size_t maxlevel_nodes = (size_t)1 << (nlev - 1);
    DevBuf<double> var3(maxlevel_nodes * 3);
    DevBuf<unsigned char> flagged(maxlevel_nodes);

    // host-side list of the large segments per level (they depend only on nu)
    std::vector<std::vector<BigNode>> big(nlev);
    {
        struct Seg { int o, l, r; };
        std::vector<Seg> cur{{1, 0, nu - 1}}, next;
        for (int lev = 0; lev < nlev && !cur.empty(); lev++) {
            next.clear();
            for (auto &g : cur) {
                int sz = g.r - g.l + 1;
                if (sz <= kSeqMax) continue;
                big[lev].push_back({g.o, g.l, g.r});
                int med = (g.l + g.r + 1) / 2;
                next.push_back({2 * g.o, g.l, med - 1});
                next.push_back({2 * g.o + 1, med + 1, g.r});
            }
            std::swap(cur, next);
        }
    }
    // every level's large segments and their chunks of kChunk points, in one
    // upload: level lev's nodes start at node_off[lev], its chunks at
    // chunk_off[lev] (chunk_node indices are local to the level)
    std::vector<int> node_off(nlev + 1, 0), chunk_off(nlev + 1, 0);
    std::vector<BigNode> all_big;
    std::vector<int> all_cn, all_co;
    int max_nbig = 0;
    for (int lev = 0; lev < nlev; lev++) {
        const auto &B = big[lev];
        for (int b = 0; b < (int)B.size(); b++) {
            const int nch = (B[b].r - B[b].l + 1 + kChunk - 1) / kChunk;
            for (int c = 0; c < nch; c++) {
                all_cn.push_back(b);
                all_co.push_back(c);
            }
        }
        all_big.insert(all_big.end(), B.begin(), B.end());
        node_off[lev + 1] = (int)all_big.size();
        chunk_off[lev + 1] = (int)all_cn.size();
        max_nbig = std::max(max_nbig, (int)B.size());
    }
    DevBuf<BigNode> dbig(std::max<size_t>(all_big.size(), 1));
    DevBuf<int> dchunk_node(std::max<size_t>(all_cn.size(), 1)), dchunk_off(std::max<size_t>(all_co.size(), 1));
    DevBuf<double> dacc((size_t)std::max(max_nbig, 1) * 9), dmean((size_t)std::max(max_nbig, 1) * 3);
    if (!all_big.empty()) {
        KD_HIP(hipMemcpyAsync(dbig.p, all_big.data(), sizeof(BigNode) * all_big.size(), hipMemcpyHostToDevice, s));
        KD_HIP(hipMemcpyAsync(dchunk_node.p, all_cn.data(), sizeof(int) * all_cn.size(), hipMemcpyHostToDevice, s));
        KD_HIP(hipMemcpyAsync(dchunk_off.p, all_co.data(), sizeof(int) * all_co.size(), hipMemcpyHostToDevice, s));
    }

    int max_seg = nu;
    for (int lev = 0; lev < nlev; lev++, max_seg /= 2) {
        int lo = 1 << lev, hi = std::min(2 * lo, memsize);
        if (lo >= memsize) break;
        int cnt = hi - lo;
        const int nbig_lev = node_off[lev + 1] - node_off[lev];
        hipLaunchKernelGGL(k_level_segs, dim3(nblk(cnt)), dim3(256), 0, s, lo, hi, nu, seg_l.p, seg_r.p, dacc.p,
                           9 * nbig_lev, flagged.p);
        UPtrs P{{X.ux.p, X.uy.p, X.uz.p}, {args[0].p, args[1].p, args[2].p}, {cv[0].p, cv[1].p, cv[2].p}};
        if (max_seg <= kSubMax) {  // this level and every one below: one workgroup per segment, in LDS
            RPtrs R{{nullptr, rank1.p, rank2.p}};
            hipLaunchKernelGGL(k_subtrees, dim3(cnt), dim3(kSubThreads), 0, s, lo, hi, seg_l.p, seg_r.p, P, R, X.uf.p,
                               X.uw.p, (unsigned long long *)T->nodes.p, T->split.p, T->nbits.p, T->perm.p,
                               T->inorder.p, X.heap_of.p, memsize, flag.p + 1);
            KD_HIP(hipGetLastError());
            break;
        }
        if (max_seg >= 4) {
            // small segments: exact sequential replay
            hipLaunchKernelGGL(k_var_seq, dim3(nblk(3ll * cnt)), dim3(256), 0, s, lo, hi, seg_l.p, seg_r.p, 4, kSeqMax,
                               (const unsigned char *)nullptr, P, var3.p);
            hipLaunchKernelGGL(k_pick_tdim, dim3(nblk(cnt)), dim3(256), 0, s, lo, hi, seg_l.p, seg_r.p, 0, kSeqMax,
                               (const unsigned char *)nullptr, var3.p, tdim.p);
            // large segments: certified parallel decision, replay where uncertain
            if (nbig_lev > 0) {
                const int nbig = nbig_lev;
                const int nchunks = chunk_off[lev + 1] - chunk_off[lev];
                const BigNode *lb = dbig.p + node_off[lev];
                const int *lcn = dchunk_node.p + chunk_off[lev], *lco = dchunk_off.p + chunk_off[lev];
                hipLaunchKernelGGL(k_big_sum, dim3(nchunks), dim3(256), 0, s, lb, lcn, lco, P, dacc.p, 0,
                                   (const double *)nullptr);
                hipLaunchKernelGGL(k_big_mean, dim3(nblk(3ll * nbig)), dim3(256), 0, s, lb, nbig, dacc.p, dmean.p);
                hipLaunchKernelGGL(k_big_sum, dim3(nchunks), dim3(256), 0, s, lb, lcn, lco, P, dacc.p, 1, dmean.p);
                hipLaunchKernelGGL(k_big_decide, dim3(nblk(nbig)), dim3(256), 0, s, lb, nbig, dacc.p, tdim.p,
                                   flagged.p, lo);
                hipLaunchKernelGGL(k_var_seq, dim3(nblk(3ll * cnt)), dim3(256), 0, s, lo, hi, seg_l.p, seg_r.p,
                                   kSeqMax + 1, 0x7fffffff, flagged.p, P, var3.p);
                hipLaunchKernelGGL(k_pick_tdim, dim3(nblk(cnt)), dim3(256), 0, s, lo, hi, seg_l.p, seg_r.p,
                                   kSeqMax + 1, 0x7fffffff, flagged.p, var3.p, tdim.p);
            }
        } else {
            hipLaunchKernelGGL(k_pick_tdim, dim3(nblk(cnt)), dim3(256), 0, s, lo, hi, seg_l.p, seg_r.p, 0, 0,
                               (const unsigned char *)nullptr, var3.p, tdim.p);
        }
        hipLaunchKernelGGL(k_write_nodes, dim3(nblk(cnt)), dim3(256), 0, s, lo, hi, seg_l.p, seg_r.p, tdim.p, P,
                           X.uf.p, X.uw.p, (unsigned long long *)T->nodes.p, T->split.p, T->nbits.p, T->perm.p,
                           T->inorder.p, X.heap_of.p);
        if (max_seg >= 4) {
            RPtrs R{{nullptr, rank1.p, rank2.p}};
            const CarriedRanks C{{crk[0].p, crk[1].p, crk[2].p, crk[3].p}}, Cn{{crn[0].p, crn[1].p, crn[2].p, crn[3].p}};
            hipLaunchKernelGGL(k_part_flags3, dim3(nblk(nu)), dim3(256), 0, s, nu, owner.p, seg_l.p, seg_r.p, tdim.p,
                               P, R, C, flags.p);
            so.excl_scan_u32(flags.p, scan.p, 3 * nu, s);
            hipLaunchKernelGGL(k_part_scatter3, dim3(nblk(nu)), dim3(256), 0, s, nu, owner.p, seg_l.p, seg_r.p,
                               tdim.p, P, C, flags.p, scan.p, nxt[0].p, nxt[1].p, nxt[2].p, cvn[0].p, cvn[1].p, cvn[2].p,
                               Cn);
            for (int d = 0; d < 3; d++) {
                std::swap(args[d], nxt[d]);
                std::swap(cv[d], cvn[d]);
            }
            for (int k = 0; k < 4; k++) std::swap(crk[k], crn[k]);
        } else {
            hipLaunchKernelGGL(k_owner_update, dim3(nblk(nu)), dim3(256), 0, s, nu, owner.p, seg_l.p, seg_r.p);
        }
    }
    KD_HIP(hipMemcpyAsync(hflag, flag.p, sizeof(hflag), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    KD_REQUIRE(!hflag[1], "tree_construct: internal error: a subtree exceeds the in-LDS build's size");
}

}  // namespace kd
