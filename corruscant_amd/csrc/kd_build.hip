// GPU build of corruscant's balanced k-d tree, node array bit-identical to
// the reference's tree_construct (src/kdbuild.c:451-537).
//
// Steps (all on the device):
//   1. lexicographic (x,y,z) order of the raw points: stable LSD radix sorts
//      over order-preserving 64-bit keys (one by x alone when no two points
//      share an x), with -0.0 keyed as +0.0 so that the IEEE == of
//      src/kdbuild.c:82-96 decides equality;
//   2. duplicate runs -> one unique point each, carrying the lowest original
//      index's value and field, weight = run length (src/kdbuild.c:484-511);
//   3. the three cyclic-superkey argsorts (src/kdbuild.c:221-243) and their
//      rank arrays;
//   4. a level-synchronous median build.  The segment [l, r] of every heap
//      node depends only on the number of unique points, so each level is one
//      pass: split dimension by maximum variance (src/kdbuild.c:308-351),
//      node records, then a stable partition of the other two argsort arrays
//      (src/kdbuild.c:251-291) done as a segmented prefix sum of
//      rank-compare flags, three launches per level; once every segment is
//      small, one launch builds the remaining levels in LDS.
//
// Split-dimension decisions reproduce the reference's sequential FP64 sums.
// Segments of up to kSeqMax points replay the sequential arithmetic exactly
// (one thread per segment and dimension).  Larger segments use a parallel sum
// whose error bound is compared with the bound on the reference's sequential
// rounding: the parallel decision is kept only when it is certain; otherwise
// the segment falls back to the exact sequential replay.

#include <algorithm>
#include <cstring>
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

// Byte fills of up to eight buffers (4-byte aligned) in one launch.
struct Fill {
    void *p;
    size_t bytes;
    unsigned char v;
};
struct Fills {
    Fill f[8];
    int n;
};

__global__ void k_fill_bufs(Fills F)
{
    const long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x, stride = (long long)gridDim.x * blockDim.x;
    for (int b = 0; b < F.n; b++) {
        unsigned *w = (unsigned *)F.f[b].p;
        const long long nw = (long long)(F.f[b].bytes / 4);
        const unsigned v = 0x01010101u * F.f[b].v;
        for (long long k = tid; k < nw; k += stride) w[k] = v;
        if (tid < (long long)(F.f[b].bytes % 4)) ((unsigned char *)F.f[b].p)[nw * 4 + tid] = F.f[b].v;
    }
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

// Level 0: every unique point belongs to the root, whose segment is [0, nu-1].
__global__ void k_init_root(int *__restrict__ owner, int nu, int *__restrict__ seg_l, int *__restrict__ seg_r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nu) owner[i] = 1;
    if (i == 0) {
        seg_l[1] = 0;
        seg_r[1] = nu - 1;
    }
}

struct UPtrs {
    const double *v[3];
    const int *a[3];
    const double *c[3];  // c[d][i] = v[d][a[d][i]]: the coordinates in args order
};

// Exact replay of src/kdbuild.c:318-338 for one (node, dim): sequential sum
// in args order, division by the size, sequential fma variance.
__device__ __forceinline__ double var_seq(const double *__restrict__ c, int l, int r)
{
    const int s = r - l + 1;
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
    return var;
}

// The split dimension of every node of a level whose segment holds s points,
// smin <= s <= smax (and is flagged, with only_flagged): the three variances
// by the sequential replay above, one thread per (node, dimension), then the
// strict-> argmax with dimension 0 as the default (src/kdbuild.c:340-348).
// Segments below four points get dimension 0.  kVarNodes nodes per workgroup.
static constexpr int kVarNodes = 85;  // per workgroup of 256 threads

__device__ __forceinline__ void var_pick_block(int blk, int lo, int hi, const int *__restrict__ seg_l,
                                               const int *__restrict__ seg_r, int smin, int smax,
                                               const unsigned char *__restrict__ only_flagged, const UPtrs &P,
                                               unsigned char *__restrict__ tdim)
{
    __shared__ double sv[kVarNodes][3];
    const int node = threadIdx.x / 3, d = threadIdx.x % 3;
    const int o = lo + blk * kVarNodes + node;
    const bool in = node < kVarNodes && o < hi;
    int l = -1, r = -2;
    if (in) {
        l = seg_l[o];
        r = seg_r[o];
    }
    const int s = r - l + 1;
    const bool ok = in && l >= 0;
    const bool take = ok && s >= smin && s <= smax && (!only_flagged || only_flagged[o - lo]);
    if (take) sv[node][d] = var_seq(P.c[d], l, r);
    __syncthreads();
    if (d != 0 || !ok) return;
    if (s < 4) {
        tdim[o] = 0;
        return;
    }
    if (!take) return;
    int dim = 0;
    double best = sv[node][0];
    for (int j = 1; j < 3; j++)
        if (sv[node][j] > best) {
            best = sv[node][j];
            dim = j;
        }
    tdim[o] = (unsigned char)dim;
}

// ---- certified parallel variance for large segments ----------------------

struct BigNode {
    int o, l, r;
};

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
__device__ void big_decide(const BigNode &nd, const double (&acc)[9], unsigned char *__restrict__ tdim,
                           unsigned char *__restrict__ flagged, int lo)
{
    const double u = 1.1102230246251565e-16;
    double s = (double)(nd.r - nd.l + 1);
    double g = (s + 16.0) * u;
    double G = g / (1.0 - g);
    double lo_b[3], hi_b[3], mid[3];
    for (int d = 0; d < 3; d++) {
        double A = acc[d * 3 + 1] * (1.0 + 2.0 * G);
        double V = acc[d * 3 + 2];
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

// acc layout per big node: [d][0] sum, [d][1] sum |v|, [d][2] sum (v-m)^2
// Pass 1's last workgroup of each node (counted in done[]) makes the node's
// decision from the finished sums (a one-chunk node's only workgroup from its
// own sum of squares, without atomics or fences).
__device__ __forceinline__ void big_sum_block(int c, const BigNode *__restrict__ nodes,
                                              const int *__restrict__ chunk_node, const int *__restrict__ chunk_off,
                                              const UPtrs &P, double *__restrict__ acc, int pass,
                                              int *__restrict__ done, unsigned char *__restrict__ tdim,
                                              unsigned char *__restrict__ flagged, int lo)
{
    int bn = chunk_node[c];
    BigNode nd = nodes[bn];
    int b0 = nd.l + chunk_off[c] * kChunk;
    int b1 = min(nd.r + 1, b0 + kChunk);
    __shared__ double red[3][2][4];
    __shared__ double vloc[3];
    const int nch = (nd.r - nd.l + 1 + kChunk - 1) / kChunk;  // chunks of this segment
    for (int d = 0; d < 3; d++) {
        const double *__restrict__ c = P.c[d];
        double s0 = 0.0, s1 = 0.0;
        // pass 1: deviations from the mean of pass 0's sum
        double m = pass ? acc[bn * 9 + d * 3] / (double)(nd.r - nd.l + 1) : 0.0;
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
            if (nch == 1) vloc[d] = s0;  // the whole segment: its sum is this workgroup's
            else atomicAdd(&acc[bn * 9 + d * 3 + 2], s0);
        } else {
            atomicAdd(&acc[bn * 9 + d * 3 + 0], s0);
            atomicAdd(&acc[bn * 9 + d * 3 + 1], s1);
        }
        if (pass && nch > 1) __threadfence();
    }
    if (!pass) return;
    __shared__ int last, s_flag;
    __shared__ double s_var[3];
    __syncthreads();
    if (threadIdx.x == 0) {
        last = nch == 1 || atomicAdd(&done[bn], 1) == nch - 1;
        s_flag = 0;
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        double a[9];
        if (nch == 1) {
            for (int k = 0; k < 9; k++) a[k] = k % 3 == 2 ? vloc[k / 3] : acc[bn * 9 + k];
        } else {
            __threadfence();
            for (int k = 0; k < 9; k++) a[k] = atomicAdd(&acc[bn * 9 + k], 0.0);  // the sums every chunk added
        }
        big_decide(nd, a, tdim, flagged, lo);
        s_flag = flagged[nd.o - lo];
    }
    __syncthreads();
    if (last && s_flag) {  // uncertain: the exact sequential replay, then the strict argmax
        if (threadIdx.x < 3) s_var[threadIdx.x] = var_seq(P.c[threadIdx.x], nd.l, nd.r);
        __syncthreads();
        if (threadIdx.x == 0) {
            int dim = 0;
            for (int j = 1; j < 3; j++)
                if (s_var[j] > s_var[dim]) dim = j;
            tdim[nd.o] = (unsigned char)dim;
        }
    }
}

__global__ __launch_bounds__(256) void k_big_sum(const BigNode *__restrict__ nodes, const int *__restrict__ chunk_node,
                                                 const int *__restrict__ chunk_off, UPtrs P, double *__restrict__ acc,
                                                 int pass, int *__restrict__ done, unsigned char *__restrict__ tdim,
                                                 unsigned char *__restrict__ flagged, int lo)
{
    big_sum_block(blockIdx.x, nodes, chunk_node, chunk_off, P, acc, pass, done, tdim, flagged, lo);
}

// One level's first launch: the split dimensions of the segments of 4 to
// kSeqMax points (workgroups below nvb) and pass 0 of the large segments'
// sums (the other workgroups, one per chunk).
__global__ __launch_bounds__(256) void k_level_sums(int nvb, int lo, int hi, const int *__restrict__ seg_l,
                                                    const int *__restrict__ seg_r, UPtrs P,
                                                    unsigned char *__restrict__ tdim, const BigNode *__restrict__ nodes,
                                                    const int *__restrict__ chunk_node,
                                                    const int *__restrict__ chunk_off, double *__restrict__ acc,
                                                    int *__restrict__ done, unsigned char *__restrict__ flagged)
{
    if ((int)blockIdx.x < nvb)
        var_pick_block(blockIdx.x, lo, hi, seg_l, seg_r, 4, kSeqMax, nullptr, P, tdim);
    else
        big_sum_block(blockIdx.x - nvb, nodes, chunk_node, chunk_off, P, acc, 0, done, tdim, flagged, lo);
}

// ---- node records, partition ---------------------------------------------

// Node records of one level, the segments [l, r] of the next level's nodes
// (children of an empty node keep the initial -1), and the clearing of the
// next level's large-segment sums (nacc doubles) and chunk counters (ndone).
// Workgroup blk of nblk (blockDim.x threads each).
struct NodeOut {
    const int *uf, *uw;
    unsigned long long *nodes;
    double *split;
    uint32_t *nbits;
    int *perm, *inorder, *heap_of;
};

__device__ __forceinline__ void write_nodes_block(int blk, int nblk, int lo, int hi, int *__restrict__ seg_l,
                                                  int *__restrict__ seg_r, int memsize,
                                                  const unsigned char *__restrict__ tdim, const UPtrs &P,
                                                  const NodeOut &N, double *__restrict__ acc, int nacc,
                                                  int *__restrict__ done, int ndone)
{
    const int *__restrict__ uf = N.uf, *__restrict__ uw = N.uw;
    unsigned long long *__restrict__ nodes = N.nodes;
    double *__restrict__ split = N.split;
    uint32_t *__restrict__ nbits = N.nbits;
    int *__restrict__ perm = N.perm, *__restrict__ inorder = N.inorder, *__restrict__ heap_of = N.heap_of;
    const int tid = blk * blockDim.x + threadIdx.x;
    for (int k = tid; k < nacc; k += nblk * blockDim.x) acc[k] = 0.0;
    for (int k = tid; k < ndone; k += nblk * blockDim.x) done[k] = 0;
    int o = lo + tid;
    if (o >= hi) return;
    int l = seg_l[o], r = seg_r[o];
    if (l < 0) return;
    int s = r - l + 1, med = (l + r + 1) / 2;
    if (2 * o < memsize) {
        seg_l[2 * o] = s >= 2 ? l : -1;
        seg_r[2 * o] = s >= 2 ? med - 1 : -1;
    }
    if (2 * o + 1 < memsize) {
        seg_l[2 * o + 1] = s >= 3 ? med + 1 : -1;
        seg_r[2 * o + 1] = s >= 3 ? r : -1;
    }
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
// argsort arrays at once, in one launch.  Element i of array d (in a segment
// of at least four points whose split dimension t is not d) is below or above
// the median by its rank in order t; its flag is 1 (below), 1 << 31 (the
// median) or 0 (above).  The sum of the flags before i inside its segment
// gives its place: the low 31 bits count the elements below the median before
// it, the top bit whether the median is among them, and the rest of the
// segment before it lies above.
//
// Single pass: tiles of kPartTile positions are taken in launch order from a
// counter (stat[0]).  A tile scans its flags, segmented at the segments' first
// positions, publishes per array the sum since its last segment start (as a
// finished prefix when it holds a start, else as an aggregate), and looks back
// over the tiles before it (64 at a time) for the sum of the segment it begins
// in.  It waits only on tiles taken before it, which publish before they look
// back themselves, so every wait ends; a wait past kLookbackSpins polls sets *err
// instead of hanging.  stat[1 + 3 tile + d] = (state << 32) | sum, state 1 =
// aggregate, 2 = prefix; the caller zeroes stat.  A place outside its segment
// (which a correct sum never gives) sets bit 1 of *err and is not written.
static constexpr int kPartThreads = 512;
static constexpr int kPartPer = 4;
static constexpr int kPartTile = kPartThreads * kPartPer;
static constexpr unsigned kLookbackSpins = 1u << 24;

__device__ __forceinline__ unsigned flag_value(unsigned code) { return code == 1u ? 1u : (code == 2u ? 0x80000000u : 0u); }

struct PartOut {
    int *a[3];     // argsort arrays of the next level
    double *c[3];  // their coordinates
};

__device__ __forceinline__ void partition_block(int nu, int *__restrict__ owner, const int *__restrict__ seg_l,
                                                const int *__restrict__ seg_r, const unsigned char *__restrict__ tdim,
                                                const UPtrs &P, const RPtrs &R, const CarriedRanks &C,
                                                const PartOut &O, const CarriedRanks &Cout,
                                                unsigned long long *__restrict__ stat, unsigned *__restrict__ err)
{
    constexpr int kWaves = kPartThreads / 64;
    __shared__ int s_tile, s_head0;
    __shared__ int s_wf[kWaves];
    __shared__ unsigned s_wv[kWaves][3];
    __shared__ unsigned s_prefix[3];
    // per position of the tile: the three flags' codes (2 bits each: 0 above,
    // 1 below, 2 the median) and bit 6 = a segment starts here; then the sums
    // inside the segment before each position, per array
    __shared__ unsigned s_code4[kPartTile / 4];
    __shared__ unsigned s_diff[3][kPartTile];
    unsigned char *s_code = (unsigned char *)s_code4;
    if (threadIdx.x == 0) s_tile = (int)atomicAdd(stat, 1ull);
    __syncthreads();
    const int tile = s_tile;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int t0 = tile * kPartTile;

    // 1. flags, with consecutive lanes on consecutive positions
    int o[kPartPer];
#pragma unroll
    for (int k = 0; k < kPartPer; k++) {
        const int j = k * kPartThreads + threadIdx.x, i = t0 + j;
        o[k] = i < nu ? owner[i] : 0;
        bool head = true;
        unsigned c3 = 0u;
        if (o[k]) {
            const int l = seg_l[o[k]], r = seg_r[o[k]];
            head = i == l;
            if (r - l + 1 >= 4) {
                const int t = tdim[o[k]], med = (l + r + 1) / 2;
                const int rm = rank_of(R, t, P.a[t][med]);  // one address per segment
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    if (d == t) continue;
                    const int rx = carried_rank(C, P, d, t, i);
                    c3 |= (rx < rm ? 1u : (rx > rm ? 0u : 2u)) << (2 * d);
                }
            }
        }
        s_code[j] = (unsigned char)(c3 | (head ? 64u : 0u));
    }
    __syncthreads();

    // 2. each thread takes kPartPer consecutive positions: their sums inside
    // the segment, from the segment's start or the thread's first position
    // (hb bit k: a segment starts at or before position k here)
    const unsigned codes = s_code4[threadIdx.x];
    unsigned ex[kPartPer][3], hb = 0u, run[3] = {0u, 0u, 0u};
    bool th = false;
#pragma unroll
    for (int k = 0; k < kPartPer; k++) {
        const unsigned c = (codes >> (8 * k)) & 0xFFu;
        if (c & 64u) {
            th = true;
            run[0] = run[1] = run[2] = 0u;
        }
        hb |= (unsigned)th << k;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            ex[k][d] = run[d];
            run[d] += flag_value((c >> (2 * d)) & 3u);
        }
    }

    // segmented scan of the threads' sums over the workgroup: (f, v) with f =
    // a segment starts inside; (f1, v1) then (f2, v2) gives (f1 | f2, f2 ? v2 : v1 + v2)
    bool f = th;
    unsigned v[3] = {run[0], run[1], run[2]};
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const bool uf = __shfl_up((int)f, off, 64) != 0;
        unsigned uv[3];
#pragma unroll
        for (int d = 0; d < 3; d++) uv[d] = __shfl_up(v[d], off, 64);
        if (lane >= off) {
            if (!f)
#pragma unroll
                for (int d = 0; d < 3; d++) v[d] += uv[d];
            f = f || uf;
        }
    }
    bool ef = __shfl_up((int)f, 1, 64) != 0;
    unsigned ev[3];
#pragma unroll
    for (int d = 0; d < 3; d++) ev[d] = __shfl_up(v[d], 1, 64);
    if (lane == 0) {
        ef = false;
        ev[0] = ev[1] = ev[2] = 0u;
    }
    if (lane == 63) {
        s_wf[wid] = f;
#pragma unroll
        for (int d = 0; d < 3; d++) s_wv[wid][d] = v[d];
    }
    if (threadIdx.x == 0) s_head0 = (int)(hb & 1u);
    __syncthreads();
    bool bf = false;
    unsigned bv[3] = {0u, 0u, 0u};
    for (int w = 0; w < wid; w++) {
        if (s_wf[w]) {
            bf = true;
#pragma unroll
            for (int d = 0; d < 3; d++) bv[d] = s_wv[w][d];
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++) bv[d] += s_wv[w][d];
        }
    }
    if (ef) {
        bf = true;
#pragma unroll
        for (int d = 0; d < 3; d++) bv[d] = ev[d];
    } else {
#pragma unroll
        for (int d = 0; d < 3; d++) bv[d] += ev[d];
    }

    // publish the tile's sums and look back for the sum its first segment
    // carries in from the tiles before (not needed when a segment starts at
    // its first position): wave d for array d, 64 tiles per step (lane j reads
    // tile - 1 - j; the sum runs up to the nearest finished prefix)
    if (wid < 3) {
        const int d = wid;
        bool tf = false;
        unsigned tv = 0u;
        for (int w = 0; w < kWaves; w++) {
            if (s_wf[w]) {
                tf = true;
                tv = s_wv[w][d];
            } else {
                tv += s_wv[w][d];
            }
        }
        unsigned long long *word = stat + 1 + (size_t)tile * 3 + d;
        const bool done = tile == 0 || tf;
        if (lane == 0)
            __hip_atomic_store(word, (done ? 2ull << 32 : 1ull << 32) | tv, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        unsigned acc = 0u;
        if (tile > 0 && !s_head0) {
            unsigned spins = 0u;
            for (int k = tile - 1; k >= 0;) {
                const int j = k - lane;
                const unsigned long long w =
                    j >= 0 ? __hip_atomic_load(stat + 1 + (size_t)j * 3 + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           : 2ull << 32;  // before the first tile: a finished prefix of 0
                const unsigned state = (unsigned)(w >> 32);
                const unsigned long long pm = __ballot(state == 2u), zm = __ballot(state == 0u);
                const int first = pm ? __ffsll((long long)pm) - 1 : 64;
                const unsigned long long need = first == 64 ? ~0ull : ((2ull << first) - 1ull);
                if (zm & need) {  // a tile up to the nearest prefix has not published yet
                    if (++spins >= kLookbackSpins) {
                        if (lane == 0) atomicOr(err, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                unsigned v = lane <= first ? (unsigned)w : 0u;
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
                acc += v;
                if (first < 64) break;
                k -= 64;
            }
            if (!done && lane == 0)
                __hip_atomic_store(word, (2ull << 32) | (acc + tv), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) s_prefix[d] = acc;
    }
    __syncthreads();
    if (!bf)
#pragma unroll
        for (int d = 0; d < 3; d++) bv[d] += s_prefix[d];
#pragma unroll
    for (int k = 0; k < kPartPer; k++) {
        const int j = threadIdx.x * kPartPer + k;
#pragma unroll
        for (int d = 0; d < 3; d++) s_diff[d][j] = ex[k][d] + (((hb >> k) & 1u) ? 0u : bv[d]);
    }
    __syncthreads();

    // 3. moves of the three arrays (indices with their coordinates and
    // carried ranks) to their places in the children's segments, and the
    // owner of every position for the next level (the median's position
    // leaves the build); consecutive lanes on consecutive positions again
    // (every input of the tile's positions is loaded before the first move,
    // so the loads overlap)
    int *const *outs = O.a;
    double *const *couts = O.c;
    int xa[kPartPer][3], ra[kPartPer][4];
    double xc[kPartPer][3];
#pragma unroll
    for (int k = 0; k < kPartPer; k++) {
        const int i = t0 + k * kPartThreads + threadIdx.x;
        const bool in = i < nu;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            xa[k][d] = in ? P.a[d][i] : 0;
            xc[k][d] = in ? P.c[d][i] : 0.0;  // the coordinate travels with its index
        }
#pragma unroll
        for (int q = 0; q < 4; q++) ra[k][q] = in ? C.r[q][i] : 0;
    }
#pragma unroll
    for (int k = 0; k < kPartPer; k++) {
        const int j = k * kPartThreads + threadIdx.x, i = t0 + j;
        if (i >= nu) break;
        const int ok = o[k];
        int l = 0, r = -1, t = 0, med = 0;
        if (ok) {
            l = seg_l[ok];
            r = seg_r[ok];
            t = tdim[ok];
            med = (l + r + 1) / 2;
        }
        const bool part = ok && r - l + 1 >= 4;
        const unsigned c = s_code[j];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int dst = i;
            if (part && t != d) {
                const unsigned diff = s_diff[d][j], code = (c >> (2 * d)) & 3u;
                const int below = (int)(diff & 0x7fffffffu), median = (int)(diff >> 31);
                if (code == 1u) dst = l + below;                                   // below the median: [l, med-1]
                else if (code == 0u) dst = med + 1 + ((i - l) - below - median);   // above: [med+1, r]
                else dst = med;  // the median itself; this slot is never read again
                if (dst < l || dst > r) {  // a wrong sum: report it, never write outside the segment
                    atomicOr(err, 2u);
                    dst = i;
                }
            }
            outs[d][dst] = xa[k][d];
            couts[d][dst] = xc[k][d];
            if (d == 0) {
                Cout.r[0][dst] = ra[k][0];
                Cout.r[1][dst] = ra[k][1];
            } else {
                Cout.r[d + 1][dst] = ra[k][d + 1];
            }
        }
        if (ok) owner[i] = i < med ? 2 * ok : (i > med ? 2 * ok + 1 : 0);
    }
}

// One level's last launch: the node records and the next level's segments
// (workgroups below nwb, one node per thread) beside the partition (the
// others, one tile each); neither reads what the other writes.
__global__ __launch_bounds__(kPartThreads) void k_nodes_partition(
    int nwb, int lo, int hi, int *__restrict__ seg_l, int *__restrict__ seg_r, int memsize,
    const unsigned char *__restrict__ tdim, UPtrs P, NodeOut N, double *__restrict__ acc, int nacc,
    int *__restrict__ done, int ndone, int nu, int *__restrict__ owner, RPtrs R, CarriedRanks C, PartOut O,
    CarriedRanks Cout, unsigned long long *__restrict__ stat, unsigned *__restrict__ err)
{
    if ((int)blockIdx.x < nwb)
        write_nodes_block(blockIdx.x, nwb, lo, hi, seg_l, seg_r, memsize, tdim, P, N, acc, nacc, done, ndone);
    else
        partition_block(nu, owner, seg_l, seg_r, tdim, P, R, C, O, Cout, stat, err);
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
//     (segment, dimension), by the exact sequential replay of var_seq, and
//     the split dimension by the strict argmax of var_pick_block (dimension 0
//     below four points);
//   - the node records of write_nodes_block;
//   - the stable partition of the argsort arrays with their coordinates by
//     rank against the median in the split dimension's order, as
//     partition_block does it: per-element flags, one block-wide
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

static inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

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
    DevBuf<unsigned> flag(3);
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
    unsigned hflag[3];
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
    for (int d = 0; d < 3; d++) {
        // zeroed once (afterwards they hold an earlier level's arrays): a slot
        // that a failed partition leaves unwritten still holds a valid index
        nxt[d].alloc(nu);
        zero_async(nxt[d].p, nxt[d].bytes(), s);
    }
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
    DevBuf<int> seg_l(memsize), seg_r(memsize), owner(nu);
    DevBuf<unsigned char> tdim(memsize);
    {
        const Fills F{{{T->nodes.p, T->nodes.bytes(), 0},
                       {T->split.p, T->split.bytes(), 0},
                       {T->nbits.p, T->nbits.bytes(), 0},
                       {T->perm.p, T->perm.bytes(), 0xFF},
                       {T->inorder.p, T->inorder.bytes(), 0xFF},
                       {seg_l.p, seg_l.bytes(), 0xFF},
                       {seg_r.p, seg_r.bytes(), 0xFF},
                       {tdim.p, tdim.bytes(), 0}},
                      8};
        hipLaunchKernelGGL(k_fill_bufs, dim3(std::min(nblk(T->nodes.bytes() / 4), 2048)), dim3(256), 0, s, F);
    }
    hipLaunchKernelGGL(k_init_root, dim3(nblk(nu)), dim3(256), 0, s, owner.p, nu, seg_l.p, seg_r.p);
    KD_REQUIRE(3LL * nu < 0x7fffffffLL, "tree_construct: too many unique points");
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
    DevBuf<unsigned char> flagged(maxlevel_nodes);
    // look-back words of the single-pass partitions, one region per level
    const int ptiles = nblk(nu, kPartTile);
    const size_t pstride = 1 + 3 * (size_t)ptiles;
    DevBuf<unsigned long long> pstat(pstride * nlev);
    zero_async(pstat.p, pstat.bytes(), s);

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
    DevBuf<double> dacc((size_t)std::max(max_nbig, 1) * 9);
    DevBuf<int> ddone(std::max(max_nbig, 1));  // pass-1 chunks finished per large segment
    zero_async(dacc.p, dacc.bytes(), s);
    zero_async(ddone.p, ddone.bytes(), s);
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
        UPtrs P{{X.ux.p, X.uy.p, X.uz.p}, {args[0].p, args[1].p, args[2].p}, {cv[0].p, cv[1].p, cv[2].p}};
        if (max_seg <= kSubMax) {  // this level and every one below: one workgroup per segment, in LDS
            RPtrs R{{nullptr, rank1.p, rank2.p}};
            hipLaunchKernelGGL(k_subtrees, dim3(cnt), dim3(kSubThreads), 0, s, lo, hi, seg_l.p, seg_r.p, P, R, X.uf.p,
                               X.uw.p, (unsigned long long *)T->nodes.p, T->split.p, T->nbits.p, T->perm.p,
                               T->inorder.p, X.heap_of.p, memsize, flag.p + 1);
            KD_HIP(hipGetLastError());
            break;
        }
        // (every segment of this level holds more than kSubMax points)
        // split dimensions: segments of 4..kSeqMax points by the exact
        // sequential replay, beside pass 0 of the large segments' sums; then
        // pass 1, whose last workgroup per segment makes the certified
        // decision or, where it is uncertain, the replay
        const int nchunks = chunk_off[lev + 1] - chunk_off[lev];
        const BigNode *lb = dbig.p + node_off[lev];
        const int *lcn = dchunk_node.p + chunk_off[lev], *lco = dchunk_off.p + chunk_off[lev];
        const int nvb = nblk(cnt, kVarNodes);
        hipLaunchKernelGGL(k_level_sums, dim3(nvb + nchunks), dim3(256), 0, s, nvb, lo, hi, seg_l.p, seg_r.p, P,
                           tdim.p, lb, lcn, lco, dacc.p, ddone.p, flagged.p);
        if (nchunks > 0)
            hipLaunchKernelGGL(k_big_sum, dim3(nchunks), dim3(256), 0, s, lb, lcn, lco, P, dacc.p, 1, ddone.p, tdim.p,
                               flagged.p, lo);
        // node records and the next level's segments, beside the partition
        const int nbig_next = lev + 1 < nlev ? node_off[lev + 2] - node_off[lev + 1] : 0;
        const int nwb = nblk(cnt, kPartThreads);
        const NodeOut N{X.uf.p, X.uw.p, (unsigned long long *)T->nodes.p, T->split.p, T->nbits.p, T->perm.p,
                        T->inorder.p, X.heap_of.p};
        const RPtrs R{{nullptr, rank1.p, rank2.p}};
        const CarriedRanks C{{crk[0].p, crk[1].p, crk[2].p, crk[3].p}}, Cn{{crn[0].p, crn[1].p, crn[2].p, crn[3].p}};
        const PartOut O{{nxt[0].p, nxt[1].p, nxt[2].p}, {cvn[0].p, cvn[1].p, cvn[2].p}};
        hipLaunchKernelGGL(k_nodes_partition, dim3(nwb + ptiles), dim3(kPartThreads), 0, s, nwb, lo, hi, seg_l.p,
                           seg_r.p, memsize, tdim.p, P, N, dacc.p, 9 * nbig_next, ddone.p, nbig_next, nu, owner.p, R,
                           C, O, Cn, pstat.p + (size_t)lev * pstride, flag.p + 2);
        for (int d = 0; d < 3; d++) {
            std::swap(args[d], nxt[d]);
            std::swap(cv[d], cvn[d]);
        }
        for (int k = 0; k < 4; k++) std::swap(crk[k], crn[k]);
    }
    KD_HIP(hipMemcpyAsync(hflag, flag.p, sizeof(hflag), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    KD_REQUIRE(!hflag[1], "tree_construct: internal error: a subtree exceeds the in-LDS build's size");
    KD_REQUIRE(!(hflag[2] & 1u), "tree_construct: internal error: a partition tile waited too long for the tiles before it");
    KD_REQUIRE(!(hflag[2] & 2u), "tree_construct: internal error: a partition place fell outside its segment");
    so.check(s);
}

}  // namespace kd
