// Spatial tiling of point sets for the pair kernels.
//
// A point set (the raw points of a catalogue for the query side, the
// deduplicated weighted nodes of a tree for the tree side) is put in Morton
// order (21 bits per axis inside the set's bounding box) and cut into tiles of
// at most `cap` consecutive points.  Tiles never straddle a Morton cell that
// holds about 8*cap points on average, which keeps every tile's bounding box
// compact.  A complete binary hierarchy of bounding boxes over the tree-side
// tiles drives the tile-pair enumeration.
//
// The reference has no counterpart: it walks its k-d tree once per query
// point and radius (src/kdquery.c:62-113).  Tiling only decides which pairs
// get evaluated; the counts stay those of the reference's node array.

#include <algorithm>
#include <cmath>
#include <cstring>

#include "kd_internal.h"

namespace kd {

static inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

// Points per occupied Morton cell, in tiles (make_tiles); measured on the
// 1M/10M and sky-patch workloads (2, 3.5 and 6 tried: 3.5 best on both)
static constexpr double kCellFill = 3.5;

__device__ __forceinline__ uint64_t ordkey_l(double v)
{
    uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

struct P3 {
    const double *p[3];
    int stride;
};

// min/max per axis as order-preserving keys (atomicMin/Max on u64)
__global__ void k_bbox(P3 P, int n, unsigned long long *__restrict__ keys6)
{
    unsigned long long mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int d = 0; d < 3; d++) {
            unsigned long long k = ordkey_l(P.p[d][(size_t)P.stride * i]);
            mn[d] = k < mn[d] ? k : mn[d];
            mx[d] = k > mx[d] ? k : mx[d];
        }
    }
    for (int d = 0; d < 3; d++) {
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long a = __shfl_down(mn[d], off, 64), b = __shfl_down(mx[d], off, 64);
            mn[d] = a < mn[d] ? a : mn[d];
            mx[d] = b > mx[d] ? b : mx[d];
        }
        if ((threadIdx.x & 63) == 0) {
            atomicMin(&keys6[d], mn[d]);
            atomicMax(&keys6[3 + d], mx[d]);
        }
    }
}

static double key_to_double(unsigned long long k)
{
    uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double v;
    memcpy(&v, &b, 8);
    return v;
}

__device__ __forceinline__ uint64_t spread3(uint32_t v)
{
    uint64_t x = v & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ void k_morton(P3 P, int n, double lx, double ly, double lz, double sx, double sy, double sz,
                         uint64_t *__restrict__ code, int *__restrict__ idx)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double lo[3] = {lx, ly, lz}, sc[3] = {sx, sy, sz};
    uint32_t q[3];
    for (int d = 0; d < 3; d++) {
        double t = (P.p[d][(size_t)P.stride * i] - lo[d]) * sc[d];
        t = t < 0.0 ? 0.0 : (t > 2097151.0 ? 2097151.0 : t);
        q[d] = (uint32_t)t;
    }
    code[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
    idx[i] = i;
}

// Histogram of the first differing bit of adjacent sorted Morton codes:
// hist[k] counts the neighbours whose codes first differ at bit k from the
// top of the 63-bit code, so a prefix of b bits cuts the points into
// 1 + hist[0] + ... + hist[b-1] occupied cells.
__global__ void k_split_bits(const uint64_t *__restrict__ code, int n, unsigned *__restrict__ hist)
{
    __shared__ unsigned h[64];
    if (threadIdx.x < 64) h[threadIdx.x] = 0u;
    __syncthreads();
    for (int j = blockIdx.x * blockDim.x + threadIdx.x + 1; j < n; j += gridDim.x * blockDim.x) {
        const uint64_t x = code[j] ^ code[j - 1];
        if (x) atomicAdd(&h[__clzll((long long)x) - 1], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

__global__ void k_cell_flags(const uint64_t *__restrict__ code, int n, int shift, int *__restrict__ flag)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    flag[j] = (j == 0) || ((code[j] >> shift) != (code[j - 1] >> shift));
}

__global__ void k_cell_starts(const int *__restrict__ flag, const int *__restrict__ incl, int n,
                              int *__restrict__ cstart)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (flag[j]) cstart[incl[j] - 1] = j;
    if (j == n - 1) cstart[incl[j]] = n;
}

__global__ void k_cell_ntiles(const int *__restrict__ cstart, int ncell, int cap, int *__restrict__ nt)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int len = cstart[c + 1] - cstart[c];
    nt[c] = (len + cap - 1) / cap;
}

__global__ void k_tile_starts(const int *__restrict__ cstart, const int *__restrict__ ntincl, int ncell, int n,
                              int *__restrict__ tstart)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int first = c ? ntincl[c - 1] : 0;
    int k = ntincl[c] - first;
    int s0 = cstart[c], len = cstart[c + 1] - s0;
    for (int t = 0; t < k; t++) tstart[first + t] = s0 + (int)(((long long)t * len) / k);
    if (c == ncell - 1) tstart[ntincl[c]] = n;
}

__global__ void k_tile_box(P3 P, const int *__restrict__ tstart, int ntiles, Box *__restrict__ box,
                           Box *__restrict__ hleaf)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    Box b;
    for (int d = 0; d < 3; d++) {
        b.lo[d] = INFINITY;
        b.hi[d] = -INFINITY;
    }
    for (int j = tstart[t]; j < tstart[t + 1]; j++)
        for (int d = 0; d < 3; d++) {
            double v = P.p[d][(size_t)P.stride * j];
            b.lo[d] = fmin(b.lo[d], v);
            b.hi[d] = fmax(b.hi[d], v);
        }
    box[t] = b;
    if (hleaf) hleaf[t] = b;
}

__global__ void k_empty_boxes(Box *__restrict__ b, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < 3; d++) {
        b[i].lo[d] = INFINITY;
        b[i].hi[d] = -INFINITY;
    }
}

__global__ void k_sub_int(int *__restrict__ a, const int *__restrict__ b, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] -= b[i];
}

__global__ void k_field_range(const int *__restrict__ f, const int *__restrict__ tstart, int ntiles,
                              int *__restrict__ fmin_, int *__restrict__ fmax_)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    int lo = 0x7fffffff, hi = (int)0x80000000;
    for (int j = tstart[t]; j < tstart[t + 1]; j++) {
        lo = min(lo, f[j]);
        hi = max(hi, f[j]);
    }
    fmin_[t] = lo;
    fmax_[t] = hi;
}

// tile-order gathers
__global__ void k_gather_query(const double *__restrict__ xyz, const int *__restrict__ fields,
                               const int *__restrict__ order, int n, double *__restrict__ qx, double *__restrict__ qy,
                               double *__restrict__ qz, int *__restrict__ qf)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int i = order[j];
    qx[j] = xyz[3 * (size_t)i];
    qy[j] = xyz[3 * (size_t)i + 1];
    qz[j] = xyz[3 * (size_t)i + 2];
    qf[j] = fields ? fields[i] : 0;
}

__global__ void k_gather_tree(const double *__restrict__ ux, const double *__restrict__ uy,
                              const double *__restrict__ uz, const int *__restrict__ uf, const int *__restrict__ uw,
                              const int *__restrict__ heap_of, const int *__restrict__ order, int n,
                              TPoint *__restrict__ pts)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int u = order[j];
    TPoint p;
    p.x = ux[u];
    p.y = uy[u];
    p.z = uz[u];
    p.meta = (uint32_t)(uw[u] & 0xFF) | ((uint32_t)(uf[u] & 0xFF) << 8);
    p.heap = heap_of[u];
    pts[j] = p;
}

// Tile centres (box midpoints, FP64) for the FP32 guarded kernel.
__global__ void k_centres(const Box *__restrict__ box, int ntiles, double *__restrict__ centre)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    for (int d = 0; d < 3; d++) centre[3 * t + d] = 0.5 * (box[t].lo[d] + box[t].hi[d]);
}

// Groups of 16 tile points in the operand layout of v_mfma_f32_16x16x4_f32
// (B: 4 x 16, lane l holds row l / 16, column l % 16): per group 64 floats,
// [k][j] = component k of point j relative to the tile centre (FP32 of the
// FP64 difference) for k < 3, and the squared norm of those FP32 components
// (summed in FP64, rounded once) for k = 3.  Points past the tile's end are
// padding: components 0 and norm +inf, so every distance to them is +inf.
__global__ void k_tile_groups(const int *__restrict__ tstart, int ntiles, int *__restrict__ ng)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    ng[t] = (tstart[t + 1] - tstart[t] + 15) / 16;
}

__global__ void k_fill_groups(const TPoint *__restrict__ pts, const int *__restrict__ tstart,
                              const int *__restrict__ gstart, int ntiles, const double *__restrict__ centre,
                              float *__restrict__ grp)
{
    const int t = blockIdx.x * (blockDim.x >> 4) + (threadIdx.x >> 4);  // 16 threads per tile
    if (t >= ntiles) return;
    const int j = threadIdx.x & 15;
    const double c0 = centre[3 * t], c1 = centre[3 * t + 1], c2 = centre[3 * t + 2];
    const int p0 = tstart[t], p1 = tstart[t + 1];
    for (int g = gstart[t]; g < gstart[t + 1]; g++) {
        const int i = p0 + 16 * (g - gstart[t]) + j;
        float bx = 0.f, by = 0.f, bz = 0.f, bn = INFINITY;
        if (i < p1) {
            const TPoint p = pts[i];
            bx = (float)(p.x - c0);
            by = (float)(p.y - c1);
            bz = (float)(p.z - c2);
            bn = (float)((double)bx * bx + (double)by * by + (double)bz * bz);
        }
        float *o = grp + 64 * (size_t)g;
        o[j] = bx;
        o[16 + j] = by;
        o[32 + j] = bz;
        o[48 + j] = bn;
    }
}

// Size class of each tree tile for the FP32 form's guard band (bits 16..19
// of the tile meta): the smallest k in [0, 15] with half-diagonal <=
// hmax 2^(k-15), so every point of a class-k tile lies within hmax 2^(k-15) of
// the tile centre.  (Runs of tiles for the FP32 form merge only within one
// class; the FP64 form's enumeration ignores these bits.)
__global__ void k_tile_class(const Box *__restrict__ box, int ntiles, double hmax, unsigned *__restrict__ meta)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    double s = 0.0;
    for (int d = 0; d < 3; d++) {
        const double w = 0.5 * (box[t].hi[d] - box[t].lo[d]);
        s += w * w;
    }
    const double h = sqrt(s) * (1.0 + 1e-12);
    int k = 0;
    while (k < 15 && h > hmax * ldexp(1.0, k - 15)) k++;
    meta[t] |= (unsigned)k << 16;
}

// Largest half-diagonal of the tile boxes: every tree point lies within this
// distance of its tile's centre (the box midpoint).
__global__ void k_max_halfdiag(const Box *__restrict__ box, int ntiles, unsigned long long *__restrict__ out)
{
    double h = 0.0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int d = 0; d < 3; d++) {
            const double w = 0.5 * (box[t].hi[d] - box[t].lo[d]);
            s += w * w;
        }
        h = fmax(h, sqrt(s) * (1.0 + 1e-12));
    }
    // non-negative doubles order like their bit patterns
    atomicMax(out, (unsigned long long)__double_as_longlong(h));
}

__global__ void k_fields_outside_byte(const int *__restrict__ f, int n, unsigned *__restrict__ flag)
{
    bool bad = false;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        bad = bad || f[i] < 0 || f[i] > 255;
    if (__ballot(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

bool fields_in_byte(const int *d_fields, int n, hipStream_t s)
{
    if (!d_fields) return true;
    DevBuf<unsigned> flag(1);
    unsigned h = 0;
    KD_HIP(hipMemsetAsync(flag.p, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_fields_outside_byte, dim3(std::min(nblk(n), 2048)), dim3(256), 0, s, d_fields, n, flag.p);
    KD_HIP(hipMemcpyAsync(&h, flag.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    return h == 0;
}

// Morton order + tiles for n points; returns the order, fills Tiles.
static void make_tiles(P3 P, int n, int cap, Tiles &T, DevBuf<int> &order, Sorter &so,
                       hipStream_t s)
{
    T.n = n;
    T.cap = cap;
    DevBuf<unsigned long long> keys6(6);
    unsigned long long init[6] = {~0ull, ~0ull, ~0ull, 0, 0, 0};
    KD_HIP(hipMemcpyAsync(keys6.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_bbox, dim3(std::min(nblk(n), 2048)), dim3(256), 0, s, P, n, keys6.p);
    unsigned long long hk[6];
    KD_HIP(hipMemcpyAsync(hk, keys6.p, sizeof(hk), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    double lo[3], sc[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = key_to_double(hk[d]);
        double hi = key_to_double(hk[3 + d]);
        double ext = hi - lo[d];
        sc[d] = ext > 0 ? 2097152.0 / ext : 0.0;
        T.cmax = std::max(T.cmax, std::max(std::fabs(lo[d]), std::fabs(hi)));
    }
    DevBuf<uint64_t> code(n), code2(n);
    DevBuf<int> idx(n);
    order.alloc(n);
    hipLaunchKernelGGL(k_morton, dim3(nblk(n)), dim3(256), 0, s, P, n, lo[0], lo[1], lo[2], sc[0], sc[1], sc[2],
                       code.p, idx.p);
    so.pairs(code.p, code2.p, idx.p, order.p, n, s);

    // Cells: the longest Morton-code prefix (any number of bits) whose
    // occupied cells still hold kCellFill * cap points on average.  Counting
    // occupied cells, not the cells of the bounding box, gives data that
    // fills only a thin part of its box (a patch of the sky, a shell) small,
    // compact cells, and uniform data the same well-filled tiles as octree
    // cells of ~4 tiles.
    int shift = 63;
    {
        DevBuf<unsigned> bh(64);
        KD_HIP(hipMemsetAsync(bh.p, 0, bh.bytes(), s));
        hipLaunchKernelGGL(k_split_bits, dim3(std::max(1, std::min(nblk(n), 1024))), dim3(256), 0, s, code2.p, n,
                           bh.p);
        unsigned hb[64];
        KD_HIP(hipMemcpyAsync(hb, bh.p, sizeof(hb), hipMemcpyDeviceToHost, s));
        so.check(s);  // waits for s
        long long cells = 1;
        for (int b = 1; b <= 63; b++) {
            cells += hb[b - 1];
            if ((double)n < kCellFill * cap * (double)cells) break;
            shift = 63 - b;
        }
    }
    DevBuf<int> flag(n), incl(n);
    hipLaunchKernelGGL(k_cell_flags, dim3(nblk(n)), dim3(256), 0, s, code2.p, n, shift, flag.p);
    so.incl_scan(flag.p, incl.p, n, s);
    int ncell = 0;
    KD_HIP(hipMemcpyAsync(&ncell, incl.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    DevBuf<int> cstart(ncell + 1), nt(ncell), ntincl(ncell);
    hipLaunchKernelGGL(k_cell_starts, dim3(nblk(n)), dim3(256), 0, s, flag.p, incl.p, n, cstart.p);
    hipLaunchKernelGGL(k_cell_ntiles, dim3(nblk(ncell)), dim3(256), 0, s, cstart.p, ncell, cap, nt.p);
    so.incl_scan(nt.p, ntincl.p, ncell, s);
    int ntiles = 0;
    KD_HIP(hipMemcpyAsync(&ntiles, ntincl.p + (ncell - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    T.ntiles = ntiles;
    T.start.alloc(ntiles + 1);
    hipLaunchKernelGGL(k_tile_starts, dim3(nblk(ncell)), dim3(256), 0, s, cstart.p, ntincl.p, ncell, n, T.start.p);
    T.box.alloc(ntiles);
    (void)P;
}

void layout_query(kd_device_tree *T, const double *d_xyz, const int *d_fields, int n, int cap, Sorter &so,
                  hipStream_t s)
{
    QuerySide &Q = T->q;
    P3 raw{{d_xyz, d_xyz + 1, d_xyz + 2}, 3};
    make_tiles(raw, n, cap, Q.t, Q.order, so, s);
    Q.x.alloc(n);
    Q.y.alloc(n);
    Q.z.alloc(n);
    Q.f.alloc(n);
    hipLaunchKernelGGL(k_gather_query, dim3(nblk(n)), dim3(256), 0, s, d_xyz, d_fields, Q.order.p, n, Q.x.p, Q.y.p,
                       Q.z.p, Q.f.p);
    P3 tiled{{Q.x.p, Q.y.p, Q.z.p}, 1};
    int nt = Q.t.ntiles;
    hipLaunchKernelGGL(k_tile_box, dim3(nblk(nt)), dim3(256), 0, s, tiled, Q.t.start.p, nt, Q.t.box.p, (Box *)nullptr);
    Q.fmin.alloc(nt);
    Q.fmax.alloc(nt);
    hipLaunchKernelGGL(k_field_range, dim3(nblk(nt)), dim3(256), 0, s, Q.f.p, Q.t.start.p, nt, Q.fmin.p, Q.fmax.p);
    KD_HIP(hipStreamSynchronize(s));
}

__global__ void k_tree_raw(const int *__restrict__ torder, const int *__restrict__ rep, int nu, int *__restrict__ raw)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nu) raw[j] = rep[torder[j]];
}

__device__ __forceinline__ unsigned long long lanes_upto(int lane)  // lanes 0..lane-1
{
    return lane ? ~0ull >> (64 - lane) : 0ull;
}

// Runs of one field among the points of a tile (grouped by field): bit k of
// the result is set when point k starts a run.  One wave per tile, lane k
// holding point k (tiles hold at most 64 points).
__device__ __forceinline__ unsigned long long field_runs(const TPoint *__restrict__ pts, int j0, int len, int lane,
                                                         unsigned &f)
{
    f = lane < len ? (pts[j0 + lane].meta >> 8) & 0xFFu : 0u;
    const unsigned prev = __shfl_up(f, 1, 64);
    return __ballot(lane < len && (lane == 0 || prev != f));
}

// Quads of each tile: every run of one field rounded up to whole quads.
__global__ __launch_bounds__(256) void k_tile_quads(const TPoint *__restrict__ pts, const int *__restrict__ tstart,
                                                    int ntiles, int *__restrict__ nq)
{
    const int lane = threadIdx.x & 63, t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ntiles) return;
    const int j0 = tstart[t], len = tstart[t + 1] - j0;
    unsigned f;
    unsigned long long runs = field_runs(pts, j0, len, lane, f);
    int q = 0;
    while (runs) {
        const int a = __ffsll((long long)runs) - 1;
        runs &= runs - 1;
        const int b = runs ? __ffsll((long long)runs) - 1 : len;
        q += (b - a + 3) >> 2;
    }
    if (lane == 0) nq[t] = q;
}

// Point k of run [a, b) of tile t goes to slot (k - a) % 4 of quad
// qs[t] + (quads of the runs before) + (k - a) / 4; the slots after a run's
// last point hold +inf.  Per quad: its field and point count, and the
// bounding box of its points.  One wave per tile.
//
// The FP32 operands of each slot (q32, 4 floats): b = p - c rounded to FP32
// from the FP64 difference (c the tile's box midpoint), and |b|^2 of those
// FP32 components summed in FP64 and rounded once; padding slots hold
// {0, 0, 0, +inf}, so every FP32 distance to them is +inf.
__global__ __launch_bounds__(256) void k_fill_quads(const TPoint *__restrict__ pts, const int *__restrict__ tstart,
                                                    const int *__restrict__ qs, int ntiles,
                                                    const double *__restrict__ centre, double *__restrict__ quad,
                                                    unsigned *__restrict__ qmeta, BoxF *__restrict__ qbox,
                                                    float *__restrict__ q32)
{
    const int lane = threadIdx.x & 63, t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ntiles) return;
    const int j0 = tstart[t], len = tstart[t + 1] - j0;
    unsigned f;
    unsigned long long runs = field_runs(pts, j0, len, lane, f);
    int a = 0, b = 0, qoff = 0, q = qs[t];
    while (runs) {  // the run holding this lane's point, and its first quad
        const int ra = __ffsll((long long)runs) - 1;
        runs &= runs - 1;
        const int rb = runs ? __ffsll((long long)runs) - 1 : len;
        if (lane >= ra && lane < rb) {
            a = ra;
            b = rb;
            qoff = q;
        }
        q += (rb - ra + 3) >> 2;
    }
    if (lane >= len) return;
    const TPoint p = pts[j0 + lane];
    const int k = lane - a, g = qoff + (k >> 2), u = k & 3;
    quad[12LL * g + u] = p.x;
    quad[12LL * g + 4 + u] = p.y;
    quad[12LL * g + 8 + u] = p.z;
    {
        const float bx = (float)(p.x - centre[3 * t]), by = (float)(p.y - centre[3 * t + 1]),
                    bz = (float)(p.z - centre[3 * t + 2]);
        const float bn = (float)((double)bx * bx + (double)by * by + (double)bz * bz);
        *(float4 *)&q32[16LL * g + 4 * u] = make_float4(bx, by, bz, bn);
    }
    if (u == 0) {
        // the quad's points are this lane's and the next up to three of the run
        const int np = min(4, b - lane);
        qmeta[g] = f | ((unsigned)np << 8);
        Box bx = {{p.x, p.y, p.z}, {p.x, p.y, p.z}};
        for (int v = 1; v < np; v++) {
            const TPoint o = pts[j0 + lane + v];
            const double c[3] = {o.x, o.y, o.z};
            for (int d = 0; d < 3; d++) {
                bx.lo[d] = fmin(bx.lo[d], c[d]);
                bx.hi[d] = fmax(bx.hi[d], c[d]);
            }
        }
        BoxF bf;
        for (int d = 0; d < 3; d++) {
            bf.lo[d] = __double2float_rd(bx.lo[d]);
            bf.hi[d] = __double2float_ru(bx.hi[d]);
        }
        qbox[g] = bf;
    }
    if (lane == b - 1)
        for (int v = u + 1; v < 4; v++) {
            quad[12LL * g + v] = INFINITY;
            quad[12LL * g + 4 + v] = INFINITY;
            quad[12LL * g + 8 + v] = INFINITY;
            *(float4 *)&q32[16LL * g + 4 * v] = make_float4(0.f, 0.f, 0.f, INFINITY);
        }
}

// The points of each tree tile grouped by their 8-bit field, in increasing
// field order and stably (Morton order within a field): a tile that straddles
// a field boundary then holds one contiguous run per field.  One wave per
// tile; the counts do not depend on the order of the points within a tile.
__global__ __launch_bounds__(256) void k_tile_field_order(const int *__restrict__ tstart, int ntiles,
                                                          const int *__restrict__ uf, const int *__restrict__ in,
                                                          int *__restrict__ out)
{
    const int lane = threadIdx.x & 63, t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ntiles) return;
    const int j0 = tstart[t], len = tstart[t + 1] - j0;
    const bool v = lane < len;
    const int u = v ? in[j0 + lane] : 0;
    const int f = v ? (uf[u] & 0xFF) : 0x7fffffff;
    int rank = 0, base = 0;
    unsigned long long rem = __ballot(v);
    while (rem) {  // the smallest field left, then its lanes in lane order
        int m = ((rem >> lane) & 1ull) ? f : 0x7fffffff;
        for (int off = 32; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off, 64));
        const unsigned long long mk = __ballot(v && f == m);
        if (v && f == m) rank = base + __popcll(mk & lanes_upto(lane));
        base += __popcll(mk);
        rem &= ~mk;
    }
    if (v) out[j0 + rank] = u;
}

__global__ void k_tile_meta(const TPoint *__restrict__ pts, const int *__restrict__ tstart, int ntiles,
                            unsigned *__restrict__ meta)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    int j0 = tstart[t], j1 = tstart[t + 1];
    unsigned f0 = j1 > j0 ? (pts[j0].meta >> 8) & 0xFF : 0;
    bool pure = true, w1 = true;
    for (int j = j0; j < j1; j++) {
        unsigned m = pts[j].meta;
        pure = pure && ((m >> 8) & 0xFF) == f0;
        w1 = w1 && (m & 0xFF) == 1;
    }
    meta[t] = (pure ? 1u : 0u) | (w1 ? 2u : 0u) | (f0 << 8);
}

// Leaves of the field/weight hierarchy: tile t's field (kHMixed when the
// tile holds several) and weight; padding leaves are empty.
__global__ void k_leaf_fields(const TPoint *__restrict__ pts, const int *__restrict__ tstart,
                              const unsigned *__restrict__ meta, int ntiles, int P, unsigned *__restrict__ hfield,
                              long long *__restrict__ hweight)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P) return;
    unsigned f = kHEmpty;
    long long w = 0;
    if (t < ntiles) {
        for (int j = tstart[t]; j < tstart[t + 1]; j++) w += pts[j].meta & 0xFF;
        f = (meta[t] & 1u) ? (meta[t] >> 8) & 0xFF : kHMixed;
        if (tstart[t + 1] == tstart[t]) f = kHEmpty;
    }
    hfield[P + t] = f;
    hweight[P + t] = w;
}

// Per-field weights of the hierarchy nodes that span 128 tile slots (the
// deepest level kept): one wave per node, an LDS histogram over its points.
__global__ __launch_bounds__(256) void k_node_fields_base(const TPoint *__restrict__ pts,
                                                          const int *__restrict__ tstart, int ntiles, int P, int nfw,
                                                          long long *__restrict__ hfw, unsigned *__restrict__ bad)
{
    __shared__ unsigned long long h[4][256];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int o = P / 128 + blockIdx.x * 4 + wid;  // node at the 128-slot level
    for (int f = lane; f < 256; f += 64) h[wid][f] = 0ull;
    __syncthreads();
    if (o < P / 64) {
        const int t0 = min(o * 128 - P, ntiles), t1 = min(o * 128 + 128 - P, ntiles);
        const int j0 = tstart[t0], j1 = tstart[t1];
        for (int j = j0 + lane; j < j1; j += 64) {
            const unsigned m = pts[j].meta;
            const int f = (int)((m >> 8) & 0xFF);
            if (f < nfw)
                atomicAdd(&h[wid][f], (unsigned long long)(m & 0xFF));
            else
                atomicOr(bad, 1u);
        }
    }
    __syncthreads();
    if (o < P / 64)
        for (int f = lane; f < nfw; f += 64) hfw[(size_t)o * nfw + f] = (long long)h[wid][f];
}

// Bottom-up reduction of a heap-ordered binary hierarchy: h[P .. 2P-1] are
// the leaves (P a power of two) and node o takes combine(h[2o], h[2o+1]).
// The lowest levels run one subtree of up to 2^kHeapLevels leaves per
// workgroup through LDS (k_heap_sub); the nodes above the subtrees' roots
// (fewer than P / 2^kHeapLevels) run in one workgroup (k_heap_top).
static constexpr int kHeapLevels = 9;

struct BoxHeap {
    Box *h;
    using T = Box;
    __device__ T load(int o) const { return h[o]; }
    __device__ void store(int o, const T &v) const { h[o] = v; }
    __device__ static T combine(const T &a, const T &b)
    {
        Box c;
        for (int d = 0; d < 3; d++) {
            c.lo[d] = fmin(a.lo[d], b.lo[d]);
            c.hi[d] = fmax(a.hi[d], b.hi[d]);
        }
        return c;
    }
};

// The field of a node (kHEmpty without points, kHMixed over several fields)
// and its summed weight.
struct FieldHeap {
    unsigned *f;
    long long *w;
    struct T {
        unsigned f;
        long long w;
    };
    __device__ T load(int o) const { return {f[o], w[o]}; }
    __device__ void store(int o, const T &v) const
    {
        f[o] = v.f;
        w[o] = v.w;
    }
    __device__ static T combine(const T &a, const T &b)
    {
        return {a.f == kHEmpty ? b.f : (b.f == kHEmpty || a.f == b.f ? a.f : kHMixed), a.w + b.w};
    }
};

template <class H>
__global__ __launch_bounds__(256) void k_heap_sub(H hp, int P, int levels)
{
    __shared__ typename H::T sv[256];
    const int tid = threadIdx.x;
    int m = 1 << (levels - 1);                                // nodes of the level above the leaves
    int base = (P >> 1) + blockIdx.x * m;                     // its first node in this subtree
    typename H::T v;
    if (tid < m) {
        v = H::combine(hp.load(2 * (base + tid)), hp.load(2 * (base + tid) + 1));
        hp.store(base + tid, v);
        sv[tid] = v;
    }
    __syncthreads();
    for (m >>= 1; m >= 1; m >>= 1) {
        base >>= 1;
        if (tid < m) v = H::combine(sv[2 * tid], sv[2 * tid + 1]);
        __syncthreads();
        if (tid < m) {
            sv[tid] = v;
            hp.store(base + tid, v);
        }
        __syncthreads();
    }
}

template <class H>
__global__ __launch_bounds__(256) void k_heap_top(H hp, int lo)
{
    for (; lo >= 1; lo >>= 1) {  // levels lo, lo/2, ..., 1
        for (int i = threadIdx.x; i < lo; i += blockDim.x) {
            const int o = lo + i;
            hp.store(o, H::combine(hp.load(2 * o), hp.load(2 * o + 1)));
        }
        __syncthreads();
    }
}

template <class H>
static void heap_up(H hp, int P, hipStream_t s)
{
    if (P < 2) return;
    int levels = 0;
    while ((1 << levels) < P && levels < kHeapLevels) levels++;
    hipLaunchKernelGGL(k_heap_sub<H>, dim3(P >> levels), dim3(256), 0, s, hp, P, levels);
    if ((P >> levels) > 1) hipLaunchKernelGGL(k_heap_top<H>, dim3(1), dim3(256), 0, s, hp, (P >> levels) / 2);
}

// Per-field weights of the hierarchy nodes above the 128-slot level, from
// it: one workgroup per field, level by level (levels lo, lo/2, ..., 1).
__global__ __launch_bounds__(256) void k_node_fields_up(long long *__restrict__ hfw, int lo, int nfw)
{
    const int f = blockIdx.x;
    for (; lo >= 1; lo >>= 1) {
        for (int i = threadIdx.x; i < lo; i += blockDim.x) {
            const size_t o = (size_t)(lo + i);
            hfw[o * nfw + f] = hfw[2 * o * nfw + f] + hfw[(2 * o + 1) * nfw + f];
        }
        __syncthreads();
    }
}

__global__ void k_fill_m1(int *a, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = -1;
}

__global__ void k_mark_reps(const int *__restrict__ rep, int nu, int *__restrict__ rep_of_raw)
{
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < nu) rep_of_raw[rep[u]] = u;
}

__global__ void k_rep_flags(const int *__restrict__ qorder, const int *__restrict__ rep_of_raw, int n,
                            int *__restrict__ flag)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) flag[j] = rep_of_raw[qorder[j]] >= 0;
}

__global__ void k_tree_order(const int *__restrict__ qorder, const int *__restrict__ rep_of_raw,
                             const int *__restrict__ excl, int n, int *__restrict__ torder)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int u = rep_of_raw[qorder[j]];
    if (u >= 0) torder[excl[j]] = u;
}

__global__ void k_tree_starts(const int *__restrict__ qstart, const int *__restrict__ excl, int ntiles, int n, int nu,
                              int *__restrict__ tstart)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    int j = qstart[t];
    tstart[t] = j < n ? excl[j] : nu;
}

// The tree side of a catalogue: each unique point sits where its
// representative raw point sits in the query-side tile order, so tree tile t
// and query tile t cover the same region (the same points when there are no
// duplicates).
void layout_tree(kd_device_tree *T, BuildExtras &X, int cap, Sorter &so, hipStream_t s)
{
    (void)cap;
    TreeSide &R = T->tr;
    const QuerySide &Q = T->q;
    const int n = T->n_raw, nu = T->n_u, nt = Q.t.ntiles;
    DevBuf<int> rep_of_raw(n), flag(n), incl(n), torder(nu);
    hipLaunchKernelGGL(k_fill_m1, dim3(nblk(n)), dim3(256), 0, s, rep_of_raw.p, n);
    hipLaunchKernelGGL(k_mark_reps, dim3(nblk(nu)), dim3(256), 0, s, X.rep.p, nu, rep_of_raw.p);
    hipLaunchKernelGGL(k_rep_flags, dim3(nblk(n)), dim3(256), 0, s, Q.order.p, rep_of_raw.p, n, flag.p);
    so.incl_scan(flag.p, incl.p, n, s);
    // exclusive scan = inclusive - flag
    hipLaunchKernelGGL(k_sub_int, dim3(nblk(n)), dim3(256), 0, s, incl.p, flag.p, n);
    hipLaunchKernelGGL(k_tree_order, dim3(nblk(n)), dim3(256), 0, s, Q.order.p, rep_of_raw.p, incl.p, n, torder.p);
    R.t.n = nu;
    R.t.cap = Q.t.cap;
    R.t.ntiles = nt;
    R.t.cmax = Q.t.cmax;
    R.t.start.alloc(nt + 1);
    hipLaunchKernelGGL(k_tree_starts, dim3(nblk(nt + 1)), dim3(256), 0, s, Q.t.start.p, incl.p, nt, n, nu,
                       R.t.start.p);
    {
        DevBuf<int> grouped(nu);
        hipLaunchKernelGGL(k_tile_field_order, dim3((nt + 3) / 4), dim3(256), 0, s, R.t.start.p, nt, X.uf.p,
                           torder.p, grouped.p);
        std::swap(torder, grouped);
    }
    R.pts.alloc(nu);
    hipLaunchKernelGGL(k_gather_tree, dim3(nblk(nu)), dim3(256), 0, s, X.ux.p, X.uy.p, X.uz.p, X.uf.p, X.uw.p,
                       X.heap_of.p, torder.p, nu, R.pts.p);
    int Pn = 1;
    while (Pn < nt) Pn <<= 1;
    R.t.P = Pn;
    R.t.box.alloc(nt);
    R.t.hbox.alloc(2 * (size_t)Pn);
    hipLaunchKernelGGL(k_empty_boxes, dim3(nblk(2ll * Pn)), dim3(256), 0, s, R.t.hbox.p, 2 * Pn);
    P3 tiled{{&R.pts.p[0].x, &R.pts.p[0].y, &R.pts.p[0].z}, 4};
    hipLaunchKernelGGL(k_tile_box, dim3(nblk(nt)), dim3(256), 0, s, tiled, R.t.start.p, nt, R.t.box.p,
                       R.t.hbox.p + Pn);
    heap_up(BoxHeap{R.t.hbox.p}, Pn, s);
    R.centre.alloc(3 * (size_t)nt);
    hipLaunchKernelGGL(k_centres, dim3(nblk(nt)), dim3(256), 0, s, R.t.box.p, nt, R.centre.p);
    {
        // the FP32 form's operand groups and the half-diagonal bound of its guard band
        DevBuf<int> ng(nt);
        DevBuf<unsigned long long> hbits(1);
        R.gstart.alloc(nt + 1);
        hipLaunchKernelGGL(k_tile_groups, dim3(nblk(nt)), dim3(256), 0, s, R.t.start.p, nt, ng.p);
        so.incl_scan(ng.p, R.gstart.p + 1, nt, s);
        KD_HIP(hipMemsetAsync(R.gstart.p, 0, sizeof(int), s));
        KD_HIP(hipMemsetAsync(hbits.p, 0, sizeof(unsigned long long), s));
        hipLaunchKernelGGL(k_max_halfdiag, dim3(std::min(nblk(nt), 1024)), dim3(256), 0, s, R.t.box.p, nt, hbits.p);
        int ngroups = 0;
        unsigned long long hb = 0;
        KD_HIP(hipMemcpyAsync(&ngroups, R.gstart.p + nt, sizeof(int), hipMemcpyDeviceToHost, s));
        KD_HIP(hipMemcpyAsync(&hb, hbits.p, sizeof(hb), hipMemcpyDeviceToHost, s));
        KD_HIP(hipStreamSynchronize(s));
        memcpy(&R.t.hmax, &hb, sizeof(double));
        R.grp.alloc(64 * (size_t)std::max(ngroups, 1));
        hipLaunchKernelGGL(k_fill_groups, dim3((nt + 15) / 16), dim3(256), 0, s, R.pts.p, R.t.start.p, R.gstart.p,
                           nt, R.centre.p, R.grp.p);
    }
    R.raw.alloc(nu);
    hipLaunchKernelGGL(k_tree_raw, dim3(nblk(nu)), dim3(256), 0, s, torder.p, X.rep.p, nu, R.raw.p);
    {
        DevBuf<int> nq(nt);
        R.quad_start.alloc(nt + 1);
        hipLaunchKernelGGL(k_tile_quads, dim3((nt + 3) / 4), dim3(256), 0, s, R.pts.p, R.t.start.p, nt, nq.p);
        so.incl_scan(nq.p, R.quad_start.p + 1, nt, s);
        KD_HIP(hipMemsetAsync(R.quad_start.p, 0, sizeof(int), s));
        int nquads = 0;
        KD_HIP(hipMemcpyAsync(&nquads, R.quad_start.p + nt, sizeof(int), hipMemcpyDeviceToHost, s));
        KD_HIP(hipStreamSynchronize(s));
        R.quad.alloc(12 * (size_t)std::max(nquads, 1));
        R.quad_meta.alloc(std::max(nquads, 1));
        R.quad_box.alloc(std::max(nquads, 1));
        R.quad32.alloc(16 * (size_t)std::max(nquads, 1));
        hipLaunchKernelGGL(k_fill_quads, dim3((nt + 3) / 4), dim3(256), 0, s, R.pts.p, R.t.start.p,
                           R.quad_start.p, nt, R.centre.p, R.quad.p, R.quad_meta.p, R.quad_box.p, R.quad32.p);
    }
    R.meta.alloc(nt);
    hipLaunchKernelGGL(k_tile_meta, dim3(nblk(nt)), dim3(256), 0, s, R.pts.p, R.t.start.p, nt, R.meta.p);
    hipLaunchKernelGGL(k_tile_class, dim3(nblk(nt)), dim3(256), 0, s, R.t.box.p, nt, R.t.hmax, R.meta.p);
    R.hfield.alloc(2 * (size_t)Pn);
    R.hweight.alloc(2 * (size_t)Pn);
    hipLaunchKernelGGL(k_leaf_fields, dim3(nblk(Pn)), dim3(256), 0, s, R.pts.p, R.t.start.p, R.meta.p, nt, Pn,
                       R.hfield.p, R.hweight.p);
    heap_up(FieldHeap{R.hfield.p, R.hweight.p}, Pn, s);
    // per-field weights of the large subtrees, for bulk items over subtrees
    // that hold several fields
    R.nfw = 0;
    unsigned hbad = 1u;
    DevBuf<unsigned> bad;
    const int nfw = std::min(256, std::max(1, T->num_fields));
    if (Pn >= 128) {
        R.hfw.alloc((size_t)(Pn / 64) * nfw);
        bad.alloc(1);
        KD_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned), s));
        hipLaunchKernelGGL(k_node_fields_base, dim3((Pn / 128 + 3) / 4), dim3(256), 0, s, R.pts.p, R.t.start.p, nt, Pn,
                           nfw, R.hfw.p, bad.p);
        if (Pn >= 256) hipLaunchKernelGGL(k_node_fields_up, dim3(nfw), dim3(256), 0, s, R.hfw.p, Pn / 256, nfw);
        KD_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    }
    KD_HIP(hipStreamSynchronize(s));
    if (Pn >= 128 && !hbad) R.nfw = nfw;
}

}  // namespace kd
