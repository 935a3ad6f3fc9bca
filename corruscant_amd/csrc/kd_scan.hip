// Prefix sums, the radix sort and buffer clears of the library, written for
// wave64.
//
// A scan of n elements runs in three launches (reduce, then scan):
//   1. every workgroup sums its tile of kScanTile elements into part[b];
//   2. one workgroup turns part[] into its exclusive prefix sums;
//   3. every workgroup scans its tile again, starting from part[b].
// Each tile passes through LDS once per launch: loads and stores stay
// coalesced (consecutive lanes, consecutive elements) while each lane scans
// kScanPer consecutive elements.  Input and output may be the same buffer.
// The temporary part[] holds one element per tile (scan_parts(n)); the
// caller provides it, so a scan can be recorded into a graph with every
// buffer fixed.
//
// zero_async clears a buffer with a kernel, which a graph replays like any
// other kernel node.
#include "kd_internal.h"

namespace kd {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanTile = kScanThreads * kScanPer;
constexpr int kScanWaves = kScanThreads / 64;

template <typename T>
__device__ inline T wave_incl(T v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// Exclusive prefix of v over the workgroup; total = the workgroup's sum.
template <typename T>
__device__ inline T block_excl(T v, T *wsum, T &total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const T inc = wave_incl(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    T before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kScanWaves; w++) {
        const T x = wsum[w];
        if (w < wid) before += x;
        total += x;
    }
    __syncthreads();
    return before + inc - v;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const T *__restrict__ in, long long n,
                                                              T *__restrict__ part)
{
    __shared__ T wsum[kScanWaves];
    const long long base = (long long)blockIdx.x * kScanTile;
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        const long long i = base + k * kScanThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    T total;
    (void)block_excl(s, wsum, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// Scan of one tile held in LDS (tile[kScanTile]): exclusive or inclusive,
// plus carry; returns the tile's sum.
template <typename T>
__device__ inline T tile_scan(T *tile, T *wsum, T carry, bool inclusive)
{
    T v[kScanPer], s = 0;
    const int b = threadIdx.x * kScanPer;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        v[k] = tile[b + k];
        s += v[k];
    }
    T total;
    T run = block_excl(s, wsum, total) + carry;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        const T x = v[k];
        tile[b + k] = inclusive ? run + x : run;
        run += x;
    }
    __syncthreads();
    return total;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_parts(T *__restrict__ part, int nb)
{
    __shared__ T tile[kScanTile];
    __shared__ T wsum[kScanWaves];
    T carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kScanTile) {
#pragma unroll
        for (int k = 0; k < kScanPer; k++) {
            const int i = b0 + k * kScanThreads + threadIdx.x;
            tile[k * kScanThreads + threadIdx.x] = i < nb ? part[i] : T(0);
        }
        __syncthreads();
        const T total = tile_scan(tile, wsum, carry, false);
#pragma unroll
        for (int k = 0; k < kScanPer; k++) {
            const int i = b0 + k * kScanThreads + threadIdx.x;
            if (i < nb) part[i] = tile[k * kScanThreads + threadIdx.x];
        }
        carry += total;
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const T *in, T *out, long long n,
                                                             const T *__restrict__ part, int inclusive)
{
    __shared__ T tile[kScanTile];
    __shared__ T wsum[kScanWaves];
    const long long base = (long long)blockIdx.x * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        const long long i = base + k * kScanThreads + threadIdx.x;
        tile[k * kScanThreads + threadIdx.x] = i < n ? in[i] : T(0);
    }
    __syncthreads();
    (void)tile_scan(tile, wsum, part[blockIdx.x], inclusive != 0);
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        const long long i = base + k * kScanThreads + threadIdx.x;
        if (i < n) out[i] = tile[k * kScanThreads + threadIdx.x];
    }
}

__global__ void k_zero32(unsigned *__restrict__ p, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

template <typename T>
void scan(const T *in, T *out, long long n, bool inclusive, T *part, hipStream_t s)
{
    if (n <= 0) return;
    KD_REQUIRE(n < (1LL << 40), "scan: too many elements");
    const long long nb = (n + kScanTile - 1) / kScanTile;
    KD_REQUIRE(nb < 0x7fffffffLL, "scan: too many tiles");
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, part);
    hipLaunchKernelGGL(k_scan_parts<T>, dim3(1), dim3(kScanThreads), 0, s, part, (int)nb);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, out, n,
                       (const T *)part, inclusive ? 1 : 0);
    KD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (64-bit key, 32-bit value) pairs, 8-bit digits.
//
// One launch counts all eight digit histograms; then one launch per digit
// moves the pairs (single pass per digit): tiles of kRsTile pairs are taken
// in launch order from a counter.  Each wave ranks its kRsPer x 64 pairs by
// digit in position order (lanes with the same digit found by eight ballots,
// per-wave digit counters in LDS), the workgroup adds up its digit counts,
// publishes them, and looks back over the tiles before it for the number of
// pairs of each digit they hold (a tile waits only on tiles taken before it,
// which publish before they look back; a wait past kRsSpins polls sets *err
// instead of hanging).  The tile is then sorted by digit in LDS and written
// out, so consecutive lanes write consecutive places of one digit.
// ---------------------------------------------------------------------------
constexpr int kRsThreads = 256;
constexpr int kRsWaves = kRsThreads / 64;
constexpr int kRsPer = 16;
constexpr int kRsTile = kRsThreads * kRsPer;
constexpr unsigned kRsSpins = 1u << 24;

__global__ __launch_bounds__(256) void k_rs_hist(const uint64_t *__restrict__ keys, int n, unsigned *__restrict__ hist)
{
    __shared__ unsigned h[8 * 256];
    for (int j = threadIdx.x; j < 8 * 256; j += blockDim.x) h[j] = 0u;
    __syncthreads();
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
#pragma unroll
        for (int p = 0; p < 8; p++) atomicAdd(&h[p * 256 + (int)((k >> (8 * p)) & 255u)], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < 8 * 256; j += blockDim.x)
        if (h[j]) atomicAdd(&hist[j], h[j]);
}

// Exclusive prefix over the workgroup of one value per thread (256 threads).
__device__ __forceinline__ unsigned rs_block_excl(unsigned v, unsigned *wsum)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned inc = wave_incl(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    unsigned before = 0u;
#pragma unroll
    for (int w = 0; w < kRsWaves; w++)
        if (w < wid) before += wsum[w];
    __syncthreads();
    return before + inc - v;
}

__global__ __launch_bounds__(kRsThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_rs_pass(const uint64_t *__restrict__ kin, const int *__restrict__ vin,
                                                        uint64_t *__restrict__ kout, int *__restrict__ vout, int n,
                                                        int shift, const unsigned *__restrict__ hist,
                                                        unsigned long long *__restrict__ stat,
                                                        unsigned *__restrict__ err)
{
    __shared__ uint64_t s_key[kRsTile];  // the tile by digit: keys, then (as ints) values
    __shared__ unsigned char s_dg[kRsTile];
    int *s_val = (int *)s_key;
    __shared__ unsigned s_wcnt[kRsWaves][256];
    __shared__ unsigned s_start[256], s_base[256];
    __shared__ unsigned s_wsum[kRsWaves];
    __shared__ int s_tile;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_tile = (int)atomicAdd(stat, 1ull);
    for (int w = 0; w < kRsWaves; w++) s_wcnt[w][threadIdx.x] = 0u;
    __syncthreads();
    const int tile = s_tile;
    const long long wbase = (long long)tile * kRsTile + wid * (kRsPer * 64);

    // rank inside the wave, in position order (positions past n sort last)
    uint64_t key[kRsPer];
    int val[kRsPer];
    unsigned rank[kRsPer];
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kRsPer; k++) {
        const long long i = wbase + k * 64 + lane;
        key[k] = i < n ? kin[i] : ~0ull;
        val[k] = i < n ? vin[i] : 0;
    }
#pragma unroll
    for (int k = 0; k < kRsPer; k++) {
        const unsigned dg = (unsigned)(key[k] >> shift) & 255u;
        unsigned long long peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (dg >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const unsigned before = s_wcnt[wid][dg];
        rank[k] = before + (unsigned)__popcll(peers & below);
        if (!(peers & below)) s_wcnt[wid][dg] = before + (unsigned)__popcll(peers);
    }
    __syncthreads();

    // thread d: digit d's count per wave -> offsets, the tile's count, and
    // where digit d starts inside the tile
    const int d = threadIdx.x;
    unsigned cnt = 0u;
#pragma unroll
    for (int w = 0; w < kRsWaves; w++) {
        const unsigned c = s_wcnt[w][d];
        s_wcnt[w][d] = cnt;
        cnt += c;
    }
    const unsigned start = rs_block_excl(cnt, s_wsum);
    const unsigned goff = rs_block_excl(hist[d], s_wsum);  // digit d's first place in the output

    // publish, look back, publish the finished prefix
    unsigned long long *word = stat + 1 + (size_t)tile * 256 + d;
    __hip_atomic_store(word, (tile == 0 ? 2ull << 32 : 1ull << 32) | cnt, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    unsigned acc = 0u;
    if (tile > 0) {
        unsigned spins = 0u;
        for (int k = tile - 1; k >= 0;) {
            const unsigned long long w =
                __hip_atomic_load(stat + 1 + (size_t)k * 256 + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned state = (unsigned)(w >> 32);
            if (state == 0u) {
                if (++spins >= kRsSpins) {
                    atomicOr(err, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            acc += (unsigned)w;
            if (state == 2u) break;
            k--;
        }
        __hip_atomic_store(word, (2ull << 32) | (acc + cnt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_start[d] = start;
    s_base[d] = goff + acc - start;  // output place of tile slot j of digit d: s_base[d] + j
    __syncthreads();

    // sort the tile by digit in LDS, then write it out in slot order: the
    // keys first, then the values through the same LDS
#pragma unroll
    for (int k = 0; k < kRsPer; k++) {
        const unsigned dg = (unsigned)(key[k] >> shift) & 255u;
        rank[k] += s_start[dg] + s_wcnt[wid][dg];
        s_key[rank[k]] = key[k];
        s_dg[rank[k]] = (unsigned char)dg;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kRsTile; j += kRsThreads) {
        const unsigned pos = s_base[s_dg[j]] + (unsigned)j;
        if (pos < (unsigned)n) kout[pos] = s_key[j];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRsPer; k++) s_val[rank[k]] = val[k];
    __syncthreads();
    for (int j = threadIdx.x; j < kRsTile; j += kRsThreads) {
        const unsigned pos = s_base[s_dg[j]] + (unsigned)j;
        if (pos < (unsigned)n) vout[pos] = s_val[j];
    }
}

template <typename T>
T *scan_temp(Sorter &so, long long n, hipStream_t)
{
    const size_t need = (size_t)scan_parts(n) * sizeof(T);
    if (so.tmp.n < need) so.tmp.alloc(need);
    return (T *)so.tmp.p;
}

}  // namespace

long long scan_parts(long long n) { return std::max<long long>(1, (n + kScanTile - 1) / kScanTile); }

void excl_scan(const long long *in, long long *out, long long n, long long *part, hipStream_t s)
{
    scan<long long>(in, out, n, false, part, s);
}

void zero_async(void *p, size_t bytes, hipStream_t s)
{
    if (!bytes) return;
    KD_REQUIRE(p && bytes % 4 == 0, "zero_async: buffer must be 4-byte words");
    const long long n = (long long)(bytes / 4);
    const int nb = (int)std::min<long long>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_zero32, dim3(nb), dim3(256), 0, s, (unsigned *)p, n);
    KD_HIP(hipGetLastError());
}

void Sorter::incl_scan(const int *in, int *out, int n, hipStream_t s)
{
    scan<int>(in, out, n, true, scan_temp<int>(*this, n, s), s);
}

void Sorter::excl_scan_i64(const long long *in, long long *out, int n, hipStream_t s)
{
    scan<long long>(in, out, n, false, scan_temp<long long>(*this, n, s), s);
}

void Sorter::pairs(const uint64_t *kin, uint64_t *kout, const int *vin, int *vout, int n, hipStream_t s)
{
    if (n <= 0) return;
    const int tiles = (n + kRsTile - 1) / kRsTile;
    const size_t words = 1 + (size_t)tiles * 256;  // look-back words of one digit's pass
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t kb = up(sizeof(uint64_t) * (size_t)n), vb = up(sizeof(int) * (size_t)n), hb = 8 * 256 * 4;
    const size_t need = kb + vb + hb + 8 * words * sizeof(unsigned long long);
    if (tmp.n < need) tmp.alloc(need);
    if (!err.p) {
        err.alloc(1);
        zero_async(err.p, err.bytes(), s);
    }
    uint64_t *kt = (uint64_t *)tmp.p;
    int *vt = (int *)(tmp.p + kb);
    unsigned *hist = (unsigned *)(tmp.p + kb + vb);
    unsigned long long *stat = (unsigned long long *)(tmp.p + kb + vb + hb);
    zero_async(hist, hb + 8 * words * sizeof(unsigned long long), s);
    hipLaunchKernelGGL(k_rs_hist, dim3(std::min(tiles, 1024)), dim3(256), 0, s, kin, n, hist);
    // digit p: from the input (p = 0), the scratch pair (odd p) or the output
    // (even p > 0), into the scratch pair (even p) or the output (odd p)
    for (int p = 0; p < 8; p++) {
        const uint64_t *ks = p == 0 ? kin : (p & 1 ? kt : kout);
        const int *vs = p == 0 ? vin : (p & 1 ? vt : vout);
        uint64_t *kd = p & 1 ? kout : kt;
        int *vd = p & 1 ? vout : vt;
        hipLaunchKernelGGL(k_rs_pass, dim3(tiles), dim3(kRsThreads), 0, s, ks, vs, kd, vd, n, 8 * p, hist + 256 * p,
                           stat + (size_t)p * words, err.p);
    }
    KD_HIP(hipGetLastError());
}

void Sorter::check(hipStream_t s)
{
    if (!err.p) return;
    unsigned h = 0u;
    KD_HIP(hipMemcpyAsync(&h, err.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    KD_REQUIRE(!h, "internal error: a radix sort tile waited too long for the tiles before it");
}


}  // namespace kd
