// Prefix sums and buffer clears of the library, written for wave64.
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

void Sorter::excl_scan_u32(const unsigned *in, unsigned *out, int n, hipStream_t s)
{
    scan<unsigned>(in, out, n, false, scan_temp<unsigned>(*this, n, s), s);
}


}  // namespace kd
