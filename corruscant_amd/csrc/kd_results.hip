// Bootstrap resampling of the jackknife fields on the GPU (results
// post-processing).
//
// corruscant/clustering.py:416-446 (results.bootstrap_pair_counts): a trial
// draws N_fields field ids with replacement; with h = bincount of the draws,
// every normalised count matrix A (N_fields x N_fields, one per radius bin and
// pair type) contributes sum_ij A[i][j] h[i] h[j], the contraction of A with
// the outer product h h^T.  The reference builds that outer product for every
// trial in Python and contracts with tensordot (N_trials x N_fields^2 doubles
// of memory); here the contraction is a quadratic form evaluated per trial.
//
// One block of four waves handles 64 trials, one per lane.  The 64 histograms
// sit in LDS as doubles, trial-minor (h[j] of all 64 trials in 512 contiguous
// bytes).  Each wave takes blocks of eight rows i, reads A[i][j] with
// wave-uniform scalar loads and accumulates a_r = sum_j A[i_r][j] h[j] for
// the eight rows, so one LDS read feeds eight FMAs; the per-wave sums
// sum_r h[i_r] a_r are added in a fixed order, so the result does not depend
// on scheduling.  FP64 throughout; every term is >= 0, so the sum's relative
// rounding error is at most about N_fields^2 * 2^-53.
#include "kd_internal.h"

namespace kd {

namespace {

constexpr int kTrials = 64;  // trials per block, one per lane
constexpr int kWaves = 4;
constexpr int kRowBlock = 8;

__device__ __forceinline__ double uload(const double *p)
{
    // wave-uniform read through the scalar cache (constant address space)
    typedef const double __attribute__((address_space(4))) *cptr;
    return *(cptr)(const void *)p;
}

__global__ __launch_bounds__(256) void k_bootstrap(const double *__restrict__ mats, int nmats, int nf,
                                                   const unsigned char *__restrict__ draws, long long ntrials,
                                                   double *__restrict__ out, unsigned *__restrict__ err)
{
    extern __shared__ double lds_d[];
    double *H = lds_d;                              // nf x kTrials
    double *part = lds_d + (size_t)nf * kTrials;    // kWaves x kTrials
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const long long t0 = (long long)blockIdx.x * kTrials;
    const int nt = (int)min((long long)kTrials, ntrials - t0);

    for (int k = tid; k < nf * kTrials; k += blockDim.x) H[k] = 0.0;
    __syncthreads();
    for (int k = tid; k < nt * nf; k += blockDim.x) {
        const int tt = k / nf;
        const int d = draws[(t0 + tt) * nf + (k - tt * nf)];
        if (d < nf)
            atomicAdd(&H[d * kTrials + tt], 1.0);  // small integers: exact
        else
            atomicOr(err, 1u);
    }
    __syncthreads();

    for (int m = 0; m < nmats; m++) {
        const double *A = mats + (size_t)m * nf * nf;
        double acc = 0.0;
        for (int i0 = wid * kRowBlock; i0 < nf; i0 += kWaves * kRowBlock) {
            const int nr = min(kRowBlock, nf - i0);
            double a[kRowBlock];
#pragma unroll
            for (int r = 0; r < kRowBlock; r++) a[r] = 0.0;
            if (nr == kRowBlock) {
                for (int j = 0; j < nf; j++) {
                    const double hj = H[j * kTrials + lane];
#pragma unroll
                    for (int r = 0; r < kRowBlock; r++) a[r] = fma(uload(A + (size_t)(i0 + r) * nf + j), hj, a[r]);
                }
            } else {
                for (int j = 0; j < nf; j++) {
                    const double hj = H[j * kTrials + lane];
#pragma unroll
                    for (int r = 0; r < kRowBlock; r++)
                        if (r < nr) a[r] = fma(uload(A + (size_t)(i0 + r) * nf + j), hj, a[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < kRowBlock; r++)
                if (r < nr) acc = fma(H[(i0 + r) * kTrials + lane], a[r], acc);
        }
        part[wid * kTrials + lane] = acc;
        __syncthreads();
        if (wid == 0 && lane < nt) {
            double s = part[lane];
            for (int w = 1; w < kWaves; w++) s += part[w * kTrials + lane];
            out[(size_t)m * ntrials + t0 + lane] = s;
        }
        __syncthreads();
    }
}

}  // namespace

void bootstrap_forms(const double *h_mats, int nmats, int nf, const unsigned char *h_draws, long long ntrials,
                     double *h_out, hipStream_t s)
{
    KD_REQUIRE(nf >= 1 && nf <= 256, "bootstrap: the number of fields must be in [1, 256]");
    KD_REQUIRE(nmats >= 0 && ntrials >= 0, "bootstrap: negative sizes");
    if (nmats == 0 || ntrials == 0) return;
    KD_REQUIRE(h_mats && h_draws && h_out, "bootstrap: NULL buffer");
    const size_t mat_n = (size_t)nmats * nf * nf;
    DevBuf<double> mats(mat_n), out((size_t)nmats * ntrials);
    DevBuf<unsigned char> draws((size_t)ntrials * nf);
    DevBuf<unsigned> err(1);
    KD_HIP(hipMemcpyAsync(mats.p, h_mats, mats.bytes(), hipMemcpyHostToDevice, s));
    KD_HIP(hipMemcpyAsync(draws.p, h_draws, draws.bytes(), hipMemcpyHostToDevice, s));
    KD_HIP(hipMemsetAsync(err.p, 0, err.bytes(), s));
    const long long nblocks = (ntrials + kTrials - 1) / kTrials;
    KD_REQUIRE(nblocks <= 0x7fffffffLL, "bootstrap: too many trials for one call");
    const size_t lds = ((size_t)nf * kTrials + kWaves * kTrials) * sizeof(double);  // <= 130 KiB
    if (lds > 65536)
        KD_HIP(hipFuncSetAttribute((const void *)k_bootstrap, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_bootstrap, dim3((unsigned)nblocks), dim3(64 * kWaves), lds, s, mats.p, nmats, nf, draws.p,
                       ntrials, out.p, err.p);
    KD_HIP(hipGetLastError());
    unsigned herr = 0;
    KD_HIP(hipMemcpyAsync(h_out, out.p, out.bytes(), hipMemcpyDeviceToHost, s));
    KD_HIP(hipMemcpyAsync(&herr, err.p, sizeof(herr), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    KD_REQUIRE(!herr, "bootstrap: a drawn field id is not below the number of fields");
}

}  // namespace kd
