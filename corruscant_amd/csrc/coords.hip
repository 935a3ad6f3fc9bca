// libcoords.so: RA/DEC conversions on the GPU (src/coordmath.c:38-80).
//
// The reference evaluates cos/sin with the host libm (GCC turns the pairs
// into sincos); here they are correctly rounded double-double evaluations
// (dd_sincos.h), and the products are the reference's FP64 products in the
// same order.
#include <hip/hip_runtime.h>

#include "dd_sincos.h"

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/corruscant_coords.h"

namespace {

std::mutex g_mu;
thread_local std::string g_err;
unsigned long long g_hard = 0;  // points of the last conversion with an uncertain last bit
constexpr double kPiOver180 = M_PI / 180.0;  // src/coordmath.c:20

__global__ void k_radec2cart(const double *__restrict__ ra, const double *__restrict__ dec,
                             const double *__restrict__ dist, double *__restrict__ xyz, int n,
                             unsigned long long *__restrict__ hard_count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sl, cl, so, co;
    bool h1, h2;
    kdmath::dd_sincos(dec[i] * kPiOver180, &sl, &cl, &h1);
    kdmath::dd_sincos(ra[i] * kPiOver180, &so, &co, &h2);
    if (h1 || h2) atomicAdd(hard_count, 1ull);
    if (dist) {
        double d = dist[i];
        xyz[3 * (size_t)i] = d * cl * co;
        xyz[3 * (size_t)i + 1] = d * cl * so;
        xyz[3 * (size_t)i + 2] = d * sl;
    } else {
        xyz[3 * (size_t)i] = cl * co;
        xyz[3 * (size_t)i + 1] = cl * so;
        xyz[3 * (size_t)i + 2] = sl;
    }
}

__global__ void k_radec2sph(const double *__restrict__ ra, const double *__restrict__ dec, double *__restrict__ th,
                            double *__restrict__ ph, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    th[i] = (90.0 - dec[i]) * kPiOver180;
    ph[i] = ra[i] * kPiOver180;
}

// HEALPix RING-scheme pixel of (theta, phi) (Gorski et al. 2005, ApJ 622,
// 759, section 4): z = cos(theta); in the equatorial zone |z| <= 2/3 the
// pixel follows from the indices of the ascending and descending edge lines
// through the point, in the polar caps from the ring index nside*sqrt(3(1-|z|))
// (written as nside*sin(theta)/sqrt((1+|z|)/3) within 0.01 rad of a pole,
// where 1-|z| loses digits).
__device__ __forceinline__ long long imodulo(long long v, long long m)
{
    long long r = v % m;
    return r < 0 ? r + m : r;
}

__global__ void k_ang2pix_ring(const double *__restrict__ theta, const double *__restrict__ phi,
                               long long *__restrict__ pix, int n, long long nside, unsigned *__restrict__ bad)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double th = theta[i], ph = phi[i];
    if (!(th >= 0.0 && th <= M_PI) || !isfinite(ph)) {
        pix[i] = -1;
        atomicOr(bad, 1u);
        return;
    }
    double sth, z;
    bool hard;
    kdmath::dd_sincos(th, &sth, &z, &hard);
    const double za = fabs(z);
    double tt = ph * (2.0 / M_PI);  // in units of pi/2, reduced to [0, 4)
    if (tt < 0.0 || tt >= 4.0) {
        tt = fmod(tt, 4.0);
        if (tt < 0.0) tt += 4.0;
        if (tt >= 4.0) tt = 0.0;
    }
    long long p;
    if (za <= 2.0 / 3.0) {
        const double t1 = (double)nside * (0.5 + tt), t2 = (double)nside * (z * 0.75);
        const long long jp = (long long)(t1 - t2);  // ascending edge line index
        const long long jm = (long long)(t1 + t2);  // descending edge line index
        const long long ir = nside + 1 + jp - jm;   // ring from z = 2/3, 1..2 nside + 1
        const long long kshift = 1 - (ir & 1);
        const long long ip = imodulo((jp + jm - nside + kshift + 1) / 2, 4 * nside);
        p = 2 * nside * (nside - 1) + (ir - 1) * 4 * nside + ip;
    } else {
        const double tp = tt - floor(tt);
        const bool near_pole = th < 0.01 || th > M_PI - 0.01;
        const double tmp = near_pole ? (double)nside * sth / sqrt((1.0 + za) / 3.0) : (double)nside * sqrt(3.0 * (1.0 - za));
        const long long jp = (long long)(tp * tmp), jm = (long long)((1.0 - tp) * tmp);
        const long long ir = jp + jm + 1;  // ring from the nearest pole
        const long long ip = imodulo((long long)(tt * (double)ir), 4 * ir);
        p = z > 0.0 ? 2 * ir * (ir - 1) + ip : 12 * nside * nside - 2 * ir * (ir + 1) + ip;
    }
    pix[i] = p;
}

bool ok(hipError_t e)
{
    if (e != hipSuccess) {
        g_err = std::string("HIP error ") + hipGetErrorString(e);
        return false;
    }
    return true;
}

void convert(const double *ra, const double *dec, const double *dist, double *xyz, int n)
{
    g_err.clear();
    if (n <= 0) return;
    std::lock_guard<std::mutex> lk(g_mu);
    double *d_in = nullptr, *d_out = nullptr;
    unsigned long long *d_hard = nullptr;
    size_t nin = dist ? 3 : 2;
    g_hard = 0;
    if (!ok(hipMalloc((void **)&d_in, sizeof(double) * nin * n + 8))) return;
    d_hard = (unsigned long long *)(d_in + nin * n);
    if (!ok(hipMalloc((void **)&d_out, sizeof(double) * 3 * n))) {
        (void)hipFree(d_in);
        return;
    }
    bool good = ok(hipMemcpy(d_in, ra, sizeof(double) * n, hipMemcpyHostToDevice)) &&
                ok(hipMemcpy(d_in + n, dec, sizeof(double) * n, hipMemcpyHostToDevice)) &&
                (!dist || ok(hipMemcpy(d_in + 2 * (size_t)n, dist, sizeof(double) * n, hipMemcpyHostToDevice)));
    good = good && ok(hipMemset(d_hard, 0, 8));
    if (good) {
        hipLaunchKernelGGL(k_radec2cart, dim3((n + 255) / 256), dim3(256), 0, 0, d_in, d_in + n,
                           dist ? d_in + 2 * (size_t)n : nullptr, d_out, n, d_hard);
        good = ok(hipGetLastError()) && ok(hipMemcpy(xyz, d_out, sizeof(double) * 3 * n, hipMemcpyDeviceToHost)) &&
               ok(hipMemcpy(&g_hard, d_hard, 8, hipMemcpyDeviceToHost));
    }
    (void)hipFree(d_in);
    (void)hipFree(d_out);
}

}  // namespace

extern "C" {

const char *coords_last_error(void) { return g_err.c_str(); }

unsigned long long coords_last_uncertain(void) { return g_hard; }

void radec2cart64(double *ra, double *dec, double *xyz, int n) { convert(ra, dec, nullptr, xyz, n); }

void radecdist2cart64(double *ra, double *dec, double *dist, double *xyz, int n) { convert(ra, dec, dist, xyz, n); }

void radec2sph64(double *ra, double *dec, double *theta, double *phi, int n)
{
    g_err.clear();
    if (n <= 0) return;
    std::lock_guard<std::mutex> lk(g_mu);
    double *d = nullptr;
    if (!ok(hipMalloc((void **)&d, sizeof(double) * 4 * n))) return;
    bool good = ok(hipMemcpy(d, ra, sizeof(double) * n, hipMemcpyHostToDevice)) &&
                ok(hipMemcpy(d + n, dec, sizeof(double) * n, hipMemcpyHostToDevice));
    if (good) {
        hipLaunchKernelGGL(k_radec2sph, dim3((n + 255) / 256), dim3(256), 0, 0, d, d + n, d + 2 * (size_t)n,
                           d + 3 * (size_t)n, n);
        good = ok(hipGetLastError()) && ok(hipMemcpy(theta, d + 2 * (size_t)n, sizeof(double) * n, hipMemcpyDeviceToHost)) &&
               ok(hipMemcpy(phi, d + 3 * (size_t)n, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    (void)hipFree(d);
}

void ang2pix_ring64(const double *theta, const double *phi, long long *pix, int n, long long nside)
{
    g_err.clear();
    if (n <= 0) return;
    if (nside < 1 || nside > (1LL << 29)) {
        g_err = "ang2pix_ring64: nside must be in [1, 2^29]";
        return;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    double *d = nullptr;
    if (!ok(hipMalloc((void **)&d, sizeof(double) * 3 * (size_t)n + 8))) return;
    long long *d_pix = (long long *)(d + 2 * (size_t)n);
    unsigned *d_bad = (unsigned *)(d + 3 * (size_t)n);
    unsigned bad = 0;
    bool good = ok(hipMemcpy(d, theta, sizeof(double) * n, hipMemcpyHostToDevice)) &&
                ok(hipMemcpy(d + n, phi, sizeof(double) * n, hipMemcpyHostToDevice)) && ok(hipMemset(d_bad, 0, 4));
    if (good) {
        hipLaunchKernelGGL(k_ang2pix_ring, dim3((n + 255) / 256), dim3(256), 0, 0, d, d + n, d_pix, n, nside, d_bad);
        good = ok(hipGetLastError()) && ok(hipMemcpy(pix, d_pix, sizeof(long long) * n, hipMemcpyDeviceToHost)) &&
               ok(hipMemcpy(&bad, d_bad, 4, hipMemcpyDeviceToHost));
    }
    (void)hipFree(d);
    if (good && bad) g_err = "ang2pix_ring64: theta outside [0, pi] or non-finite phi (pixel -1)";
}

}  // extern "C"
