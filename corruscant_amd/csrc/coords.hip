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

}  // extern "C"
