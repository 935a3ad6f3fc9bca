// sin and cos of a double, correctly rounded to nearest, in double-double
// arithmetic (about 106 bits) built only from IEEE FP64 add, mul and fma, so
// host and device give the same bits.
//
// Method: Cody-Waite reduction x = k*pi/2 + r with pi/2 split into four
// parts (the first two with 33 significant bits, so k*C1 and k*C2 are exact
// for |k| < 2^20), then the Taylor series of sin and cos for |r| <= pi/4 in
// double-double Horner form up to the 1/29! and 1/30! terms (truncation below
// 2^-110 relative), then the quadrant's sign/swap.  The constants are
// double-double splits of pi/2 and of 1/n!, computed exactly from their
// definitions.
//
// `hard` is set when the double-double result lies within 2^-98 relative of
// a rounding boundary (the last bit is then not certain) or when |x| >= 2^20
// (the reduction is not exact there); callers report those cases.
#pragma once

#include <hip/hip_runtime.h>

namespace kdmath {

struct dd {
    double hi, lo;
};

__host__ __device__ inline dd two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

__host__ __device__ inline dd fast_two_sum(double a, double b)
{
    double s = a + b;
    return {s, b - (s - a)};
}

__host__ __device__ inline dd two_prod(double a, double b)
{
    double p = a * b;
    return {p, fma(a, b, -p)};
}

__host__ __device__ inline dd dd_add(dd x, dd y)
{
    dd s = two_sum(x.hi, y.hi);
    dd t = two_sum(x.lo, y.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}

__host__ __device__ inline dd dd_mul(dd x, dd y)
{
    dd p = two_prod(x.hi, y.hi);
    p.lo = fma(x.hi, y.lo, p.lo);
    p.lo = fma(x.lo, y.hi, p.lo);
    return fast_two_sum(p.hi, p.lo);
}

// 1/n! as double-double, n = 0..31 (index n)
__host__ __device__ inline dd inv_fact(int n)
{
    switch (n) {
    case 0: return {1.0, 0.0};
    case 1: return {1.0, 0.0};
    case 2: return {0x1.0000000000000p-1, 0.0};
    case 3: return {0x1.5555555555555p-3, 0x1.5555555555555p-57};
    case 4: return {0x1.5555555555555p-5, 0x1.5555555555555p-59};
    case 5: return {0x1.1111111111111p-7, 0x1.1111111111111p-63};
    case 6: return {0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65};
    case 7: return {0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73};
    case 8: return {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76};
    case 9: return {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73};
    case 10: return {0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76};
    case 11: return {0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80};
    case 12: return {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83};
    case 13: return {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87};
    case 14: return {0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92};
    case 15: return {0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97};
    case 16: return {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101};
    case 17: return {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103};
    case 18: return {0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107};
    case 19: return {0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112};
    case 20: return {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120};
    case 21: return {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120};
    case 22: return {0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124};
    case 23: return {0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130};
    case 24: return {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135};
    case 25: return {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139};
    case 26: return {0x1.88e85fc6a4e5ap-89, -0x1.71c37ebd16540p-143};
    case 27: return {0x1.d1ab1c2dccea3p-94, 0x1.054d0c78aea14p-149};
    case 28: return {0x1.0a18a2635085dp-98, 0x1.b9e2e28e1aa54p-153};
    case 29: return {0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157};
    case 30: return {0x1.3932c5047d60ep-108, 0x1.832b7b530a627p-162};
    default: return {0x1.434d2e783f5bcp-113, 0x1.0b87b91be9affp-167};
    }
}

// Is the double-double value v within `rel` (relative) of a point halfway
// between two doubles (where its rounding is not certain)?
__host__ __device__ inline bool near_midpoint(dd v, double rel)
{
    double h = v.hi;
    if (h == 0.0) return false;
    int e;
    (void)frexp(h, &e);
    double ulp = ldexp(1.0, e - 53);  // spacing above |h| (below a power of two it halves)
    double a = fabs(v.lo);
    double tol = rel * fabs(h);
    double half = 0.5 * ulp;
    if (fabs(a - half) <= tol) return true;
    if (fabs(a - 0.5 * half) <= tol) return true;  // |h| a power of two: spacing below is ulp/2
    return false;
}

__host__ __device__ inline void dd_sincos(double x, double *s_out, double *c_out, bool *hard)
{
    *hard = false;
    if (!(fabs(x) < 1e300)) {  // inf or NaN
        *s_out = x - x;
        *c_out = x - x;
        return;
    }
    if (x == 0.0) {  // sin(+-0) = +-0, cos(+-0) = 1
        *s_out = x;
        *c_out = 1.0;
        return;
    }
    if (fabs(x) >= 0x1p20) *hard = true;
    const double C1 = 0x1.921fb54400000p+0, C2 = 0x1.0b4611a600000p-34, C3 = 0x1.3198a2e037073p-69,
                 C4 = 0x1.129024e088a68p-123, TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    double k = rint(x * TWO_OVER_PI);
    // r = x - k*(C1 + C2 + C3 + C4) as double-double
    double t1 = x - k * C1;  // exact (Sterbenz, k*C1 exact)
    dd r = two_sum(t1, -(k * C2));
    r = dd_add(r, {-two_prod(k, C3).hi, -two_prod(k, C3).lo});
    r = dd_add(r, {-(k * C4), 0.0});
    if (k == 0.0) r = {x, 0.0};  // keeps the sign of -0.0
    dd r2 = dd_mul(r, r);
    // sin r = r * sum (-1)^i r^(2i) / (2i+1)!, cos r = sum (-1)^i r^(2i) / (2i)!
    dd ps = inv_fact(29);
    for (int n = 27; n >= 1; n -= 2) {
        dd c = inv_fact(n);
        if ((n >> 1) & 1) c = {-c.hi, -c.lo};
        ps = dd_add(dd_mul(ps, r2), c);
    }
    dd sinr = dd_mul(ps, r);
    dd pc = inv_fact(30);
    pc = {-pc.hi, -pc.lo};  // (-1)^15
    for (int n = 28; n >= 0; n -= 2) {
        dd c = inv_fact(n);
        if ((n >> 1) & 1) c = {-c.hi, -c.lo};
        pc = dd_add(dd_mul(pc, r2), c);
    }
    dd cosr = pc;
    long long q = (long long)k & 3;
    dd sv, cv;
    switch (q) {
    case 0: sv = sinr; cv = cosr; break;
    case 1: sv = cosr; cv = {-sinr.hi, -sinr.lo}; break;
    case 2: sv = {-sinr.hi, -sinr.lo}; cv = {-cosr.hi, -cosr.lo}; break;
    default: sv = {-cosr.hi, -cosr.lo}; cv = sinr; break;
    }
    const double rel = 0x1p-98;
    if (near_midpoint(sv, rel) || near_midpoint(cv, rel)) *hard = true;
    *s_out = sv.hi + sv.lo;
    *c_out = cv.hi + cv.lo;
}

}  // namespace kdmath
