// Correctly rounded powers on the device, as the reference's deterministic
// maths library defines them (src/SourceSharp.MapFormats/Numerics/):
//   pow_f32      = DetMathF.Pow(float, float)       (DetMathF.cs:265-292): x^y
//                  rounded once to the nearest float, ties to even;
//   pow_to_single = DetMath.PowToSingle(double, double) (DetMath.cs:196-222):
//                  x^y rounded to the nearest double, then to float.
//
// Method: the device pow in double is within a few ulps, so when every double
// within 2^-49 (relative) of it rounds to one float, that float is the
// answer.  Otherwise (about one argument in 2^25) the power is recomputed as
// exp(y log x) in double-double -- exact products by fma, exp by a 9-term
// Taylor series of r/512 after reduction by ln 2 and nine squarings, log by
// one Newton step from the double log -- to about 2^-90, and rounded against
// the one float midpoint the interval holds; a power that close to the
// midpoint is one (e.g. (29/32)^5) and ties to even.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vh {
namespace detmath {

struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD dd_two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}

__device__ __forceinline__ DD dd_fast(double a, double b)
{
    double s = a + b;
    return {s, b - (s - a)};
}

__device__ __forceinline__ DD dd_prod(double a, double b)
{
    double p = a * b;
    return {p, fma(a, b, -p)};
}

__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    DD s = dd_two_sum(a.hi, b.hi);
    DD t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_fast(s.hi, s.lo);
    s.lo += t.lo;
    return dd_fast(s.hi, s.lo);
}

__device__ __forceinline__ DD dd_mul(DD a, DD b)
{
    DD p = dd_prod(a.hi, b.hi);
    p.lo += (a.hi * b.lo) + (a.lo * b.hi);
    return dd_fast(p.hi, p.lo);
}

__device__ __forceinline__ DD dd_mul_d(DD a, double b)
{
    DD p = dd_prod(a.hi, b);
    p.lo += a.lo * b;
    return dd_fast(p.hi, p.lo);
}

__device__ __forceinline__ DD dd_div_d(DD a, double b)
{
    double q1 = a.hi / b;
    DD p = dd_prod(q1, b);
    DD r = dd_add(a, {-p.hi, -p.lo});
    return dd_fast(q1, r.hi / b);
}

__device__ __noinline__ DD dd_exp(DD a)
{
    const double ln2_hi = 6.93147180559945286227e-01;  // ln 2 = ln2_hi + ln2_lo
    const double ln2_lo = 2.31904681384629955842e-17;
    double k = rint(a.hi / ln2_hi);
    DD t = dd_prod(k, ln2_hi);
    DD r = dd_add(a, {-t.hi, -t.lo});
    r = dd_add(r, {-(k * ln2_lo), 0.0});
    DD s = {ldexp(r.hi, -9), ldexp(r.lo, -9)};
    DD p = {1.0, 0.0};
    for (int n = 9; n >= 1; n--) p = dd_add({1.0, 0.0}, dd_div_d(dd_mul(p, s), (double)n));
    for (int i = 0; i < 9; i++) p = dd_mul(p, p);
    return {ldexp(p.hi, (int)k), ldexp(p.lo, (int)k)};
}

__device__ __forceinline__ DD dd_log(double x)
{
    double l0 = log(x);
    DD e = dd_mul_d(dd_exp({-l0, 0.0}), x);
    return dd_add({l0, 0.0}, dd_add(e, {-1.0, 0.0}));
}

// True when x^y is settled by the double result r alone (special values,
// or every double near r rounds to the same float); lo/hi bracket it if not.
__device__ __forceinline__ bool settled(double x, double y, double r, float &lo, float &hi)
{
    if (x <= 0.0 || x == 1.0 || y == 0.0 || isnan(r) || isinf(x) || isinf(y) || r == 0.0 ||
        isinf(r))
        return true;
    double m = fabs(r) * 0x1p-49;
    lo = (float)(r - m);
    hi = (float)(r + m);
    return lo == hi;
}

// The lighting path only raises positive bases (cone multipliers in [0, 1],
// luxel intensities): x <= 0 takes the device pow's own special values.
__device__ __noinline__ float pow_f32(float xf, float yf)
{
    double x = xf, y = yf;
    double r = pow(x, y);
    float lo = 0.0f, hi = 0.0f;
    if (settled(x, y, r, lo, hi)) return (float)r;
    DD v = dd_exp(dd_mul_d(dd_log(x), y));
    double mid = isinf(hi) ? 0x1.ffffffp127 : ((double)lo + (double)hi) * 0.5;
    DD d = dd_add(v, {-mid, 0.0});
    double diff = d.hi + d.lo;
    if (fabs(diff) <= mid * 0x1p-90) return (__float_as_uint(lo) & 1u) ? hi : lo;
    return diff > 0.0 ? hi : lo;
}

__device__ __noinline__ float pow_to_single(double x, double y)
{
    double r = pow(x, y);
    float lo = 0.0f, hi = 0.0f;
    if (settled(x, y, r, lo, hi)) return (float)r;
    DD v = dd_exp(dd_mul_d(dd_log(x), y));
    return (float)(v.hi + v.lo);
}

}  // namespace detmath
}  // namespace vh
