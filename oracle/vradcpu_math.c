/*
 * vradcpu -- CPU oracle: correctly rounded powers.  TEST INFRASTRUCTURE ONLY
 * (see vradcpu.h).
 *
 * Restates two functions of the reference's deterministic maths library
 * (paths under the reference's src/SourceSharp.MapFormats/Numerics/):
 *   DetMathF.Pow(float x, float y)       (DetMathF.cs:265-292): x^y rounded
 *       once, to the nearest float (ties to even);
 *   DetMath.PowToSingle(double x, double y) (DetMath.cs:196-222): x^y rounded
 *       to the nearest double, then that double to the nearest float.
 * The reference decides most arguments on a fast path and falls back to an
 * exact tier; this file does the same with its own method:
 *   1. r = pow(x, y) in double (glibc, < 1 ulp);
 *   2. if every double within 2^-49 (relative) of r rounds to the same float,
 *      that float is the answer (the exact power lies in that interval);
 *   3. otherwise the power is recomputed as exp(y log x) in double-double
 *      (about 2^-90 relative here) and compared with the one float midpoint
 *      the interval holds; a power within 2^-90 of it is an exact midpoint
 *      (x^2 of a 12-bit-fraction x, say) and ties to even.
 * Pinned by the reference's own table of mpmath results (DetMathGoldens.cs:
 * the F2 Pow and S2 PowToSingle rows) and by mpmath on arguments found to take
 * step 3.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "vradcpu.h"

typedef struct {
    double hi, lo;
} dd;

static inline dd two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    dd r = {s, (a - (s - bb)) + (b - bb)};
    return r;
}

static inline dd fast_two_sum(double a, double b)
{
    double s = a + b;
    dd r = {s, b - (s - a)};
    return r;
}

static inline dd two_prod(double a, double b)
{
    double p = a * b;
    dd r = {p, fma(a, b, -p)};
    return r;
}

static dd dd_add(dd a, dd b)
{
    dd s = two_sum(a.hi, b.hi);
    dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}

static dd dd_mul(dd a, dd b)
{
    dd p = two_prod(a.hi, b.hi);
    p.lo += (a.hi * b.lo) + (a.lo * b.hi);
    return fast_two_sum(p.hi, p.lo);
}

static dd dd_mul_d(dd a, double b)
{
    dd p = two_prod(a.hi, b);
    p.lo += a.lo * b;
    return fast_two_sum(p.hi, p.lo);
}

static dd dd_div_d(dd a, double b)
{
    double q1 = a.hi / b;
    dd p = two_prod(q1, b);
    dd r = dd_add(a, (dd){-p.hi, -p.lo});
    double q2 = r.hi / b;
    return fast_two_sum(q1, q2);
}

/* ln 2 as a double-double: the double nearest ln 2 and the rest. */
static const double kLn2Hi = 6.93147180559945286227e-01;
static const double kLn2Lo = 2.31904681384629955842e-17;

/* exp of a double-double: k = round(a / ln 2), r = a - k ln 2, exp(r / 512)
 * by its Taylor series to the 9th power (|r / 512| < 6.8e-4, so the next term
 * is under 2^-112), squared nine times, scaled by 2^k. */
static dd dd_exp(dd a)
{
    double k = nearbyint(a.hi / kLn2Hi);
    dd t = two_prod(k, kLn2Hi);
    dd r = dd_add(a, (dd){-t.hi, -t.lo});
    r = dd_add(r, (dd){-(k * kLn2Lo), 0.0});
    dd s = {ldexp(r.hi, -9), ldexp(r.lo, -9)};
    dd p = {1.0, 0.0};
    for (int n = 9; n >= 1; n--) p = dd_add((dd){1.0, 0.0}, dd_div_d(dd_mul(p, s), (double)n));
    for (int i = 0; i < 9; i++) p = dd_mul(p, p);
    dd out = {ldexp(p.hi, (int)k), ldexp(p.lo, (int)k)};
    return out;
}

/* log of a positive double: one Newton step from the double log,
 * l = l0 + (x e^-l0 - 1), which squares the double's error away. */
static dd dd_log(double x)
{
    double l0 = log(x);
    dd e = dd_mul_d(dd_exp((dd){-l0, 0.0}), x);
    dd eps = dd_add(e, (dd){-1.0, 0.0});
    return dd_add((dd){l0, 0.0}, eps);
}

/* |x|^y in double-double, x finite and nonzero. */
static dd dd_pow(double x, double y) { return dd_exp(dd_mul_d(dd_log(fabs(x)), y)); }

/* The pair of adjacent floats around r when r's 2^-49 neighbourhood rounds
 * to two floats; returns 1 then, 0 when it rounds to one. */
static int float_bracket(double r, float *lo, float *hi)
{
    double m = fabs(r) * 0x1p-49;
    float a = (float)(r - m), b = (float)(r + m);
    if (a == b) return 0;
    *lo = a;
    *hi = b;
    return 1;
}

/* Integer-valued y: the sign of (-x)^y. */
static int odd_integer(double y)
{
    if (!(fabs(y) < 0x1p53) || y != nearbyint(y)) return 0;
    return fmod(fabs(y), 2.0) == 1.0;
}

/* Rounds the double-double v to the float between lo and hi (adjacent), ties
 * (|v - mid| under 2^-90 of mid) to the even one. */
static float pick(dd v, float lo, float hi)
{
    double mid;
    if (isinf(hi) || isinf(lo)) {
        /* the midpoint between FLT_MAX and the next power of two */
        mid = copysign(0x1.ffffffp127, isinf(hi) ? 1.0 : -1.0);
    } else {
        mid = ((double)lo + (double)hi) * 0.5;
    }
    dd diff = dd_add(v, (dd){-mid, 0.0});
    double d = diff.hi + diff.lo;
    if (fabs(d) <= fabs(mid) * 0x1p-90) {
        uint32_t bl;
        memcpy(&bl, &lo, 4);
        return (bl & 1u) ? hi : lo;
    }
    return d > 0 ? hi : lo;
}

float vc_pow_f32(float xf, float yf)
{
    double x = xf, y = yf;
    double r = pow(x, y);
    if (x == 0.0 || x == 1.0 || y == 0.0 || isnan(r) || isinf(x) || isinf(y) || r == 0.0 ||
        isinf(r))
        return (float)r;
    float lo, hi;
    if (!float_bracket(r, &lo, &hi)) return (float)r;
    dd v = dd_pow(x, y);
    if (x < 0 && odd_integer(y)) v = (dd){-v.hi, -v.lo};
    return pick(v, lo, hi);
}

float vc_pow_to_single(double x, double y)
{
    double r = pow(x, y);
    if (x == 0.0 || x == 1.0 || y == 0.0 || isnan(r) || isinf(x) || isinf(y) || r == 0.0 ||
        isinf(r))
        return (float)r;
    float lo, hi;
    if (!float_bracket(r, &lo, &hi)) return (float)r;
    dd v = dd_pow(x, y);
    if (x < 0 && odd_integer(y)) v = (dd){-v.hi, -v.lo};
    return (float)(v.hi + v.lo);
}

/* 1 when vc_pow_f32 (kind 0) or vc_pow_to_single (kind 1) needs the
 * double-double step for these arguments. */
int32_t vc_pow_needs_exact(int32_t kind, double x, double y)
{
    if (kind == 0) {
        x = (float)x;
        y = (float)y;
    }
    double r = pow(x, y);
    if (x == 0.0 || x == 1.0 || y == 0.0 || isnan(r) || isinf(x) || isinf(y) || r == 0.0 ||
        isinf(r))
        return 0;
    float lo, hi;
    return float_bracket(r, &lo, &hi);
}

void vc_pow_f32_array(const float *x, const float *y, int64_t n, float *out)
{
    for (int64_t i = 0; i < n; i++) out[i] = vc_pow_f32(x[i], y[i]);
}

void vc_pow_to_single_array(const double *x, const double *y, int64_t n, float *out)
{
    for (int64_t i = 0; i < n; i++) out[i] = vc_pow_to_single(x[i], y[i]);
}

/* Scans count consecutive float bit patterns from first_bits for arguments
 * that take the double-double step: kind 0, x = that float under the float
 * exponent y; kind 1, x = that float / 256 (double) under the double y.  With
 * misrounded set, only those where the plain (float)pow(double) differs from
 * the correctly rounded answer.  Writes up to max_found bit patterns to found;
 * returns how many. */
int32_t vc_pow_scan(int32_t kind, double y, uint32_t first_bits, int64_t count, int32_t misrounded,
                    uint32_t *found, int32_t max_found)
{
    int32_t n = 0;
    for (int64_t i = 0; i < count && n < max_found; i++) {
        uint32_t b = first_bits + (uint32_t)i;
        float xf;
        memcpy(&xf, &b, 4);
        double x = kind == 0 ? (double)xf : (double)xf / 256.0;
        if (!vc_pow_needs_exact(kind, x, y)) continue;
        if (misrounded) {
            float cr = kind == 0 ? vc_pow_f32(xf, (float)y) : vc_pow_to_single(x, y);
            if ((float)pow(x, y) == cr) continue;
        }
        found[n++] = b;
    }
    return n;
}
