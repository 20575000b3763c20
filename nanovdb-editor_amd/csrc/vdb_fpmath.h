// vdb_fpmath.h -- deterministic FP32 arithmetic shared by host and device code of libvdbraster.
//
// The raster3d contract evaluates every expression of the reference shaders in IEEE FP32 with
// the operation order written in the sources and no FMA contraction (the library is compiled
// with -ffp-contract=off). Transcendentals are correctly rounded: exp/log/pow are evaluated in
// double precision (error ~1e-16 relative) and rounded once to float, so they match any
// correctly rounded libm (glibc's expf/logf/powf included) except for inputs whose exact
// result lies within ~1e-16 of a float rounding boundary.
//
// exp: Cody-Waite reduction x = k*ln2 + r (ln2 split in a 32-bit head and a tail), Taylor
//      polynomial of degree 13 on |r| <= 0.3466, scaling by 2^k through the exponent field.
// log: x = m * 2^e with m in [sqrt(1/2), sqrt(2)), log(m) = 2*atanh(s), s = (m-1)/(m+1)
//      carried as a double-double (exact residual of the division through fma), series to s^23.
// pow: exp(y * log(x)) with log and the product as double-doubles (x > 0).
// log and pow round their double-double result to float through round-to-odd, so the float
// result has a single rounding. fma() is exact on both sides (libm on the host, v_fma_f64 on
// the device).
#ifndef VDB_FPMATH_H
#define VDB_FPMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VDB_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#include <string.h>
#define VDB_HD static inline
#endif

VDB_HD double vdb_bits_to_double(uint64_t u)
{
#if defined(__HIPCC__)
    return __builtin_bit_cast(double, u);
#else
    double d;
    memcpy(&d, &u, 8);
    return d;
#endif
}

VDB_HD uint64_t vdb_double_to_bits(double d)
{
#if defined(__HIPCC__)
    return __builtin_bit_cast(uint64_t, d);
#else
    uint64_t u;
    memcpy(&u, &d, 8);
    return u;
#endif
}

VDB_HD uint32_t vdb_float_to_bits(float f)
{
#if defined(__HIPCC__)
    return __builtin_bit_cast(uint32_t, f);
#else
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
#endif
}

VDB_HD float vdb_bits_to_float(uint32_t u)
{
#if defined(__HIPCC__)
    return __builtin_bit_cast(float, u);
#else
    float f;
    memcpy(&f, &u, 4);
    return f;
#endif
}

// round-half-even to integer, exact for |x| < 2^52 (adding and removing 1.5*2^52)
VDB_HD double vdb_rint(double x)
{
    const double shifter = 6755399441055744.0; // 1.5 * 2^52
    double t = x + shifter;
    return t - shifter;
}

// exp(x) in double for |x| <= 709
VDB_HD double vdb_exp_d(double x)
{
    const double inv_ln2 = 1.4426950408889634074;
    const double ln2_hi = 6.93147180369123816490e-01; // 32 significant bits: k*ln2_hi is exact
    const double ln2_lo = 1.90821492927058770002e-10;
    double kd = vdb_rint(x * inv_ln2);
    double r = x - kd * ln2_hi;
    r = r - kd * ln2_lo;
    double p = 1.0 / 6227020800.0;    // 1/13!
    p = p * r + 1.0 / 479001600.0;     // 1/12!
    p = p * r + 1.0 / 39916800.0;      // 1/11!
    p = p * r + 1.0 / 3628800.0;       // 1/10!
    p = p * r + 1.0 / 362880.0;        // 1/9!
    p = p * r + 1.0 / 40320.0;         // 1/8!
    p = p * r + 1.0 / 5040.0;          // 1/7!
    p = p * r + 1.0 / 720.0;           // 1/6!
    p = p * r + 1.0 / 120.0;           // 1/5!
    p = p * r + 1.0 / 24.0;            // 1/4!
    p = p * r + 1.0 / 6.0;             // 1/3!
    p = p * r + 0.5;                   // 1/2!
    p = p * r + 1.0;
    p = p * r + 1.0;
    int64_t k = (int64_t)kd;
    // split the scale so each factor is a normal double
    int64_t k1 = k / 2;
    int64_t k2 = k - k1;
    double s1 = vdb_bits_to_double((uint64_t)(k1 + 1023) << 52);
    double s2 = vdb_bits_to_double((uint64_t)(k2 + 1023) << 52);
    return (p * s1) * s2;
}

// a + b = s + err exactly (Knuth TwoSum)
VDB_HD void vdb_two_sum(double a, double b, double* s, double* err)
{
    double t = a + b;
    double bb = t - a;
    *err = (a - (t - bb)) + (b - bb);
    *s = t;
}

// Rounds the double-double value hi + lo (|lo| <= ulp(hi)/2, hi normal) to float with a
// single rounding: hi is first rounded to odd at double precision using the sign of lo.
VDB_HD float vdb_dd_to_float(double hi, double lo)
{
    if (lo == 0.0)
        return (float)hi;
    uint64_t u = vdb_double_to_bits(hi);
    int same_sign = (hi > 0.0) == (lo > 0.0);
    if (!same_sign)
        u = u - 1u; // magnitude of hi + lo is below |hi|: truncate toward zero
    u = u | 1u;     // sticky bit: inexact
    return (float)vdb_bits_to_double(u);
}

// log(x) for finite x > 0 as a double-double (hi, lo), relative error ~1e-19
VDB_HD void vdb_log_dd(double x, double* out_hi, double* out_lo)
{
    uint64_t u = vdb_double_to_bits(x);
    int64_t e = (int64_t)((u >> 52) & 0x7FF) - 1023;
    double m = vdb_bits_to_double((u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951)
    {
        m = m * 0.5;
        e = e + 1;
    }
    double a = m - 1.0; // exact (Sterbenz)
    double b_hi, b_lo;
    vdb_two_sum(m, 1.0, &b_hi, &b_lo);
    double s_hi = a / b_hi;
    double r = fma(-s_hi, b_hi, a); // exact residual
    r = r - s_hi * b_lo;
    double s_lo = r / b_hi;
    double s2 = s_hi * s_hi;
    double t = 1.0 / 23.0;
    t = t * s2 + 1.0 / 21.0;
    t = t * s2 + 1.0 / 19.0;
    t = t * s2 + 1.0 / 17.0;
    t = t * s2 + 1.0 / 15.0;
    t = t * s2 + 1.0 / 13.0;
    t = t * s2 + 1.0 / 11.0;
    t = t * s2 + 1.0 / 9.0;
    t = t * s2 + 1.0 / 7.0;
    t = t * s2 + 1.0 / 5.0;
    t = t * s2 + 1.0 / 3.0;
    // log(m) = 2 s + 2 s^3 t(s^2)
    double lm_hi = 2.0 * s_hi;
    double lm_lo = 2.0 * s_lo + 2.0 * s_hi * s2 * t;
    const double ln2_hi = 6.93147180369123816490e-01; // 32 significant bits: e*ln2_hi is exact
    const double ln2_lo = 1.90821492927058770002e-10;
    double ed = (double)e;
    double h, l;
    vdb_two_sum(ed * ln2_hi, lm_hi, &h, &l);
    l = l + (ed * ln2_lo + lm_lo);
    double hh = h + l;
    double ll = l - (hh - h);
    *out_hi = hh;
    *out_lo = ll;
}

// log(x) in double for finite x > 0
VDB_HD double vdb_log_d(double x)
{
    double h, l;
    vdb_log_dd(x, &h, &l);
    return h + l;
}

// exp(zh + zl) as a double-double, |zh| <= 709, |zl| <= ulp(zh)
VDB_HD void vdb_exp_dd(double zh, double zl, double* out_hi, double* out_lo)
{
    const double inv_ln2 = 1.4426950408889634074;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double kd = vdb_rint(zh * inv_ln2);
    double r_hi = zh - kd * ln2_hi; // exact
    double r_lo = zl - kd * ln2_lo;
    double r, rl;
    vdb_two_sum(r_hi, r_lo, &r, &rl);
    // exp(r) - 1 - r = r^2 * q(r)
    double q = 1.0 / 6227020800.0;
    q = q * r + 1.0 / 479001600.0;
    q = q * r + 1.0 / 39916800.0;
    q = q * r + 1.0 / 3628800.0;
    q = q * r + 1.0 / 362880.0;
    q = q * r + 1.0 / 40320.0;
    q = q * r + 1.0 / 5040.0;
    q = q * r + 1.0 / 720.0;
    q = q * r + 1.0 / 120.0;
    q = q * r + 1.0 / 24.0;
    q = q * r + 1.0 / 6.0;
    q = q * r + 0.5;
    double r2 = r * r;
    double r2_lo = fma(r, r, -r2);
    double tail = r2 * q;
    tail = tail + (r2_lo * q + rl * (1.0 + r)); // d/dr exp(r) ~ 1 + r for the rl correction
    double h, l;
    vdb_two_sum(1.0, r, &h, &l);
    l = l + tail;
    double hh = h + l;
    double ll = l - (hh - h);
    int64_t k = (int64_t)kd;
    int64_t k1 = k / 2;
    int64_t k2 = k - k1;
    double s1 = vdb_bits_to_double((uint64_t)(k1 + 1023) << 52);
    double s2 = vdb_bits_to_double((uint64_t)(k2 + 1023) << 52);
    *out_hi = (hh * s1) * s2; // exact scalings while the result stays a normal double
    *out_lo = (ll * s1) * s2;
}

VDB_HD float vdb_expf(float x)
{
    if (x != x)
        return x;
    if (x > 89.0f)
        return vdb_bits_to_float(0x7F800000u);
    if (x < -104.0f)
        return 0.0f;
    return (float)vdb_exp_d((double)x);
}

VDB_HD float vdb_logf(float x)
{
    if (x != x || x < 0.0f)
        return vdb_bits_to_float(0x7FC00000u);
    if (x == 0.0f)
        return vdb_bits_to_float(0xFF800000u);
    if (vdb_float_to_bits(x) == 0x7F800000u)
        return x;
    if (x == 1.0f)
        return 0.0f;
    double h, l;
    vdb_log_dd((double)x, &h, &l);
    return vdb_dd_to_float(h, l);
}

// pow for x >= 0 and finite y (the raster3d value pass only raises values in [0,1] to a
// positive finite power)
VDB_HD float vdb_powf_slow(float x, float y)
{
    double lh, ll;
    vdb_log_dd((double)x, &lh, &ll);
    double yd = (double)y;
    double zh = yd * lh;
    double zl = fma(yd, lh, -zh) + yd * ll;
    if (zh > 89.0)
        return vdb_bits_to_float(0x7F800000u);
    if (zh < -104.0)
        return 0.0f;
    double eh, el;
    vdb_exp_dd(zh, zl, &eh, &el);
    return vdb_dd_to_float(eh, el);
}

// pow for x >= 0 and finite y (the raster3d value pass only raises values in [0,1] to a
// positive finite power). Fast path: p = exp(y * log(x)) in double, relative error below
// (|y log x| + 4) * 5e-16; its float rounding is returned when p is farther than that bound
// from both rounding midpoints around it, otherwise the double-double evaluation decides.
VDB_HD float vdb_powf(float x, float y)
{
    if (x != x || y != y)
        return vdb_bits_to_float(0x7FC00000u);
    if (x == 1.0f || y == 0.0f)
        return 1.0f;
    if (x == 0.0f)
        return y > 0.0f ? 0.0f : vdb_bits_to_float(0x7F800000u);
    if (x < 0.0f)
        return vdb_bits_to_float(0x7FC00000u);
    double z = (double)y * vdb_log_d((double)x);
    if (z > 88.0 || z < -86.0)
        return vdb_powf_slow(x, y); // overflow / subnormal territory
    double p = vdb_exp_d(z);
    float f = (float)p;
    uint32_t fb = vdb_float_to_bits(f);
    double lo = 0.5 * ((double)f + (double)vdb_bits_to_float(fb - 1u));
    double hi = 0.5 * ((double)f + (double)vdb_bits_to_float(fb + 1u));
    double eps = (fabs(z) + 4.0) * 5e-16 * p;
    if (p - lo > eps && hi - p > eps)
        return f;
    return vdb_powf_slow(x, y);
}

#endif
