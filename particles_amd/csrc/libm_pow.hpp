// x ** y for the adaptive step controller, bit-identical to the host libm the
// reference links.
//
// The reference's step-size factors are Fortran real ** real
// (sources/integrations.F90:177,186,466,486,581), which flang lowers to the C
// library's pow().  On the reference's platform that is glibc >= 2.28, whose
// pow is the table-driven algorithm published in ARM optimized-routines
// (Szabolcs Nagy, 2018): log(x) to ~68 bits as k ln2 + log(c) + log1p(z/c-1)
// with a 128-entry table, then exp(y log x) with a 128-entry 2^(k/N) table
// and a degree-5 polynomial.  On x86-64 CPUs with FMA, glibc runs that source
// built with -mfma, where GCC fuses every multiply whose only uses are
// additions; the fma() calls below are exactly those fusions.  Any other
// rounding makes a step size differ by an ulp now and then, and trajectories
// in turbulent fields are chaotic enough that such differences grow to O(1)
// within a few gyro-periods -- hence a restatement rather than the device
// libm's own pow.
//
// Compiles for host and device: PTC_POW_HD marks the functions and
// PTC_POW_TABLE the tables (device constant memory in device passes).
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIP_DEVICE_COMPILE__)
#define PTC_POW_TABLE static __device__ __constant__
#else
#define PTC_POW_TABLE static
#endif
#if defined(__HIPCC__)
#define PTC_POW_HD __host__ __device__ __forceinline__
#else
#define PTC_POW_HD static inline
#endif

#include "pow_tables.h"

namespace ptc_pow {

PTC_POW_HD uint64_t bits(double x)
{
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}
PTC_POW_HD double from_bits(uint64_t u)
{
    double x;
    memcpy(&x, &u, 8);
    return x;
}
PTC_POW_HD uint32_t top12(double x) { return (uint32_t)(bits(x) >> 52); }

// log part: y + *tail = log(x), |tail| ~ 2^-68 |y|
PTC_POW_HD double log_inline(uint64_t ix, double *tail, const double (*logt)[3])
{
    const double ln2hi = 0x1.62e42fefa3800p-1, ln2lo = 0x1.ef35793c76730p-45;
    // log1p(r) - r + r^2/2 minimax coefficients, pre-scaled for the
    // evaluation scheme below (A0 = -1/2)
    const double A0 = -0x1p-1, A1 = 0x1.555555555556p-2 * -2, A2 = -0x1.0000000000006p-2 * -2,
                 A3 = 0x1.999999959554ep-3 * 4, A4 = -0x1.555555529a47ap-3 * 4,
                 A5 = 0x1.2495b9b4845e9p-3 * -8, A6 = -0x1.0002b8b263fc3p-3 * -8;
    const uint64_t OFF = 0x3fe6955500000000ULL;
    const uint64_t tmp = ix - OFF;
    const int i = (int)((tmp >> 45) % 128);
    const int k = (int)((int64_t)tmp >> 52);
    const uint64_t iz = ix - (tmp & 0xfffULL << 52);
    const double z = from_bits(iz), kd = (double)k;
    const double invc = logt[i][0], logc = logt[i][1], logctail = logt[i][2];
    const double r = fma(z, invc, -1.0);  // exact: 1/c has few bits
    const double t1 = fma(kd, ln2hi, logc);
    const double t2 = t1 + r;
    const double lo1 = fma(kd, ln2lo, logctail);
    const double lo2 = t1 - t2 + r;
    const double ar = A0 * r, ar2 = r * ar, ar3 = r * ar2;
    const double hi = t2 + ar2;
    const double lo3 = fma(ar, r, -ar2);
    const double lo4 = t2 - hi + ar2;
    const double p = ar3 * fma(ar2, fma(ar2, fma(r, A6, A5), fma(r, A4, A3)), fma(r, A2, A1));
    const double lo = lo1 + lo2 + lo3 + lo4 + p;
    const double y = hi + lo;
    *tail = hi - y + lo;
    return y;
}

// exp for results outside the normal scale range (k/N out of [-1022, 1023])
PTC_POW_HD double exp_special(double tmp, uint64_t sbits, uint64_t ki)
{
    if ((ki & 0x80000000) == 0) {
        sbits -= 1009ULL << 52;
        const double scale = from_bits(sbits);
        return 0x1p1009 * fma(scale, tmp, scale);
    }
    sbits += 1022ULL << 52;
    const double scale = from_bits(sbits);
    double y = fma(scale, tmp, scale);
    if (fabs(y) < 1.0) {
        const double one = y < 0.0 ? -1.0 : 1.0;
        double lo = fma(scale, tmp, scale - y);
        const double hi = one + y;
        lo = one - hi + y + lo;
        y = (hi + lo) - one;
        if (y == 0) y = from_bits(sbits & 0x8000000000000000ULL);
    }
    return 0x1p-1022 * y;
}

// sign_bias * exp(x + xtail)
PTC_POW_HD double exp_inline(double x, double xtail, uint64_t sign_bias,
                             const unsigned long long *expt)
{
    const double inv_ln2_n = 0x1.71547652b82fep0 * 128, neg_ln2hi_n = -0x1.62e42fefa0000p-8,
                 neg_ln2lo_n = -0x1.cf79abc9e3b3ap-47, shift = 0x1.8p52;
    const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3, C4 = 0x1.55555cf172b91p-5,
                 C5 = 0x1.1111167a4d017p-7;
    uint32_t abstop = top12(x) & 0x7ff;
    if (abstop - top12(0x1p-54) >= top12(512.0) - top12(0x1p-54)) {
        if (abstop - top12(0x1p-54) >= 0x80000000u) {
            const double one = 1.0 + x;
            return sign_bias ? -one : one;
        }
        if (abstop >= top12(1024.0)) {
            const double big = (bits(x) >> 63) ? 0x1p-767 * 0x1p-767 : 0x1p769 * 0x1p769;
            return sign_bias ? -big : big;
        }
        abstop = 0;  // large |x| handled by exp_special
    }
    double kd = fma(inv_ln2_n, x, shift);
    const uint64_t ki = bits(kd);
    kd -= shift;
    double r = fma(kd, neg_ln2lo_n, fma(kd, neg_ln2hi_n, x));
    r += xtail;
    const uint64_t idx = 2 * (ki % 128);
    const uint64_t top = (ki + sign_bias) << 45;
    const double tail = from_bits(expt[idx]);
    const uint64_t sbits = expt[idx + 1] + top;
    const double r2 = r * r;
    const double tmp = fma(r2 * r2, fma(r, C5, C4), fma(r2, fma(r, C3, C2), tail + r));
    if (abstop == 0) return exp_special(tmp, sbits, ki);
    const double scale = from_bits(sbits);
    return fma(scale, tmp, scale);
}

// 0: not an integer, 1: odd integer, 2: even integer (iy finite, nonzero)
PTC_POW_HD int checkint(uint64_t iy)
{
    const int e = (int)(iy >> 52 & 0x7ff);
    if (e < 0x3ff) return 0;
    if (e > 0x3ff + 52) return 2;
    if (iy & ((1ULL << (0x3ff + 52 - e)) - 1)) return 0;
    if (iy & (1ULL << (0x3ff + 52 - e))) return 1;
    return 2;
}

PTC_POW_HD bool zeroinfnan(uint64_t i)
{
    return 2 * i - 1 >= 2 * bits(__builtin_inf()) - 1;
}

}  // namespace ptc_pow

// Table locations: the constant tables of pow_tables.h, or copies of them
// (device code keeps them in LDS).
struct PowTables {
    const double (*log)[3];
    const unsigned long long *exp;
};

// x ** y with glibc's results, including its special cases (zeros, infinities,
// NaN, negative x with integer y, subnormal x, |y| tiny or huge).
PTC_POW_HD double ptc_pow_with(double x, double y, const double (*logt)[3],
                               const unsigned long long *expt)
{
    using namespace ptc_pow;
    uint64_t sign_bias = 0;
    uint64_t ix = bits(x);
    const uint64_t iy = bits(y);
    uint32_t topx = top12(x);
    const uint32_t topy = top12(y);
    if (topx - 0x001 >= 0x7ff - 0x001 || (topy & 0x7ff) - 0x3be >= 0x43e - 0x3be) {
        if (zeroinfnan(iy)) {
            if (2 * iy == 0) return 1.0;
            if (ix == bits(1.0)) return 1.0;
            if (2 * ix > 2 * bits(__builtin_inf()) || 2 * iy > 2 * bits(__builtin_inf())) return x + y;
            if (2 * ix == 2 * bits(1.0)) return 1.0;
            if ((2 * ix < 2 * bits(1.0)) == !(iy >> 63)) return 0.0;
            return y * y;
        }
        if (zeroinfnan(ix)) {
            double x2 = x * x;
            if ((ix >> 63) && checkint(iy) == 1) x2 = -x2;
            return (iy >> 63) ? 1.0 / x2 : x2;
        }
        if (ix >> 63) {
            const int yint = checkint(iy);
            if (yint == 0) return (x - x) / (x - x);
            if (yint == 1) sign_bias = 0x800ULL << 7;
            ix &= 0x7fffffffffffffffULL;
            topx &= 0x7ff;
        }
        if ((topy & 0x7ff) - 0x3be >= 0x43e - 0x3be) {
            if (ix == bits(1.0)) return 1.0;
            if ((topy & 0x7ff) < 0x3be) return ix > bits(1.0) ? 1.0 + y : 1.0 - y;
            return (ix > bits(1.0)) == (topy < 0x800) ? 0x1p769 * 0x1p769 : 0x1p-767 * 0x1p-767;
        }
        if (topx == 0) {  // subnormal x: normalise
            ix = bits(x * 0x1p52);
            ix &= 0x7fffffffffffffffULL;
            ix -= 52ULL << 52;
        }
    }
    double lo;
    const double hi = log_inline(ix, &lo, logt);
    const double ehi = y * hi;
    const double elo = fma(y, lo, fma(y, hi, -ehi));
    return exp_inline(ehi, elo, sign_bias, expt);
}

PTC_POW_HD double ptc_libm_pow(double x, double y)
{
    return ptc_pow_with(x, y, ptc_pow_log_tab, ptc_pow_exp_tab);
}
