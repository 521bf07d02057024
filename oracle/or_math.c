/*
 * or_math.c -- spec math of the oracle (test infrastructure only).
 *
 * The reference evaluates powf / logf / sinf / cosf / expf either with the
 * host libm or with CUDA fast-math intrinsics, so its results depend on the
 * build.  The hot path here fixes one definition that the oracle and the
 * GPU kernels both implement:
 *
 *   f(x) = (float) F((double) x)
 *
 * where F is evaluated in IEEE double precision from +, -, *, / and floor
 * only (no fused multiply-add), by the following published methods:
 *   log  : x = m * 2^e, m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
 *          ln m = 2 * atanh(s) by its odd power series (Abramowitz & Stegun
 *          4.1.27), ln x = e * ln2 + ln m;
 *   exp  : Cody-Waite reduction z = k ln2 + r with a two-part ln2,
 *          e^r by its Taylor series to degree 15, scaling by 2^k;
 *   pow  : x^y = exp(y * ln x)  (x > 0);  0^y = 0 for y > 0, 1 for y = 0;
 *   sin/cos : reduction by pi/2 with a two-part constant (fdlibm's
 *          pio2_1 / pio2_1t split), Taylor polynomials on [-pi/4, pi/4],
 *          quadrant selection.
 * Series truncation error is below 1e-17 relative, so after rounding to
 * float the results equal the correctly rounded value except in rare
 * near-halfway cases.
 *
 * Pinned by tests/test_oracle_math.py: against numpy / libm (<= 1 ulp over
 * 10^5 arguments each), against closed forms (pow(8, 1/3), sin(pi/6), ...),
 * and the Philox4x32-10 known-answer vectors of Salmon et al., "Parallel
 * random numbers: as easy as 1, 2, 3" (SC'11), plus an independent
 * pure-Python Philox.
 */
#include <string.h>
#include "or_internal.h"

#define D_LN2_HI  6.93147180369123816490e-01   /* 0x3FE62E42FEE00000 */
#define D_LN2_LO  1.90821492927058770002e-10
#define D_INV_LN2 1.44269504088896338700e+00
#define D_SQRT2   1.41421356237309514547e+00
#define D_PIO2_1  1.57079632673412561417e+00   /* 0x3FF921FB54400000 */
#define D_PIO2_1T 6.07710050650619224932e-11
#define D_2_PI    6.36619772367581382433e-01

static double d_log(double x)
{
    /* x is a positive normal double (all callers pass a converted float) */
    uint64_t b;
    memcpy(&b, &x, sizeof b);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    b = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    memcpy(&m, &b, sizeof m);
    if (m > D_SQRT2) { m = m * 0.5; e = e + 1; }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 2.0 / 23.0;
    p = p * s2 + 2.0 / 21.0;
    p = p * s2 + 2.0 / 19.0;
    p = p * s2 + 2.0 / 17.0;
    p = p * s2 + 2.0 / 15.0;
    p = p * s2 + 2.0 / 13.0;
    p = p * s2 + 2.0 / 11.0;
    p = p * s2 + 2.0 / 9.0;
    p = p * s2 + 2.0 / 7.0;
    p = p * s2 + 2.0 / 5.0;
    p = p * s2 + 2.0 / 3.0;
    p = p * s2 + 2.0;
    const double lnm = s * p;
    const double de = (double)e;
    return de * D_LN2_HI + (de * D_LN2_LO + lnm);
}

static double d_exp(double z)
{
    /* only called where the float result is finite and non-zero or where
     * the clamps below decide it */
    if (z > 90.0) return INFINITY;
    if (z < -110.0) return 0.0;
    const double k = floor(z * D_INV_LN2 + 0.5);
    const double r = (z - k * D_LN2_HI) - k * D_LN2_LO;
    double p = 1.0 / 1307674368000.0;           /* 1/15! */
    p = p * r + 1.0 / 87178291200.0;            /* 1/14! */
    p = p * r + 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    const int ki = (int)k;
    const uint64_t sb = (uint64_t)(ki + 1023) << 52;
    double scale;
    memcpy(&scale, &sb, sizeof scale);
    return p * scale;
}

float or_logf(float x)
{
    if (x != x) return x;
    if (x < 0.f) return NAN;
    if (x == 0.f) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    return (float)d_log((double)x);
}

float or_expf(float x)
{
    if (x != x) return x;
    return (float)d_exp((double)x);
}

float or_powf(float x, float y)
{
    if (x != x || y != y) return NAN;
    if (y == 0.f) return 1.f;
    if (x < 0.f) return NAN;
    if (x == 0.f) return y > 0.f ? 0.f : INFINITY;
    if (x == INFINITY) return y > 0.f ? INFINITY : 0.f;
    return (float)d_exp((double)y * d_log((double)x));
}

void or_sincosf(float a, float* s, float* c)
{
    const double x = (double)a;
    const double k = floor(x * D_2_PI + 0.5);
    const double r = (x - k * D_PIO2_1) - k * D_PIO2_1T;
    const double r2 = r * r;
    /* sin r = r (1 - r^2/3! + r^4/5! - ... - r^16/17!) */
    double ps = -1.0 / 355687428096000.0;        /* -1/17! */
    ps = ps * r2 + 1.0 / 1307674368000.0;        /* 1/15! */
    ps = ps * r2 - 1.0 / 6227020800.0;           /* 1/13! */
    ps = ps * r2 + 1.0 / 39916800.0;             /* 1/11! */
    ps = ps * r2 - 1.0 / 362880.0;               /* 1/9!  */
    ps = ps * r2 + 1.0 / 5040.0;                 /* 1/7!  */
    ps = ps * r2 - 1.0 / 120.0;                  /* 1/5!  */
    ps = ps * r2 + 1.0 / 6.0;                    /* 1/3!  */
    const double sr = r - r * (r2 * ps) ;
    /* cos r = 1 - r^2/2! + r^4/4! - ... + r^16/16! */
    double pc = 1.0 / 20922789888000.0;          /* 1/16! */
    pc = pc * r2 - 1.0 / 87178291200.0;          /* 1/14! */
    pc = pc * r2 + 1.0 / 479001600.0;            /* 1/12! */
    pc = pc * r2 - 1.0 / 3628800.0;              /* 1/10! */
    pc = pc * r2 + 1.0 / 40320.0;                /* 1/8!  */
    pc = pc * r2 - 1.0 / 720.0;                  /* 1/6!  */
    pc = pc * r2 + 1.0 / 24.0;                   /* 1/4!  */
    pc = pc * r2 - 0.5;                          /* 1/2!  */
    const double cr = 1.0 + r2 * pc;
    const int q = ((int)(k - 4.0 * floor(k * 0.25))) & 3;
    double sv, cv;
    switch (q) {
    case 0: sv = sr; cv = cr; break;
    case 1: sv = cr; cv = -sr; break;
    case 2: sv = -sr; cv = -cr; break;
    default: sv = -cr; cv = sr; break;
    }
    *s = (float)sv;
    *c = (float)cv;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo)
{
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n1 = lo1;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        const uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

float or_u01(uint32_t x)
{
    return (float)(x >> 8) * (1.0f / 16777216.0f);
}
