// hr_math.h -- single-precision vector helpers, the written-spec
// transcendental functions and the counter-based RNG, for gfx950 kernels and
// the host side of the library.
//
// Arithmetic contract: every helper keeps the evaluation order of the
// reference's host algebra (src/Core/Algebra.hpp:38-310) and the library is
// compiled with -ffp-contract=off, so a kernel result equals the IEEE result
// of the same expression on the CPU.  Division and square root are the
// correctly rounded ones (HIP's default for float).
//
// Transcendentals follow one definition, f(x) = (float)F((double)x), with F
// evaluated in double precision from +, -, *, / and floor:
//   log: m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh((m-1)/(m+1)) by its odd
//        series to s^23 (Abramowitz & Stegun 4.1.27);
//   exp: Cody-Waite reduction by a two-part ln 2, Taylor series to r^15;
//   pow: exp(y ln x);
//   sin/cos: reduction by a two-part pi/2 (fdlibm's pio2_1/pio2_1t split),
//        Taylor polynomials to r^17 / r^16.
// RNG: Philox4x32-10 (Salmon et al., SC'11), float = (u >> 8) * 2^-24.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HR_HD __host__ __device__ __forceinline__

namespace hr {

constexpr float kEps = 0.00001f;          // src/CUDAStdAfx.h:117
constexpr float kPi = 3.14159265358979323846f;
constexpr float kIsectEps = 0.0001f;      // src/RT/Algorithm/RayTriangleIntersector.h:36
constexpr float kFltMax = 3.402823466e+38f;

struct f3 {
    float x, y, z;
};

HR_HD f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
HR_HD f3 rep(float v) { return mk(v, v, v); }
HR_HD f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
HR_HD f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
HR_HD f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
HR_HD f3 operator/(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
HR_HD f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
HR_HD f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
HR_HD f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
HR_HD f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
HR_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HR_HD f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
HR_HD f3 vmin(f3 a, f3 b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
HR_HD f3 vmax(f3 a, f3 b) { return mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
HR_HD float hmax(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
HR_HD f3 vabs(f3 a) { return mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
HR_HD float len(f3 a) { return sqrtf(dot(a, a)); }
HR_HD f3 normalize(f3 a) { return a * (1.f / sqrtf(dot(a, a))); }
HR_HD float get(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// getLocalCoordinates (src/Core/Algebra.hpp:281-298), tangent scaled by
// 1/(n_c0^2 + n_z^2) exactly as the reference writes it.
HR_HD void local_frame(f3 n, f3& t, f3& b)
{
    const bool xdom = fabsf(n.x) > fabsf(n.y);
    const float sig = xdom ? -1.f : 1.f;
    const float nc0 = xdom ? n.x : n.y;
    const float inv_len = 1.f / (nc0 * nc0 + n.z * n.z);
    const float tc0 = n.z * sig * inv_len;
    t.x = xdom ? tc0 : 0.f;
    t.y = xdom ? 0.f : tc0;
    t.z = nc0 * -1.f * sig * inv_len;
    b = cross(n, t);
}

// ---------------------------------------------------------------------------
// double-precision kernels of the transcendental definition
// ---------------------------------------------------------------------------

HR_HD double d_floor(double x) { return floor(x); }

HR_HD double d_ln(double x)
{
    uint64_t b = __builtin_bit_cast(uint64_t, x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    b = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m = __builtin_bit_cast(double, b);
    if (m > 1.41421356237309514547e+00) { m = m * 0.5; e = e + 1; }
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
    const double de = (double)e;
    return de * 6.93147180369123816490e-01 + (de * 1.90821492927058770002e-10 + s * p);
}

HR_HD double d_exp(double z)
{
    if (z > 90.0) return __builtin_huge_val();
    if (z < -110.0) return 0.0;
    const double k = d_floor(z * 1.44269504088896338700e+00 + 0.5);
    const double r = (z - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
    double p = 1.0 / 1307674368000.0;
    p = p * r + 1.0 / 87178291200.0;
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
    const uint64_t sb = (uint64_t)((int)k + 1023) << 52;
    return p * __builtin_bit_cast(double, sb);
}

HR_HD float spec_powf(float x, float y)
{
    if (x != x || y != y) return __builtin_nanf("");
    if (y == 0.f) return 1.f;
    if (x < 0.f) return __builtin_nanf("");
    if (x == 0.f) return y > 0.f ? 0.f : __builtin_huge_valf();
    if (x == __builtin_huge_valf()) return y > 0.f ? __builtin_huge_valf() : 0.f;
    return (float)d_exp((double)y * d_ln((double)x));
}

HR_HD float spec_logf(float x)
{
    if (x != x) return x;
    if (x < 0.f) return __builtin_nanf("");
    if (x == 0.f) return -__builtin_huge_valf();
    if (x == __builtin_huge_valf()) return x;
    return (float)d_ln((double)x);
}

HR_HD float spec_expf(float x)
{
    if (x != x) return x;
    return (float)d_exp((double)x);
}

HR_HD void spec_sincosf(float a, float& s, float& c)
{
    const double x = (double)a;
    const double k = d_floor(x * 6.36619772367581382433e-01 + 0.5);
    const double r = (x - k * 1.57079632673412561417e+00) - k * 6.07710050650619224932e-11;
    const double r2 = r * r;
    double ps = -1.0 / 355687428096000.0;
    ps = ps * r2 + 1.0 / 1307674368000.0;
    ps = ps * r2 - 1.0 / 6227020800.0;
    ps = ps * r2 + 1.0 / 39916800.0;
    ps = ps * r2 - 1.0 / 362880.0;
    ps = ps * r2 + 1.0 / 5040.0;
    ps = ps * r2 - 1.0 / 120.0;
    ps = ps * r2 + 1.0 / 6.0;
    const double sr = r - r * (r2 * ps);
    double pc = 1.0 / 20922789888000.0;
    pc = pc * r2 - 1.0 / 87178291200.0;
    pc = pc * r2 + 1.0 / 479001600.0;
    pc = pc * r2 - 1.0 / 3628800.0;
    pc = pc * r2 + 1.0 / 40320.0;
    pc = pc * r2 - 1.0 / 720.0;
    pc = pc * r2 + 1.0 / 24.0;
    pc = pc * r2 - 0.5;
    const double cr = 1.0 + r2 * pc;
    const int q = ((int)(k - 4.0 * d_floor(k * 0.25))) & 3;
    double sv = sr, cv = cr;
    if (q == 1) { sv = cr; cv = -sr; }
    else if (q == 2) { sv = -sr; cv = -cr; }
    else if (q == 3) { sv = -cr; cv = sr; }
    s = (float)sv;
    c = (float)cv;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 counter RNG
// ---------------------------------------------------------------------------

struct u4 {
    uint32_t a, b, c, d;
};

HR_HD u4 philox4x32_10(u4 ctr, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)ctr.a;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)ctr.c;
        u4 n;
        n.a = (uint32_t)(p1 >> 32) ^ ctr.b ^ k0;
        n.b = (uint32_t)p1;
        n.c = (uint32_t)(p0 >> 32) ^ ctr.d ^ k1;
        n.d = (uint32_t)p0;
        ctr = n;
    }
    return ctr;
}

HR_HD float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// stream identifiers: pixel / image / stream / tag form the counter
constexpr uint32_t kStreamPrimary = 0u;
constexpr uint32_t kStreamAO = 16u;      // + sample index
constexpr uint32_t kStreamPT = 1024u;    // + 2 * vertex + {0, 1}
constexpr uint32_t kStreamPathMK = 2048u;     // multi-kernel path tracer: + 2 * vertex + {0, 1}
constexpr uint32_t kStreamPathLight = 4096u;  // its shadow-ray light sample
constexpr uint32_t kCounterTag = 0x68617070u;

HR_HD void rand4(uint32_t seed_lo, uint32_t seed_hi, uint32_t pixel, uint32_t image,
                 uint32_t stream, float u[4])
{
    u4 c;
    c.a = pixel; c.b = image; c.c = stream; c.d = kCounterTag;
    const u4 o = philox4x32_10(c, seed_lo, seed_hi);
    u[0] = u01(o.a); u[1] = u01(o.b); u[2] = u01(o.c); u[3] = u01(o.d);
}

}  // namespace hr
