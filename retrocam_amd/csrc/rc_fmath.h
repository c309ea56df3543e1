/*
 * rc_fmath.h — single-precision elementary functions with a fixed operation
 * sequence, compiled identically for gfx950 device code and for x86-64 host code.
 *
 * Why: GLSL leaves the precision of sin/cos/exp/log/pow to the driver
 * (pathtrace.comp:267-268,299,511,558; atrous_denoise.frag; p01..p16), so no
 * bit pattern is "the" reference result. The render chain needs host and device
 * results that agree bit for bit (primary hit IDs, PCG streams) and to 1e-5 on
 * the post chain, so both sides evaluate these routines with the same IEEE-754
 * binary32 operations (no FMA contraction: every file including this header is
 * compiled with -ffp-contract=off; '/' and sqrtf are correctly rounded on both).
 *
 * Methods (published, restated here):
 *   sin/cos — Cody & Waite (1980) additive range reduction by pi/2 with a
 *             four-part constant, then the minimax polynomials of the Cephes
 *             library's sinf/cosf (S. L. Moshier, 1984-1992) on [-pi/4, pi/4].
 *   exp     — Cody-Waite reduction by ln 2, Cephes expf polynomial, ldexpf.
 *   log     — frexpf + Cephes logf polynomial on [sqrt(1/2)-1, sqrt(2)-1].
 *   pow     — exp(y * log(x)) for x > 0 (the GLSL definition for x > 0).
 * Accuracy (checked against float64 in the tests): sin/cos <= 2 ulp plus
 * 2^-24 absolute for |x| <= 1e5; exp <= 2 ulp on [-87, 88]; log <= 2 ulp.
 */
#ifndef RC_FMATH_H
#define RC_FMATH_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define RC_FN __host__ __device__ static inline
#else
#define RC_FN static inline
#endif

#define RC_PI_F 3.14159265359f

RC_FN uint32_t rc_f2u(float f) { union { float f; uint32_t u; } c; c.f = f; return c.u; }
RC_FN float rc_u2f(uint32_t u) { union { float f; uint32_t u; } c; c.u = u; return c.f; }
RC_FN int32_t rc_f2i_bits(float f) { union { float f; int32_t i; } c; c.f = f; return c.i; }

/* pi/2 = P1 + P2 + P3 + P4; P1..P3 carry 8 significant bits each, so k*P1,
 * k*P2 and k*P3 are exact for |k| < 2^16 (|x| < ~1e5). */
#define RC_PIO2_P1  1.5703125f
#define RC_PIO2_P2  0.000484466552734375f
#define RC_PIO2_P3 (-6.407499313354492e-07f)
#define RC_PIO2_P4  9.92093629470503e-10f
#define RC_2_OVER_PI 0.6366197466850281f

RC_FN float rc_poly_sin(float r) {
    float r2 = r * r;
    float p = -1.9515295891e-4f;
    p = p * r2 + 8.3321608736e-3f;
    p = p * r2 + -1.6666654611e-1f;
    return r + (r * r2) * p;
}

RC_FN float rc_poly_cos(float r) {
    float r2 = r * r;
    float p = 2.443315711809948e-5f;
    p = p * r2 + -1.388731625493765e-3f;
    p = p * r2 + 4.166664568298827e-2f;
    return (1.0f - 0.5f * r2) + (r2 * r2) * p;
}

/* Reduces x to r in [-pi/4, pi/4] and quadrant q (x = r + q*pi/2, mod 4). */
RC_FN float rc_reduce_pio2(float x, int *q) {
    float k = rintf(x * RC_2_OVER_PI);
    float r = x - k * RC_PIO2_P1;
    r = r - k * RC_PIO2_P2;
    r = r - k * RC_PIO2_P3;
    r = r - k * RC_PIO2_P4;
    *q = ((int)k) & 3;
    return r;
}

RC_FN float rc_sinf(float x) {
    int q;
    float r = rc_reduce_pio2(x, &q);
    float s = rc_poly_sin(r), c = rc_poly_cos(r);
    switch (q) {
        case 0: return s;
        case 1: return c;
        case 2: return -s;
        default: return -c;
    }
}

RC_FN float rc_cosf(float x) {
    int q;
    float r = rc_reduce_pio2(x, &q);
    float s = rc_poly_sin(r), c = rc_poly_cos(r);
    switch (q) {
        case 0: return c;
        case 1: return -s;
        case 2: return -c;
        default: return s;
    }
}

#define RC_LOG2E_F  1.4426950216293335f
#define RC_LN2_HI   0.693359375f
#define RC_LN2_LO  (-2.12194440e-4f)

RC_FN float rc_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283905206835f) return INFINITY;
    if (x < -103.972077083991796f) return 0.0f;
    float n = rintf(x * RC_LOG2E_F);
    float r = x - n * RC_LN2_HI;
    r = r - n * RC_LN2_LO;
    float z = r * r;
    float p = 1.9875691500e-4f;
    p = p * r + 1.3981999507e-3f;
    p = p * r + 8.3334519073e-3f;
    p = p * r + 4.1665795894e-2f;
    p = p * r + 1.6666665459e-1f;
    p = p * r + 5.0000001201e-1f;
    float y = (p * z + r) + 1.0f;
    return ldexpf(y, (int)n);
}

#define RC_SQRTHF 0.707106781186547524f

RC_FN float rc_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    int e;
    float m = frexpf(x, &e);
    if (m < RC_SQRTHF) { e -= 1; m = (m + m) - 1.0f; }
    else { m = m - 1.0f; }
    float z = m * m;
    float p = 7.0376836292e-2f;
    p = p * m + -1.1514610310e-1f;
    p = p * m + 1.1676998740e-1f;
    p = p * m + -1.2420140846e-1f;
    p = p * m + 1.4249322787e-1f;
    p = p * m + -1.6668057665e-1f;
    p = p * m + 2.0000714765e-1f;
    p = p * m + -2.4999993993e-1f;
    p = p * m + 3.3333331174e-1f;
    float fe = (float)e;
    float y = (p * m) * z;
    y = y + fe * RC_LN2_LO;
    y = y - 0.5f * z;
    float r = m + y;
    return r + fe * RC_LN2_HI;
}

RC_FN float rc_powf(float x, float y) {
    if (x > 0.0f) return rc_expf(y * rc_logf(x));
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : (y == 0.0f ? 1.0f : INFINITY);
    return NAN;
}

/* IEEE minNum/maxNum (GLSL min/max on non-NaN operands). */
RC_FN float rc_minf(float a, float b) { return fminf(a, b); }
RC_FN float rc_maxf(float a, float b) { return fmaxf(a, b); }
RC_FN float rc_clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

/* float -> binary16 bits, round to nearest even (GL RGBA16F image stores). */
RC_FN uint16_t rc_f32_to_f16(float f) {
    uint32_t x = rc_f2u(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) {                      /* inf or NaN */
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    }
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);   /* rounds to inf */
    if (ax < 0x33000001u) return (uint16_t)sign;              /* <= 2^-25: rounds to 0 */
    int32_t e = (int32_t)(ax >> 23) - 127;
    uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
    if (e < -14) {                                /* subnormal half */
        int shift = -14 - e + 13;                 /* 14..24 */
        uint32_t q = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (q & 1u))) q += 1u;
        return (uint16_t)(sign | q);
    }
    uint32_t q = mant >> 13;
    uint32_t rem = mant & 0x1fffu;
    uint32_t hbits = ((uint32_t)(e + 15) << 10) | (q & 0x3ffu);
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) hbits += 1u;
    return (uint16_t)(sign | hbits);
}

RC_FN float rc_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t m = (uint32_t)h & 0x3ffu;
    if (e == 0) {
        if (m == 0) return rc_u2f(sign);
        float v = (float)m * 5.9604644775390625e-8f;   /* m * 2^-24, exact */
        return sign ? -v : v;
    }
    if (e == 31) return rc_u2f(sign | 0x7f800000u | (m << 13));
    return rc_u2f(sign | ((e + 112u) << 23) | (m << 13));
}

#endif /* RC_FMATH_H */
