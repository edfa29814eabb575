/*
 * ORACLE (test infrastructure only) -- deterministic FP32 transcendental set.
 *
 * The reference calls CUDA's acosf/asinf/atan2f/sinf/cosf
 * (include/snch_lbvh/core/cone.cuh:94,102,116,185,193,207,270-271,290-291,
 * 446; scene.cuh:153,952).  Vendor libm results differ by an ulp between
 * platforms, so the library evaluates these functions with its own
 * polynomial approximations made only of IEEE-exact operations
 * (+ - * / sqrt, comparisons, floor; no fused multiply-add).  This header is
 * the oracle's independent copy of that method; the coefficients are the ones
 * printed by snch_lbvh_amd/csrc/gen_detmath.py (least-squares fit on Chebyshev
 * nodes).  Accuracy is pinned by tests/test_oracle_math.py against float64
 * libm (<= 3 ulp or 2e-7 absolute near zero).
 *
 * Build with -ffp-contract=off: the expression order below is the contract.
 */
#ifndef SNCH_ORACLE_DETMATH_H
#define SNCH_ORACLE_DETMATH_H
#include <math.h>
#include <stdint.h>
#include <string.h>

#define ODM_PI_HI 3.14159274101257324219f    /* float(pi) */
#define ODM_PI_LO (-8.742278012618954e-08f) /* pi - float(pi) */
#define ODM_PIO2_HI 1.57079637050628662109f
#define ODM_PIO2_LO (-4.371139006309477e-08f)
#define ODM_PIO4_HI 0.785398185253143310547f
#define ODM_PIO4_LO (-2.1855695031547384e-08f)
#define ODM_TAN_PI8 0.414213562373095f
#define ODM_TWO_OVER_PI 0.636619772367581343f
/* Cody-Waite split of pi/2: A and B carry few significant bits so k*A, k*B
 * are exact for |k| < 2^12. */
#define ODM_PIO2_A 1.5703125f
#define ODM_PIO2_B 4.837512969970703125e-04f
#define ODM_PIO2_C 7.549790126404332e-08f

static inline float odm_asin_poly(float z)
{
    float p = 0.03820645064115524f;
    p = p * z + 0.026494277641177177f;
    p = p * z + 0.045010700821876526f;
    p = p * z + 0.07498808950185776f;
    p = p * z + 0.16666673123836517f;
    return p;
}

static inline float odm_atan_poly(float z)
{
    float q = -0.06459376215934753f;
    q = q * z + 0.10746298730373383f;
    q = q * z + -0.1426423192024231f;
    q = q * z + 0.1999955028295517f;
    q = q * z + -0.3333333134651184f;
    return q;
}

static inline float odm_sin_poly(float z)
{
    float s = 2.724996420511161e-06f;
    s = s * z + -0.00019840087043121457f;
    s = s * z + 0.008333331905305386f;
    s = s * z + -0.1666666716337204f;
    return s;
}

static inline float odm_cos_poly(float z)
{
    float c = -2.730098742631526e-07f;
    c = c * z + 2.480059993104078e-05f;
    c = c * z + -0.00138888880610466f;
    c = c * z + 0.0416666679084301f;
    return c;
}

static inline int odm_signbit(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    return (int)(u >> 31);
}

/* asin(a) for a in [0, 0.5] */
static inline float odm_asin_small(float a)
{
    float z = a * a;
    return a + (a * z) * odm_asin_poly(z);
}

static inline float odm_asinf(float x)
{
    float a = fabsf(x);
    float r;
    if (!(a <= 1.0f))
        return (x - x) / (x - x);
    if (a <= 0.5f) {
        r = odm_asin_small(a);
    } else {
        float z = (1.0f - a) * 0.5f;
        float s = sqrtf(z);
        float t = s + (s * z) * odm_asin_poly(z);
        r = (ODM_PIO2_HI - 2.0f * t) + ODM_PIO2_LO;
    }
    return x < 0.0f ? -r : r;
}

static inline float odm_acosf(float x)
{
    if (!(fabsf(x) <= 1.0f))
        return (x - x) / (x - x);
    if (x > 0.5f) {
        float z = (1.0f - x) * 0.5f;
        float s = sqrtf(z);
        float t = s + (s * z) * odm_asin_poly(z);
        return 2.0f * t;
    }
    if (x < -0.5f) {
        float z = (1.0f + x) * 0.5f;
        float s = sqrtf(z);
        float t = s + (s * z) * odm_asin_poly(z);
        return (ODM_PI_HI - 2.0f * t) + ODM_PI_LO;
    }
    {
        float a = fabsf(x);
        float t = odm_asin_small(a);
        if (x < 0.0f)
            t = -t;
        return (ODM_PIO2_HI - t) + ODM_PIO2_LO;
    }
}

/* atan(a) for |a| <= tan(pi/8) */
static inline float odm_atan_small(float a)
{
    float z = a * a;
    return a + (a * z) * odm_atan_poly(z);
}

static inline float odm_atan2f(float y, float x)
{
    float ax, ay, mn, mx, t;
    if (x != x || y != y)
        return x + y;
    ax = fabsf(x);
    ay = fabsf(y);
    mx = ax < ay ? ay : ax;
    mn = ax < ay ? ax : ay;
    if (mx == 0.0f) {
        t = 0.0f;
    } else if (mx == INFINITY) {
        t = (mn == INFINITY) ? ODM_PIO4_HI : 0.0f;
    } else {
        float a = mn / mx;
        if (a > ODM_TAN_PI8) {
            float a2 = (a - 1.0f) / (a + 1.0f);
            t = (ODM_PIO4_HI + odm_atan_small(a2)) + ODM_PIO4_LO;
        } else {
            t = odm_atan_small(a);
        }
    }
    if (ay > ax)
        t = (ODM_PIO2_HI - t) + ODM_PIO2_LO;
    if (odm_signbit(x))
        t = (ODM_PI_HI - t) + ODM_PI_LO;
    if (odm_signbit(y))
        t = -t;
    return t;
}

/* sin and cos together; accurate for |x| < 2^12 * pi/2 */
static inline void odm_sincosf(float x, float *sn, float *cs)
{
    float kf = floorf(x * ODM_TWO_OVER_PI + 0.5f);
    int k = (int)kf;
    float r = ((x - kf * ODM_PIO2_A) - kf * ODM_PIO2_B) - kf * ODM_PIO2_C;
    float z = r * r;
    float s = r + (r * z) * odm_sin_poly(z);
    float c = (1.0f - 0.5f * z) + (z * z) * odm_cos_poly(z);
    switch (k & 3) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
    }
}

/* log(x) for positive normal x: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
 * log(m) = 2 atanh(s) = 2 s + s^3 P(s^2), s = (m - 1) / (m + 1) */
#define ODM_LN2_HI 0.6931471824645996f
#define ODM_LN2_LO (-1.904654323148236e-09f)
static inline float odm_logf(float x)
{
    uint32_t u;
    int e;
    float m, s, z, p;
    memcpy(&u, &x, 4);
    e = (int)((u >> 23) & 0xFFu) - 127;
    u = (u & 0x7FFFFFu) | 0x3F800000u;
    memcpy(&m, &u, 4);
    if (m > 1.41421356237f) {
        m = m * 0.5f;
        e += 1;
    }
    s = (m - 1.0f) / (m + 1.0f);
    z = s * s;
    p = 0.2333136796951294f;
    p = p * z + 0.28550779819488525f;
    p = p * z + 0.4000012278556824f;
    p = p * z + 0.6666666865348816f;
    return (float)e * ODM_LN2_HI + ((2.0f * s + (s * z) * p) + (float)e * ODM_LN2_LO);
}

#endif
