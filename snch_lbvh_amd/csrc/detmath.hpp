// Deterministic FP32 transcendentals: acos, asin, atan2, sin/cos.
//
// The reference calls the CUDA math library here (core/cone.cuh:94, 102, 116,
// 185, 193, 207, 290-291, 446; scene.cuh:153, 952).  Vendor math libraries
// differ by an ulp from one platform to the next, so a GPU-built tree and a
// CPU-built tree would disagree in the last bit of their cones, and a rare
// borderline pruning decision with them.  These functions are built only
// from IEEE-exact operations (+ - * / sqrt floor and comparisons; the whole
// library is compiled with -ffp-contract=off), so the GPU kernels, the
// library's host code and any IEEE single-precision CPU produce identical
// bits.  Method: range reduction to |x| <= 0.5 (asin), |x| <= tan(pi/8)
// (atan) and |r| <= pi/4 with a Cody-Waite split of pi/2 (sin, cos), then
// polynomials fitted by least squares on Chebyshev nodes
// (csrc/gen_detmath.py).  Accuracy: <= 3 ulp over the ranges used.
#pragma once
#include <cstdint>

#ifndef SNCH_HD
#define SNCH_HD __host__ __device__ __forceinline__
#endif

namespace snch {
namespace dm {

constexpr float PI_HI = 3.14159274101257324219f;
constexpr float PI_LO = -8.742278012618954e-08f;
constexpr float PIO2_HI = 1.57079637050628662109f;
constexpr float PIO2_LO = -4.371139006309477e-08f;
constexpr float PIO4_HI = 0.785398185253143310547f;
constexpr float PIO4_LO = -2.1855695031547384e-08f;
constexpr float TAN_PI8 = 0.414213562373095f;
constexpr float TWO_OVER_PI = 0.636619772367581343f;
constexpr float PIO2_A = 1.5703125f;
constexpr float PIO2_B = 4.837512969970703125e-04f;
constexpr float PIO2_C = 7.549790126404332e-08f;

SNCH_HD float fabs_(float x) { return __builtin_fabsf(x); }
SNCH_HD float sqrt_(float x) { return __builtin_sqrtf(x); }
SNCH_HD bool signbit_(float x) { return (__builtin_bit_cast(uint32_t, x) >> 31) != 0; }

SNCH_HD float asin_poly(float z)
{
    float p = 0.03820645064115524f;
    p = p * z + 0.026494277641177177f;
    p = p * z + 0.045010700821876526f;
    p = p * z + 0.07498808950185776f;
    p = p * z + 0.16666673123836517f;
    return p;
}

SNCH_HD float atan_poly(float z)
{
    float q = -0.06459376215934753f;
    q = q * z + 0.10746298730373383f;
    q = q * z + -0.1426423192024231f;
    q = q * z + 0.1999955028295517f;
    q = q * z + -0.3333333134651184f;
    return q;
}

SNCH_HD float sin_poly(float z)
{
    float s = 2.724996420511161e-06f;
    s = s * z + -0.00019840087043121457f;
    s = s * z + 0.008333331905305386f;
    s = s * z + -0.1666666716337204f;
    return s;
}

SNCH_HD float cos_poly(float z)
{
    float c = -2.730098742631526e-07f;
    c = c * z + 2.480059993104078e-05f;
    c = c * z + -0.00138888880610466f;
    c = c * z + 0.0416666679084301f;
    return c;
}

SNCH_HD float asin_small(float a)
{
    float z = a * a;
    return a + (a * z) * asin_poly(z);
}

SNCH_HD float asinf_(float x)
{
    float a = fabs_(x);
    float r;
    if (!(a <= 1.0f))
        return (x - x) / (x - x);
    if (a <= 0.5f) {
        r = asin_small(a);
    } else {
        float z = (1.0f - a) * 0.5f;
        float s = sqrt_(z);
        float t = s + (s * z) * asin_poly(z);
        r = (PIO2_HI - 2.0f * t) + PIO2_LO;
    }
    return x < 0.0f ? -r : r;
}

SNCH_HD float acosf_(float x)
{
    if (!(fabs_(x) <= 1.0f))
        return (x - x) / (x - x);
    if (x > 0.5f) {
        float z = (1.0f - x) * 0.5f;
        float s = sqrt_(z);
        float t = s + (s * z) * asin_poly(z);
        return 2.0f * t;
    }
    if (x < -0.5f) {
        float z = (1.0f + x) * 0.5f;
        float s = sqrt_(z);
        float t = s + (s * z) * asin_poly(z);
        return (PI_HI - 2.0f * t) + PI_LO;
    }
    float t = asin_small(fabs_(x));
    if (x < 0.0f)
        t = -t;
    return (PIO2_HI - t) + PIO2_LO;
}

SNCH_HD float atan_small(float a)
{
    float z = a * a;
    return a + (a * z) * atan_poly(z);
}

SNCH_HD float atan2f_(float y, float x)
{
    if (x != x || y != y)
        return x + y;
    float ax = fabs_(x), ay = fabs_(y);
    float mx = ax < ay ? ay : ax;
    float mn = ax < ay ? ax : ay;
    float t;
    if (mx == 0.0f) {
        t = 0.0f;
    } else if (mx == __builtin_inff()) {
        t = (mn == __builtin_inff()) ? PIO4_HI : 0.0f;
    } else {
        float a = mn / mx;
        if (a > TAN_PI8) {
            float a2 = (a - 1.0f) / (a + 1.0f);
            t = (PIO4_HI + atan_small(a2)) + PIO4_LO;
        } else {
            t = atan_small(a);
        }
    }
    if (ay > ax)
        t = (PIO2_HI - t) + PIO2_LO;
    if (signbit_(x))
        t = (PI_HI - t) + PI_LO;
    if (signbit_(y))
        t = -t;
    return t;
}

// accurate for |x| < 2^12 * pi/2
SNCH_HD void sincosf_(float x, float* sn, float* cs)
{
    float kf = __builtin_floorf(x * TWO_OVER_PI + 0.5f);
    int k = (int)kf;
    float r = ((x - kf * PIO2_A) - kf * PIO2_B) - kf * PIO2_C;
    float z = r * r;
    float s = r + (r * z) * sin_poly(z);
    float c = (1.0f - 0.5f * z) + (z * z) * cos_poly(z);
    switch (k & 3) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
    }
}

// log(x) for positive normal x: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log(m) = 2 atanh(s) = 2 s + s^3 P(s^2), s = (m - 1) / (m + 1)
constexpr float LN2_HI = 0.6931471824645996f;
constexpr float LN2_LO = -1.904654323148236e-09f;
SNCH_HD float logf_(float x)
{
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    int e = (int)((u >> 23) & 0xFFu) - 127;
    float m = __builtin_bit_cast(float, (u & 0x7FFFFFu) | 0x3F800000u);
    if (m > 1.41421356237f) {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float z = s * s;
    float p = 0.2333136796951294f;
    p = p * z + 0.28550779819488525f;
    p = p * z + 0.4000012278556824f;
    p = p * z + 0.6666666865348816f;
    return (float)e * LN2_HI + ((2.0f * s + (s * z) * p) + (float)e * LN2_LO);
}

} // namespace dm
} // namespace snch
