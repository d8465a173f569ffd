// Single-precision sin/cos, asin/acos, atan2 and x^5 for the shading kernels.
//
// The same source is compiled for gfx950 and for the host, so the directions, lobe choices and
// weights of a path come out as the same bits on the device and in a host build. Every build of
// this file runs with FP contraction off, and every multiply-add below is an explicit fma. fma,
// sqrt and division are correctly rounded on gfx950 (the library is built with correctly rounded
// divide/sqrt) and in the C library alike, so each function is one fixed sequence of correctly
// rounded operations.
//
// Method.
//   sin/cos: Cody-Waite reduction by pi/2 with pi/2 split into three floats (hi + mid + lo,
//     about 72 bits; each step one fma, so x - j hi is exact for the small j of the samplers'
//     angles and a result near a zero of sin or cos keeps its relative accuracy), then
//     odd/even polynomials on [-pi/4, pi/4] and the quadrant j mod 4.
//   asin/acos: x + x^3 P(x^2) on [0, 1/2]; above 1/2, asin(x) = pi/2 - 2 asin(sqrt((1 - x)/2))
//     and acos(x) = 2 asin(sqrt((1 - x)/2)).
//   atan2: the ratio min/max of |x|, |y| in [0, 1]; above tan(pi/8) the identity
//     atan(z) = pi/4 + atan((z - 1)/(z + 1)); then z + z^3 P(z^2); octant by |y| > |x| and by the
//     signs (signbit, so signed zeros follow C's atan2).
// The polynomial coefficients are near-minimax fits of the relative error (Lawson's iteratively
// reweighted least squares on 20001 points in float64), rounded to float. Constants near pi are
// split into a float and its float remainder. The worst-case error of each function against a
// float64 evaluation is checked by the tests (within 2 ulp for sin, cos, asin and acos, 3 for atan2).
//
// Arguments: sincos is accurate for |x| up to a few thousand and deterministic beyond; asin/acos expect
// [-1, 1] (callers clamp); atan2 takes any finite pair. sin(-0) returns +0.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define WF_FM __host__ __device__ __forceinline__
#else
#define WF_FM inline
#endif

namespace wf {
namespace fm {

WF_FM float mad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

constexpr float kPio2Hi = 1.57079637f, kPio2Lo = -4.37113883e-08f;
constexpr float kPiHi = 3.14159274f, kPiLo = -8.74227766e-08f;
constexpr float kPio4Hi = 0.785398185f, kPio4Lo = -2.18556941e-08f;

WF_FM void sincos(float x, float& s, float& c) {
    const float j = __builtin_rintf(x * 0.636619747f);
    float r = mad(-j, kPio2Hi, x);
    r = mad(-j, kPio2Lo, r);
    r = mad(-j, -1.71512451e-15f, r);
    const float t = r * r;
    float ps = mad(t, 2.718122e-06f, -0.000198393129f);
    ps = mad(t, ps, 0.00833332911f);
    ps = mad(t, ps, -0.166666672f);
    const float sr = mad(r * t, ps, r);
    float pc = mad(t, -2.72102369e-07f, 2.47995195e-05f);
    pc = mad(t, pc, -0.00138888834f);
    pc = mad(t, pc, 0.0416666679f);
    const float cr = mad(t * t, pc, mad(t, -0.5f, 1.0f));
    switch ((int)j & 3) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
    }
}

WF_FM float sin(float x) {
    float s, c;
    sincos(x, s, c);
    return s;
}

WF_FM float cos(float x) {
    float s, c;
    sincos(x, s, c);
    return c;
}

// x + x^3 P(x^2) ~ asin(x) for 0 <= x <= 1/2; t = x^2.
WF_FM float asin_core(float x, float t) {
    float p = mad(t, 0.0421663076f, 0.0241795201f);
    p = mad(t, p, 0.0454703756f);
    p = mad(t, p, 0.0749529749f);
    p = mad(t, p, 0.166667521f);
    return mad(x * t, p, x);
}

WF_FM float asin(float x) {
    const float a = __builtin_fabsf(x);
    float r;
    if (a <= 0.5f) {
        r = asin_core(a, a * a);
    } else {
        const float z = 0.5f * (1.0f - a);
        const float s = __builtin_sqrtf(z);
        r = mad(-2.0f, asin_core(s, z), kPio2Hi) + kPio2Lo;
    }
    return __builtin_copysignf(r, x);
}

WF_FM float acos(float x) {
    const float a = __builtin_fabsf(x);
    if (a <= 0.5f) return kPio2Hi - (asin_core(x, x * x) - kPio2Lo);
    const float z = 0.5f * (1.0f - a);
    const float s = __builtin_sqrtf(z);
    const float h = 2.0f * asin_core(s, z);
    return x > 0.0f ? h : kPiHi - (h - kPiLo);
}

// z + z^3 P(z^2) ~ atan(z) for |z| <= tan(pi/8).
WF_FM float atan_core(float z) {
    const float t = z * z;
    float p = mad(t, -0.0607822388f, 0.105938151f);
    p = mad(t, p, -0.142435342f);
    p = mad(t, p, 0.199984714f);
    p = mad(t, p, -0.333333164f);
    return mad(z * t, p, z);
}

WF_FM float atan2(float y, float x) {
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    const float hi = ax > ay ? ax : ay;
    const float lo = ax > ay ? ay : ax;
    const float z = hi == 0.0f ? 0.0f : lo / hi;
    float r;
    if (z > 0.414213568f) {
        r = kPio4Hi + (atan_core((z - 1.0f) / (z + 1.0f)) + kPio4Lo);
    } else {
        r = atan_core(z);
    }
    if (ay > ax) r = (kPio2Hi - r) + kPio2Lo;
    if (__builtin_signbit(x)) r = (kPiHi - r) + kPiLo;
    return __builtin_copysignf(r, y);
}

// x^5 by two squarings (Schlick's (1 - cos)^5).
WF_FM float pow5(float x) {
    const float x2 = x * x;
    return x2 * x2 * x;
}

}  // namespace fm
}  // namespace wf
