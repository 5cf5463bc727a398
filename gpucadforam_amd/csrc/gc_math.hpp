// Device math of the field kernels: sine/cosine in revolutions and atan2.
//
// Arguments are in revolutions (one period = 1), so the range reduction
// r - rint(r) is exact for |r| < 2^22 and no 2*pi multiply is needed.
//   cos_rev(r) = sin(2*pi*d),  d = 1/4 - |r - rint(r)|  in [-1/4, 1/4]
//   sin_rev(r) = sin(2*pi*d),  d = a or sign(a)/2 - a,  a = r - rint(r)
// sin(2*pi*d) ~= d * P(d^2): minimax (Remez exchange, absolute error) odd
// polynomial of degree 9; coefficients derived by derive_poly.py.  Measured
// binary32 evaluation error <= 1.8e-7 absolute.
// atan(a), a in [0,1]: minimax odd polynomial of degree 21 (derive_poly.py),
// binary32 evaluation error <= 1e-7 absolute; octant reduction for atan2.
//
// Every operation is an IEEE binary32 add/mul/fma/div/rint/abs/min/max, so the
// same sequence evaluated on the host reproduces these results bit for bit.
#pragma once

#include <hip/hip_runtime.h>

namespace gc {

#define GC_SINREV_C1 0x1.921fb4p+2f
#define GC_SINREV_C3 -0x1.4abbb6p+5f
#define GC_SINREV_C5 0x1.46676ep+6f
#define GC_SINREV_C7 -0x1.3232fap+6f
#define GC_SINREV_C9 0x1.3c4b2cp+5f

__device__ __forceinline__ float sin2pi_reduced(float d) {
    float d2 = d * d;
    float p = __builtin_fmaf(GC_SINREV_C9, d2, GC_SINREV_C7);
    p = __builtin_fmaf(p, d2, GC_SINREV_C5);
    p = __builtin_fmaf(p, d2, GC_SINREV_C3);
    p = __builtin_fmaf(p, d2, GC_SINREV_C1);
    return d * p;
}

__device__ __forceinline__ float cos_rev(float r) {
    float a = r - __builtin_rintf(r);
    float d = 0.25f - __builtin_fabsf(a);
    return sin2pi_reduced(d);
}

// sin(2*pi*a), a = r - rint(r) in [-1/2, 1/2], reflected into [-1/4, 1/4]:
// sin(2*pi*a) = sin(2*pi*(sign(a)/2 - a)); both steps are exact (Sterbenz).
__device__ __forceinline__ float sin_rev(float r) {
    float a = r - __builtin_rintf(r);
    float d = (__builtin_fabsf(a) > 0.25f) ? __builtin_copysignf(0.5f, a) - a : a;
    return sin2pi_reduced(d);
}

#define GC_ATAN_C1 0x1p+0f
#define GC_ATAN_C3 -0x1.55554p-2f
#define GC_ATAN_C5 0x1.99950ap-3f
#define GC_ATAN_C7 -0x1.245722p-3f
#define GC_ATAN_C9 0x1.c3bbccp-4f
#define GC_ATAN_C11 -0x1.64eb78p-4f
#define GC_ATAN_C13 0x1.0bef6p-4f
#define GC_ATAN_C15 -0x1.58c94ep-5f
#define GC_ATAN_C17 0x1.4f589ap-6f
#define GC_ATAN_C19 -0x1.a316f6p-8f
#define GC_ATAN_C21 0x1.eb8dep-11f
#define GC_HALF_PI 0x1.921fb6p+0f
#define GC_PI 0x1.921fb6p+1f

// atan2(y, x) in radians; atan2(0, 0) = 0.
__device__ __forceinline__ float atan2_poly(float y, float x) {
    float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    float mx = __builtin_fmaxf(ax, ay), mn = __builtin_fminf(ax, ay);
    float a = (mx == 0.0f) ? 0.0f : mn / mx;
    float a2 = a * a;
    float q = __builtin_fmaf(GC_ATAN_C21, a2, GC_ATAN_C19);
    q = __builtin_fmaf(q, a2, GC_ATAN_C17);
    q = __builtin_fmaf(q, a2, GC_ATAN_C15);
    q = __builtin_fmaf(q, a2, GC_ATAN_C13);
    q = __builtin_fmaf(q, a2, GC_ATAN_C11);
    q = __builtin_fmaf(q, a2, GC_ATAN_C9);
    q = __builtin_fmaf(q, a2, GC_ATAN_C7);
    q = __builtin_fmaf(q, a2, GC_ATAN_C5);
    q = __builtin_fmaf(q, a2, GC_ATAN_C3);
    q = __builtin_fmaf(q, a2, GC_ATAN_C1);
    float t = a * q;
    if (ay > ax) t = GC_HALF_PI - t;
    if (x < 0.0f) t = GC_PI - t;
    if (y < 0.0f) t = -t;
    return t;
}

// Correctly rounded square root of x >= 2^-96 (and of +inf and NaN).
// v_sqrt_f32 is within 1 ulp on that range; Tuckerman's rounding test
// (Markstein 1990) then picks the correctly rounded neighbour from the fma
// residuals of s - ulp and s + ulp:
//   s - ulp  if x - (s - ulp) s <= 0,   s + ulp  if x - (s + ulp) s > 0.
// This equals sqrtf on that range (and on +inf, where both residuals are
// NaN), and at +0 and -0 (v_sqrt_f32 returns them; the residual of s - ulp
// is NaN and that of s + ulp is not positive).
__device__ __forceinline__ float sqrt_rn_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    const float rd = __builtin_fmaf(-sd, s, x);
    const float ru = __builtin_fmaf(-su, s, x);
    const float r0 = rd <= 0.0f ? sd : s;
    return ru > 0.0f ? su : r0;
}

// Correctly rounded square root of any x: an input below 2^-96 (subnormals,
// zeros and negatives included) is scaled by 2^64 first and its root by
// 2^-32; both products are exact and correct rounding commutes with them, so
// the result is still sqrtf(x) (zeros keep their sign, negatives give NaN).
__device__ __forceinline__ float sqrt_rn(float x) {
    const bool tiny = x < 0x1p-96f;
    const float r = sqrt_rn_normal(tiny ? x * 0x1p64f : x);
    return tiny ? r * 0x1p-32f : r;
}

}  // namespace gc
