// exact_math.h -- arithmetic shortcuts that return exactly what the reference's FP32 operations
// return, shared by the HIP kernels and the host (so the CPU tests can check them exhaustively).
//
//  birp_div_rn(a, b, rb):   RN(a / b) for a precomputed rb = RN(1/b), by two residual corrections
//                           (q1 = q0 + (a - q0 b) rb is faithful; Markstein's theorem makes the
//                           second correction correctly rounded). Replaces the per-sample
//                           divisions of DataCube.cpp:206,209,212 and the ~12-instruction IEEE
//                           division sequence with 1 multiply + 4 FMAs.
//  birp_div_rn64(a, rb64):  RN(a / b) for rb64 = RN64(1/b): the float of the double product
//                           a * rb64. That product is within 2^-52 (relative) of a/b, while a
//                           quotient of two normal floats is never a float rounding midpoint and
//                           lies at least 2^-49 (relative) from every one (for a midpoint m,
//                           a - b m is a non-zero multiple of ulp(b) times half an ulp of the
//                           quotient, and b < 2^24 ulp(b)), so rounding the
//                           product to float gives RN(a/b) for every normal quotient. 3 VALU
//                           operations (two conversions and a double multiply).
//  birp_scale_sample(vw):   f32( f64(vw) * 637810000 / (4/pi) / 10000 ), the chain of
//                           Camera.cpp:201, as one double multiply by the chain's constant plus a
//                           guard: the two results differ by < 8 double ulps, so their float
//                           roundings agree unless the product lies within 16 double ulps of a
//                           float rounding midpoint (or outside the normal float range); those
//                           rare inputs take the three-operation chain.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define BIRP_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#include <string.h>
#define BIRP_HD static inline
#endif

#define BIRP_SCALE_K (1000 * 100 * 6378.1) /* 637810000.0 in double, as Camera.cpp:201 folds it */
#define BIRP_SCALE_C (4 / 3.14159265358979323846)

BIRP_HD float birp_fmaf(float a, float b, float c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fmaf(a, b, c);
#else
    return fmaf(a, b, c);
#endif
}

BIRP_HD uint64_t birp_f64_bits(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint64_t)__double_as_longlong(x);
#else
    uint64_t u;
    memcpy(&u, &x, sizeof(u));
    return u;
#endif
}

BIRP_HD float birp_div_rn(float a, float b, float rb) {
    float q0 = a * rb;
    float e0 = birp_fmaf(-q0, b, a);
    float q1 = birp_fmaf(e0, rb, q0);
    float e1 = birp_fmaf(-q1, b, a);
    return birp_fmaf(e1, rb, q1);
}

BIRP_HD float birp_div_rn64(float a, double rb64) { return (float)((double)a * rb64); }

// The reference chain, operation by operation (Camera.cpp:201).
BIRP_HD float birp_scale_chain(float vw) {
    return (float)(((double)vw * BIRP_SCALE_K / BIRP_SCALE_C) / 10000);
}

// The chain's constant: the chain applied to 1.0 (the first product is exact).
#define BIRP_SCALE_FOLDED ((1.0 * BIRP_SCALE_K / BIRP_SCALE_C) / 10000)

// 1 when the folded product y cannot be trusted to round like the chain.
BIRP_HD int birp_scale_needs_chain(double y) {
    uint64_t u = birp_f64_bits(y);
    uint32_t lo = (uint32_t)u;
    uint32_t ex = (uint32_t)(u >> 52) & 0x7FFu;
    uint32_t m29 = lo & 0x1FFFFFFFu;
    int near_mid = (m29 - 0x0FFFFFF0u) < 0x21u;   /* |m29 - 2^28| <= 16 */
    int normal = (ex - 898u) <= 251u;             /* 2^-125 <= |y| < 2^127 */
    return (near_mid || !normal) && y != 0.0;
}

BIRP_HD float birp_scale_sample(float vw) {
    double y = (double)vw * BIRP_SCALE_FOLDED;
    if (birp_scale_needs_chain(y)) return birp_scale_chain(vw);
    return (float)y;
}
