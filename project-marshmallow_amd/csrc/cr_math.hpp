// Correctly rounded single-precision square root and reciprocal for gfx950 device code.
//
// v_sqrt_f32 and v_rcp_f32 are accurate to 1 ulp, and the compiler lowers sqrtf / __fsqrt_rn to the
// former (with denormal scaling only). The scalar restatements round every square root and division
// correctly (IEEE 754), and in the Earth-radius geometry a 1-ulp difference in a normalisation moves
// a sample by ~0.5 m, so the kernels use these instead. tests/c/rn_math_check.c checks the selection
// and correction steps on the CPU over every mantissa from every 1-ulp start.
#pragma once
#include <hip/hip_runtime.h>

// RN(sqrt(x)) for x = 0 or x >= 2^-96: take the hardware result s and test the residuals x - s' * s of
// its two neighbours s' (exact through FMA) to pick the correctly rounded value.
__device__ __forceinline__ float sqrt_rn_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float down = __int_as_float(__float_as_int(s) - 1);
    const float up = __int_as_float(__float_as_int(s) + 1);
    const float vd = __builtin_fmaf(-down, s, x);
    const float vu = __builtin_fmaf(-up, s, x);
    const float r = vd <= 0.0f ? down : s;
    return vu > 0.0f ? up : r;
}

// RN(sqrt(x)) for any x >= 0: inputs below 2^-96 are scaled by 2^32 first (result by 2^-16) so
// v_sqrt_f32 never sees a denormal.
__device__ __forceinline__ float sqrt_rn(float x) {
    const bool tiny = x < 0x1p-96f;
    const float r = sqrt_rn_normal(tiny ? x * 0x1p32f : x);
    return tiny ? r * 0x1p-16f : r;
}

// RN(1 / s) for |s| in [2^-125, 2^125]: v_rcp_f32, then one Newton step y + y (1 - s y) with FMA, which
// reaches the correctly rounded reciprocal from every start within 1 ulp except when the mantissa of s
// is all ones: 1/s = 2^k (1 + 2^-24 + 2^-48 + ...) lies just above the midpoint of 2^k and its upper
// neighbour, and the step can stop at 2^k, so that case is set to 2^k (1 + 2^-23) directly.
__device__ __forceinline__ float rcp_rn(float s) {
    const float y0 = __builtin_amdgcn_rcpf(s);
    const float e = __builtin_fmaf(-s, y0, 1.0f);
    const float y = __builtin_fmaf(e, y0, y0);
    const bool ones = (__float_as_uint(s) & 0x7fffffu) == 0x7fffffu;
    return ones ? __uint_as_float((__float_as_uint(y) & 0xff800000u) | 1u) : y;
}
