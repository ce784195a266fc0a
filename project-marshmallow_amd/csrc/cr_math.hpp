// Correctly rounded single-precision square root for gfx950 device code.
//
// v_sqrt_f32 is accurate to 1 ulp, and the compiler lowers sqrtf / __fsqrt_rn to it (with denormal
// scaling only). The scalar restatements round every square root correctly (IEEE 754), and in the
// Earth-radius geometry a 1-ulp difference in a normalisation moves a sample by ~0.5 m, so the
// kernels use this instead: take the hardware result s, and test the residuals x - s' * s of its
// two neighbours s' (exact through FMA) to pick the correctly rounded value. Inputs below 2^-96
// are scaled by 2^32 first (result by 2^-16) so v_sqrt_f32 never sees a denormal.
#pragma once
#include <hip/hip_runtime.h>

__device__ __forceinline__ float sqrt_rn(float x) {
    const bool tiny = x < 0x1p-96f;
    const float xs = tiny ? x * 0x1p32f : x;
    const float s = __builtin_amdgcn_sqrtf(xs);
    const float down = __int_as_float(__float_as_int(s) - 1);
    const float up = __int_as_float(__float_as_int(s) + 1);
    const float vd = __builtin_fmaf(-down, s, xs);
    const float vu = __builtin_fmaf(-up, s, xs);
    float r = vd <= 0.0f ? down : s;
    r = vu > 0.0f ? up : r;
    return tiny ? r * 0x1p-16f : r;
}
