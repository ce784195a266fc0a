/* CPU check of the correctly rounded square root and reciprocal used by the HIP kernels
 * (project-marshmallow_amd/csrc/cr_math.hpp). The hardware v_sqrt_f32 / v_rcp_f32 are accurate to 1 ulp;
 * for every float mantissa (two binades for the square root, one for the reciprocal, whose behaviour
 * repeats in every binade) and every start within 1 ulp of the true value, the kernel's selection /
 * correction step must return the correctly rounded result. Built with gcc -mfma; prints the counts
 * and exits non-zero on any miss. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bf(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* the kernel's sqrt_rn_normal selection, from a start s */
static float sqrt_select(float x, float s) {
    float down = fb(bf(s) - 1), up = fb(bf(s) + 1);
    float vd = fmaf(-down, s, x), vu = fmaf(-up, s, x);
    float r = vd <= 0.0f ? down : s;
    return vu > 0.0f ? up : r;
}

/* the kernel's rcp_rn correction, from a start y0 */
static float rcp_correct(float s, float y0) {
    float e = fmaf(-s, y0, 1.0f);
    float y = fmaf(e, y0, y0);
    int ones = (bf(s) & 0x7fffffu) == 0x7fffffu;
    return ones ? fb((bf(y) & 0xff800000u) | 1u) : y;
}

/* ulp of the binade holding v > 0 */
static double ulp_of(double v) { int e; frexp(v, &e); return ldexp(1.0, e - 24); }
/* a start within 1 ulp: the distance is at most one float spacing at the smaller of the two values
 * (below a power of two the spacing halves) */
static int within_1ulp(double start, double exact) {
    return fabs(start - exact) <= ulp_of(start < exact ? start : exact);
}

int main(void) {
    long bad = 0, n = 0;
    /* reciprocal: s in [1, 2) */
    for (uint32_t m = 0; m < (1u << 23); m++) {
        float s = fb(0x3f800000u | m);
        double exact = 1.0 / (double)s; /* 1/s is never within 2^-49 of a float midpoint: float(exact) is RN */
        float rn = (float)exact;
        for (int d = -3; d <= 3; d++) {
            float y0 = fb(bf(rn) + d);
            if (!within_1ulp((double)y0, exact)) continue;
            n++;
            if (rcp_correct(s, y0) != rn) {
                if (bad++ < 5) printf("rcp miss: s=%a y0=%a\n", s, y0);
            }
        }
    }
    printf("reciprocal: %ld starts, %ld misses\n", n, bad);
    long bad2 = 0, n2 = 0;
    /* square root: x in [1, 4) */
    for (uint32_t m = 0; m < (1u << 24); m++) {
        float x = fb(0x3f800000u + m);
        double exact = sqrt((double)x);
        float rn = sqrtf(x); /* libm's sqrtf is correctly rounded (IEEE 754) */
        for (int d = -2; d <= 2; d++) {
            float s = fb(bf(rn) + d);
            if (!within_1ulp((double)s, exact)) continue;
            n2++;
            if (sqrt_select(x, s) != rn) {
                if (bad2++ < 5) printf("sqrt miss: x=%a s=%a\n", x, s);
            }
        }
    }
    printf("square root: %ld starts, %ld misses\n", n2, bad2);
    return (bad || bad2) ? 1 : 0;
}
