// Oracle: restatement of the reference's curl-noise generator, ImageUtils.cpp:7-223.
// Test infrastructure only (see oracle.h). Compiled with -ffp-contract=off: every float/double
// operation below rounds exactly where the reference's (MSVC x64, SSE2, no FMA) build rounds.
#include <math.h>
#include <float.h>
#include <string.h>
#include <stdlib.h>
#include "oracle.h"

namespace {

// ImageUtils.cpp:10-23: the twelve edge-midpoint gradients, written as double literals and
// narrowed to float by the glm::vec3 constructor.
const float kBasis[12][3] = {
    {(float)0.7071, (float)0.7071, 0.f},  {(float)0.7071, (float)-0.7071, 0.f},
    {(float)-0.7071, (float)0.7071, 0.f}, {(float)-0.7071, (float)-0.7071, 0.f},
    {(float)0.7071, 0.f, (float)0.7071},  {(float)0.7071, 0.f, (float)-0.7071},
    {(float)-0.7071, 0.f, (float)0.7071}, {(float)-0.7071, 0.f, (float)-0.7071},
    {0.f, (float)0.7071, (float)0.7071},  {0.f, (float)0.7071, (float)-0.7071},
    {0.f, (float)-0.7071, (float)0.7071}, {0.f, (float)-0.7071, (float)-0.7071},
};

const double kEps = 0.0005;  // ImageUtils.cpp:8, a double #define

// Planted-mistake switch for the fixture's sensitivity tests (0 = the faithful restatement):
// 1 lerp entirely in float, 2 row/col transposed, 3 curl.x sign flipped, 4 FBM persistence 0.5,
// 5 gradient index from round() instead of floor().
int g_mutation = 0;

}  // namespace

extern "C" void orc_set_curl_mutation(int m) { g_mutation = m; }

// sin by Cody-Waite reduction modulo pi/2 (two-part split of pi/2 as in fdlibm's __ieee754_rem_pio2
// medium-size path) and Taylor polynomials of sin/cos on [-pi/4, pi/4] evaluated in Horner form.
// Accurate to a few double ulps for |x| < 2^20, so rounding the result to float gives the
// correctly rounded float sine except within ~1e-16 relative of a float rounding tie.
extern "C" double orc_sin_poly(double x) {
    const double invpio2 = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079632673412561417e+00;   // first 33 bits of pi/2
    const double pio2_1t = 6.07710050650619224932e-11;  // pi/2 - pio2_1
    double k = nearbyint(x * invpio2);
    double r = (x - k * pio2_1) - k * pio2_1t;
    double r2 = r * r;
    double s = r + r * r2 * (-1.0 / 6.0 + r2 * (1.0 / 120.0 + r2 * (-1.0 / 5040.0 + r2 * (1.0 / 362880.0
               + r2 * (-1.0 / 39916800.0 + r2 * (1.0 / 6227020800.0 + r2 * (-1.0 / 1307674368000.0)))))));
    double c = 1.0 + r2 * (-0.5 + r2 * (1.0 / 24.0 + r2 * (-1.0 / 720.0 + r2 * (1.0 / 40320.0
               + r2 * (-1.0 / 3628800.0 + r2 * (1.0 / 479001600.0 + r2 * (-1.0 / 87178291200.0
               + r2 * (1.0 / 20922789888000.0))))))));
    long q = ((long)k) & 3;
    if (q == 0) return s;
    if (q == 1) return c;
    if (q == 2) return -s;
    return -c;
}

static inline float ref_sinf(float x, int sin_kind) {
    if (sin_kind == 1) return sinf(x);
    return (float)orc_sin_poly((double)x);
}

// ImageUtils.cpp:25-29 hashNoise. glm's vec3 dot rounds the three products, then sums left to right.
extern "C" float orc_hash_noise(float x, float y, float z, int sin_kind) {
    float px = x * 12.9898f;
    float py = y * 78.233f;
    float pz = z * (float)47.387;
    float d = (px + py) + pz;
    float n = ref_sinf(d, sin_kind) * 43758.5453f;
    n = n - floorf(n);
    return n;
}

// ImageUtils.cpp:31-34 hashVec.
static inline const float* hash_vec(float x, float y, float z, int sin_kind) {
    float a = orc_hash_noise(x, y, z, sin_kind);
    if (g_mutation == 5) return kBasis[((int)roundf(12.f * a)) % 12];
    return kBasis[(int)floorf(12.f * a)];
}

// ImageUtils.cpp:36-38: (1.0 - t) * a is a double product, t * b a float product; the double sum
// is narrowed to float on return.
static inline float ref_lerp(float a, float b, float t) {
    if (g_mutation == 1) return (1.0f - t) * a + t * b;
    return (float)((1.0 - (double)t) * (double)a + (double)(t * b));
}

static inline float dot3(const float* g, float dx, float dy, float dz) {
    float a = g[0] * dx, b = g[1] * dy, c = g[2] * dz;
    return (a + b) + c;
}

// ImageUtils.cpp:41-73 perlinNoise. The "force tiling boundaries" block (:47-51) only ever assigns
// f1 = freq when f1 already equals the integer freq, so f1 = f + 1 throughout.
extern "C" float orc_perlin(float x, float y, float z, float freq, int sin_kind) {
    x *= freq; y *= freq; z *= freq;
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    float rx = x - fx, ry = y - fy, rz = z - fz;
    // u = r*r*r * (r*(r*6 - 15) + 10), each operation rounded componentwise (glm vec3 ops)
    float ux = ((rx * rx) * rx) * ((rx * ((rx * 6.0f) - 15.0f)) + 10.0f);
    float uy = ((ry * ry) * ry) * ((ry * ((ry * 6.0f) - 15.0f)) + 10.0f);
    float uz = ((rz * rz) * rz) * ((rz * ((rz * 6.0f) - 15.0f)) + 10.0f);
    float f1x = fx + 1.0f, f1y = fy + 1.0f, f1z = fz + 1.0f;

    float nnn = dot3(hash_vec(fx, fy, fz, sin_kind), rx, ry, rz);
    float nnp = dot3(hash_vec(fx, fy, f1z, sin_kind), rx - 0.f, ry - 0.f, rz - 1.f);
    float npn = dot3(hash_vec(fx, f1y, fz, sin_kind), rx - 0.f, ry - 1.f, rz - 0.f);
    float npp = dot3(hash_vec(fx, f1y, f1z, sin_kind), rx - 0.f, ry - 1.f, rz - 1.f);
    float pnn = dot3(hash_vec(f1x, fy, fz, sin_kind), rx - 1.f, ry - 0.f, rz - 0.f);
    float pnp = dot3(hash_vec(f1x, fy, f1z, sin_kind), rx - 1.f, ry - 0.f, rz - 1.f);
    float ppn = dot3(hash_vec(f1x, f1y, fz, sin_kind), rx - 1.f, ry - 1.f, rz - 0.f);
    float ppp = dot3(hash_vec(f1x, f1y, f1z, sin_kind), rx - 1.f, ry - 1.f, rz - 1.f);

    float nn = ref_lerp(nnn, pnn, ux);
    float np = ref_lerp(nnp, pnp, ux);
    float pn = ref_lerp(npn, ppn, ux);
    float pp = ref_lerp(npp, ppp, ux);
    float n = ref_lerp(nn, pn, uy);
    float p = ref_lerp(np, pp, uy);
    return ref_lerp(n, p, uz);
}

// ImageUtils.cpp:75-87 FBM (persistence 0.4).
extern "C" float orc_fbm(float x, float y, float z, float freq, int octaves, int sin_kind) {
    float noise = 0.0f, weight = 1.0f, totalWeight = 0.0f;
    const float persistence = g_mutation == 4 ? 0.5f : 0.4f;
    for (int i = 0; i < octaves; i++) {
        totalWeight += weight;
        noise += weight * orc_perlin(x, y, z, freq, sin_kind);
        freq *= 2.0f;
        weight *= persistence;
    }
    return noise / totalWeight;
}

// Central difference (b - a) / (2 * EPS): float difference, double quotient, narrowed to float.
static inline float central(float a, float b) {
    return (float)((double)(b - a) / (2.0 * kEps));
}

// ImageUtils.cpp:130-169 curlNoiseFBM. pt +- EPS is evaluated in double and narrowed by the
// glm::vec3 constructor.
extern "C" void orc_curl_fbm(float px, float py, float freq, int octaves, int sin_kind, float out[3]) {
    const float xm = (float)((double)px - kEps), xp = (float)((double)px + kEps);
    const float ym = (float)((double)py - kEps), yp = (float)((double)py + kEps);
    float a, b;
    a = orc_fbm(xm, py, 0.5f, freq, octaves, sin_kind);
    b = orc_fbm(xp, py, 0.5f, freq, octaves, sin_kind);
    float dydx = central(a, b);
    a = orc_fbm(px, ym, 0.5f, freq, octaves, sin_kind);
    b = orc_fbm(px, yp, 0.5f, freq, octaves, sin_kind);
    float dxdy = central(a, b);
    a = orc_fbm(px, 0.5f, ym, freq, octaves, sin_kind);
    b = orc_fbm(px, 0.5f, yp, freq, octaves, sin_kind);
    float dxdz = central(a, b);
    a = orc_fbm(xm, 0.5f, py, freq, octaves, sin_kind);
    b = orc_fbm(xp, 0.5f, py, freq, octaves, sin_kind);
    float dzdx = central(a, b);
    a = orc_fbm(0.5f, ym, px, freq, octaves, sin_kind);
    b = orc_fbm(0.5f, yp, px, freq, octaves, sin_kind);
    float dzdy = central(a, b);
    a = orc_fbm(0.5f, py, xm, freq, octaves, sin_kind);
    b = orc_fbm(0.5f, py, xp, freq, octaves, sin_kind);
    float dydz = central(a, b);
    out[0] = g_mutation == 3 ? dydz - dzdy : dzdy - dydz;
    out[1] = dxdz - dzdx;
    out[2] = dydx - dxdy;
}

// ImageUtils.cpp:176-223 GenerateCurlNoise: curlNoiseFBM(col/128, row/128, 3, 4) per texel,
// per-channel global min/max, remap to [0,1], (uchar)(int)roundf(x * 255), A = 255.
extern "C" void orc_curl_generate(int dim, int sin_kind, uint8_t* rgba, float* raw_xyz, float* minmax) {
    float* curls = (float*)malloc(sizeof(float) * 3 * (size_t)dim * dim);
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
#pragma omp parallel for schedule(dynamic, 4)
    for (int row = 0; row < dim; row++) {
        for (int col = 0; col < dim; col++) {
            float* c = curls + 3 * ((size_t)row * dim + col);
            if (g_mutation == 2)
                orc_curl_fbm((float)row / (float)dim, (float)col / (float)dim, 3.f, 4, sin_kind, c);
            else
                orc_curl_fbm((float)col / (float)dim, (float)row / (float)dim, 3.f, 4, sin_kind, c);
        }
    }
    for (size_t i = 0; i < (size_t)dim * dim; i++) {
        for (int k = 0; k < 3; k++) {
            float v = curls[3 * i + k];
            lo[k] = v < lo[k] ? v : lo[k];      // std::min(lowerBound, v)
            hi[k] = hi[k] < v ? v : hi[k];      // std::max(upperBound, v)
        }
    }
    if (raw_xyz) memcpy(raw_xyz, curls, sizeof(float) * 3 * (size_t)dim * dim);
    if (minmax) {
        for (int k = 0; k < 3; k++) { minmax[2 * k] = lo[k]; minmax[2 * k + 1] = hi[k]; }
    }
    for (size_t i = 0; i < (size_t)dim * dim; i++) {
        for (int k = 0; k < 3; k++) {
            // ImageUtils.cpp:171-174 remap(x, lo, hi, 0, 1)
            float m = 0.f + ((curls[3 * i + k] - lo[k]) / (hi[k] - lo[k]) * (1.f - 0.f));
            rgba[4 * i + k] = (uint8_t)(int)roundf(m * 255.f);
        }
        rgba[4 * i + 3] = 255;
    }
    free(curls);
}
