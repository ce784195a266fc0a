// CPU restatement of the post chain that consumes the cloud image (test infrastructure: the
// checker of the HIP kernels, never linked into the product).
//
// The offscreen pass records (VulkanApplication.cpp:938-966) background -> god rays -> radial blur
// (-> meshes, outside this path) -> tonemap, every framebuffer RGBA32F (VulkanApplication.cpp:1306-
// 1308) sampled LINEAR / CLAMP_TO_EDGE (:1290-1302) by a full-screen quad whose UV is (0, 0) at the
// top-left pixel (Geometry.cpp:91-99). The background pass samples the cloud image at pixel
// centres, i.e. copies it, so the chain starts from the cloud image itself:
//   god rays    god-ray.frag:41-76   alpha = 0.5 a + sum of 8 decayed alpha taps toward the sun
//   radial blur radialBlur.frag:36-63 rgb = sun.color * intensity * mean(10 alpha taps * 1.1)
//                                     + 0.5 rgb
//   tonemap     tonemap.frag:11-34   Uncharted-2 curve, white point, gamma, vignette; UNORM8 out
// Choices where GLSL leaves latitude, shared with the HIP kernels:
//   * fragUV = ((x + 0.5) / W, (y + 0.5) / H), one rounding per division;
//   * the sampler: x = u * n - 0.5 rounded to 1/2^bits, i0 = floor, i1 = i0 + 1, both clamped to
//     [0, n - 1] (the float floor is clamped to [-1, n] before the integer conversion), weights
//     applied as a + f (b - a) with the product rounded;
//   * proj * view * location is (proj * view) * location, sums in index order;
//   * pow(x, y) = exp2(y log2 x) with both written out as float series (pow_f below; GLSL defines pow
//     through exp2/log2 and leaves their last bits to the implementation); UNORM8 =
//     floor(clamp(v, 0, 1) * 255 + 0.5), NaN -> 0;
//   * sub-texel precision of 1..8 bits (8 is the hardware's): a pixel-centre tap then returns the
//     texel itself, so the tonemap's read of framebuffer 2 is the radial-blur value of that pixel.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

namespace {

inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

inline void axis_clamp(float u, int n, int bits, int& i0, int& i1, float& f) {
    float x = u * (float)n - 0.5f;
    if (bits > 0) {
        const float s = (float)(1 << bits);
        x = floorf(x * s + 0.5f) / s;
    }
    const float fl = floorf(x);
    f = x - fl;
    const int i = (int)clampf(fl, -1.0f, (float)n);
    i0 = i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    i1 = i + 1 < 0 ? 0 : (i + 1 > n - 1 ? n - 1 : i + 1);
}

inline float lerp_r(float a, float b, float t) { return a + t * (b - a); }

// texture(tex, uv).c of a W x H image with `stride` floats between rows and `step` floats per texel.
inline float sample_ch(const float* img, int w, int h, int stride, int step, float u, float v, int bits) {
    int x0, x1, y0, y1;
    float fx, fy;
    axis_clamp(u, w, bits, x0, x1, fx);
    axis_clamp(v, h, bits, y0, y1, fy);
    const float a = lerp_r(img[(size_t)y0 * stride + (size_t)x0 * step], img[(size_t)y0 * stride + (size_t)x1 * step], fx);
    const float b = lerp_r(img[(size_t)y1 * stride + (size_t)x0 * step], img[(size_t)y1 * stride + (size_t)x1 * step], fx);
    return lerp_r(a, b, fy);
}

// pow(x, y) = exp2(y log2 x), both written out as float operations (the HIP kernel evaluates the
// same operations in the same order). log2: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log2 m by the
// atanh series in s = (m - 1) / (m + 1) to s^9; exp2: t = n + f, 2^f = sqrt(2) 2^(f - 1/2) by the
// exponential series in g = (f - 1/2) ln 2 to g^7, scaled by 2^n exactly. exp2 within 3 ulp, log2
// within 1e-6 absolute; pow within 1e-6 relative over the tonemap's inputs (the float rounding of
// y log2 x dominates).
inline float log2_series(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (x == INFINITY) return INFINITY;
    int e = 0;
    if (x < 1.17549435e-38f) {  // subnormal: scale into the normal range
        x = x * 16777216.0f;
        e = -24;
    }
    uint32_t b;
    memcpy(&b, &x, 4);
    e += (int)((b >> 23) & 255u) - 127;
    b = (b & 0x007fffffu) | 0x3f800000u;
    float m;
    memcpy(&m, &b, 4);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 0.320598898f;
    p = p * s2 + 0.412198583f;
    p = p * s2 + 0.577078016f;
    p = p * s2 + 0.961796694f;
    p = p * s2 + 2.88539008f;
    return (float)e + s * p;
}

inline float exp2_series(float t) {
    if (t != t) return t;
    if (t >= 128.0f) return INFINITY;
    if (t < -150.0f) return 0.0f;
    const float n = floorf(t);
    const float g = ((t - n) - 0.5f) * 0.693147181f;
    float p = 1.98412698e-4f;
    p = p * g + 1.38888889e-3f;
    p = p * g + 8.33333333e-3f;
    p = p * g + 4.16666667e-2f;
    p = p * g + 0.166666667f;
    p = p * g + 0.5f;
    p = p * g + 1.0f;
    p = p * g + 1.0f;
    return ldexpf(p * 1.41421356f, (int)n);
}

inline float pow_f(float x, float y) { return exp2_series(y * log2_series(x)); }

inline float uc2(float x) {
    const float A = 0.15f, k05 = 0.1f * 0.5f, k004 = 0.2f * 0.02f, k06 = 0.2f * 0.3f, k0067 = 0.02f / 0.3f;
    return ((x * (A * x + k05) + k004) / (x * (A * x + 0.5f) + k06)) - k0067;
}

inline uint8_t unorm8(float v) {
    if (!(v == v)) return 0;
    return (uint8_t)floorf(clampf(v, 0.0f, 1.0f) * 255.0f + 0.5f);
}

}  // namespace

extern "C" void orc_post_sun_ndc(const mm_camera_ubo* cam, const mm_sun_ubo* sun, float out_xy[2]) {
    const float* P = cam->proj;
    const float* V = cam->view;
    float M[16];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++)
            M[c * 4 + r] = ((P[0 * 4 + r] * V[c * 4 + 0] + P[1 * 4 + r] * V[c * 4 + 1]) + P[2 * 4 + r] * V[c * 4 + 2]) +
                           P[3 * 4 + r] * V[c * 4 + 3];
    float s[4];
    for (int r = 0; r < 4; r++)
        s[r] = ((M[0 * 4 + r] * sun->location[0] + M[1 * 4 + r] * sun->location[1]) + M[2 * 4 + r] * sun->location[2]) +
               M[3 * 4 + r] * sun->location[3];
    out_xy[0] = s[0] / s[3];
    out_xy[1] = s[1] / s[3];
}

extern "C" float orc_post_uc2(float x) { return uc2(x); }
extern "C" float orc_post_pow(float x, float y) { return pow_f(x, y); }

extern "C" int orc_post_process(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const float* image, int width,
                                int height, int pitch, float* god_alpha, float* hdr, uint8_t* rgba8,
                                const mm_post_params* p) {
    if (!cam || !sun || !image || !god_alpha || !p || width <= 0 || height <= 0 || pitch < width) return -1;
    float sp[2];
    orc_post_sun_ndc(cam, sun, sp);
    const bool below = sun->direction[1] < 0.0f;
    const float W = (float)width, H = (float)height;
    const int bits = p->subtexel_bits;
    const int ns = p->god_ray_samples;
    const float sw = 1.0f / (float)ns;
    // god rays (god-ray.frag:41-76): the alpha plane of framebuffer 1
#pragma omp parallel for schedule(static)
    for (int y = 0; y < height; y++) {
        for (int x = 0; x < width; x++) {
            const float u = ((float)x + 0.5f) / W, v = ((float)y + 0.5f) / H;
            float a;
            if (below) {
                a = 1.0f;
            } else {
                float cx = u * 2.0f - 1.0f, cy = v * 2.0f - 1.0f;
                const float k = sw * p->god_ray_density;
                const float dx = (cx - sp[0]) * k, dy = (cy - sp[1]) * k;
                float acc = sample_ch(image + 3, width, height, 4 * pitch, 4, u, v, bits) * 0.5f;
                float decay = 1.0f;
                for (int i = 0; i < ns; i++) {
                    cx = cx - dx;
                    cy = cy - dy;
                    float s = sample_ch(image + 3, width, height, 4 * pitch, 4, cx * 0.5f + 0.5f, cy * 0.5f + 0.5f, bits) * 0.5f;
                    s = s * (sw * decay);
                    acc = acc + s;
                    decay = decay * p->god_ray_decay;
                }
                a = acc * p->god_ray_exposure;
            }
            god_alpha[(size_t)y * pitch + x] = a;
        }
    }
    static const float kTaps[10] = {-0.08f, -0.05f, -0.03f, -0.02f, -0.01f, 0.01f, 0.02f, 0.03f, 0.05f, 0.08f};
    const float gc[3] = {sun->color[0] * sun->intensity, sun->color[1] * sun->intensity, sun->color[2] * sun->intensity};
    const float whitemap = 1.0f / uc2(p->white_point);
    // radial blur (radialBlur.frag:36-63) and tonemap (tonemap.frag:11-34)
#pragma omp parallel for schedule(static)
    for (int y = 0; y < height; y++) {
        for (int x = 0; x < width; x++) {
            const float u = ((float)x + 0.5f) / W, v = ((float)y + 0.5f) / H;
            float rgb[3];
            for (int c = 0; c < 3; c++) rgb[c] = sample_ch(image + c, width, height, 4 * pitch, 4, u, v, bits);
            float o[3];
            if (below) {
                for (int c = 0; c < 3; c++) o[c] = rgb[c];
            } else {
                const float sx = u * 2.0f - 1.0f, sy = v * 2.0f - 1.0f;
                float lx = sp[0] - sx, ly = sp[1] - sy;
                const float dist = sqrtf(lx * lx + ly * ly);
                lx = lx / dist;
                ly = ly / dist;
                float acc = 0.0f;
                for (int i = 0; i < 10; i++) {
                    const float ox = ((kTaps[i] * lx) * p->blur_scale) * dist, oy = ((kTaps[i] * ly) * p->blur_scale) * dist;
                    acc = acc + sample_ch(god_alpha, width, height, pitch, 1, (sx + ox) * 0.5f + 0.5f, (sy + oy) * 0.5f + 0.5f,
                                          bits) * p->blur_gain;
                }
                acc = acc / 10.0f;
                for (int c = 0; c < 3; c++) o[c] = gc[c] * acc + 0.5f * rgb[c];
            }
            if (hdr) {
                float* h = hdr + ((size_t)y * pitch + x) * 4;
                h[0] = o[0]; h[1] = o[1]; h[2] = o[2]; h[3] = 1.0f;
            }
            if (rgba8) {
                const float vx = u - 0.5f, vy = v - 0.5f;
                const float vig = vx * vx + vy * vy;
                uint8_t* q = rgba8 + ((size_t)y * pitch + x) * 4;
                for (int c = 0; c < 3; c++) {
                    // the framebuffer is sampled at the pixel centre: the radial-blur value itself
                    float t = uc2(p->exposure * o[c]);
                    t = t * whitemap;
                    t = pow_f(t, p->inv_gamma);
                    t = t * (1.0f - vig) + p->vignette[c] * vig;
                    q[c] = unorm8(t);
                }
                q[3] = 255;
            }
        }
    }
    return 0;
}
