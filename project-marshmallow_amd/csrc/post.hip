// Post chain for gfx950: the HIP counterpart of the offscreen passes that consume the cloud image,
// god-ray.frag:41-76 -> radialBlur.frag:36-63 -> tonemap.frag:11-34 (VulkanApplication.cpp:938-966).
//
// Three launches of 16x16-pixel blocks (four wave64 rows of 16x4), one contiguous band of tiles per XCD:
//   alpha_plane_kernel  the cloud image's alpha as a plane of floats (16 B read, 4 B written);
//   god_ray_kernel      one work-item per pixel: 9 bilinear taps of that plane (centre + 8 toward
//                       the sun), writes the alpha plane of framebuffer 1 (4 B/pixel);
//   blur_tonemap_kernel one work-item per pixel: bilinear centre tap of the image's RGB, 10 bilinear
//                       taps of that alpha plane toward the sun, the radial-blur colour, then the
//                       Uncharted-2 tonemap + vignette, written as RGBA8 (and optionally the HDR
//                       radial-blur value).
// The taps of neighbouring pixels overlap, so after the first touch they come from L2. The algorithmic
// traffic is 16 B read + 4 B written (alpha plane), 4 B read + 4 B written (god rays) and 16 + 4 B
// read + 4 (or 20) B written (radial blur + tonemap) per pixel; the two tap kernels issue more VALU
// work per pixel (9 and 11 bilinear taps, exact division) than that traffic takes at HBM rate.
// Arithmetic follows the shaders' evaluation order with IEEE division/sqrt and no contraction; the
// sampler and the other choices GLSL leaves open are the scalar restatement's, so the result is the
// same bit for bit.
#include <hip/hip_runtime.h>
#include <math.h>

#include "mm_internal.hpp"
#include "cr_math.hpp"

namespace {

__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// LINEAR / CLAMP_TO_EDGE axis: x = u n - 0.5 rounded to 1/2^bits, i0 = floor(x) and i0 + 1 clamped.
__device__ __forceinline__ void axis_clamp(float u, int n, int bits, int& i0, int& i1, float& f) {
    float x = u * (float)n - 0.5f;
    // division by 2^bits is the exact multiplication by 2^-bits (one rounding either way)
    const float s = (float)(1 << bits), inv = __int_as_float((127 - bits) << 23);
    x = floorf(x * s + 0.5f) * inv;
    const float fl = floorf(x);
    f = x - fl;
    const int i = (int)clampf(fl, -1.0f, (float)n);
    i0 = min(max(i, 0), n - 1);
    i1 = min(max(i + 1, 0), n - 1);
}

__device__ __forceinline__ float lerp_r(float a, float b, float t) { return a + t * (b - a); }

// Bilinear tap of one float channel: texel (x, y) at byte offset ((y * pitch + x) * step + ch) * 4.
template <int STEP>
__device__ __forceinline__ float tap(const float* __restrict__ img, int ch, const mm::PostConsts& c, float u, float v) {
    int x0, x1, y0, y1;
    float fx, fy;
    axis_clamp(u, c.width, c.bits, x0, x1, fx);
    axis_clamp(v, c.height, c.bits, y0, y1, fy);
    const uint32_t r0 = (uint32_t)y0 * (uint32_t)c.pitch, r1 = (uint32_t)y1 * (uint32_t)c.pitch;
    const float t00 = img[(r0 + (uint32_t)x0) * STEP + ch], t01 = img[(r0 + (uint32_t)x1) * STEP + ch];
    const float t10 = img[(r1 + (uint32_t)x0) * STEP + ch], t11 = img[(r1 + (uint32_t)x1) * STEP + ch];
    return lerp_r(lerp_r(t00, t01, fx), lerp_r(t10, t11, fx), fy);
}

// The RGB of one bilinear tap of the RGBA image (one 16-byte load per texel).
__device__ __forceinline__ void tap_rgb(const float4* __restrict__ img, const mm::PostConsts& c, float u, float v,
                                        float rgb[3]) {
    int x0, x1, y0, y1;
    float fx, fy;
    axis_clamp(u, c.width, c.bits, x0, x1, fx);
    axis_clamp(v, c.height, c.bits, y0, y1, fy);
    const uint32_t r0 = (uint32_t)y0 * (uint32_t)c.pitch, r1 = (uint32_t)y1 * (uint32_t)c.pitch;
    const float4 t00 = img[r0 + (uint32_t)x0], t01 = img[r0 + (uint32_t)x1];
    const float4 t10 = img[r1 + (uint32_t)x0], t11 = img[r1 + (uint32_t)x1];
    rgb[0] = lerp_r(lerp_r(t00.x, t01.x, fx), lerp_r(t10.x, t11.x, fx), fy);
    rgb[1] = lerp_r(lerp_r(t00.y, t01.y, fx), lerp_r(t10.y, t11.y, fx), fy);
    rgb[2] = lerp_r(lerp_r(t00.z, t01.z, fx), lerp_r(t10.z, t11.z, fx), fy);
}

// Uncharted-2 curve (tonemap.frag:11-14) with the shader's constant products folded in float.
__device__ __forceinline__ float uc2(float x) {
    const float A = 0.15f, k05 = 0.1f * 0.5f, k004 = 0.2f * 0.02f, k06 = 0.2f * 0.3f, k0067 = 0.02f / 0.3f;
    return __fdiv_rn(x * (A * x + k05) + k004, x * (A * x + 0.5f) + k06) - k0067;
}

// pow(x, y) = exp2(y log2 x), both written out as float operations (the scalar restatement
// evaluates the same operations in the same order). log2: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log2 m by the atanh series in s = (m - 1) / (m + 1) to s^9; exp2: t = n + f,
// 2^f = sqrt(2) 2^(f - 1/2) by the exponential series in g = (f - 1/2) ln 2 to g^7, scaled by 2^n
// exactly. exp2 within 3 ulp, log2 within 1e-6 absolute; pow within 1e-6 relative over the
// tonemap's inputs (the float rounding of y log2 x dominates); the result is quantised to 8 bits.
__device__ __forceinline__ float log2_series(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : __int_as_float(0x7fc00000);
    if (x == INFINITY) return INFINITY;
    int e = 0;
    if (x < 1.17549435e-38f) {  // subnormal: scale into the normal range
        x = x * 16777216.0f;
        e = -24;
    }
    uint32_t b = __float_as_uint(x);
    e += (int)((b >> 23) & 255u) - 127;
    float m = __uint_as_float((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float s = __fdiv_rn(m - 1.0f, m + 1.0f);
    const float s2 = s * s;
    float p = 0.320598898f;
    p = p * s2 + 0.412198583f;
    p = p * s2 + 0.577078016f;
    p = p * s2 + 0.961796694f;
    p = p * s2 + 2.88539008f;
    return (float)e + s * p;
}

__device__ __forceinline__ float exp2_series(float t) {
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
    return __builtin_amdgcn_ldexpf(p * 1.41421356f, (int)n);
}

// Block -> 16x16 tile. Workgroups are dealt round-robin to the 8 XCDs (block b runs on XCD b mod 8),
// each with its own L2; the taps of a pixel reach far toward the sun, so tiles that share cache
// lines should share an XCD. Block b takes tile (b mod 8) * per + b / 8 of the row-major tile list
// (per = ceil(tiles / 8)): each XCD marches one contiguous band of tile rows. Returns false for the
// padding blocks past the last tile.
__device__ __forceinline__ bool xcd_tile(const mm::PostConsts& c, int& x, int& y) {
    const int gx = (c.width + 15) >> 4, tiles = gx * ((c.height + 15) >> 4);
    const int per = (tiles + 7) >> 3;
    const int t = (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
    if (t >= tiles) return false;
    x = (t % gx) * 16 + (threadIdx.x & 15);
    y = (t / gx) * 16 + (threadIdx.x >> 4);
    return true;
}

__device__ __forceinline__ uint32_t unorm8(float v) {
    if (!(v == v)) return 0u;
    return (uint32_t)floorf(clampf(v, 0.0f, 1.0f) * 255.0f + 0.5f);
}

// The image's alpha as a plane of its own: the god-ray taps then read 4-byte texels (32 per
// 128-byte line instead of 8 RGBA32F ones), which is what keeps their far-reaching taps in cache.
__global__ __launch_bounds__(256) void alpha_plane_kernel(const mm::PostConsts c, const float4* __restrict__ img,
                                                          float* __restrict__ plane) {
    int x, y;
    if (!xcd_tile(c, x, y) || x >= c.width || y >= c.height) return;
    const uint32_t i = (uint32_t)y * (uint32_t)c.pitch + (uint32_t)x;
    plane[i] = img[i].w;
}

__global__ __launch_bounds__(256) void god_ray_kernel(const mm::PostConsts c, const float* __restrict__ a0,
                                                      float* __restrict__ alpha) {
    int x, y;
    if (!xcd_tile(c, x, y) || x >= c.width || y >= c.height) return;
    float a = 1.0f;  // sun below the horizon: the pass writes alpha 1 (god-ray.frag:55-57)
    if (!c.below) {
        const float u = __fdiv_rn((float)x + 0.5f, (float)c.width), v = __fdiv_rn((float)y + 0.5f, (float)c.height);
        float cx = u * 2.0f - 1.0f, cy = v * 2.0f - 1.0f;
        const float dx = (cx - c.sunX) * c.k, dy = (cy - c.sunY) * c.k;
        float acc = tap<1>(a0, 0, c, u, v) * 0.5f;
        float decay = 1.0f;
        for (int i = 0; i < c.samples; i++) {
            cx = cx - dx;
            cy = cy - dy;
            float s = tap<1>(a0, 0, c, cx * 0.5f + 0.5f, cy * 0.5f + 0.5f) * 0.5f;
            s = s * (c.sw * decay);
            acc = acc + s;
            decay = decay * c.decay;
        }
        a = acc * c.godExposure;
    }
    alpha[(uint32_t)y * (uint32_t)c.pitch + (uint32_t)x] = a;
}

__constant__ float kBlurTaps[10] = {-0.08f, -0.05f, -0.03f, -0.02f, -0.01f, 0.01f, 0.02f, 0.03f, 0.05f, 0.08f};

__global__ __launch_bounds__(256) void blur_tonemap_kernel(const mm::PostConsts c, const float4* __restrict__ img,
                                                           const float* __restrict__ alpha, float4* __restrict__ hdr,
                                                           uint32_t* __restrict__ rgba8) {
    int x, y;
    if (!xcd_tile(c, x, y) || x >= c.width || y >= c.height) return;
    const float u = __fdiv_rn((float)x + 0.5f, (float)c.width), v = __fdiv_rn((float)y + 0.5f, (float)c.height);
    float rgb[3];
    tap_rgb(img, c, u, v, rgb);
    float o[3] = {rgb[0], rgb[1], rgb[2]};
    if (!c.below) {
        const float sx = u * 2.0f - 1.0f, sy = v * 2.0f - 1.0f;
        float lx = c.sunX - sx, ly = c.sunY - sy;
        const float dist = sqrt_rn(lx * lx + ly * ly);
        lx = __fdiv_rn(lx, dist);
        ly = __fdiv_rn(ly, dist);
        float acc = 0.0f;
#pragma unroll
        for (int i = 0; i < 10; i++) {
            const float ox = ((kBlurTaps[i] * lx) * c.blurScale) * dist, oy = ((kBlurTaps[i] * ly) * c.blurScale) * dist;
            acc = acc + tap<1>(alpha, 0, c, (sx + ox) * 0.5f + 0.5f, (sy + oy) * 0.5f + 0.5f) * c.blurGain;
        }
        acc = __fdiv_rn(acc, 10.0f);
        for (int k = 0; k < 3; k++) o[k] = c.gc[k] * acc + 0.5f * rgb[k];
    }
    const uint32_t i = (uint32_t)y * (uint32_t)c.pitch + (uint32_t)x;
    if (hdr) hdr[i] = make_float4(o[0], o[1], o[2], 1.0f);
    if (rgba8) {
        const float vx = u - 0.5f, vy = v - 0.5f;
        const float vig = vx * vx + vy * vy;
        uint32_t q = 255u << 24;
        for (int k = 0; k < 3; k++) {
            // framebuffer 2 sampled at the pixel centre returns this pixel's radial-blur value
            float t = uc2(c.exposure * o[k]);
            t = t * c.whitemap;
            t = exp2_series(c.invGamma * log2_series(t));  // pow (tonemap.frag:22)
            t = t * (1.0f - vig) + c.vignette[k] * vig;
            q |= unorm8(t) << (8 * k);
        }
        rgba8[i] = q;
    }
}

}  // namespace

namespace mm {

int launch_post(const PostConsts& c, const float* image, float* scratch, float* hdr, uint8_t* rgba8, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const long tiles = (long)((c.width + 15) / 16) * ((c.height + 15) / 16);
    dim3 grid((unsigned)(8 * ((tiles + 7) / 8))), block(256);  // 1-D: xcd_tile maps blocks to tiles
    // scratch: plane 0 = the image's alpha, plane 1 = framebuffer 1's alpha (the god-ray output)
    float* a0 = scratch;
    float* a1 = scratch + (size_t)c.pitch * c.height;
    if (!c.below) hipLaunchKernelGGL(alpha_plane_kernel, grid, block, 0, s, c, (const float4*)image, a0);
    hipLaunchKernelGGL(god_ray_kernel, grid, block, 0, s, c, (const float*)a0, a1);
    hipLaunchKernelGGL(blur_tonemap_kernel, grid, block, 0, s, c, (const float4*)image, (const float*)a1,
                       (float4*)hdr, (uint32_t*)rgba8);
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

}  // namespace mm
