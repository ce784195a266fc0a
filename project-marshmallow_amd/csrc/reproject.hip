// Temporal reprojection for gfx950: the HIP counterpart of the dispatch of
// SkyEngine/SkyEngine/Shaders/reproject.comp:91-153.
//
// One work-item per target pixel, 16x16-pixel blocks (four wave64 rows of 16x4). Each pixel makes
// one ray/sphere hit, one 3x4 transform and 11 nearest RGBA32F reads along its motion vector from
// the previous frame; the taps of neighbouring pixels overlap, so after the first touch the reads
// come from L2/MALL and the pass is bound by the 16 B write + 16 B first read per pixel of HBM.
// Arithmetic follows the shader's evaluation order with IEEE division/sqrt and no contraction, so it
// matches the scalar restatement bit for bit; GLSL latitude is fixed the same way on both sides
// (round half away from zero, saturating float -> int with NaN -> 0).
#include <hip/hip_runtime.h>
#include <math.h>

#include "mm_internal.hpp"
#include "cr_math.hpp"

namespace {

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 scl(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ F3 nrm(F3 a) {
    const float r = __fdiv_rn(1.0f, sqrt_rn(dot(a, a)));
    return f3(a.x * r, a.y * r, a.z * r);
}

// Correctly rounded division and square root in either precision (the oracle's `/` and `sqrt`).
__device__ __forceinline__ float dv(float a, float b) { return __fdiv_rn(a, b); }
__device__ __forceinline__ double dv(double a, double b) { return __ddiv_rn(a, b); }
__device__ __forceinline__ float sq(float a) { return sqrt_rn(a); }
__device__ __forceinline__ double sq(double a) { return __dsqrt_rn(a); }

// `.point` of raySphereIntersection (reproject.comp:56-87), zero on a miss.
template <class I>
__device__ __forceinline__ void sphere_point(I rox, I roy, I roz, I rdx, I rdy, I rdz, I cx, I cy, I cz, I w, I& px,
                                             I& py, I& pz) {
    rox = dv(rox - cx, w);
    roy = dv(roy - cy, w);
    roz = dv(roz - cz, w);
    const I A = (rdx * rdx + rdy * rdy) + rdz * rdz;
    const I B = I(2) * ((rdx * rox + rdy * roy) + rdz * roz);
    const I C = ((rox * rox + roy * roy) + roz * roz) - I(0.25);
    const I disc = (B * B) - ((I(4) * A) * C);
    px = py = pz = I(0);
    if (disc < I(0)) return;
    const I r = sq(disc);
    I t = dv(-r - B, A) * I(0.5);
    if (t < I(0)) t = dv(r - B, A) * I(0.5);
    if (!(t >= I(0))) return;
    px = (rox + rdx * t) * w + cx;
    py = (roy + rdy * t) * w + cy;
    pz = (roz + rdz * t) * w + cz;
}

// clamp(roundf(v), 0, n - 1) as an index (NaN gives 0), without the rounding call: clamping first
// changes nothing (roundf is monotone and keeps the integer bounds), and in [0, n - 1] (n <= 2^15)
// f - trunc(f) is exact, so rounding half away from zero is trunc(f) + (f - trunc(f) >= 0.5).
__device__ __forceinline__ int round_index(float v, int n) {
    const float f = v != v ? 0.0f : fminf(fmaxf(v, 0.0f), (float)(n - 1));
    const int i = (int)f;
    return i + (f - (float)i >= 0.5f ? 1 : 0);
}

__global__ __launch_bounds__(256) void reproject_kernel(const mm::ReprojConsts c, const float4* __restrict__ src,
                                                        float4* __restrict__ dst) {
    const int px = blockIdx.x * 16 + (threadIdx.x & 15), py = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (px >= c.width || py >= c.height) return;
    if (c.skipPhase >= 0 && (px & 3) == (c.skipPhase & 3) && (py & 3) == (c.skipPhase >> 2)) return;  // the march writes it
    const float dimx = (float)c.width, dimy = (float)c.height;
    const float uvx = __fdiv_rn((float)px, dimx), uvy = __fdiv_rn((float)py, dimy);
    const float sx = uvx * 2.0f - 1.0f, sy = uvy * 2.0f - 1.0f;
    const F3 camPos = f3(c.camPos[0], c.camPos[1], c.camPos[2]);
    const F3 p = sub(add(f3(c.refPoint[0], c.refPoint[1], c.refPoint[2]),
                         scl((c.aspect * sx) * c.tanH, f3(c.camRight[0], c.camRight[1], c.camRight[2]))),
                     scl(sy * c.tanH, f3(c.camUp[0], c.camUp[1], c.camUp[2])));
    const F3 rd = nrm(sub(p, camPos));
    float hx, hy, hz;
    if (c.fp64) {
        double x, y, z;
        sphere_point<double>(camPos.x, camPos.y, camPos.z, rd.x, rd.y, rd.z, camPos.x, c.centerY, camPos.z,
                             (double)c.innerDiam, x, y, z);
        hx = (float)x;
        hy = (float)y;
        hz = (float)z;
    } else {
        sphere_point<float>(camPos.x, camPos.y, camPos.z, rd.x, rd.y, rd.z, camPos.x, c.centerY, camPos.z,
                            c.innerDiam, hx, hy, hz);
    }
    const float* P = c.prevView;
    const F3 q = f3(((P[0] * hx + P[4] * hy) + P[8] * hz) + P[12], ((P[1] * hx + P[5] * hy) + P[9] * hz) + P[13],
                    ((P[2] * hx + P[6] * hy) + P[10] * hz) + P[14]);
    F3 d = nrm(q);
    const float nz = -d.z;
    d = f3(__fdiv_rn(d.x, nz), __fdiv_rn(d.y, nz), __fdiv_rn(d.z, nz));
    float oldU = __fdiv_rn(__fdiv_rn(d.x, c.tanH), c.aspect);
    oldU = oldU * 0.5f + 0.5f;
    float oldV = __fdiv_rn(-d.y, c.tanH);
    oldV = oldV * 0.5f + 0.5f;
    const float bx = oldU - uvx, by = oldV - uvy;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int s = 0; s < 10; s++) {
        const float k = __fdiv_rn((float)s, 9.0f) - 0.5f;
        const int ix = round_index((oldU - bx * k) * dimx, c.width);
        const int iy = round_index((oldV - by * k) * dimy, c.height);
        const float4 t = src[(size_t)iy * c.pitch + ix];
        acc.x += t.x;
        acc.y += t.y;
        acc.z += t.z;
        acc.w += t.w;
    }
    const int cx = round_index(oldU * dimx, c.width), cy = round_index(oldV * dimy, c.height);
    const float a = src[(size_t)cy * c.pitch + cx].w;
    dst[(size_t)py * c.pitch + px] = make_float4(__fdiv_rn(acc.x, 10.0f), __fdiv_rn(acc.y, 10.0f),
                                                 __fdiv_rn(acc.z, 10.0f), a);
}

}  // namespace

namespace mm {

int launch_reproject(const ReprojConsts& c, const float* src, float* dst, void* stream) {
    dim3 grid((unsigned)((c.width + 15) / 16), (unsigned)((c.height + 15) / 16)), block(256);
    hipLaunchKernelGGL(reproject_kernel, grid, block, 0, (hipStream_t)stream, c, (const float4*)src, (float4*)dst);
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

}  // namespace mm
