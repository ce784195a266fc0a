// Noise-texture generators for gfx950: curl map, low-res Perlin-Worley/Worley volume, hi-res
// Worley volume and weather map, written directly in the march's storage formats and layouts.
//
// Built with -ffp-contract=off and IEEE division/sqrt so every float/double operation rounds as in
// the CPU restatement: outputs are equal byte for byte.
//
// Curl map: restates GenerateCurlNoise (ImageUtils.cpp:176-223). hashNoise (ImageUtils.cpp:25-29)
// only ever sees integer lattice coordinates in [-1, 24] (pt in [-EPS, 1) times freq <= 24), so the
// gradient index of every lattice point is tabulated once per call (27^3 entries) and the 48
// Perlin evaluations per texel read the table. The global per-channel min/max of the curl field
// is reduced with atomics on order-preserving integer encodings, then a second pass quantises.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <float.h>
#include <mutex>
#include "mm_internal.hpp"
#include "cr_math.hpp"

namespace {

// ------------------------------------------------------------------ curl (ImageUtils.cpp)
constexpr int kLat = 27;       // lattice coordinates -1 .. 25
constexpr int kLatOff = 1;

__device__ __forceinline__ double sin_poly(double x) {
    const double invpio2 = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079632673412561417e+00;
    const double pio2_1t = 6.07710050650619224932e-11;
    double k = rint(x * invpio2);
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

__global__ void curl_table_kernel(uint8_t* __restrict__ table) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kLat * kLat * kLat) return;
    const float x = (float)(i % kLat - kLatOff), y = (float)((i / kLat) % kLat - kLatOff),
                z = (float)(i / (kLat * kLat) - kLatOff);
    float px = x * 12.9898f, py = y * 78.233f, pz = z * (float)47.387;
    float d = (px + py) + pz;
    float n = (float)sin_poly((double)d) * 43758.5453f;
    n = n - floorf(n);
    table[i] = (uint8_t)(int)floorf(12.f * n);
}

__constant__ float kBasis[12][3] = {
    {(float)0.7071, (float)0.7071, 0.f},  {(float)0.7071, (float)-0.7071, 0.f},
    {(float)-0.7071, (float)0.7071, 0.f}, {(float)-0.7071, (float)-0.7071, 0.f},
    {(float)0.7071, 0.f, (float)0.7071},  {(float)0.7071, 0.f, (float)-0.7071},
    {(float)-0.7071, 0.f, (float)0.7071}, {(float)-0.7071, 0.f, (float)-0.7071},
    {0.f, (float)0.7071, (float)0.7071},  {0.f, (float)0.7071, (float)-0.7071},
    {0.f, (float)-0.7071, (float)0.7071}, {0.f, (float)-0.7071, (float)-0.7071},
};

__device__ __forceinline__ float ref_lerp(float a, float b, float t) {
    return (float)((1.0 - (double)t) * (double)a + (double)(t * b));
}

__device__ __forceinline__ float grad_dot(const uint8_t* __restrict__ tab, float x, float y, float z, float dx, float dy,
                                          float dz) {
    const int i = ((int)x + kLatOff) + kLat * (((int)y + kLatOff) + kLat * ((int)z + kLatOff));
    const float* g = kBasis[tab[i]];
    float a = g[0] * dx, b = g[1] * dy, c = g[2] * dz;
    return (a + b) + c;
}

// perlinNoise (ImageUtils.cpp:41-73)
__device__ float perlin_ref(const uint8_t* __restrict__ tab, float x, float y, float z, float freq) {
    x *= freq; y *= freq; z *= freq;
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    float rx = x - fx, ry = y - fy, rz = z - fz;
    float ux = ((rx * rx) * rx) * ((rx * ((rx * 6.0f) - 15.0f)) + 10.0f);
    float uy = ((ry * ry) * ry) * ((ry * ((ry * 6.0f) - 15.0f)) + 10.0f);
    float uz = ((rz * rz) * rz) * ((rz * ((rz * 6.0f) - 15.0f)) + 10.0f);
    float gx = fx + 1.0f, gy = fy + 1.0f, gz = fz + 1.0f;
    float nnn = grad_dot(tab, fx, fy, fz, rx, ry, rz);
    float nnp = grad_dot(tab, fx, fy, gz, rx - 0.f, ry - 0.f, rz - 1.f);
    float npn = grad_dot(tab, fx, gy, fz, rx - 0.f, ry - 1.f, rz - 0.f);
    float npp = grad_dot(tab, fx, gy, gz, rx - 0.f, ry - 1.f, rz - 1.f);
    float pnn = grad_dot(tab, gx, fy, fz, rx - 1.f, ry - 0.f, rz - 0.f);
    float pnp = grad_dot(tab, gx, fy, gz, rx - 1.f, ry - 0.f, rz - 1.f);
    float ppn = grad_dot(tab, gx, gy, fz, rx - 1.f, ry - 1.f, rz - 0.f);
    float ppp = grad_dot(tab, gx, gy, gz, rx - 1.f, ry - 1.f, rz - 1.f);
    float nn = ref_lerp(nnn, pnn, ux), np = ref_lerp(nnp, pnp, ux);
    float pn = ref_lerp(npn, ppn, ux), pp = ref_lerp(npp, ppp, ux);
    float n = ref_lerp(nn, pn, uy), p = ref_lerp(np, pp, uy);
    return ref_lerp(n, p, uz);
}

// FBM (ImageUtils.cpp:75-87)
__device__ float fbm_ref(const uint8_t* __restrict__ tab, float x, float y, float z, float freq, int octaves) {
    float noise = 0.0f, weight = 1.0f, total = 0.0f;
    for (int i = 0; i < octaves; i++) {
        total += weight;
        noise += weight * perlin_ref(tab, x, y, z, freq);
        freq *= 2.0f;
        weight *= 0.4f;
    }
    return noise / total;
}

__device__ __forceinline__ float central(float a, float b) { return (float)((double)(b - a) / (2.0 * 0.0005)); }

// Order-preserving float <-> uint32 encoding for atomic min/max.
__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// curlNoiseFBM (ImageUtils.cpp:130-169) per texel + block min/max.
__global__ __launch_bounds__(256) void curl_field_kernel(const uint8_t* __restrict__ tab, int dim,
                                                         float* __restrict__ field, uint32_t* __restrict__ mm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    if (i < dim * dim) {
        const int row = i / dim, col = i % dim;
        const float px = (float)col / (float)dim, py = (float)row / (float)dim;
        const double eps = 0.0005;
        const float xm = (float)((double)px - eps), xp = (float)((double)px + eps);
        const float ym = (float)((double)py - eps), yp = (float)((double)py + eps);
        const float f = 3.f;
        const int o = 4;
        float dydx = central(fbm_ref(tab, xm, py, 0.5f, f, o), fbm_ref(tab, xp, py, 0.5f, f, o));
        float dxdy = central(fbm_ref(tab, px, ym, 0.5f, f, o), fbm_ref(tab, px, yp, 0.5f, f, o));
        float dxdz = central(fbm_ref(tab, px, 0.5f, ym, f, o), fbm_ref(tab, px, 0.5f, yp, f, o));
        float dzdx = central(fbm_ref(tab, xm, 0.5f, py, f, o), fbm_ref(tab, xp, 0.5f, py, f, o));
        float dzdy = central(fbm_ref(tab, 0.5f, ym, px, f, o), fbm_ref(tab, 0.5f, yp, px, f, o));
        float dydz = central(fbm_ref(tab, 0.5f, py, xm, f, o), fbm_ref(tab, 0.5f, py, xp, f, o));
        float v[3] = {dzdy - dydz, dxdz - dzdx, dydx - dxdy};
        for (int k = 0; k < 3; k++) {
            field[3 * i + k] = v[k];
            lo[k] = v[k];
            hi[k] = v[k];
        }
    }
    // wave reduction, then one atomic per wave and channel
    for (int k = 0; k < 3; k++) {
        for (int off = 32; off > 0; off >>= 1) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], off));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off));
        }
    }
    if ((threadIdx.x & 63) == 0) {
        for (int k = 0; k < 3; k++) {
            atomicMin(&mm[2 * k], f2ord(lo[k]));
            atomicMax(&mm[2 * k + 1], f2ord(hi[k]));
        }
    }
}

__global__ void curl_minmax_init_kernel(uint32_t* mm) {
    if (threadIdx.x < 3) {
        mm[2 * threadIdx.x] = 0xffffffffu;
        mm[2 * threadIdx.x + 1] = 0u;
    }
}

// remap to [0,1] (ImageUtils.cpp:171-174, 200-204) and quantise (:209-212), written in the
// destination format/layout. FP16x4 keeps the remapped value unquantised.
__global__ __launch_bounds__(256) void curl_store_kernel(const float* __restrict__ field, const uint32_t* __restrict__ mm,
                                                         int dim, int apron, int fmt, uint8_t* __restrict__ dst) {
    const int sd = dim + apron;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sd * sd) return;
    const int x = i % sd, y = i / sd;
    const int src = (y % dim) * dim + (x % dim);
    float m[3];
    for (int k = 0; k < 3; k++) {
        const float lo = ord2f(mm[2 * k]), hi = ord2f(mm[2 * k + 1]);
        m[k] = 0.f + ((field[3 * src + k] - lo) / (hi - lo) * (1.f - 0.f));
    }
    if (fmt == MM_FMT_RGBA8) {
        uint32_t r = (uint32_t)(uint8_t)(int)roundf(m[0] * 255.f);
        uint32_t g = (uint32_t)(uint8_t)(int)roundf(m[1] * 255.f);
        uint32_t b = (uint32_t)(uint8_t)(int)roundf(m[2] * 255.f);
        reinterpret_cast<uint32_t*>(dst)[i] = r | (g << 8) | (b << 16) | (255u << 24);
    } else {  // MM_FMT_FP16X4
        uint2 v;
        v.x = (uint32_t)__half_as_ushort(__float2half_rn(m[0])) | ((uint32_t)__half_as_ushort(__float2half_rn(m[1])) << 16);
        v.y = (uint32_t)__half_as_ushort(__float2half_rn(m[2]));
        reinterpret_cast<uint2*>(dst)[i] = v;
    }
}

// ------------------------------------------------------------------ build-defined generators
// PCG3D hash (Jarzynski & Olano 2020, "Hash Functions for GPU Rendering").
__device__ __forceinline__ void hash3(uint32_t x, uint32_t y, uint32_t z, uint32_t seed, uint32_t& a, uint32_t& b,
                                      uint32_t& c) {
    uint32_t v0 = x + seed, v1 = y ^ (seed * 0x85EBCA6Bu), v2 = z + (seed * 0xC2B2AE35u);
    v0 = v0 * 1664525u + 1013904223u;
    v1 = v1 * 1664525u + 1013904223u;
    v2 = v2 * 1664525u + 1013904223u;
    v0 += v1 * v2; v1 += v2 * v0; v2 += v0 * v1;
    v0 ^= v0 >> 16; v1 ^= v1 >> 16; v2 ^= v2 >> 16;
    v0 += v1 * v2; v1 += v2 * v0; v2 += v0 * v1;
    a = v0; b = v1; c = v2;
}
__device__ __forceinline__ float u01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }
__device__ __forceinline__ int wrapi(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}
__device__ __forceinline__ float clamp01(float v) { return v < 0.f ? 0.f : (v > 1.f ? 1.f : v); }
__device__ __forceinline__ float lerp_(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ uint32_t layer_seed(uint32_t seed, uint32_t layer) {
    return seed * 0x9E3779B9u + layer * 0x632BE5ABu;
}

// Sources of per-cell data for the Worley and Perlin bodies, addressed by unwrapped cell
// coordinates. HashCells hashes the wrapped cell on the spot; LdsCells reads what a block staged
// in LDS (the same hash of the same wrapped cell, computed once per block), so both feed the same
// arithmetic and give the same bits.
struct HashCells {
    int cells;
    uint32_t seed;
    __device__ __forceinline__ void point(int x, int y, int z, float& ux, float& uy, float& uz) const {
        uint32_t h0, h1, h2;
        hash3((uint32_t)wrapi(x, cells), (uint32_t)wrapi(y, cells), (uint32_t)wrapi(z, cells), seed, h0, h1, h2);
        ux = u01(h0);
        uy = u01(h1);
        uz = u01(h2);
    }
};

struct LdsCells {
    const float4* cell;      // (x, y, z, unused) per staged cell: one 16-byte LDS read per lookup
    int bx, by, bz, nx, ny;  // first staged cell and the staged extent in x and y
    __device__ __forceinline__ void point(int x, int y, int z, float& ux, float& uy, float& uz) const {
        const float4 v = cell[((z - bz) * ny + (y - by)) * nx + (x - bx)];
        ux = v.x;
        uy = v.y;
        uz = v.z;
    }
};

// Inverted F1 cellular noise on a periodic grid of `cells`^3 cells.
template <class Cells>
__device__ __forceinline__ float worley_at(float px, float py, float pz, int cells, const Cells& src) {
    float qx = px * (float)cells, qy = py * (float)cells, qz = pz * (float)cells;
    int ix = (int)floorf(qx), iy = (int)floorf(qy), iz = (int)floorf(qz);
    float best = 1e30f;
#pragma unroll
    for (int oz = -1; oz <= 1; oz++) {
        const float bz = qz - (float)(iz + oz);
#pragma unroll
        for (int oy = -1; oy <= 1; oy++) {
            const float by = qy - (float)(iy + oy);
#pragma unroll
            for (int ox = -1; ox <= 1; ox++) {
                float ux, uy, uz;
                src.point(ix + ox, iy + oy, iz + oz, ux, uy, uz);
                float dx = (qx - (float)(ix + ox)) - ux;
                float dy = by - uy;
                float dz = bz - uz;
                float d2 = ((dx * dx) + (dy * dy)) + (dz * dz);
                best = __builtin_fminf(d2, best);  // = d2 < best ? d2 : best (no NaN: d2 >= 0)
            }
        }
    }
    float d = sqrt_rn(best);
    return 1.0f - (d < 1.0f ? d : 1.0f);
}

__device__ float worley(float px, float py, float pz, int cells, uint32_t seed) {
    return worley_at(px, py, pz, cells, HashCells{cells, seed});
}

__device__ float worley_fbm(float px, float py, float pz, int cells, uint32_t seed) {
    float a = worley(px, py, pz, cells, seed);
    float b = worley(px, py, pz, cells * 2, seed + 1u);
    float c = worley(px, py, pz, cells * 4, seed + 2u);
    return (0.625f * a + 0.25f * b) + 0.125f * c;
}

__constant__ float kGrad[12][3] = {{1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1},
                                   {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};

// Gradient of a lattice point of the periodic Perlin grid, from the hash of the wrapped point.
struct HashGrads {
    int period;
    uint32_t seed;
    __device__ __forceinline__ void point(int x, int y, int z, float& gx, float& gy, float& gz) const {
        uint32_t h0, h1, h2;
        hash3((uint32_t)wrapi(x, period), (uint32_t)wrapi(y, period), (uint32_t)wrapi(z, period), seed, h0, h1, h2);
        const float* g = kGrad[h0 % 12u];
        gx = g[0];
        gy = g[1];
        gz = g[2];
    }
};

__device__ __forceinline__ float fade(float t) { return ((t * t) * t) * ((t * ((t * 6.0f) - 15.0f)) + 10.0f); }

template <class Grads>
__device__ __forceinline__ float perlin_at(float px, float py, float pz, int period, const Grads& src) {
    float qx = px * (float)period, qy = py * (float)period, qz = pz * (float)period;
    float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
    float rx = qx - fx, ry = qy - fy, rz = qz - fz;
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    float n[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
        float g0, g1, g2;
        src.point(ix + dx, iy + dy, iz + dz, g0, g1, g2);
        float ax = rx - (float)dx, ay = ry - (float)dy, az = rz - (float)dz;
        n[c] = ((g0 * ax) + (g1 * ay)) + (g2 * az);
    }
    float ux = fade(rx), uy = fade(ry), uz = fade(rz);
    float x00 = lerp_(n[0], n[1], ux), x10 = lerp_(n[2], n[3], ux);
    float x01 = lerp_(n[4], n[5], ux), x11 = lerp_(n[6], n[7], ux);
    float y0 = lerp_(x00, x10, uy), y1 = lerp_(x01, x11, uy);
    return lerp_(y0, y1, uz);
}

__device__ float perlin_tiled(float px, float py, float pz, int period, uint32_t seed) {
    return perlin_at(px, py, pz, period, HashGrads{period, seed});
}

__device__ float perlin_fbm(float px, float py, float pz, int period, uint32_t seed) {
    float s = 0.f, w = 1.f;
    for (int o = 0; o < 4; o++) {
        s = s + w * perlin_tiled(px, py, pz, period << o, seed + (uint32_t)o);
        w = w * 0.5f;
    }
    return s * (1.0f / 1.875f);
}

__device__ float value2(float px, float py, int period, uint32_t seed) {
    float qx = px * (float)period, qy = py * (float)period;
    float fx = floorf(qx), fy = floorf(qy);
    float tx = qx - fx, ty = qy - fy;
    int ix = (int)fx, iy = (int)fy;
    float sx = (tx * tx) * (3.0f - 2.0f * tx), sy = (ty * ty) * (3.0f - 2.0f * ty);
    float v[4];
    for (int c = 0; c < 4; c++) {
        uint32_t h0, h1, h2;
        hash3((uint32_t)wrapi(ix + (c & 1), period), (uint32_t)wrapi(iy + (c >> 1), period), 0u, seed, h0, h1, h2);
        v[c] = u01(h0);
    }
    return lerp_(lerp_(v[0], v[1], sx), lerp_(v[2], v[3], sx), sy);
}

__device__ float value_fbm(float px, float py, int period, uint32_t seed) {
    float s = 0.f, w = 1.f;
    for (int o = 0; o < 5; o++) {
        s = s + w * value2(px, py, period << o, seed + (uint32_t)o);
        w = w * 0.5f;
    }
    return s * (1.0f / 1.9375f);
}

__device__ __forceinline__ uint32_t q8(float v) {
    v = clamp01(v);
    return (uint32_t)(int)floorf(v * 255.0f + 0.5f);
}

__device__ __forceinline__ uint32_t half_bits(float v) { return (uint32_t)__half_as_ushort(__float2half_rn(v)); }

// Round-to-nearest-even FP8 e4m3 (OCP: bias 7, max finite 448, no infinities) of a value >= 0.
__device__ __forceinline__ uint32_t fp8_e4m3(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 448.0f) return 0x7Eu;
    uint32_t u = __float_as_uint(v);
    int e = (int)((u >> 23) & 0xff) - 127;
    if (e < -6) {  // subnormal: multiples of 2^-9
        float q = rintf(v * 512.0f);
        return (uint32_t)q;  // q in [0, 8]; 8 encodes the smallest normal
    }
    uint32_t mant = u & 0x7fffffu;
    uint32_t keep = mant >> 20, rest = mant & 0xfffffu;
    if (rest > 0x80000u || (rest == 0x80000u && (keep & 1u))) keep++;
    uint32_t ee = (uint32_t)(e + 7);
    if (keep == 8u) { keep = 0u; ee++; }
    uint32_t r = (ee << 3) | keep;
    return r > 0x7Eu ? 0x7Eu : r;
}

struct GenArgs {
    uint8_t* dst;
    int dim, apron, fmt;
    uint32_t seed;
    float coverage;
    float inv_scale[4];
};

// ---- low-res volume: 4 Perlin octaves (periods 4..32) and 4 Worley layers of 3 octaves
// (cells 4..128). "Level" l < 4 is Perlin octave l, l >= 4 is Worley layer (l - 4) / 3, octave
// (l - 4) % 3; both sources below hand out the lattice data of a level.
constexpr int kLowLevels = 16;
__host__ __device__ constexpr int low_level_cells(int l) {
    return l < 4 ? 4 << l : (4 << ((l - 4) / 3)) << ((l - 4) % 3);
}
__device__ __forceinline__ uint32_t low_level_seed(uint32_t seed, int l) {
    return l < 4 ? layer_seed(seed, 1) + (uint32_t)l : layer_seed(seed, 2 + (uint32_t)(l - 4) / 3) + (uint32_t)((l - 4) % 3);
}

// Bricks of the low-res kernel (texels per axis) and the lattice cells a block can stage in LDS.
constexpr int kBrickX = 8, kBrickY = 8, kBrickZ = 4;
constexpr int kStageCap = 1536;  // staged cells per block: 24 KB of LDS

// Sources of the low-res levels: get<l>() hands out the lattice data of level l.
struct LowResHash {
    uint32_t seed;
    template <int L>
    __device__ __forceinline__ auto get() const {
        if constexpr (L < 4) return HashGrads{low_level_cells(L), low_level_seed(seed, L)};
        else return HashCells{low_level_cells(L), low_level_seed(seed, L)};
    }
};

// Lattice cells of one level staged for a brick: cells [b, b + n) per axis, at offset `off`.
struct StageMeta {
    int bx, by, bz, nx, ny, nz, off;
};

struct LowResLds {
    const float4* cells;
    const StageMeta* meta;
    __device__ __forceinline__ LdsCells level(int l) const {
        const StageMeta& m = meta[l];
        return LdsCells{cells + m.off, m.bx, m.by, m.bz, m.nx, m.ny};
    }
    template <int L>
    __device__ __forceinline__ LdsCells get() const { return level(L); }
};

// The staged boxes of a 128^3 volume. Bricks start at multiples of 8 (x, y) and 4 (z), and texel t
// sits at q = (2t + 1) c / 256 in cell units (exact), so a brick of B texels spans
// floor(q1) - floor(q0) = max(B c / 128 - 1, 0) cells at every position: the boxes that axis_span
// computes have the same size for every brick, kDim128N cells per axis (pad 1 for Worley levels).
constexpr int kDim128Span(int l, int b) { return b * low_level_cells(l) / 128 >= 1 ? b * low_level_cells(l) / 128 - 1 : 0; }
constexpr int kDim128N(int l, int b) { return kDim128Span(l, b) + 2 + (l < 4 ? 0 : 1); }
constexpr int kDim128Cells(int l) { return kDim128N(l, kBrickX) * kDim128N(l, kBrickY) * kDim128N(l, kBrickZ); }
constexpr int kDim128Off(int l) { return l == 0 ? 0 : kDim128Off(l - 1) + kDim128Cells(l - 1); }
static_assert(kDim128N(15, 8) == 10 && kDim128N(15, 4) == 6 && kDim128N(0, 8) == 2, "128^3 staged box sizes");
static_assert(kDim128Off(kLowLevels) <= kStageCap, "a 128^3 brick's lattice cells fit the stage");

// LDS cells with a compile-time box: the 27 neighbour reads of a Worley evaluation become one base
// address plus immediate offsets.
template <int NX, int NY>
struct LdsCellsC {
    const float4* cell;
    int bx, by, bz;
    __device__ __forceinline__ void point(int x, int y, int z, float& ux, float& uy, float& uz) const {
        const float4 v = cell[((z - bz) * NY + (y - by)) * NX + (x - bx)];
        ux = v.x;
        uy = v.y;
        uz = v.z;
    }
    __device__ __forceinline__ const float4* home(int x, int y, int z) const {
        return cell + (((z - bz) * NY + (y - by)) * NX + (x - bx));
    }
};

// worley_at for a compile-time box: the same arithmetic, the neighbour cells read at constant
// offsets from the home cell.
template <int NX, int NY>
__device__ __forceinline__ float worley_at(float px, float py, float pz, int cells, const LdsCellsC<NX, NY>& src) {
    float qx = px * (float)cells, qy = py * (float)cells, qz = pz * (float)cells;
    int ix = (int)floorf(qx), iy = (int)floorf(qy), iz = (int)floorf(qz);
    const float4* h = src.home(ix, iy, iz);
    float best = 1e30f;
#pragma unroll
    for (int oz = -1; oz <= 1; oz++) {
        const float bz = qz - (float)(iz + oz);
#pragma unroll
        for (int oy = -1; oy <= 1; oy++) {
            const float by = qy - (float)(iy + oy);
#pragma unroll
            for (int ox = -1; ox <= 1; ox++) {
                const float4 u = h[(oz * NY + oy) * NX + ox];
                float dx = (qx - (float)(ix + ox)) - u.x;
                float dy = by - u.y;
                float dz = bz - u.z;
                float d2 = ((dx * dx) + (dy * dy)) + (dz * dz);
                best = __builtin_fminf(d2, best);
            }
        }
    }
    float d = sqrt_rn(best);
    return 1.0f - (d < 1.0f ? d : 1.0f);
}

struct LowResLds128 {
    const float4* cells;
    const StageMeta* meta;
    template <int L>
    __device__ __forceinline__ LdsCellsC<kDim128N(L, kBrickX), kDim128N(L, kBrickY)> get() const {
        // wave-uniform box origin (the brick's), in scalar registers
        const int bx = __builtin_amdgcn_readfirstlane(meta[L].bx), by = __builtin_amdgcn_readfirstlane(meta[L].by),
                  bz = __builtin_amdgcn_readfirstlane(meta[L].bz);
        constexpr int off = kDim128Off(L);  // evaluated here, not by a device-side recursive call
        return {cells + off, bx, by, bz};
    }
};

template <int LY, class Src>
__device__ __forceinline__ float worley_fbm_at(float px, float py, float pz, const Src& src) {
    constexpr int cells = 4 << LY;
    float a = worley_at(px, py, pz, cells, src.template get<4 + 3 * LY>());
    float b = worley_at(px, py, pz, cells * 2, src.template get<5 + 3 * LY>());
    float c = worley_at(px, py, pz, cells * 4, src.template get<6 + 3 * LY>());
    return (0.625f * a + 0.25f * b) + 0.125f * c;
}

// The four channels of the low-res texel at (px, py, pz): Perlin-Worley, then Worley FBMs of
// base frequency 8, 16 and 32.
template <class Src>
__device__ __forceinline__ void low_res_values(float px, float py, float pz, const Src& src, float v[4]) {
    float s = 0.f, w = 1.f;
    s = s + w * perlin_at(px, py, pz, 4, src.template get<0>());
    w = w * 0.5f;
    s = s + w * perlin_at(px, py, pz, 8, src.template get<1>());
    w = w * 0.5f;
    s = s + w * perlin_at(px, py, pz, 16, src.template get<2>());
    w = w * 0.5f;
    s = s + w * perlin_at(px, py, pz, 32, src.template get<3>());
    const float perlin = s * (1.0f / 1.875f);
    float p01 = clamp01(0.5f + 1.5f * perlin);
    float wf = worley_fbm_at<0>(px, py, pz, src);
    float pw = (p01 - (wf - 1.0f)) / (1.0f - (wf - 1.0f));
    v[0] = clamp01(0.8f * pw - 0.04f);
    v[1] = clamp01(0.26f + 0.9f * worley_fbm_at<1>(px, py, pz, src));
    v[2] = clamp01(0.26f + 0.9f * worley_fbm_at<2>(px, py, pz, src));
    v[3] = clamp01(0.26f + 0.9f * worley_fbm_at<3>(px, py, pz, src));
}

__device__ __forceinline__ void store_low_res(const GenArgs& a, long i, const float v[4]) {
    if (a.fmt == MM_FMT_RGBA8) {
        reinterpret_cast<uint32_t*>(a.dst)[i] = q8(v[0]) | (q8(v[1]) << 8) | (q8(v[2]) << 16) | (q8(v[3]) << 24);
    } else {  // MM_FMT_FP8X4
        uint32_t r = 0;
        for (int k = 0; k < 4; k++) r |= fp8_e4m3(v[k] * a.inv_scale[k]) << (8 * k);
        reinterpret_cast<uint32_t*>(a.dst)[i] = r;
    }
}

// Brick kernel: a block of 256 work-items takes an 8 x 8 x 4 brick of logical texels. The lattice
// cells the brick's texels touch at each level (about 1.2k cells for a 128^3 volume, against 356
// hashes per texel) are hashed once into LDS; the texels then read them there. When a brick
// touches more cells than LDS holds (small volumes, where cells outnumber texels) it hashes in
// place. Texels whose coordinate is below the apron width are also written to the apron copy
// (x + dim, ...).

// Cells [floor(q(t0)) - pad, floor(q(t1)) + 1] of one axis, q(t) = ((2t + 1) / (2 dim)) * cells,
// evaluated as the texel evaluates it (q is monotone in t, so every texel's cells lie inside).
__device__ __forceinline__ void axis_span(int t0, int t1, float inv, int cells, int pad, int& base, int& n) {
    const float q0 = ((float)(2 * t0 + 1) * inv) * (float)cells;
    const float q1 = ((float)(2 * t1 + 1) * inv) * (float)cells;
    base = (int)floorf(q0) - pad;
    n = ((int)floorf(q1) + 1) - base + 1;
}

// DIM128: the volume is 128^3 (asserted by the host), whose staged boxes have compile-time sizes.
template <bool DIM128>
__global__ __launch_bounds__(256) void low_res_brick_kernel(GenArgs a) {
    __shared__ float4 cells[kStageCap];
    __shared__ StageMeta meta[kLowLevels];
    __shared__ int total;
    const int tid = threadIdx.x;
    const int nbx = (a.dim + kBrickX - 1) / kBrickX, nby = (a.dim + kBrickY - 1) / kBrickY;
    const int x0 = (int)(blockIdx.x % nbx) * kBrickX, y0 = (int)((blockIdx.x / nbx) % nby) * kBrickY;
    const int z0 = (int)(blockIdx.x / (nbx * nby)) * kBrickZ;
    const int x1 = min(x0 + kBrickX, a.dim) - 1, y1 = min(y0 + kBrickY, a.dim) - 1, z1 = min(z0 + kBrickZ, a.dim) - 1;
    const float inv = 1.0f / (float)(2 * a.dim);
    if (tid < kLowLevels) {
        const int cells = low_level_cells(tid), pad = tid < 4 ? 0 : 1;
        StageMeta m;
        axis_span(x0, x1, inv, cells, pad, m.bx, m.nx);
        axis_span(y0, y1, inv, cells, pad, m.by, m.ny);
        axis_span(z0, z1, inv, cells, pad, m.bz, m.nz);
        m.off = m.nx * m.ny * m.nz;  // count for now; prefix-summed below
        meta[tid] = m;
    }
    __syncthreads();
    if (tid == 0) {
        int off = 0;
        for (int l = 0; l < kLowLevels; l++) {
            const int cnt = meta[l].off;
            meta[l].off = off;
            off += cnt;
        }
        total = off;
    }
    __syncthreads();
    const int nst = total;
    const bool staged = DIM128 || nst <= kStageCap;  // uniform across the block
    if (staged) {
        for (int e = tid; e < nst; e += 256) {
            int l = 0;
            while (l + 1 < kLowLevels && meta[l + 1].off <= e) l++;
            const StageMeta m = meta[l];
            const int r = e - m.off, jx = r % m.nx, jy = (r / m.nx) % m.ny, jz = r / (m.nx * m.ny);
            float u0, u1, u2;
            if (l < 4)
                HashGrads{low_level_cells(l), low_level_seed(a.seed, l)}.point(m.bx + jx, m.by + jy, m.bz + jz, u0, u1, u2);
            else
                HashCells{low_level_cells(l), low_level_seed(a.seed, l)}.point(m.bx + jx, m.by + jy, m.bz + jz, u0, u1, u2);
            cells[e] = make_float4(u0, u1, u2, 0.0f);
        }
    }
    __syncthreads();
    const int x = x0 + (tid & (kBrickX - 1)), y = y0 + ((tid / kBrickX) & (kBrickY - 1)), z = z0 + tid / (kBrickX * kBrickY);
    if (x > x1 || y > y1 || z > z1) return;
    const float px = (float)(2 * x + 1) * inv, py = (float)(2 * y + 1) * inv, pz = (float)(2 * z + 1) * inv;
    float v[4];
    if constexpr (DIM128)
        low_res_values(px, py, pz, LowResLds128{cells, meta}, v);
    else if (staged)
        low_res_values(px, py, pz, LowResLds{cells, meta}, v);
    else
        low_res_values(px, py, pz, LowResHash{a.seed}, v);
    const long sd = a.dim + a.apron;
    for (int cz = 0; cz < 2; cz++) {
        const int zs = z + cz * a.dim;
        if (cz && zs >= sd) break;
        for (int cy = 0; cy < 2; cy++) {
            const int ys = y + cy * a.dim;
            if (cy && ys >= sd) break;
            for (int cx = 0; cx < 2; cx++) {
                const int xs = x + cx * a.dim;
                if (cx && xs >= sd) break;
                store_low_res(a, ((long)zs * sd + ys) * sd + xs, v);
            }
        }
    }
}

__global__ __launch_bounds__(256) void hi_res_kernel(GenArgs a) {
    const int sd = a.dim + a.apron;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sd * sd * sd) return;
    const int x = (i % sd) % a.dim, y = ((i / sd) % sd) % a.dim, z = (i / (sd * sd)) % a.dim;
    const float inv = 1.0f / (float)(2 * a.dim);
    float px = (float)(2 * x + 1) * inv, py = (float)(2 * y + 1) * inv, pz = (float)(2 * z + 1) * inv;
    float v[3];
    v[0] = clamp01(0.26f + 0.9f * worley_fbm(px, py, pz, 4, layer_seed(a.seed, 6)));
    v[1] = clamp01(0.26f + 0.9f * worley_fbm(px, py, pz, 8, layer_seed(a.seed, 7)));
    v[2] = clamp01(0.26f + 0.9f * worley_fbm(px, py, pz, 16, layer_seed(a.seed, 8)));
    if (a.fmt == MM_FMT_RGBA8) {
        reinterpret_cast<uint32_t*>(a.dst)[i] = q8(v[0]) | (q8(v[1]) << 8) | (q8(v[2]) << 16);
    } else {  // MM_FMT_FP16X4
        uint2 o;
        o.x = half_bits(v[0]) | (half_bits(v[1]) << 16);
        o.y = half_bits(v[2]);
        reinterpret_cast<uint2*>(a.dst)[i] = o;
    }
}

__global__ __launch_bounds__(256) void weather_kernel(GenArgs a) {
    const int sd = a.dim + a.apron;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sd * sd) return;
    const int x = (i % sd) % a.dim, y = (i / sd) % a.dim;
    const float inv = 1.0f / (float)(2 * a.dim);
    float px = (float)(2 * x + 1) * inv, py = (float)(2 * y + 1) * inv;
    float va = value_fbm(px, py, 4, layer_seed(a.seed, 11));
    float vb = value_fbm(px, py, 8, layer_seed(a.seed, 12));
    float vc = value_fbm(px, py, 8, layer_seed(a.seed, 13));
    float r = 0.5f + 2.0f * (va - 0.5f);
    float R = clamp01(r + (a.coverage - 0.5f));
    float G = clamp01(0.5f + 1.5f * ((0.5f * va + 0.5f * vc) - 0.5f));
    float B = clamp01(0.5f + 2.5f * (vb - 0.5f));
    if (a.fmt == MM_FMT_RGBA8) {
        reinterpret_cast<uint32_t*>(a.dst)[i] = q8(R) | (q8(G) << 8) | (q8(B) << 16) | (255u << 24);
    } else {  // MM_FMT_RG8: (coverage, type)
        reinterpret_cast<uint16_t*>(a.dst)[i] = (uint16_t)(q8(R) | (q8(B) << 8));
    }
}

// ------------------------------------------------------------------ repacking of RGBA8 linear textures
struct PackArgs {
    const uint8_t* src;
    uint8_t* dst;
    int w, h, d;        // logical dims (d = 1 for 2D)
    int apron, fmt, kind;
    float inv_scale[4];
};

__global__ __launch_bounds__(256) void pack_kernel(PackArgs a) {
    const int sw = a.w + a.apron, sh = a.h + a.apron, sdp = (a.d > 1) ? a.d + a.apron : 1;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)sw * sh * sdp) return;
    const int x = (int)(i % sw) % a.w, y = (int)((i / sw) % sh) % a.h, z = (int)(i / ((long)sw * sh)) % a.d;
    const uint32_t t = reinterpret_cast<const uint32_t*>(a.src)[((long)z * a.h + y) * a.w + x];
    const uint32_t c0 = t & 0xff, c1 = (t >> 8) & 0xff, c2 = (t >> 16) & 0xff, c3 = t >> 24;
    switch (a.fmt) {
        case MM_FMT_RGBA8: reinterpret_cast<uint32_t*>(a.dst)[i] = t; break;
        case MM_FMT_RG8: reinterpret_cast<uint16_t*>(a.dst)[i] = (uint16_t)(c0 | (c2 << 8)); break;
        case MM_FMT_FP16X4: {
            uint2 o;
            o.x = half_bits((float)c0 / 255.0f) | (half_bits((float)c1 / 255.0f) << 16);
            o.y = half_bits((float)c2 / 255.0f) | (half_bits((float)c3 / 255.0f) << 16);
            reinterpret_cast<uint2*>(a.dst)[i] = o;
            break;
        }
        case MM_FMT_FP8X4: {
            const uint32_t c[4] = {c0, c1, c2, c3};
            uint32_t r = 0;
            for (int k = 0; k < 4; k++) r |= fp8_e4m3(((float)c[k] / 255.0f) * a.inv_scale[k]) << (8 * k);
            reinterpret_cast<uint32_t*>(a.dst)[i] = r;
            break;
        }
    }
}

inline unsigned blocks_for(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

namespace mm {

size_t noise_workspace_bytes() {
    // gradient table (27^3 B, rounded up) + min/max (6 x u32) + curl field for up to 512^2 texels
    return 20480 + 256 + (size_t)512 * 512 * 3 * sizeof(float);
}

}  // namespace mm

namespace {
// Device scratch of the allocation-free generator entry point: the curl chain's gradient table,
// min/max cells and float curl field for curl maps up to kStaticCurlMax^2 texels. One copy per GPU,
// held until the process ends; every call on that GPU shares it, so calls that overlap on one GPU
// race. The min/max cells are reset by a kernel at the start of every curl generation.
constexpr size_t kScratchBytes = 20480 + 256 + (size_t)MM_STATIC_CURL_MAX * MM_STATIC_CURL_MAX * 3 * sizeof(float);
__device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

namespace mm {

size_t static_scratch_bytes() { return kScratchBytes; }

// The device address of this GPU's copy of `scratch` (the array's name in host code is only a
// stand-in for the symbol), looked up once per device with that device current.
int static_scratch(void** ptr) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return MM_E_DEVICE;
    if (dev < 0 || dev >= kMaxDevices) return MM_E_DEVICE;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return MM_E_DEVICE;
        }
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return MM_OK;
}

namespace {
// A second stream for work that runs beside the caller's (the curl chain's small grids, the hi-res
// volume and the weather map beside the low-res volume; the reprojection beside a phase march):
// forked from the caller's stream and joined back into it with events, so a call stays ordered on
// that stream and can be captured into a graph. There is one side stream per (device, caller
// stream): a side stream joins whatever capture its caller's stream is in and stays in it until that
// capture ends, so work forked from another caller stream must never reach it. Created on first use
// for each caller stream, up to kMaxSides per process; beyond that the work stays on the caller's
// stream. The caller holds side_lock() from the fork to the join.
struct SideStream {
    int dev = -1;
    hipStream_t main = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
constexpr int kMaxSides = 256;
std::mutex g_side_mu;
SideStream g_side[kMaxSides];
int g_side_n = 0;

SideStream* side_for(int dev, hipStream_t main) {
    for (int i = 0; i < g_side_n; i++)
        if (g_side[i].dev == dev && g_side[i].main == main) return &g_side[i];
    if (g_side_n == kMaxSides) return nullptr;
    hipStream_t st = nullptr;
    hipEvent_t f = nullptr, j = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&f, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&j, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();  // the work then stays on the caller's stream
        if (st) (void)hipStreamDestroy(st);
        if (f) (void)hipEventDestroy(f);
        return nullptr;
    }
    SideStream& d = g_side[g_side_n++];
    d.dev = dev;
    d.main = main;
    d.s = st;
    d.fork = f;
    d.join = j;
    return &d;
}
}  // namespace

std::unique_lock<std::mutex> side_lock() { return std::unique_lock<std::mutex>(g_side_mu); }

hipStream_t side_fork(hipStream_t main) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    SideStream* d = side_for(dev, main);
    if (!d) return nullptr;
    if (hipEventRecord(d->fork, main) != hipSuccess || hipStreamWaitEvent(d->s, d->fork, 0) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return d->s;
}

int side_join(hipStream_t main, hipStream_t side) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return MM_E_DEVICE;
    SideStream* d = side_for(dev, main);
    if (!d || d->s != side) return MM_E_DEVICE;
    if (hipEventRecord(d->join, side) != hipSuccess || hipStreamWaitEvent(main, d->join, 0) != hipSuccess) {
        // the caller's stream could not be ordered after the side work: wait for it here, so nothing
        // the caller enqueues next (such as freeing a workspace the side work reads) overtakes it
        (void)hipGetLastError();
        (void)hipStreamSynchronize(side);
        return MM_E_DEVICE;
    }
    return MM_OK;
}

int launch_generate(const mm_noise_params* p, const mm_noise_set* out, void* workspace, void* stream,
                    bool allow_side) {
    hipStream_t main = (hipStream_t)stream;
    const uint32_t side_work = p->what & (MM_GEN_CURL | MM_GEN_HI_RES | MM_GEN_WEATHER);
    std::unique_lock<std::mutex> lock;
    hipStream_t side = nullptr;
    if (allow_side && (p->what & MM_GEN_LOW_RES) && side_work) {
        lock = side_lock();
        side = side_fork(main);
    }
    hipStream_t s = side ? side : main;
    if (p->what & MM_GEN_CURL) {
        const mm_texture& t = out->curl;
        uint8_t* ws = (uint8_t*)workspace;
        uint8_t* table = ws;
        uint32_t* mmx = (uint32_t*)(ws + 20480);
        float* field = (float*)(ws + 20480 + 256);
        const int dim = t.width;
        hipLaunchKernelGGL(curl_table_kernel, dim3(blocks_for(kLat * kLat * kLat)), dim3(256), 0, s, table);
        hipLaunchKernelGGL(curl_minmax_init_kernel, dim3(1), dim3(64), 0, s, mmx);
        hipLaunchKernelGGL(curl_field_kernel, dim3(blocks_for((long)dim * dim)), dim3(256), 0, s, table, dim, field, mmx);
        const int apron = t.layout == MM_LAYOUT_APRON ? 1 : 0;
        const long n = (long)(dim + apron) * (dim + apron);
        hipLaunchKernelGGL(curl_store_kernel, dim3(blocks_for(n)), dim3(256), 0, s, field, mmx, dim, apron, t.format,
                           (uint8_t*)t.data);
    }
    if (p->what & MM_GEN_HI_RES) {
        const mm_texture& t = out->hi_res;
        GenArgs a{(uint8_t*)t.data, t.width, t.layout == MM_LAYOUT_APRON ? 1 : 0, t.format, p->seed, 0.f, {1, 1, 1, 1}};
        const long sd = t.width + a.apron;
        hipLaunchKernelGGL(hi_res_kernel, dim3(blocks_for(sd * sd * sd)), dim3(256), 0, s, a);
    }
    if (p->what & MM_GEN_WEATHER) {
        const mm_texture& t = out->weather;
        GenArgs a{(uint8_t*)t.data, t.width, t.layout == MM_LAYOUT_APRON ? 1 : 0, t.format, p->weather_seed, p->coverage,
                  {1, 1, 1, 1}};
        const long sd = t.width + a.apron;
        hipLaunchKernelGGL(weather_kernel, dim3(blocks_for(sd * sd)), dim3(256), 0, s, a);
    }
    if (p->what & MM_GEN_LOW_RES) {
        const mm_texture& t = out->low_res;
        GenArgs a{(uint8_t*)t.data, t.width, t.layout == MM_LAYOUT_APRON ? 1 : 0, t.format, p->seed, 0.f, {1, 1, 1, 1}};
        if (t.format == MM_FMT_FP8X4)
            for (int k = 0; k < 4; k++) a.inv_scale[k] = 1.0f / t.scale[k];
        const long bricks = (long)((t.width + kBrickX - 1) / kBrickX) * ((t.width + kBrickY - 1) / kBrickY) *
                            ((t.width + kBrickZ - 1) / kBrickZ);
        if (t.width == 128)
            hipLaunchKernelGGL(low_res_brick_kernel<true>, dim3((unsigned)bricks), dim3(256), 0, main, a);
        else
            hipLaunchKernelGGL(low_res_brick_kernel<false>, dim3((unsigned)bricks), dim3(256), 0, main, a);
    }
    if (side) {
        const int e = side_join(main, side);
        if (e) return e;
    }
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

int launch_pack(int kind, const mm_texture* src, const mm_texture* dst, void* stream) {
    (void)kind;
    PackArgs a;
    a.src = (const uint8_t*)src->data;
    a.dst = (uint8_t*)dst->data;
    a.w = src->width; a.h = src->height; a.d = src->depth;
    a.apron = dst->layout == MM_LAYOUT_APRON ? 1 : 0;
    a.fmt = dst->format;
    a.kind = kind;
    for (int k = 0; k < 4; k++) a.inv_scale[k] = (dst->format == MM_FMT_FP8X4) ? 1.0f / dst->scale[k] : 1.0f;
    const long n = (long)(a.w + a.apron) * (a.h + a.apron) * (a.d > 1 ? a.d + a.apron : 1);
    hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, a);
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

}  // namespace mm
