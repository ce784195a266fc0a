// Oracle: the build-defined noise-volume and weather-map generators (test infrastructure only).
//
// The reference ships these textures as baked Houdini assets (README.md:68; loaded by
// Texture3D::initFromFile, Texture.cpp:502-538), so their generator is specified by this build.
// The consumer contract comes from compute-clouds.comp:
//   low-res 128^3: R = Perlin-Worley base shape, GBA = Worley FBMs read as 0.625g+0.25b+0.125a
//                  (comp:238-247);
//   hi-res 32^3:   RGB = Worley FBMs read as 0.625r+0.25g+0.125b, A unused (comp:223-224);
//   weather 512^2: R = coverage (comp:245), B = cloud type (comp:237), G unused by the march.
// Method: Worley cellular noise (Worley 1996, "A cellular texture basis function"), Perlin's
// improved gradient noise (Perlin 2002) and Perlin-Worley dilation as described by Schneider
// (2015, "The real-time volumetric cloudscapes of Horizon Zero Dawn"), all periodic so REPEAT
// sampling is seamless, hashed with PCG3D (Jarzynski & Olano 2020) so CPU and GPU agree bit for bit.
// Every float operation is written out in evaluation order; -ffp-contract=off on both sides.
#include <math.h>
#include <stdint.h>
#include "oracle.h"

namespace {

// Planted mistakes for the tests that check each generator against its definition
// (orc_set_noise_mutation); 0 = none.
int g_mutation = 0;

inline void pcg3d(uint32_t v[3]) {
    const uint32_t inc = g_mutation == 1 ? 1013904225u : 1013904223u;
    for (int i = 0; i < 3; i++) v[i] = v[i] * 1664525u + inc;
    v[0] += v[1] * v[2]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1];
    v[0] ^= v[0] >> 16; v[1] ^= v[1] >> 16; v[2] ^= v[2] >> 16;
    v[0] += v[1] * v[2]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1];
}

inline float u01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

inline int wrapi(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }
// mutation 2: a cell index that is not wrapped at the period (the Worley grid stops tiling)
inline int wrap_worley(int i, int n) { return g_mutation == 2 ? i : wrapi(i, n); }

// Perlin 2002 gradient set (the 12 cube-edge directions).
const float kGrad[12][3] = {
    {1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1},
    {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};

inline float fade(float t) {
    if (g_mutation == 4) return (t * t) * (3.0f - 2.0f * t);  // cubic fade instead of Perlin 2002's quintic
    return ((t * t) * t) * ((t * ((t * 6.0f) - 15.0f)) + 10.0f);
}
inline float lerpf(float a, float b, float t) { return a + t * (b - a); }
inline float clamp01(float v) { return v < 0.f ? 0.f : (v > 1.f ? 1.f : v); }

// Seeds of the individual noise layers.
inline uint32_t layer_seed(uint32_t seed, uint32_t layer) { return seed * 0x9E3779B9u + layer * 0x632BE5ABu; }

}  // namespace

extern "C" float orc_worley_fbm(float px, float py, float pz, int cells, uint32_t seed) {
    float a = orc_worley(px, py, pz, cells, seed);
    float b = orc_worley(px, py, pz, cells * 2, seed + 1u);
    float c = orc_worley(px, py, pz, cells * 4, seed + 2u);
    return (0.625f * a + 0.25f * b) + 0.125f * c;
}

extern "C" float orc_perlin_fbm(float px, float py, float pz, int period, uint32_t seed) {
    float s = 0.f, w = 1.f;
    for (int o = 0; o < 4; o++) {
        s = s + w * orc_perlin_tiled(px, py, pz, period << o, seed + (uint32_t)o);
        w = w * 0.5f;
    }
    return s * (1.0f / 1.875f);
}

extern "C" float orc_value2(float px, float py, int period, uint32_t seed) {
    float qx = px * (float)period, qy = py * (float)period;
    float fx = floorf(qx), fy = floorf(qy);
    float tx = qx - fx, ty = qy - fy;
    int ix = (int)fx, iy = (int)fy;
    float sx = (tx * tx) * (3.0f - 2.0f * tx), sy = (ty * ty) * (3.0f - 2.0f * ty);
    uint32_t h[3];
    float v[4];
    for (int c = 0; c < 4; c++) {
        const int vx = ix + (c & 1), vy = iy + (c >> 1);
        if (g_mutation == 6) orc_hash3((uint32_t)vx, (uint32_t)vy, 0u, seed, h);  // lattice not wrapped
        else orc_hash3((uint32_t)wrapi(vx, period), (uint32_t)wrapi(vy, period), 0u, seed, h);
        v[c] = u01(h[0]);
    }
    return lerpf(lerpf(v[0], v[1], sx), lerpf(v[2], v[3], sx), sy);
}

extern "C" float orc_value_fbm(float px, float py, int period, uint32_t seed) {
    float s = 0.f, w = 1.f;
    for (int o = 0; o < 5; o++) {
        s = s + w * orc_value2(px, py, period << o, seed + (uint32_t)o);
        w = w * 0.5f;
    }
    return s * (1.0f / 1.9375f);
}

extern "C" void orc_set_noise_mutation(int m) { g_mutation = m; }

extern "C" uint32_t orc_hash3(uint32_t x, uint32_t y, uint32_t z, uint32_t seed, uint32_t out[3]) {
    uint32_t v[3] = {x + seed, y ^ (seed * 0x85EBCA6Bu), z + (seed * 0xC2B2AE35u)};
    pcg3d(v);
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
    return v[0];
}

// Inverted F1 cellular noise on a periodic grid of `cells`^3 cells, one feature point per cell:
// 1 - min(1, distance to the nearest feature point in cell units).
extern "C" float orc_worley(float px, float py, float pz, int cells, uint32_t seed) {
    float qx = px * (float)cells, qy = py * (float)cells, qz = pz * (float)cells;
    float cx = floorf(qx), cy = floorf(qy), cz = floorf(qz);
    int ix = (int)cx, iy = (int)cy, iz = (int)cz;
    float best = 1e30f;
    uint32_t h[3];
    const int reach = g_mutation == 3 ? 0 : 1;  // mutation 3: only the query's own cell (not F1)
    for (int oz = -reach; oz <= reach; oz++)
        for (int oy = -reach; oy <= reach; oy++)
            for (int ox = -reach; ox <= reach; ox++) {
                orc_hash3((uint32_t)wrap_worley(ix + ox, cells), (uint32_t)wrap_worley(iy + oy, cells),
                          (uint32_t)wrap_worley(iz + oz, cells), seed, h);
                float dx = (qx - (float)(ix + ox)) - u01(h[0]);
                float dy = (qy - (float)(iy + oy)) - u01(h[1]);
                float dz = (qz - (float)(iz + oz)) - u01(h[2]);
                float d2 = ((dx * dx) + (dy * dy)) + (dz * dz);
                best = d2 < best ? d2 : best;
            }
    float d = sqrtf(best);
    return 1.0f - (d < 1.0f ? d : 1.0f);
}

// Periodic improved gradient noise, roughly in [-1, 1].
extern "C" float orc_perlin_tiled(float px, float py, float pz, int period, uint32_t seed) {
    float qx = px * (float)period, qy = py * (float)period, qz = pz * (float)period;
    float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
    float rx = qx - fx, ry = qy - fy, rz = qz - fz;
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    float n[8];
    uint32_t h[3];
    for (int c = 0; c < 8; c++) {
        int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
        orc_hash3((uint32_t)wrapi(ix + dx, period), (uint32_t)wrapi(iy + dy, period),
                  (uint32_t)wrapi(iz + dz, period), seed, h);
        const float* g = kGrad[(g_mutation == 5 ? h[1] : h[0]) % 12u];
        float ax = g_mutation == 7 ? rx + (float)dx : rx - (float)dx;  // mutation 7: corner offset sign
        float ay = ry - (float)dy, az = rz - (float)dz;
        n[c] = ((g[0] * ax) + (g[1] * ay)) + (g[2] * az);
    }
    float ux = fade(rx), uy = fade(ry), uz = fade(rz);
    float x00 = lerpf(n[0], n[1], ux), x10 = lerpf(n[2], n[3], ux);
    float x01 = lerpf(n[4], n[5], ux), x11 = lerpf(n[6], n[7], ux);
    float y0 = lerpf(x00, x10, uy), y1 = lerpf(x01, x11, uy);
    return lerpf(y0, y1, uz);
}

extern "C" uint8_t orc_quantize_unorm8(float v) {
    v = clamp01(v);
    return (uint8_t)(int)floorf(v * 255.0f + 0.5f);
}

// Low-res texel: R = Perlin-Worley, G/B/A = Worley FBM at base 8, 16, 32 cells.
extern "C" void orc_low_res_texel(int dim, int x, int y, int z, uint32_t seed, float out[4]) {
    const float inv = 1.0f / (float)(2 * dim);
    float px = (float)(2 * x + 1) * inv, py = (float)(2 * y + 1) * inv, pz = (float)(2 * z + 1) * inv;
    float perlin = orc_perlin_fbm(px, py, pz, 4, layer_seed(seed, 1));
    float p01 = clamp01(0.5f + 1.5f * perlin);
    float w = orc_worley_fbm(px, py, pz, 4, layer_seed(seed, 2));
    // Perlin-Worley: remap(p01, w - 1, 1, 0, 1) (Schneider's dilation of Perlin by Worley)
    float pw = (p01 - (w - 1.0f)) / (1.0f - (w - 1.0f));
    out[0] = clamp01(0.8f * pw - 0.04f);
    out[1] = clamp01(0.26f + 0.9f * orc_worley_fbm(px, py, pz, 8, layer_seed(seed, 3)));
    out[2] = clamp01(0.26f + 0.9f * orc_worley_fbm(px, py, pz, 16, layer_seed(seed, 4)));
    out[3] = clamp01(0.26f + 0.9f * orc_worley_fbm(px, py, pz, 32, layer_seed(seed, 5)));
}

// Hi-res texel: R/G/B = Worley FBM at base 4, 8, 16 cells; A = 0 like the reference asset.
extern "C" void orc_hi_res_texel(int dim, int x, int y, int z, uint32_t seed, float out[4]) {
    const float inv = 1.0f / (float)(2 * dim);
    float px = (float)(2 * x + 1) * inv, py = (float)(2 * y + 1) * inv, pz = (float)(2 * z + 1) * inv;
    out[0] = clamp01(0.26f + 0.9f * orc_worley_fbm(px, py, pz, 4, layer_seed(seed, 6)));
    out[1] = clamp01(0.26f + 0.9f * orc_worley_fbm(px, py, pz, 8, layer_seed(seed, 7)));
    out[2] = clamp01(0.26f + 0.9f * orc_worley_fbm(px, py, pz, 16, layer_seed(seed, 8)));
    out[3] = 0.0f;
}

// Weather texel: R = coverage value-noise FBM shifted by (coverage - 0.5), G = a field correlated
// with R, B = cloud type. A = 1.
extern "C" void orc_weather_texel(int dim, int x, int y, uint32_t seed, float coverage, float out[4]) {
    const float inv = 1.0f / (float)(2 * dim);
    float px = (float)(2 * x + 1) * inv, py = (float)(2 * y + 1) * inv;
    float a = orc_value_fbm(px, py, 4, layer_seed(seed, 11));
    float b = orc_value_fbm(px, py, 8, layer_seed(seed, 12));
    float c = orc_value_fbm(px, py, 8, layer_seed(seed, 13));
    float r = 0.5f + 2.0f * (a - 0.5f);
    out[0] = clamp01(r + (coverage - 0.5f));
    out[1] = clamp01(0.5f + 1.5f * ((0.5f * a + 0.5f * c) - 0.5f));
    out[2] = clamp01(0.5f + 2.5f * (b - 0.5f));
    out[3] = 1.0f;
}

extern "C" void orc_generate_low_res(int dim, uint32_t seed, uint8_t* rgba) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = 0; z < dim; z++)
        for (int y = 0; y < dim; y++)
            for (int x = 0; x < dim; x++) {
                float v[4];
                orc_low_res_texel(dim, x, y, z, seed, v);
                uint8_t* o = rgba + 4 * (((size_t)z * dim + y) * dim + x);
                for (int k = 0; k < 4; k++) o[k] = orc_quantize_unorm8(v[k]);
            }
}

extern "C" void orc_generate_hi_res(int dim, uint32_t seed, uint8_t* rgba) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = 0; z < dim; z++)
        for (int y = 0; y < dim; y++)
            for (int x = 0; x < dim; x++) {
                float v[4];
                orc_hi_res_texel(dim, x, y, z, seed, v);
                uint8_t* o = rgba + 4 * (((size_t)z * dim + y) * dim + x);
                for (int k = 0; k < 4; k++) o[k] = orc_quantize_unorm8(v[k]);
            }
}

extern "C" void orc_generate_weather(int dim, uint32_t seed, float coverage, uint8_t* rgba) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < dim; y++)
        for (int x = 0; x < dim; x++) {
            float v[4];
            orc_weather_texel(dim, x, y, seed, coverage, v);
            uint8_t* o = rgba + 4 * ((size_t)y * dim + x);
            for (int k = 0; k < 4; k++) o[k] = orc_quantize_unorm8(v[k]);
        }
}

extern "C" void orc_generate_low_res_f32(int dim, uint32_t seed, float* rgba) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = 0; z < dim; z++)
        for (int y = 0; y < dim; y++)
            for (int x = 0; x < dim; x++)
                orc_low_res_texel(dim, x, y, z, seed, rgba + 4 * (((size_t)z * dim + y) * dim + x));
}

extern "C" void orc_generate_hi_res_f32(int dim, uint32_t seed, float* rgba) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = 0; z < dim; z++)
        for (int y = 0; y < dim; y++)
            for (int x = 0; x < dim; x++)
                orc_hi_res_texel(dim, x, y, z, seed, rgba + 4 * (((size_t)z * dim + y) * dim + x));
}
