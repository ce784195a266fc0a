/*
 * CPU oracle of the volumetric-cloud path. TEST INFRASTRUCTURE ONLY: the tests, the smoke check
 * and the benchmark's CPU-baseline leg load it as the checker; the product library never links it.
 *
 * It is a scalar restatement of the reference written from its sources:
 *   - orc_curl_*      ImageUtils.cpp:7-223 (GenerateCurlNoise and its Perlin/FBM/curl helpers)
 *   - orc_noise_*     the build-defined Perlin-Worley / Worley / weather generators (the reference
 *                     ships baked assets, README.md:68, so these follow the build's own spec)
 *   - orc_march_*     compute-clouds.comp:65-500 (the cloud march), with the sampler semantics of
 *                     Texture.cpp:29-52, 315-338 (LINEAR filter, REPEAT wrap, RGBA8 UNORM)
 *   - orc_reproject   reproject.comp:56-153 (temporal reprojection of the previous frame)
 *   - orc_cloud_shadow model.frag:52-129, 240-283 (the mesh pass's cloud-shadow march)
 *   - orc_post_*      god-ray.frag:41-76, radialBlur.frag:36-63, tonemap.frag:11-34 (post chain)
 * Built with -ffp-contract=off and no fast-math so every float operation rounds as written.
 */
#ifndef MM_ORACLE_H
#define MM_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/mm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- curl generator (ImageUtils.cpp) ---- */
/* sin used by hashNoise: 0 = correctly rounded float sin (double evaluation rounded once),
 * 1 = the C library's sinf. */
float orc_hash_noise(float x, float y, float z, int sin_kind);
float orc_perlin(float x, float y, float z, float freq, int sin_kind);
float orc_fbm(float x, float y, float z, float freq, int octaves, int sin_kind);
void orc_curl_fbm(float px, float py, float freq, int octaves, int sin_kind, float out_xyz[3]);
/* GenerateCurlNoise: dim x dim RGBA8 (A = 255), rows top-down. raw_xyz (optional) receives the
 * un-normalised curl vectors, minmax (optional) the 6 bounds (xmin, xmax, ymin, ymax, zmin, zmax). */
void orc_curl_generate(int dim, int sin_kind, uint8_t* rgba, float* raw_xyz, float* minmax);
double orc_sin_poly(double x);   /* the reduction + polynomial sin the GPU generator also uses */
void orc_set_curl_mutation(int m); /* planted mistakes for fixture-sensitivity tests; 0 = none */

/* ---- build-defined volume / weather generators (RGBA8, linear layout) ---- */
uint32_t orc_hash3(uint32_t x, uint32_t y, uint32_t z, uint32_t seed, uint32_t out[3]);
float orc_worley(float px, float py, float pz, int cells, uint32_t seed);
float orc_perlin_tiled(float px, float py, float pz, int period, uint32_t seed);
float orc_worley_fbm(float px, float py, float pz, int cells, uint32_t seed);   /* cells, 2 cells, 4 cells */
float orc_perlin_fbm(float px, float py, float pz, int period, uint32_t seed);  /* 4 octaves */
float orc_value2(float px, float py, int period, uint32_t seed);               /* 2D value noise */
float orc_value_fbm(float px, float py, int period, uint32_t seed);            /* 5 octaves */
void orc_set_noise_mutation(int m); /* planted mistakes for the generator tests; 0 = none */
void orc_low_res_texel(int dim, int x, int y, int z, uint32_t seed, float out_rgba[4]);
void orc_hi_res_texel(int dim, int x, int y, int z, uint32_t seed, float out_rgba[4]);
void orc_weather_texel(int dim, int x, int y, uint32_t seed, float coverage, float out_rgba[4]);
void orc_generate_low_res(int dim, uint32_t seed, uint8_t* rgba);       /* dim^3 * 4 bytes */
void orc_generate_hi_res(int dim, uint32_t seed, uint8_t* rgba);        /* dim^3 * 4 bytes */
void orc_generate_weather(int dim, uint32_t seed, float coverage, uint8_t* rgba); /* dim^2 * 4 */
uint8_t orc_quantize_unorm8(float v);
/* Unquantised float RGBA of the same generators (dim^3 * 4 or dim^2 * 4 floats). */
void orc_generate_low_res_f32(int dim, uint32_t seed, float* rgba);
void orc_generate_hi_res_f32(int dim, uint32_t seed, float* rgba);

/* ---- march ---- */
typedef struct orc_textures {
    const uint8_t* weather; int weather_w, weather_h;     /* RGBA8 linear */
    const uint8_t* curl; int curl_w, curl_h;               /* RGBA8 linear */
    const uint8_t* low; int low_dim;                       /* RGBA8 linear, dim^3 */
    const uint8_t* hi; int hi_dim;                         /* RGBA8 linear, dim^3 */
    const uint8_t* night; int night_w, night_h;            /* RGBA8 linear, may be NULL */
} orc_textures;

/* Sampler restatement. Coordinates are normalised; returns RGBA in [0,1]. */
void orc_sample2d(const uint8_t* tex, int w, int h, float u, float v, int subtexel_bits, float out[4]);
void orc_sample3d(const uint8_t* tex, int n, float u, float v, float w, int subtexel_bits, float out[4]);

/* March pixels [y0, y1) of the frame (all x) for the given phase (-1 = all).
 * Writes RGBA into out (width*height*4) for marched pixels only. counts (optional, per pixel
 * int32[2]: density samples, detail samples). threads: OpenMP threads (0 = default). */
int orc_march_f32(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                  const orc_textures* tex, const mm_march_params* prm, int width, int height,
                  int phase, int y0, int y1, float* out, int32_t* counts, int threads);
int orc_march_f64(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                  const orc_textures* tex, const mm_march_params* prm, int width, int height,
                  int phase, int y0, int y1, double* out, int32_t* counts, int threads);
/* March an explicit list of pixels (x, y pairs); out has 4 values per listed pixel. */
int orc_march_pixels_f32(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                         const orc_textures* tex, const mm_march_params* prm, int width, int height,
                         const int32_t* xy, int n, float* out, int threads);
int orc_march_pixels_f64(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                         const orc_textures* tex, const mm_march_params* prm, int width, int height,
                         const int32_t* xy, int n, double* out, int threads);
/* Every pixel (or the listed pixels) with its two transmittance forms: state[2] per pixel = the
 * opacity accumulator after the horizon fade and 0.999 cap (comp:484-486; 1 - state[0] is the
 * opacity complement) and the light-energy accumulator `transmittance` (comp:464). */
int orc_march_state_f32(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                        const orc_textures* tex, const mm_march_params* prm, int width, int height, float* out,
                        float* state, int threads);
int orc_march_state_f64(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                        const orc_textures* tex, const mm_march_params* prm, int width, int height, double* out,
                        double* state, int threads);
int orc_march_pixels_state_f64(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                               const orc_textures* tex, const mm_march_params* prm, int width, int height,
                               const int32_t* xy, int n, double* out, double* state, int threads);

int orc_max_threads(void);

/* One pixel's march loop, trip by trip: trips[max_trips][4] = (t, step size, low-res density, event),
 * event 0 coarse miss, 1 first hit (back-step, fine step, uncounted continue), 2 fine trip whose
 * detail density is below 1e-4 (uncounted continue), 3 fine hit (light cone, accumulation), 4 fine
 * miss, 5 the miss that reverts to coarse steps. counts = (density, detail samples); out = RGBA.
 * Returns the number of trips (entries past max_trips are not written) or a negative MM_E_*. */
int orc_march_trace_f32(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                        const orc_textures* tex, const mm_march_params* prm, int width, int height, int px, int py,
                        float* trips, int max_trips, int32_t* counts, float* out);

/* ---- temporal reprojection (reproject.comp) ---- */
/* Target rows [y0, y1) from the previous frame src (both RGBA32F, pitch pixels per row). prm gives
 * the inner sphere (inner_diameter, center_y) and the FP64 flag; it must not be NULL. */
int orc_reproject(const mm_camera_ubo* cam, const mm_camera_ubo* prev, const float* src, float* dst,
                  int width, int height, int pitch, const mm_march_params* prm, int y0, int y1);

/* ---- cloud shadows (model.frag) ---- */
/* RGBA8 linear weather (ww x wh) and low-res (ln^3); points [n][3]; out [n]. */
float orc_shadow_layer_density(float h, float type);   /* model.frag:88-98 */
int orc_cloud_shadow(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky, const uint8_t* weather,
                     int ww, int wh, const uint8_t* low, int ln, const float* points, int n, float* out,
                     const mm_shadow_params* p);

/* ---- post chain (god-ray.frag, radialBlur.frag, tonemap.frag) ---- */
void orc_post_sun_ndc(const mm_camera_ubo* cam, const mm_sun_ubo* sun, float out_xy[2]); /* (proj*view)*location / w */
float orc_post_uc2(float x);                                                         /* tonemap.frag:11-14 */
float orc_post_pow(float x, float y);                                 /* exp2(y log2 x) as float series */
int orc_post_process(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const float* image, int width, int height,
                     int pitch, float* god_alpha, float* hdr, uint8_t* rgba8, const mm_post_params* p);

#ifdef __cplusplus
}
#endif

#endif
