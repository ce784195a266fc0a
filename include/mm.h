/*
 * mm.h - C ABI of the MI355X volumetric-cloud path (noise generation + cloud raymarch).
 *
 * Drop-in for the reference's compute dispatch of compute-clouds.comp:
 *   - the operator it replaces is ComputeShader (SkyEngine/SkyEngine/Shader.h:286-377), whose
 *     pipeline layout is {storage image, storage image, UBOs + samplers}
 *     (SkyEngine/SkyEngine/Shader.cpp:662-693, 911-954);
 *   - the uniform blocks below are byte-identical to the std140 blocks the shader declares
 *     (SkyEngine/SkyEngine/Shaders/compute-clouds.comp:12-41) and the host structs that fill them
 *     (SkyEngine/SkyEngine/Shader.h:24-28, SkyEngine/SkyEngine/SkyManager.h:8-14, 28-36);
 *   - the output is the RGBA32F storage image the shader writes with imageStore
 *     (compute-clouds.comp:9, 499): row-major, 4 floats per pixel, row 0 = top of the view.
 *
 * Every entry point returns MM_OK (0) or a negative MM_E_* code and never throws.
 * Device pointers are caller-owned; work is enqueued on the given stream (a hipStream_t passed as
 * void*, NULL = the null stream) and is asynchronous: the caller synchronises.
 */
#ifndef MM_H
#define MM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MM_ABI_VERSION 1

/* ---------------------------------------------------------------- errors */
enum {
    MM_OK = 0,
    MM_E_NULL = -1,      /* a required pointer is NULL */
    MM_E_DIM = -2,       /* zero/negative/unsupported image or texture dimensions */
    MM_E_PHASE = -3,     /* phase outside [-2, 15] or sun.color.a outside [0, 16) */
    MM_E_PITCH = -4,     /* row pitch smaller than the width */
    MM_E_FORMAT = -5,    /* unknown or unsupported texture format / layout */
    MM_E_PARAM = -6,     /* non-finite uniforms, zero-length sun direction, bad params */
    MM_E_DEVICE = -7,    /* HIP runtime error or no device */
    MM_E_SIZE = -8,      /* a texture allocation is smaller than its format requires */
    MM_E_VERSION = -9,   /* params.size does not match this library's struct */
    MM_E_IO = -10        /* a file could not be opened or read */
};
const char* mm_strerror(int code);
int mm_abi_version(void);

/* ---------------------------------------------------------------- uniform blocks (std140) */
/* UniformCameraObject: Shader.h:24-28, compute-clouds.comp:12-17. 160 bytes. The previous frame's
 * camera travels in the same host struct (VulkanApplication.cpp:354-358); the shaders' block
 * UniformCameraObjectPrev (compute-clouds.comp:19-23, reproject.comp:19-23) is its first 144 bytes.
 * view and proj are column-major 4x4 (glm/GLSL order: m[col*4 + row]).
 * cameraParams = (aspect, tan(0.5 * DEG2RAD * fov), 0, 0) (VulkanApplication.cpp:365-366). */
typedef struct mm_camera_ubo {
    float view[16];
    float proj[16];
    float cameraPosition[4];
    float cameraParams[4];
} mm_camera_ubo;

/* UniformSunObject: SkyManager.h:8-14, compute-clouds.comp:26-32. 116 bytes used, padded to 128.
 * color[3] carries the 1/16 pixel phase (VulkanApplication.cpp:381-384).
 * directionBasis columns: 0 = N, 1 = sun direction, 2 = B (SkyManager.cpp:60-69). */
typedef struct mm_sun_ubo {
    float location[4];
    float direction[4];
    float color[4];
    float directionBasis[16];
    float intensity;
    float _pad[3];
} mm_sun_ubo;

/* UniformSkyObject: SkyManager.h:28-36, compute-clouds.comp:35-41. 52 bytes used, padded to 64.
 * wind.xyz = wind direction, wind.w = time. */
typedef struct mm_sky_ubo {
    float betaR[4];
    float betaV[4];
    float wind[4];
    float mie_directional;
    float _pad[3];
} mm_sky_ubo;

/* ---------------------------------------------------------------- textures */
/* Texel formats. The reference binds every texture as RGBA8 UNORM (Texture.h:29, 85).
 * Channel use by the march: weather R = coverage, B = cloud type (comp:237, 245; G unused);
 * curl .xyz (comp:218); low-res R = base shape, GBA = erosion FBMs (comp:238-247);
 * hi-res RGB = erosion FBMs (comp:223-224). */
enum {
    MM_FMT_RGBA8 = 1,       /* 4 x UNORM8, 4 B/texel: the reference format, every texture */
    MM_FMT_RG8 = 2,         /* 2 x UNORM8, 2 B/texel: weather only, (R = coverage, G = type) */
    /* 3 and 4 are not assigned */
    MM_FMT_FP8X4 = 5,       /* low-res only, 4 x FP8 e4m3 with per-channel FP32 scale (mm_texture.scale) */
    MM_FMT_FP16X4 = 6       /* 4 x FP16, 8 B/texel: curl (x,y,z,0) or hi-res (r,g,b,0) */
};

/* Texture sets the march accepts (any other combination returns MM_E_FORMAT):
 *   REF : every texture MM_FMT_RGBA8, MM_LAYOUT_LINEAR (the reference's bindings)
 *   FAST: MM_LAYOUT_APRON; weather MM_FMT_RG8, curl / low-res / hi-res MM_FMT_RGBA8. Same arithmetic
 *         as REF (same image), fewer weather bytes and no index wrapping in the filters
 *   PREC: MM_LAYOUT_APRON; weather MM_FMT_RG8, curl MM_FMT_FP16X4, low-res MM_FMT_FP8X4,
 *         hi-res MM_FMT_FP16X4 (the precision variant) */

/* Memory layouts. */
enum {
    MM_LAYOUT_LINEAR = 0,   /* reference: z = slice, rows top-down, x fastest (Texture.cpp:502-538) */
    MM_LAYOUT_APRON = 1     /* (N+1) per axis, index N repeats index 0 so every filter footprint is
                               contiguous in x: REPEAT wrap without index wrapping in the kernel */
};

typedef struct mm_texture {
    void* data;             /* device pointer */
    size_t bytes;           /* bytes available at data (validated against format and dims) */
    int32_t width, height, depth;  /* depth = 1 for 2D textures */
    int32_t format;         /* MM_FMT_* */
    int32_t layout;         /* MM_LAYOUT_* */
    int32_t _pad;
    float scale[4];         /* per-channel dequantisation scale (FP8 formats); ignored otherwise */
} mm_texture;

/* The four textures compute-clouds.comp samples (set 2, bindings 4, 6, 7, 8: Shader.cpp:771-908). */
typedef struct mm_noise_set {
    mm_texture weather;     /* binding 4 cloudPlacement, 2D */
    mm_texture curl;        /* binding 6 curlNoise, 2D */
    mm_texture low_res;     /* binding 7 lowResCloudShape, 3D */
    mm_texture hi_res;      /* binding 8 hiResCloudShape, 3D */
} mm_noise_set;

/* Bytes a texture of the given shape, format and layout occupies (0 if unsupported). */
size_t mm_texture_bytes(int32_t width, int32_t height, int32_t depth, int32_t format, int32_t layout);

/* ---------------------------------------------------------------- noise generation */
enum {
    MM_GEN_WEATHER = 1, MM_GEN_CURL = 2, MM_GEN_LOW_RES = 4, MM_GEN_HI_RES = 8, MM_GEN_ALL = 15
};

typedef struct mm_noise_params {
    uint32_t size;          /* sizeof(mm_noise_params) */
    uint32_t what;          /* MM_GEN_* mask */
    uint32_t seed;          /* seeds the low-res and hi-res volumes */
    uint32_t weather_seed;  /* seeds the weather map */
    float coverage;         /* 0..1: shifts the weather coverage (R) field; 0.5 = unbiased */
    int32_t _reserved[7];
} mm_noise_params;

void mm_noise_params_default(mm_noise_params* p);

/* Generate the textures selected by p->what into out (dims, format and layout are read from
 * out->*; data must be allocated by the caller with at least mm_texture_bytes()).
 * workspace: device scratch of at least mm_noise_workspace_bytes() bytes (used by the curl
 * generator's gradient table, float curl field and global min/max), or NULL to let the library
 * allocate it stream-ordered for this call. With the low-res volume and any other texture selected,
 * the other generators run on a second stream the library keeps per device, forked from `stream`
 * and joined back into it with events: the call is ordered on `stream` like a single launch (and
 * can be captured into a HIP graph).
 * The curl map restates ImageUtils.cpp:176-223 (GenerateCurlNoise) and does not depend on the seed.
 * Replaces: the baked TGA/PNG assets loaded by VulkanApplication.cpp:253-262 and the offline
 * curl tool GenerateCurlNoise (ImageUtils.h:12). */
size_t mm_noise_workspace_bytes(const mm_noise_set* out);
int mm_generate_noise(const mm_noise_params* p, const mm_noise_set* out, void* workspace, void* stream);

/* Allocation-free variant of mm_generate_noise. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The curl chain's scratch lives in a fixed device array inside the library
 * (MM_STATIC_SCRATCH_BYTES, one copy per GPU, held until the process ends); the call makes no
 * allocation and creates no stream: every generator runs on `stream` in order.
 * Limit: curl maps up to MM_STATIC_CURL_MAX x MM_STATIC_CURL_MAX texels (larger ones return
 * MM_E_SIZE); other textures as mm_generate_noise. All calls on one GPU share that scratch array and
 * it keeps the last call's contents, so calls that overlap on one GPU (other streams, other threads)
 * race: serialise them. Results are identical to mm_generate_noise's. */
#define MM_STATIC_CURL_MAX 256
#define MM_STATIC_SCRATCH_BYTES (20480 + 256 + MM_STATIC_CURL_MAX * MM_STATIC_CURL_MAX * 3 * 4)
int mm_generate_noise_noalloc(const mm_noise_params* p, const mm_noise_set* out, void* stream);

/* Repack a linear RGBA8 texture (the reference's asset layout) into another format/layout.
 * kind selects the channel mapping: 0 = weather, 1 = curl, 2 = low-res, 3 = hi-res. */
int mm_pack_texture(int32_t kind, const mm_texture* src, const mm_texture* dst, void* stream);

/* ---------------------------------------------------------------- march */
enum {
    MM_MARCH_ISECT_FP64 = 1u << 0,   /* ray/shell intersection in FP64 (same formula as comp:147-177) */
    MM_MARCH_BOUNDED = 1u << 1,      /* every fine-mode sample counts toward max_fine_steps (the
                                        reference's `continue` at comp:437 does not count) */
    MM_MARCH_NO_BATCH = 1u << 2      /* evaluate light cones in place (no wave-level cone batching) */
};

enum { MM_CONE_PLAIN = 0, MM_CONE_SURROGATE = 1 };

typedef struct mm_march_params {
    uint32_t size;              /* sizeof(mm_march_params) */
    uint32_t flags;             /* MM_MARCH_* */
    /* shell: spheres are given by their diameter like the reference's vec4(center, w) (comp:359-361) */
    float inner_diameter;       /* ATMOSPHERE_RADIUS = 2e6 */
    float outer_diameter;       /* ATMOSPHERE_RADIUS * 1.02 */
    float center_y;             /* earthCenter.y = -ATMOSPHERE_RADIUS * 0.5 * 0.995 (comp:358) */
    float thickness;            /* 0.5 * ATMOSPHERE_RADIUS * 0.02 (comp:360) */
    float proj_radius;          /* 0.5 * ATMOSPHERE_RADIUS (getProjectedShellPoint, comp:181) */
    /* stepping (comp:390, 426-433, 468-481) */
    float coarse_step_frac;     /* 0.05: coarse step = frac * thickness */
    float fine_step_mul;        /* 0.3 */
    int32_t miss_limit;         /* 10 */
    float opacity_cutoff;       /* 0.99 */
    int32_t max_steps;          /* 100: break when ++steps > max_steps (counted steps); < 0 = off */
    int32_t max_coarse_steps;   /* counted coarse steps limit, same `++n > limit` form; < 0 = off */
    int32_t max_fine_steps;     /* fine steps limit; < 0 = off */
    /* sampling (comp:218, 223, 236, 238, 279, 414) */
    float wind_strength;        /* 20 */
    float weather_scale;        /* 0.000009 */
    float low_scale;            /* 0.00002 */
    float curl_scale;           /* 0.0001 */
    float hi_scale;             /* 0.0004 */
    int32_t subtexel_bits;      /* sampler weight precision: 8 = fixed-point like the hardware, 0 = float */
    /* light cone */
    int32_t cone_mode;          /* MM_CONE_* */
    float surrogate_threshold;  /* jobs whose weighted error bound is >= this take the plain cone */
    const void* surrogate;      /* device model block of mm_surrogate_bytes() (MM_CONE_SURROGATE);
                                   REF and FAST texture sets only */
    int32_t tile_width;         /* pixels per wave tile row: 8 (8x8, default when 0), 16 (16x4), 32 (32x2);
                                   the image does not depend on it */
    int32_t lanes_per_ray;      /* 1, 2, 4 or 8 lanes march one ray, sampling the next trips ahead
                                   and committing them in order (0 = default: 8 for one phase, else
                                   2); plain cones only (other modes use 1); the image does not
                                   depend on it */
    int32_t _reserved[6];
} mm_march_params;

/* Preset with the reference's constants (compute-clouds.comp, bit-for-bit). */
void mm_march_params_reference(mm_march_params* p);
/* Preset of the scenario configs: Earth radius 6.371e6 m, shell 1500-4000 m, <=128 coarse and
 * <=256 fine steps, FP64 intersection. */
void mm_march_params_scenario(mm_march_params* p);

/* Hard bound on march loop trips per ray (counted or not). The reference's uncounted `continue`
 * paths (comp:433, 437) are only bounded by t reaching the outer shell; this cap keeps every ray
 * finite for degenerate parameters. Never reached by the presets. */
#define MM_MAX_MARCH_ITERATIONS 16384

#define MM_PHASE_ALL (-1)       /* march every pixel */
#define MM_PHASE_FROM_UBO (-2)  /* phase = int(sun.color.a) like the reference (comp:292) */

/* One cloud dispatch. Replaces ComputeShader::bindShader + vkCmdDispatch of compute-clouds.comp
 * (Shader.h:358-376, VulkanApplication.cpp:1061-1071; the shader body is comp:288-500).
 * out_rgba32f: device, row-major with row_pitch_px pixels per row, 4 floats per pixel.
 * phase 0..15 marches pixels (4i + phase%4, 4j + phase/4); other pixels are left untouched.
 * night_sky: binding 5 (RGBA8 linear 2D), may be NULL (black night sky).
 * params: NULL = mm_march_params_reference. */
int mm_march_clouds(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                    const mm_noise_set* textures, const mm_texture* night_sky,
                    float* out_rgba32f, int32_t width, int32_t height, int32_t row_pitch_px,
                    int32_t phase, const mm_march_params* params, void* stream);

/* The march with work counters, accumulated (added) into counts: device uint64[4] =
 * {density samples (cloudTest evaluations), detail samples (cloudHiRes evaluations),
 *  light-cone jobs (fine-mode hits that light a sample, comp:440-464),
 *  light-cone jobs evaluated by the plain six-sample cone (all of them unless MM_CONE_SURROGATE)}. */
int mm_march_clouds_counted(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                            const mm_noise_set* textures, const mm_texture* night_sky,
                            float* out_rgba32f, int32_t width, int32_t height, int32_t row_pitch_px,
                            int32_t phase, const mm_march_params* params,
                            unsigned long long* counts, void* stream);

/* mm_march_clouds_counted with six counters: counts[0..3] as above, counts[4] = density samples that
 * read the weather map, counts[5] = density samples that read the low-res volume (the others end
 * before their texel reads: outside the layer heights, comp:193-204, or on a zero layer profile). */
int mm_march_clouds_counted_ex(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                               const mm_noise_set* textures, const mm_texture* night_sky,
                               float* out_rgba32f, int32_t width, int32_t height, int32_t row_pitch_px,
                               int32_t phase, const mm_march_params* params,
                               unsigned long long* counts, void* stream);

/* The march with each pixel's two transmittance forms written beside its colour, for error
 * reporting against an FP64 evaluation: state_out (device, 2 floats per pixel, indexed like
 * out_rgba32f: (y * row_pitch_px + x) * 2) receives the opacity accumulator after the horizon fade
 * and the 0.999 cap (comp:484-486; 1 - it is the opacity complement the alpha uses, comp:496) and the
 * light-energy accumulator `transmittance` (comp:464). Runs the lock-step kernel (the image is the
 * same as every other kernel shape's); plain light cones only (MM_E_PARAM otherwise).
 * No counterpart in the reference, whose shader keeps both values in registers. */
int mm_march_clouds_state(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                          const mm_noise_set* textures, const mm_texture* night_sky,
                          float* out_rgba32f, int32_t width, int32_t height, int32_t row_pitch_px,
                          int32_t phase, const mm_march_params* params, float* state_out, void* stream);

/* Tile schedule. No counterpart in the reference, whose dispatch runs its workgroups in launch
 * order: a frame's time is set as much by when its longest tiles start as by the total work (the
 * rays that graze the shell near the horizon or cross deep cloud take many times the median's
 * trips). A schedule records each tile's cost (the most loop trips any of its rays took) in one
 * launch; the next launch of the same shape (tile count, tiles per row, marched grid, kernel shape)
 * first orders its tiles longest first from those costs (a counting sort over log-scale cost
 * buckets, three small launches on the stream) and then marches in that order, recording the new
 * costs. Costs are indexed by the tile's
 * place in the image, so they stay valid as the camera moves; a launch of another shape marches in
 * the default order and starts a new record. The image does not depend on the order, bit for bit:
 * the schedule only changes which tiles start first. The lock-step kernel (MM_MARCH_NO_BATCH)
 * ignores it.
 * mm_march_schedule_create allocates device buffers for up to max_tiles tiles (a full 1920x1080
 * frame has 32400 at the scheduled default of one lane per ray in 16 x 4 tiles, 64800 at two lanes
 * per ray); launches with more tiles return MM_E_SIZE, and
 * a schedule is used on the device that created it. Launches that share a schedule must be ordered
 * (one stream): the order buffer is rewritten by each. */
typedef struct mm_march_schedule mm_march_schedule;
int mm_march_schedule_create(int32_t max_tiles, mm_march_schedule** schedule);
void mm_march_schedule_destroy(mm_march_schedule* schedule);
/* mm_march_clouds with a schedule; counts: NULL or device uint64[4] as for mm_march_clouds_counted. */
int mm_march_clouds_scheduled(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                              const mm_noise_set* textures, const mm_texture* night_sky,
                              float* out_rgba32f, int32_t width, int32_t height, int32_t row_pitch_px,
                              int32_t phase, const mm_march_params* params, mm_march_schedule* schedule,
                              unsigned long long* counts, void* stream);

/* ---------------------------------------------------------------- temporal reprojection
 * Replaces the dispatch of reproject.comp (VulkanApplication.cpp:1051-1057; the shader body is
 * reproject.comp:91-153): every target pixel is fetched from the previous frame `source` along the
 * pixel's motion vector between camera_prev and camera (10 nearest taps averaged, alpha from the
 * centre tap), the motion being that of the ray's hit on the inner shell sphere. params gives the
 * sphere (inner_diameter, center_y, MM_MARCH_ISECT_FP64); NULL = the reference preset, whose sphere
 * is reproject.comp's ATMOSPHERE_RADIUS 2e6. source and target: RGBA32F, width x height,
 * row_pitch_px pixels per row; they must not overlap. GLSL latitude fixed here: round() is half
 * away from zero; a NaN coordinate reads texel 0. */
int mm_reproject(const mm_camera_ubo* camera, const mm_camera_ubo* camera_prev, const float* source,
                 float* target, int32_t width, int32_t height, int32_t row_pitch_px, const mm_march_params* params,
                 void* stream);

/* One frame of the reference's compute command buffer (VulkanApplication.cpp:1046-1071):
 * mm_reproject(source -> target) then the cloud march of the pixel phase int(sun.color.a) into
 * target. The caller swaps source and target for the next frame. The reprojection leaves out the
 * phase's pixels (the march writes all of them; the image is the same) and runs beside the march on
 * a second stream the library keeps per device, forked from and joined back into `stream`. */
int mm_render_frame(const mm_camera_ubo* camera, const mm_camera_ubo* camera_prev, const mm_sun_ubo* sun,
                    const mm_sky_ubo* sky, const mm_noise_set* textures, const mm_texture* night_sky,
                    const float* source, float* target, int32_t width, int32_t height, int32_t row_pitch_px,
                    const mm_march_params* params, void* stream);

/* ---------------------------------------------------------------- cloud shadows on surfaces
 * Replaces the shadow march inside the mesh pass (model.frag:240-283 with its own cloud functions,
 * model.frag:52-129): for each world-space point, NUM_SHADOW_STEPS low-res density samples toward
 * the sun, keeping the maximum (the mesh colour is multiplied by 1 - 2 * result, model.frag:282).
 * Textures: the march's weather and low-res volume (REF or FAST formats). points: device float
 * [n][3] world positions (fragPositionWC); out: device float [n]. */
typedef struct mm_shadow_params {
    uint32_t size;              /* sizeof(mm_shadow_params) */
    uint32_t flags;             /* MM_MARCH_ISECT_FP64 */
    float inner_diameter;       /* ATMOSPHERE_RADIUS = 1e6 (model.frag:61) */
    float center_y;             /* -ATMOSPHERE_RADIUS * 0.5 * 0.99 (model.frag:247) */
    float thickness;            /* 0.5 * ATMOSPHERE_RADIUS * 0.02 (model.frag:249) */
    float proj_radius;          /* 0.5 * ATMOSPHERE_RADIUS (model.frag:75) */
    float step;                 /* 0.1 * thickness (model.frag:256) */
    int32_t steps;              /* NUM_SHADOW_STEPS = 6 */
    float origin_scale;         /* shadowRayOrigin = fragPositionWC * 4 (model.frag:260) */
    float wind_strength;        /* 20 */
    float weather_scale;        /* 0.00001 (model.frag:110) */
    float low_scale;            /* 0.000057 (model.frag:113) */
    int32_t subtexel_bits;      /* 8 = fixed-point sampler weights, 0 = float */
    int32_t _reserved[7];
} mm_shadow_params;

void mm_shadow_params_reference(mm_shadow_params* p);
int mm_cloud_shadow(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                    const mm_texture* weather, const mm_texture* low_res, const float* points, int32_t n,
                    float* out, const mm_shadow_params* params, void* stream);

/* ---------------------------------------------------------------- post chain
 * Replaces the offscreen post passes that consume the cloud image (VulkanApplication.cpp:938-966:
 * background copy -> god rays -> radial blur -> [meshes, not part of this path] -> tonemap; every
 * framebuffer RGBA32F, sampled LINEAR / CLAMP_TO_EDGE, VulkanApplication.cpp:1290-1308):
 *   god rays    (god-ray.frag:41-76)    alpha = 0.5 a + 8 decayed alpha taps toward the sun
 *   radial blur (radialBlur.frag:36-63) rgb = sun.color * intensity * mean(10 alpha taps * 1.1) + 0.5 rgb
 *   tonemap     (tonemap.frag:11-34)    Uncharted-2 curve, white point, gamma 1/2.2, vignette
 * image: device RGBA32F cloud image (the march or reprojection output), row_pitch_px pixels per row
 * (the same pitch for every image here). god_alpha: device float scratch of two planes of
 * row_pitch_px * height floats (the image's alpha, then framebuffer 1's alpha, which the god-ray
 * pass produces). out_hdr_rgba32f (optional): the radial-blur output (framebuffer 2 before meshes).
 * out_rgba8 (optional): the tonemapped frame as UNORM8 R, G, B, A = 255 (the swapchain is
 * B8G8R8A8_UNORM, VulkanApplication.cpp:1439; the channel order here is RGBA). */
typedef struct mm_post_params {
    uint32_t size;              /* sizeof(mm_post_params) */
    uint32_t flags;             /* none defined: 0 */
    int32_t subtexel_bits;      /* sampler sub-texel precision, 1..8 (8 = the hardware's) */
    int32_t god_ray_samples;    /* NUM_SAMPLES = 8 (god-ray.frag:31) */
    float god_ray_density;      /* DENSITY = 0.75 */
    float god_ray_decay;        /* DECAY = 0.99 */
    float god_ray_exposure;     /* EXPOSURE = 0.9 */
    float blur_scale;           /* 1.5 (radialBlur.frag:59) */
    float blur_gain;            /* 1.1 */
    float exposure;             /* 0.7 (tonemap.frag:30) */
    float inv_gamma;            /* 1 / 2.2 */
    float white_point;          /* 50.2 */
    float vignette[3];          /* (0.1, 0.05, 0.13) */
    int32_t _reserved[7];
} mm_post_params;

void mm_post_params_reference(mm_post_params* p);
int mm_post_process(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const float* image_rgba32f,
                    int32_t width, int32_t height, int32_t row_pitch_px, float* god_alpha,
                    float* out_hdr_rgba32f, uint8_t* out_rgba8, const mm_post_params* params, void* stream);

/* ---------------------------------------------------------------- asset ingestion
 * Replaces the reference's texture loading: stb_image 2.16 with STBI_rgb_alpha for 2D images
 * (Texture.cpp:212-246) and for volumes stored as one TGA per z slice, "<prefix>(<i>).tga"
 * (Texture.cpp:502-538). Decodes on the host into caller memory as RGBA8, rows top first, slice i
 * at byte offset i * width * height * 4; upload it into an RGBA8 LINEAR texture and convert it to
 * another format set with mm_pack_texture. TGA: types 1/2/3/9/10/11, 8/15/16/24/32-bit pixels.
 * PNG: all colour types, bit depths 1-16 (16-bit samples keep their high byte, as stb's 8-bit
 * interface does), tRNS; interlaced PNGs are rejected (MM_E_FORMAT), images above 2^28 pixels
 * with MM_E_DIM, and a failed allocation returns MM_E_SIZE. */
int mm_image_info(const char* path, int32_t* width, int32_t* height);
int mm_load_image_rgba8(const char* path, uint8_t* rgba, size_t bytes, int32_t* width, int32_t* height);
int mm_load_volume_rgba8(const char* prefix, int32_t depth, uint8_t* rgba, size_t bytes, int32_t* width,
                         int32_t* height);

/* ---------------------------------------------------------------- many views in one launch
 * No counterpart in the reference (which dispatches one view per frame, VulkanApplication.cpp:
 * 1061-1071): many views of any resolution marched by one launch over the tiles of all views (one
 * wave each), handed to the CUs by the hardware dispatcher as slots free up. Views share the curl
 * and noise volumes and bring their own uniforms, weather map and output image. Each view's image is identical to mm_march_clouds of that view with the same
 * params. A plan holds the per-view constants in device memory; mm_march_views enqueues the march
 * (after the plan's first launch, preceded by three small launches that order all views' tiles
 * longest first by the previous launch's tile costs; the images do not depend on the order) with
 * no allocation, no host transfer and no memset, so it can be captured into a HIP graph and
 * replayed (a capture of the plan's first launch keeps the launch order). */
typedef struct mm_view {
    mm_camera_ubo camera;
    mm_sun_ubo sun;
    mm_sky_ubo sky;
    mm_texture weather;         /* this view's weather map (format of the shared set's weather) */
    float* out_rgba32f;         /* device, row_pitch_px pixels per row */
    int32_t width, height, row_pitch_px;
    int32_t phase;              /* 0..15, MM_PHASE_ALL or MM_PHASE_FROM_UBO */
} mm_view;

typedef struct mm_view_plan mm_view_plan;

/* shared: curl, low_res and hi_res are used (its weather entry is ignored). params applies to every
 * view (NULL = reference preset); every view is marched with params->lanes_per_ray lanes per ray
 * (0 = 1, whatever the views' phases), with one lane in 64-ray tiles params->tile_width wide
 * (0 = 16). */
int mm_view_plan_create(const mm_view* views, int32_t n_views, const mm_noise_set* shared,
                        const mm_march_params* params, mm_view_plan** plan, void* stream);
/* Replace the views' uniforms / outputs (same count) without reallocating. Synchronises the stream.
 * The launch shape must not change: an update whose views, phases, sizes or params give a different
 * tile layout (any view's tile count), a different lanes-per-ray or a different kernel (params
 * switching between the bounded-step loop -- MM_MARCH_BOUNDED, max_steps < 0, both budgets >= 0 --
 * and the general one) returns MM_E_PARAM and leaves
 * the plan as it was, so a graph that captured mm_march_views can be replayed after any accepted
 * update. A different shape needs a new plan (and a new capture). */
int mm_view_plan_update(mm_view_plan* plan, const mm_view* views, int32_t n_views, const mm_march_params* params,
                        void* stream);
/* Total tiles of the plan and the launch's grid (one wave per tile). */
int mm_view_plan_info(const mm_view_plan* plan, int32_t* total_tiles, int32_t* grid_waves);
/* counts: NULL or device uint64[4] accumulators as for mm_march_clouds_counted. */
int mm_march_views(const mm_view_plan* plan, unsigned long long* counts, void* stream);
void mm_view_plan_destroy(mm_view_plan* plan);

/* ---------------------------------------------------------------- learned light-cone surrogate
 * No counterpart in the reference: an optional replacement of the six-sample light cone
 * (comp:440-453) by a 16-64-64-64-2 MLP (hard-SiLU hidden units, z * clamp(z / 6 + 1/2, 0, 1))
 * trained with BF16 MFMA and evaluated with FP16 MFMA inside the march. For each
 * cone job it predicts the density along the light D and an error bound e on exp(-D); the job's
 * result enters the pixel's transmittance with weight (1 - accum) * inScatter * HG (comp:464), and
 * jobs whose weighted bound e * (1 - accum) * |inScatter * HG| is >= params.surrogate_threshold
 * take the plain cone, so threshold 0 is the plain march.
 * A model is one device block of mm_surrogate_bytes() bytes (FP32 parameters, Adam state, feature
 * statistics, packed BF16 training and FP16 inference fragments). Training samples are rows of
 * MM_SURROGATE_ROW floats:
 * 16 features followed by the cone's D, recorded by mm_march_clouds_collect. */
#define MM_SURROGATE_FEATURES 16
#define MM_SURROGATE_ROW 17
#define MM_SURROGATE_PARAMS 9538   /* W1[16][64], b1[64], W2[64][64], b2, W3[64][64], b3, W4[64][2], b4[2] */

size_t mm_surrogate_bytes(void);
/* Random (seeded) weights, cleared optimiser state, identity feature normalisation. */
int mm_surrogate_init(void* model, uint32_t seed, void* stream);
/* Per-feature mean and inverse standard deviation over n sample rows (device). */
int mm_surrogate_fit_stats(void* model, const float* samples, int32_t n, void* stream);
/* `steps` Adam steps of `batch` rows (a multiple of 64) drawn from the n rows with a seeded hash.
 * loss_out: NULL or device float[steps], the summed per-row loss of each step. Synchronises. */
int mm_surrogate_train(void* model, const float* samples, int32_t n, int32_t batch, int32_t steps, float lr,
                       uint32_t seed, float* loss_out, void* stream);
/* Replace the FP32 parameters (device float[MM_SURROGATE_PARAMS]) and/or the feature statistics
 * (device float[32]: mean[16], inverse std[16]) of a model, either may be NULL, and repack its
 * BF16 and FP16 fragments: loading a saved model, or setting weights for a test. */
int mm_surrogate_load(void* model, const float* params, const float* stats, void* stream);
/* Exact mean gradient of the training loss over rows 0..n-1 (no sampling, no update): grad_out =
 * device float[MM_SURROGATE_PARAMS] in the parameter layout, loss_out = device float, the summed
 * per-row loss. For gradient checks against a host implementation. */
int mm_surrogate_gradient(const void* model, const float* samples, int32_t n, float* grad_out, float* loss_out,
                          void* stream);
/* Batch inference over n rows of stride MM_SURROGATE_ROW with the march's FP16 path: out[2i] = -D
 * estimate, out[2i+1] = e. */
int mm_surrogate_predict(const void* model, const float* samples, int32_t n, float* out, void* stream);

/* The march with plain light cones that also records cone jobs as surrogate training rows: a job is
 * kept when a hash of its position and `seed` is divisible by keep_one_in (1 = every job). Rows go
 * to samples[capacity][MM_SURROGATE_ROW]; *count (device uint32, zeroed by the caller) receives the
 * number of kept jobs, which may exceed capacity (extra rows are dropped). The image is identical
 * to mm_march_clouds. REF and FAST texture sets. */
int mm_march_clouds_collect(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                            const mm_noise_set* textures, const mm_texture* night_sky,
                            float* out_rgba32f, int32_t width, int32_t height, int32_t row_pitch_px,
                            int32_t phase, const mm_march_params* params, float* samples, uint32_t capacity,
                            uint32_t keep_one_in, uint32_t seed, uint32_t* count, void* stream);

/* ---------------------------------------------------------------- uniform builder (caller side) */
/* Scene description from which the three UBOs are derived exactly as VulkanApplication.cpp:351-395
 * does through Camera (camera.cpp:27-39, camera.h:72) and SkyManager (SkyManager.cpp:15-70). */
typedef struct mm_scene {
    float camera_pos[3];        /* world position */
    float pitch;                /* Camera::m_pitch (radians); the view looks along -forward */
    float yaw;                  /* Camera::m_yaw (radians) */
    float fov_deg;              /* vertical field of view in degrees (Camera::m_fov) */
    float aspect;               /* width / height */
    float sun_elevation;        /* SkyManager elevation parameter (SkyManager.cpp:46-54) */
    float sun_azimuth;          /* SkyManager azimuth parameter */
    float time;                 /* sky.wind.w */
    float wind[3];              /* sky.wind.xyz, reference default (1, 0.05, 1) */
    float rayleigh;             /* 2 */
    float mie;                  /* 0.005 */
    float mie_directional;      /* 0.8 */
    float turbidity;            /* uninitialised in the reference; 10 here */
    int32_t phase;              /* written to sun.color.a */
} mm_scene;

void mm_scene_default(mm_scene* s);
/* Scene of the scenario configs: camera at `altitude` m, view pitched up `pitch_deg` and facing -z,
 * `fov_deg` vertical FOV, sun `sun_elevation_deg` above the horizon towards -z. */
void mm_scene_scenario(mm_scene* s, float altitude, float pitch_deg, float fov_deg, float aspect,
                       float sun_elevation_deg, float time);
int mm_build_ubos(const mm_scene* s, mm_camera_ubo* camera, mm_sun_ubo* sun, mm_sky_ubo* sky);

#ifdef __cplusplus
}
#endif

#endif /* MM_H */
