// C ABI of the cloud path (include/mm.h): argument validation, presets, hoisting of the
// uniform-derived per-dispatch constants, and dispatch to the HIP launchers. No exception crosses
// the boundary; every failure is a negative MM_E_* code.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <new>
#include <vector>
#include "mm_internal.hpp"

namespace {

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

bool finite_n(const float* v, int n) {
    for (int i = 0; i < n; i++)
        if (!isfinite(v[i])) return false;
    return true;
}

int texel_bytes(int fmt) {
    switch (fmt) {
        case MM_FMT_RGBA8: return 4;
        case MM_FMT_RG8: return 2;
        case MM_FMT_FP8X4: return 4;
        case MM_FMT_FP16X4: return 8;
    }
    return 0;
}

// Which formats each texture slot accepts (generation, packing and march).
bool slot_format_ok(int slot, int fmt) {
    switch (slot) {
        case 0: return fmt == MM_FMT_RGBA8 || fmt == MM_FMT_RG8;                              // weather
        case 1: return fmt == MM_FMT_RGBA8 || fmt == MM_FMT_FP16X4;                           // curl
        case 2: return fmt == MM_FMT_RGBA8 || fmt == MM_FMT_FP8X4;     // low-res
        case 3: return fmt == MM_FMT_RGBA8 || fmt == MM_FMT_FP16X4;    // hi-res
    }
    return false;
}

// Shape/format/size validation of one texture slot. dims = 2 or 3.
int check_texture(const mm_texture* t, int slot, int dims, bool need_pow2) {
    if (!t->data) return MM_E_NULL;
    if (t->layout != MM_LAYOUT_LINEAR && t->layout != MM_LAYOUT_APRON) return MM_E_FORMAT;
    if (!slot_format_ok(slot, t->format)) return MM_E_FORMAT;
    if (t->width <= 0 || t->height <= 0 || t->depth <= 0) return MM_E_DIM;
    if (dims == 2 && t->depth != 1) return MM_E_DIM;
    if (dims == 3 && (t->width != t->height || t->width != t->depth)) return MM_E_DIM;
    if (need_pow2 && (!is_pow2(t->width) || !is_pow2(t->height))) return MM_E_DIM;
    if (t->width > 8192 || t->height > 8192 || t->depth > 1024) return MM_E_DIM;
    if (t->format == MM_FMT_FP8X4) {
        for (int k = 0; k < 4; k++)
            if (!(t->scale[k] > 0.f) || !isfinite(t->scale[k])) return MM_E_PARAM;
    }
    size_t need = mm_texture_bytes(t->width, t->height, t->depth, t->format, t->layout);
    if (need == 0) return MM_E_FORMAT;
    if (need > 0xffffffffull) return MM_E_DIM;  // the kernels address texels with 32-bit byte offsets
    if (t->bytes < need) return MM_E_SIZE;
    return MM_OK;
}

// Format sets the march kernel is instantiated for.
int format_set(const mm_noise_set* n) {
    const mm_texture &w = n->weather, &c = n->curl, &l = n->low_res, &h = n->hi_res;
    auto all = [&](int layout) {
        return w.layout == layout && c.layout == layout && l.layout == layout && h.layout == layout;
    };
    if (all(MM_LAYOUT_LINEAR) && w.format == MM_FMT_RGBA8 && c.format == MM_FMT_RGBA8 && l.format == MM_FMT_RGBA8 &&
        h.format == MM_FMT_RGBA8)
        return mm::FSET_REF;
    if (all(MM_LAYOUT_APRON) && w.format == MM_FMT_RG8 && c.format == MM_FMT_RGBA8 && l.format == MM_FMT_RGBA8 &&
        h.format == MM_FMT_RGBA8)
        return mm::FSET_FAST;
    if (all(MM_LAYOUT_APRON) && w.format == MM_FMT_RG8 && c.format == MM_FMT_FP16X4 && l.format == MM_FMT_FP8X4 &&
        h.format == MM_FMT_FP16X4)
        return mm::FSET_PREC;
    return -1;
}

mm::TexView view_of(const mm_texture& t) {
    mm::TexView v;
    v.data = (const uint8_t*)t.data;
    v.w = t.width; v.h = t.height; v.d = t.depth;
    const int a = t.layout == MM_LAYOUT_APRON ? 1 : 0;
    v.stride = t.width + a;
    v.sliceStride = (t.width + a) * (t.height + a);
    for (int k = 0; k < 4; k++) v.scale[k] = t.scale[k];
    v.lw = is_pow2(t.width) ? __builtin_ctz((unsigned)t.width) : 0;
    v.lh = is_pow2(t.height) ? __builtin_ctz((unsigned)t.height) : 0;
    return v;
}

// hipGetLastError() reports the last error of any earlier runtime call on this thread, including
// benign ones made by the caller's own runtime set-up; launch status is read with it, so it is
// cleared before this library's launches.
void reset_error_state() { (void)hipGetLastError(); }

int check_params(const mm_march_params* p) {
    if (p->size != sizeof(mm_march_params)) return MM_E_VERSION;
    const float f[] = {p->inner_diameter, p->outer_diameter, p->center_y, p->thickness, p->proj_radius,
                       p->coarse_step_frac, p->fine_step_mul, p->opacity_cutoff, p->wind_strength, p->weather_scale,
                       p->low_scale, p->curl_scale, p->hi_scale};
    if (!finite_n(f, (int)(sizeof(f) / sizeof(f[0])))) return MM_E_PARAM;
    if (!(p->inner_diameter > 0.f) || !(p->outer_diameter > 0.f) || !(p->thickness > 0.f)) return MM_E_PARAM;
    if (!(p->coarse_step_frac > 0.f) || !(p->fine_step_mul > 0.f) || p->fine_step_mul > 1.f) return MM_E_PARAM;
    if (p->miss_limit < 1) return MM_E_PARAM;
    if (p->subtexel_bits != 0 && p->subtexel_bits != 8) return MM_E_PARAM;
    if (p->cone_mode != MM_CONE_PLAIN && p->cone_mode != MM_CONE_SURROGATE) return MM_E_PARAM;
    if (p->tile_width != 0 && p->tile_width != 8 && p->tile_width != 16 && p->tile_width != 32) return MM_E_PARAM;
    if (p->lanes_per_ray != 0 && p->lanes_per_ray != 1 && p->lanes_per_ray != 2 && p->lanes_per_ray != 4 &&
        p->lanes_per_ray != 8)
        return MM_E_PARAM;
    if (p->cone_mode == MM_CONE_SURROGATE && (!p->surrogate || !(p->surrogate_threshold >= 0.f)))
        return MM_E_PARAM;
    return MM_OK;
}

}  // namespace

extern "C" {

const char* mm_strerror(int code) {
    switch (code) {
        case MM_OK: return "ok";
        case MM_E_NULL: return "required pointer is NULL";
        case MM_E_DIM: return "unsupported image or texture dimensions";
        case MM_E_PHASE: return "phase out of range";
        case MM_E_PITCH: return "row pitch smaller than width";
        case MM_E_FORMAT: return "unsupported texture format or layout";
        case MM_E_PARAM: return "invalid or non-finite parameters";
        case MM_E_DEVICE: return "HIP runtime error";
        case MM_E_SIZE: return "texture allocation too small";
        case MM_E_VERSION: return "struct size mismatch";
        case MM_E_IO: return "file could not be opened or read";
    }
    return "unknown error";
}

int mm_abi_version(void) { return MM_ABI_VERSION; }

size_t mm_texture_bytes(int32_t width, int32_t height, int32_t depth, int32_t format, int32_t layout) {
    const int tb = texel_bytes(format);
    if (tb == 0 || width <= 0 || height <= 0 || depth <= 0) return 0;
    if (layout != MM_LAYOUT_LINEAR && layout != MM_LAYOUT_APRON) return 0;
    const size_t a = layout == MM_LAYOUT_APRON ? 1 : 0;
    const size_t d = depth > 1 ? (size_t)depth + a : 1;
    return (size_t)tb * ((size_t)width + a) * ((size_t)height + a) * d;
}

void mm_noise_params_default(mm_noise_params* p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->size = sizeof(mm_noise_params);
    p->what = MM_GEN_ALL;
    p->seed = 1;
    p->weather_seed = 1;
    p->coverage = 0.5f;
}

void mm_march_params_reference(mm_march_params* p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->size = sizeof(mm_march_params);
    p->flags = 0;
    const float AR = 2000000.0f;                 // ATMOSPHERE_RADIUS (comp:56)
    p->inner_diameter = AR;
    p->outer_diameter = AR * 1.02f;              // comp:361
    p->center_y = -AR * 0.5f * 0.995f;           // comp:358
    p->thickness = 0.5f * AR * 0.02f;            // comp:360
    p->proj_radius = 0.5f * AR;                  // comp:181
    p->coarse_step_frac = 0.05f;                 // comp:390
    p->fine_step_mul = 0.3f;                     // comp:431, 472
    p->miss_limit = 10;                          // comp:470
    p->opacity_cutoff = 0.99f;                   // comp:476
    p->max_steps = 100;                          // comp:286, 481
    p->max_coarse_steps = -1;
    p->max_fine_steps = -1;
    p->wind_strength = 20.0f;                    // comp:279
    p->weather_scale = 0.000009f;                // comp:236
    p->low_scale = 0.00002f;                     // comp:238
    p->curl_scale = 0.0001f;                     // comp:217
    p->hi_scale = 0.0004f;                       // comp:223
    p->subtexel_bits = 8;
    p->cone_mode = MM_CONE_PLAIN;
    p->surrogate_threshold = 0.f;
    p->surrogate = nullptr;
}

void mm_march_params_scenario(mm_march_params* p) {
    if (!p) return;
    mm_march_params_reference(p);
    const float R = 6371000.0f;  // Earth radius
    p->inner_diameter = 2.0f * (R + 1500.0f);
    p->outer_diameter = 2.0f * (R + 4000.0f);
    p->center_y = -R;
    p->thickness = 2500.0f;
    p->proj_radius = R + 1500.0f;
    p->flags = MM_MARCH_ISECT_FP64 | MM_MARCH_BOUNDED;
    p->max_steps = -1;
    p->max_coarse_steps = 128;
    p->max_fine_steps = 256;
}

size_t mm_noise_workspace_bytes(const mm_noise_set* out) {
    (void)out;
    return mm::noise_workspace_bytes();
}

static int check_noise_call(const mm_noise_params* p, const mm_noise_set* out) {
    if (!p || !out) return MM_E_NULL;
    if (p->size != sizeof(mm_noise_params)) return MM_E_VERSION;
    if ((p->what & ~(uint32_t)MM_GEN_ALL) != 0) return MM_E_PARAM;
    if (!isfinite(p->coverage) || p->coverage < 0.f || p->coverage > 1.f) return MM_E_PARAM;
    int e;
    if ((p->what & MM_GEN_WEATHER) && (e = check_texture(&out->weather, 0, 2, false))) return e;
    if ((p->what & MM_GEN_CURL) && (e = check_texture(&out->curl, 1, 2, false))) return e;
    if ((p->what & MM_GEN_LOW_RES) && (e = check_texture(&out->low_res, 2, 3, false))) return e;
    if ((p->what & MM_GEN_HI_RES) && (e = check_texture(&out->hi_res, 3, 3, false))) return e;
    if (p->what & MM_GEN_CURL) {
        if (out->curl.width != out->curl.height || out->curl.width > 512) return MM_E_DIM;
    }
    return MM_OK;
}

int mm_generate_noise(const mm_noise_params* p, const mm_noise_set* out, void* workspace, void* stream) {
    const int e = check_noise_call(p, out);
    if (e) return e;
    reset_error_state();
    void* ws = workspace;
    hipStream_t s = (hipStream_t)stream;
    bool own = false;
    if (!ws && (p->what & MM_GEN_CURL)) {
        if (hipMallocAsync(&ws, mm::noise_workspace_bytes(), s) != hipSuccess) return MM_E_DEVICE;
        own = true;
    }
    int r = mm::launch_generate(p, out, ws, stream);
    if (own) (void)hipFreeAsync(ws, s);
    return r;
}

int mm_generate_noise_noalloc(const mm_noise_params* p, const mm_noise_set* out, void* stream) {
    static_assert(MM_STATIC_SCRATCH_BYTES >= 20480 + 256, "scratch holds the gradient table and min/max cells");
    const int e = check_noise_call(p, out);
    if (e) return e;
    if ((p->what & MM_GEN_CURL) && out->curl.width > MM_STATIC_CURL_MAX) return MM_E_SIZE;
    if (mm::static_scratch_bytes() != (size_t)MM_STATIC_SCRATCH_BYTES) return MM_E_VERSION;
    reset_error_state();
    void* ws = nullptr;
    if (p->what & MM_GEN_CURL) {
        const int s = mm::static_scratch(&ws);
        if (s) return s;
    }
    return mm::launch_generate(p, out, ws, stream, /*allow_side=*/false);
}

int mm_pack_texture(int32_t kind, const mm_texture* src, const mm_texture* dst, void* stream) {
    if (!src || !dst) return MM_E_NULL;
    if (kind < 0 || kind > 3) return MM_E_PARAM;
    const int dims = kind >= 2 ? 3 : 2;
    int e;
    if ((e = check_texture(src, kind, dims, false))) return e;
    if (src->format != MM_FMT_RGBA8 || src->layout != MM_LAYOUT_LINEAR) return MM_E_FORMAT;
    if ((e = check_texture(dst, kind, dims, false))) return e;
    if (dst->width != src->width || dst->height != src->height || dst->depth != src->depth) return MM_E_DIM;
    reset_error_state();
    return mm::launch_pack(kind, src, dst, stream);
}

struct CollectArgs {
    float* samples;
    uint32_t capacity, keep, seed;
    uint32_t* count;
};

// Validates one dispatch and derives its per-dispatch constants and texture views. `empty` is set
// when the phase selects no pixel of the image (nothing to launch).
// Default lanes per ray (see ray_parallel_tiles), measured on the 1080p scenario frame: every pixel
// 1 / 2 / 4 / 8 lanes 1.85 / 1.73 / 2.07 / 2.91 ms; one 1/16 phase 0.93 / 0.56 / 0.40 / 0.33 ms.
constexpr int kDefaultFrameLanesPerRay = 2;
constexpr int kDefaultPhaseLanesPerRay = 8;
// A views launch has many more waves (64 views of 256x256 to 1920x1080: 33.6 M pixels, 1 / 2 / 4
// lanes 20.4 / 21.7 / 29.1 ms): latency is hidden already and one lane per ray wastes no samples.
constexpr int kDefaultViewsLanesPerRay = 1;
// ... in 16 x 4 tiles (8 x 8 / 16 x 4 / 32 x 2: 20.4 / 20.0 / 20.0 ms).
constexpr int kDefaultViewsTileWidth = 16;

// Launch order of the tile rows: from the row just above the horizon up the sky (the longest rays
// first), then the rows below the horizon, whose rays end at once. The sign of each tile row's ray
// elevation is taken at the image's centre column (comp:305-322 with sx = 0). Any order gives the
// same image; without a row above the horizon the order is top-down.
static void horizon_first_rows(mm::MarchConsts& c) {
    c.rowStart = 0;
    c.rowDir = 1;
    const int th = c.lanesPerRay > 1 ? 64 / c.lanesPerRay / 8 : (64 >> c.tileShift);
    auto up = [&](int tr) {  // y of the unnormalised ray direction at the tile row's middle
        int gj = tr * th + th / 2;
        if (gj > c.gridH - 1) gj = c.gridH - 1;
        const int py = c.phase >= 0 ? gj * 4 + (c.phase >> 2) : gj;
        const double sy = 2.0 * py / c.height - 1.0;
        return (double)c.refPoint[1] - sy * c.tanH * c.camUp[1] - c.camPos[1];
    };
    if (c.tilesY < 2) return;
    const bool upright = up(0) >= up(c.tilesY - 1);  // elevation falls from the top row down
    int start = -1;
    for (int tr = 0; tr < c.tilesY; tr++)
        if (up(tr) > 0.0 && (start < 0 || upright)) start = tr;  // upright: the lowest such row
    if (start < 0) return;
    c.rowStart = start;
    c.rowDir = upright ? -1 : 1;
}

static int build_march(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                       const mm_noise_set* tex, const mm_texture* night_sky, const float* out, int32_t width,
                       int32_t height, int32_t pitch, int32_t phase, const mm_march_params* params,
                       const CollectArgs* collect, mm::MarchConsts& c, mm::MarchTextures& t, int& fset_out,
                       bool& empty) {
    if (!camera || !sun || !sky || !tex || !out) return MM_E_NULL;
    if (width <= 0 || height <= 0 || width > 32768 || height > 32768) return MM_E_DIM;
    if (pitch < width) return MM_E_PITCH;
    mm_march_params ref;
    if (!params) {
        mm_march_params_reference(&ref);
        params = &ref;
    }
    int e = check_params(params);
    if (e) return e;
    if (phase == MM_PHASE_FROM_UBO) {
        float a = sun->color[3];
        if (!(a >= 0.f && a < 16.f)) return MM_E_PHASE;
        phase = (int)a;  // GLSL int() truncates (comp:292)
    }
    if (phase < -1 || phase > 15) return MM_E_PHASE;
    if (!finite_n(camera->view, 16) || !finite_n(camera->cameraPosition, 4) || !finite_n(camera->cameraParams, 2))
        return MM_E_PARAM;
    if (!finite_n(sun->direction, 4) || !finite_n(sun->color, 3) || !finite_n(sun->directionBasis, 16) ||
        !isfinite(sun->intensity))
        return MM_E_PARAM;
    if (!finite_n(sky->betaR, 3) || !finite_n(sky->betaV, 3) || !finite_n(sky->wind, 4) ||
        !isfinite(sky->mie_directional))
        return MM_E_PARAM;
    const float* B = sun->directionBasis;
    if (B[4] * B[4] + B[5] * B[5] + B[6] * B[6] == 0.f) return MM_E_PARAM;
    if ((e = check_texture(&tex->weather, 0, 2, true))) return e;
    if ((e = check_texture(&tex->curl, 1, 2, true))) return e;
    if ((e = check_texture(&tex->low_res, 2, 3, true))) return e;
    if ((e = check_texture(&tex->hi_res, 3, 3, true))) return e;
    const int fset = format_set(tex);
    if (fset < 0) return MM_E_FORMAT;
    if (night_sky) {
        if (!night_sky->data) return MM_E_NULL;
        if (night_sky->format != MM_FMT_RGBA8 || night_sky->layout != MM_LAYOUT_LINEAR) return MM_E_FORMAT;
        if (night_sky->depth != 1 || night_sky->width <= 0 || night_sky->height <= 0) return MM_E_DIM;
        if (night_sky->bytes < mm_texture_bytes(night_sky->width, night_sky->height, 1, MM_FMT_RGBA8, MM_LAYOUT_LINEAR))
            return MM_E_SIZE;
    }

    memset(&c, 0, sizeof(c));
    const float* V = camera->view;
    for (int k = 0; k < 3; k++) {
        c.camPos[k] = camera->cameraPosition[k];
        c.camLook[k] = V[4 * k + 2];   // (view[0][2], view[1][2], view[2][2]) (comp:312)
        c.camRight[k] = V[4 * k + 0];  // comp:313
        c.camUp[k] = V[4 * k + 1];     // comp:314
        c.refPoint[k] = c.camPos[k] - c.camLook[k];
    }
    c.aspect = camera->cameraParams[0];
    c.tanH = camera->cameraParams[1];
    {
        float x = B[4], y = B[5], z = B[6];
        float s = 1.0f / sqrtf((x * x + y * y) + z * z);
        c.sunDir[0] = x * s; c.sunDir[1] = y * s; c.sunDir[2] = z * s;
    }
    const float cone[6][3] = {{0, 0.6f, 0}, {0, 0.5f, 0.05f}, {0.1f, 0.75f, 0}, {0.2f, 2.5f, 0.3f}, {0, 6, 0},
                              {-0.1f, 1, -0.2f}};
    for (int i = 0; i < 6; i++)
        for (int k = 0; k < 3; k++)
            c.coneOffsets[i][k] = ((B[k] * cone[i][0]) + (B[4 + k] * cone[i][1])) + (B[8 + k] * cone[i][2]);
    const float probe[3] = {0.6f, 2.5f, 6.0f};  // surrogate features: axial offsets of the cone table
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) c.probeOffsets[i][k] = B[4 + k] * probe[i];
    c.earthCenter[0] = c.camPos[0];
    c.earthCenter[1] = params->center_y;
    c.earthCenter[2] = c.camPos[2];
    c.innerDiam = params->inner_diameter;
    c.outerDiam = params->outer_diameter;
    c.thickness = params->thickness;
    c.invThickness = 1.0f / params->thickness;
    c.projRadius = params->proj_radius;
    c.stepSize0 = params->coarse_step_frac * params->thickness;
    c.fineMul = params->fine_step_mul;
    for (int k = 0; k < 3; k++) {
        c.betaR[k] = sky->betaR[k];
        c.betaM[k] = sky->betaV[k];
        c.sunColor[k] = sun->color[k];
        c.wind[k] = sky->wind[k];
    }
    c.sunIntensity = sun->intensity;
    c.sunDirY = sun->direction[1];
    c.mieG = sky->mie_directional;
    c.time = sky->wind[3];
    c.windStrength = params->wind_strength;
    c.weatherScale = params->weather_scale;
    c.lowScale = params->low_scale;
    c.curlScale = params->curl_scale;
    c.hiScale = params->hi_scale;
    c.opacityCutoff = params->opacity_cutoff;
    {
        // fromAngleAxis(normalize(vec3(1, 0, 1)), sun.direction.y * 0.5) (comp:256-277, 369)
        float s = 1.0f / sqrtf((1.0f * 1.0f + 0.0f * 0.0f) + 1.0f * 1.0f);
        float ax = 1.0f * s, ay = 0.0f * s, az = 1.0f * s;
        float ang = sun->direction[1] * 0.5f;
        float ct = cosf(ang), st = sinf(ang), omc = 1.f - ct;
        float m[3][3] = {{ct + ax * ax * omc, ay * ax * omc + az * st, az * ax * omc - ay * st},
                         {ax * ay * omc - az * st, ct + ay * ay * omc, az * ay * omc + ax * st},
                         {ax * az * omc + ay * st, ay * az * omc - ax * st, ct + az * az * omc}};
        memcpy(c.nightRot, m, sizeof(m));
    }
    c.missLimit = params->miss_limit;
    c.maxSteps = params->max_steps;
    c.maxCoarse = params->max_coarse_steps;
    c.maxFine = params->max_fine_steps;
    c.flags = (int)params->flags;
    c.subtexelBits = params->subtexel_bits;
    c.width = width;
    c.height = height;
    c.pitch = pitch;
    c.phase = phase;
    if (phase >= 0) {
        c.gridW = (width - (phase & 3) + 3) / 4;
        c.gridH = (height - (phase >> 2) + 3) / 4;
    } else {
        c.gridW = width;
        c.gridH = height;
    }
    c.coneMode = params->cone_mode;
    if (c.coneMode == MM_CONE_SURROGATE) {
        if (fset == mm::FSET_PREC) return MM_E_FORMAT;
        c.surrThreshold = params->surrogate_threshold;
        c.surrParams = mm::surrogate_params(params->surrogate);
        c.surrStats = mm::surrogate_stats(params->surrogate);
        c.surrFrags = mm::surrogate_infer_frags(params->surrogate);
    }
    if (collect) {
        if (fset == mm::FSET_PREC) return MM_E_FORMAT;
        c.coneMode = 2;  // plain cones + recorder
        c.collectBuf = collect->samples;
        c.collectCap = collect->capacity;
        c.collectKeep = collect->keep;
        c.collectSeed = collect->seed;
        c.collectCount = collect->count;
    }
    empty = c.gridW <= 0 || c.gridH <= 0;  // no pixel of this phase in the image
    c.lanesPerRay = 1;
    c.tileShift = params->tile_width == 16 ? 4 : params->tile_width == 32 ? 5 : 3;
    c.tilesX = (c.gridW + (1 << c.tileShift) - 1) >> c.tileShift;
    c.tilesY = (c.gridH + (64 >> c.tileShift) - 1) / (64 >> c.tileShift);
    c.tiles = c.tilesX * c.tilesY;
    horizon_first_rows(c);

    memset(&t, 0, sizeof(t));
    t.weather = view_of(tex->weather);
    t.curl = view_of(tex->curl);
    t.low = view_of(tex->low_res);
    t.hi = view_of(tex->hi_res);
    if (night_sky) {
        t.night = view_of(*night_sky);
        t.hasNight = 1;
    }
    fset_out = fset;
    return MM_OK;
}

// Lanes per ray of a plain march (params.lanes_per_ray; 0 = the default: 8 for a single 1/16 phase,
// where there are few waves and each lasts as long as its longest ray, else 2); 1 for the other cone
// modes and the lock-step kernel. The image does not depend on it.
static int lanes_for(const mm::MarchConsts& c, const mm_march_params* params, int32_t phase,
                     bool scheduled = false) {
    int lpr = params ? params->lanes_per_ray : 0;
    const bool one_phase = phase >= 0 || phase == MM_PHASE_FROM_UBO;
    if (lpr == 0)
        lpr = one_phase ? kDefaultPhaseLanesPerRay : (scheduled ? 1 : kDefaultFrameLanesPerRay);
    if (c.flags & MM_MARCH_NO_BATCH) return 1;
    // the surrogate-lit march runs with one or two lanes per ray; the sample recorder with one
    if (c.coneMode == MM_CONE_SURROGATE) return lpr >= 2 ? 2 : 1;
    if (c.coneMode != MM_CONE_PLAIN) return 1;
    return lpr;
}

// The ray-parallel kernel's tiles of 8 x (8 / lanes) rays (lanes 1 keeps the 64-ray tiles) and the
// launch order of their rows.
static void set_lanes(mm::MarchConsts& c, int lpr) {
    c.lanesPerRay = lpr;
    if (lpr > 1) {
        const int th = 64 / lpr / 8;
        c.tileShift = 3;
        c.tilesX = (c.gridW + 7) >> 3;
        c.tilesY = (c.gridH + th - 1) / th;
        c.tiles = c.tilesX * c.tilesY;
    }
    horizon_first_rows(c);
}

// Per-tile cost history of one launch shape (mm_march_clouds_scheduled).
struct mm_march_schedule {
    int capacity = 0;              // tiles the buffers hold
    unsigned* d_sched = nullptr;   // device record: order state, then each physical tile's cost in the last launch
    int* d_order = nullptr;        // device: the order built for the current launch
    bool history = false;          // d_sched holds the costs of a launch of shape `key`
    int key[6] = {0, 0, 0, 0, 0, 0};
    int device = 0;
};

int mm_march_schedule_create(int32_t max_tiles, mm_march_schedule** schedule) {
    if (!schedule) return MM_E_NULL;
    *schedule = nullptr;
    if (max_tiles < 1 || max_tiles > (1 << 26)) return MM_E_PARAM;
    mm_march_schedule* s = new (std::nothrow) mm_march_schedule;
    if (!s) return MM_E_SIZE;
    const size_t record = sizeof(unsigned) * ((size_t)mm::kSchedHeader + (size_t)max_tiles);
    if (hipGetDevice(&s->device) != hipSuccess || hipMalloc((void**)&s->d_sched, record) != hipSuccess) {
        delete s;
        return MM_E_DEVICE;
    }
    // a clear record (no bucket counts), complete before any stream uses it
    if (hipMalloc((void**)&s->d_order, sizeof(int) * (size_t)max_tiles) != hipSuccess ||
        hipMemsetAsync(s->d_sched, 0, record, nullptr) != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess) {
        (void)hipFree(s->d_sched);
        (void)hipFree(s->d_order);
        delete s;
        return MM_E_DEVICE;
    }
    s->capacity = max_tiles;
    *schedule = s;
    return MM_OK;
}

void mm_march_schedule_destroy(mm_march_schedule* schedule) {
    if (!schedule) return;
    (void)hipFree(schedule->d_sched);
    (void)hipFree(schedule->d_order);
    delete schedule;
}

// The launch shape a schedule's costs belong to.
static void schedule_key(const mm::MarchConsts& c, int key[6]) {
    key[0] = c.tiles; key[1] = c.tilesX; key[2] = c.gridW; key[3] = c.gridH;
    key[4] = c.lanesPerRay; key[5] = c.tileShift;
}

// Hooks a schedule into a built launch: the march records each tile's cost, and when the schedule
// holds the costs of a launch of the same shape, a longest-first order built from them on the stream
// first. The lock-step kernel takes no schedule (it keeps the shader's structure).
static int use_schedule(mm::MarchConsts& c, mm_march_schedule* sched, void* stream) {
    if (!sched || (c.flags & MM_MARCH_NO_BATCH)) return MM_OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return MM_E_DEVICE;
    if (dev != sched->device) return MM_E_PARAM;
    if (c.tiles > sched->capacity) return MM_E_SIZE;
    int key[6];
    schedule_key(c, key);
    if (sched->history && memcmp(key, sched->key, sizeof(key)) == 0) {
        const int e = mm::launch_tile_order(sched->d_sched, c.tiles, sched->d_order, stream);
        if (e) {
            sched->history = false;  // the next launch marches unordered and clears the record's counts
            return e;
        }
        c.tileOrder = sched->d_order;
    }
    c.tileSched = sched->d_sched;
    return MM_OK;
}

// After the march's launch: its costs are the history of its shape, if it was enqueued.
static void commit_schedule(const mm::MarchConsts& c, mm_march_schedule* sched, int launch_error) {
    if (!sched || !c.tileSched) return;
    schedule_key(c, sched->key);
    sched->history = launch_error == MM_OK;
}

static int march_impl(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                      const mm_noise_set* tex, const mm_texture* night_sky, float* out, int32_t width, int32_t height,
                      int32_t pitch, int32_t phase, const mm_march_params* params, unsigned long long* counts,
                      const CollectArgs* collect, void* stream, float* state = nullptr, int count_slots = 4,
                      mm_march_schedule* sched = nullptr) {
    mm::MarchConsts c;
    mm::MarchTextures t;
    int fset = 0;
    bool empty = false;
    const int e = build_march(camera, sun, sky, tex, night_sky, out, width, height, pitch, phase, params, collect, c,
                              t, fset, empty);
    if (e || empty) return e;
    c.countSlots = count_slots;
    if (state) {
        // the instrumented launch: the lock-step kernel (same image as every other kernel shape)
        // writes each pixel's two accumulators beside its colour
        if (c.coneMode != MM_CONE_PLAIN) return MM_E_PARAM;
        c.flags |= MM_MARCH_NO_BATCH;
        c.stateOut = state;
    }
    const int lanes = lanes_for(c, params, phase, sched != nullptr);
    if (sched && lanes == 1 && (!params || params->tile_width == 0) && phase == MM_PHASE_ALL) {
        // scheduled full frames default to 16 x 4 tiles (with longest-first order the measured best
        // shape is one lane per ray in 16-wide tiles)
        c.tileShift = 4;
        c.tilesX = (c.gridW + 15) >> 4;
        c.tilesY = (c.gridH + 3) / 4;
        c.tiles = c.tilesX * c.tilesY;
    }
    set_lanes(c, lanes);
    reset_error_state();
    const int es = use_schedule(c, sched, stream);
    if (es) return es;
    const int el = mm::launch_march(fset, c, t, out, counts, stream);
    commit_schedule(c, sched, el);
    return el;
}

int mm_march_clouds_scheduled(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                              const mm_noise_set* textures, const mm_texture* night_sky, float* out_rgba32f,
                              int32_t width, int32_t height, int32_t row_pitch_px, int32_t phase,
                              const mm_march_params* params, mm_march_schedule* schedule,
                              unsigned long long* counts, void* stream) {
    if (!schedule) return MM_E_NULL;
    return march_impl(camera, sun, sky, textures, night_sky, out_rgba32f, width, height, row_pitch_px, phase, params,
                      counts, nullptr, stream, nullptr, 4, schedule);
}

int mm_march_clouds(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                    const mm_noise_set* textures, const mm_texture* night_sky, float* out_rgba32f, int32_t width,
                    int32_t height, int32_t row_pitch_px, int32_t phase, const mm_march_params* params, void* stream) {
    return march_impl(camera, sun, sky, textures, night_sky, out_rgba32f, width, height, row_pitch_px, phase, params,
                      nullptr, nullptr, stream);
}

int mm_march_clouds_counted(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                            const mm_noise_set* textures, const mm_texture* night_sky, float* out_rgba32f,
                            int32_t width, int32_t height, int32_t row_pitch_px, int32_t phase,
                            const mm_march_params* params, unsigned long long* counts, void* stream) {
    if (!counts) return MM_E_NULL;
    return march_impl(camera, sun, sky, textures, night_sky, out_rgba32f, width, height, row_pitch_px, phase, params,
                      counts, nullptr, stream);
}

int mm_march_clouds_counted_ex(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                               const mm_noise_set* textures, const mm_texture* night_sky, float* out_rgba32f,
                               int32_t width, int32_t height, int32_t row_pitch_px, int32_t phase,
                               const mm_march_params* params, unsigned long long* counts, void* stream) {
    if (!counts) return MM_E_NULL;
    return march_impl(camera, sun, sky, textures, night_sky, out_rgba32f, width, height, row_pitch_px, phase, params,
                      counts, nullptr, stream, nullptr, 6);
}

int mm_march_clouds_state(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                          const mm_noise_set* textures, const mm_texture* night_sky, float* out_rgba32f,
                          int32_t width, int32_t height, int32_t row_pitch_px, int32_t phase,
                          const mm_march_params* params, float* state_out, void* stream) {
    if (!state_out) return MM_E_NULL;
    return march_impl(camera, sun, sky, textures, night_sky, out_rgba32f, width, height, row_pitch_px, phase, params,
                      nullptr, nullptr, stream, state_out);
}

int mm_march_clouds_collect(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                            const mm_noise_set* textures, const mm_texture* night_sky, float* out_rgba32f,
                            int32_t width, int32_t height, int32_t row_pitch_px, int32_t phase,
                            const mm_march_params* params, float* samples, uint32_t capacity, uint32_t keep_one_in,
                            uint32_t seed, uint32_t* count, void* stream) {
    if (!samples || !count) return MM_E_NULL;
    if (capacity == 0 || keep_one_in == 0) return MM_E_PARAM;
    if (params && params->cone_mode != MM_CONE_PLAIN) return MM_E_PARAM;
    CollectArgs a{samples, capacity, keep_one_in, seed, count};
    return march_impl(camera, sun, sky, textures, night_sky, out_rgba32f, width, height, row_pitch_px, phase, params,
                      nullptr, &a, stream);
}

// Validates one reprojection and derives its constants (every pixel written: skipPhase -1).
static int build_reproject(const mm_camera_ubo* camera, const mm_camera_ubo* camera_prev, const float* source,
                           float* target, int32_t width, int32_t height, int32_t row_pitch_px,
                           const mm_march_params* params, mm::ReprojConsts& c) {
    if (!camera || !camera_prev || !source || !target) return MM_E_NULL;
    if (width <= 0 || height <= 0 || width > 32768 || height > 32768) return MM_E_DIM;
    if (row_pitch_px < width) return MM_E_PITCH;
    if (source == target) return MM_E_PARAM;
    mm_march_params ref;
    if (!params) {
        mm_march_params_reference(&ref);
        params = &ref;
    }
    int e = check_params(params);
    if (e) return e;
    if (!finite_n(camera->view, 16) || !finite_n(camera->cameraPosition, 4) || !finite_n(camera->cameraParams, 2) ||
        !finite_n(camera_prev->view, 16))
        return MM_E_PARAM;
    if (!(camera->cameraParams[0] != 0.f) || !(camera->cameraParams[1] != 0.f)) return MM_E_PARAM;
    memset(&c, 0, sizeof(c));
    const float* V = camera->view;
    for (int k = 0; k < 3; k++) {
        c.camPos[k] = camera->cameraPosition[k];
        c.camRight[k] = V[4 * k + 0];
        c.camUp[k] = V[4 * k + 1];
        c.refPoint[k] = camera->cameraPosition[k] - V[4 * k + 2];  // cameraPos - camLook
    }
    c.aspect = camera->cameraParams[0];
    c.tanH = camera->cameraParams[1];
    memcpy(c.prevView, camera_prev->view, sizeof(c.prevView));
    c.centerY = params->center_y;
    c.innerDiam = params->inner_diameter;
    c.fp64 = (params->flags & MM_MARCH_ISECT_FP64) ? 1 : 0;
    c.width = width;
    c.height = height;
    c.pitch = row_pitch_px;
    c.skipPhase = -1;
    return MM_OK;
}

int mm_reproject(const mm_camera_ubo* camera, const mm_camera_ubo* camera_prev, const float* source,
                 float* target, int32_t width, int32_t height, int32_t row_pitch_px, const mm_march_params* params,
                 void* stream) {
    mm::ReprojConsts c;
    const int e = build_reproject(camera, camera_prev, source, target, width, height, row_pitch_px, params, c);
    if (e) return e;
    reset_error_state();
    return mm::launch_reproject(c, source, target, stream);
}

// The reprojection writes every pixel except those of the phase the march then writes (the
// reference's order, reproject then march, gives the same image), so the two run side by side: the
// reprojection on the library's second stream, forked from and joined back into `stream`.
int mm_render_frame(const mm_camera_ubo* camera, const mm_camera_ubo* camera_prev, const mm_sun_ubo* sun,
                    const mm_sky_ubo* sky, const mm_noise_set* textures, const mm_texture* night_sky,
                    const float* source, float* target, int32_t width, int32_t height, int32_t row_pitch_px,
                    const mm_march_params* params, void* stream) {
    mm::ReprojConsts rc;
    int e = build_reproject(camera, camera_prev, source, target, width, height, row_pitch_px, params, rc);
    if (e) return e;
    mm::MarchConsts c;
    mm::MarchTextures t;
    int fset = 0;
    bool empty = false;
    e = build_march(camera, sun, sky, textures, night_sky, target, width, height, row_pitch_px, MM_PHASE_FROM_UBO,
                    params, nullptr, c, t, fset, empty);
    if (e) return e;
    set_lanes(c, lanes_for(c, params, MM_PHASE_FROM_UBO));
    if (!empty) rc.skipPhase = c.phase;
    reset_error_state();
    hipStream_t main = (hipStream_t)stream;
    auto lock = mm::side_lock();
    hipStream_t side = empty ? nullptr : mm::side_fork(main);
    e = mm::launch_reproject(rc, source, target, side ? side : main);
    if (!e && !empty) e = mm::launch_march(fset, c, t, target, nullptr, stream);
    if (side) {
        const int j = mm::side_join(main, side);
        if (!e) e = j;
    }
    return e;
}

void mm_shadow_params_reference(mm_shadow_params* p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->size = sizeof(mm_shadow_params);
    const float R = 1000000.0f;                   // ATMOSPHERE_RADIUS (model.frag:61)
    p->inner_diameter = R;
    p->center_y = -R * 0.5f * 0.99f;              // model.frag:247
    p->thickness = 0.5f * R * 0.02f;              // model.frag:249
    p->proj_radius = 0.5f * R;                    // model.frag:75
    p->step = 0.1f * p->thickness;                // model.frag:256
    p->steps = 6;                                 // NUM_SHADOW_STEPS
    p->origin_scale = 4.0f;                       // model.frag:260
    p->wind_strength = 20.0f;                     // WIND_STRENGTH
    p->weather_scale = 0.00001f;                  // model.frag:110
    p->low_scale = 0.000057f;                     // model.frag:113
    p->subtexel_bits = 8;
}

int mm_cloud_shadow(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                    const mm_texture* weather, const mm_texture* low_res, const float* points, int32_t n, float* out,
                    const mm_shadow_params* params, void* stream) {
    if (!camera || !sun || !sky || !weather || !low_res) return MM_E_NULL;
    if (n < 0) return MM_E_PARAM;
    if (n > 0 && (!points || !out)) return MM_E_NULL;  // an empty batch may pass NULL buffers
    mm_shadow_params ref;
    if (!params) {
        mm_shadow_params_reference(&ref);
        params = &ref;
    }
    if (params->size != sizeof(mm_shadow_params)) return MM_E_VERSION;
    const float f[] = {params->inner_diameter, params->center_y, params->thickness, params->proj_radius, params->step,
                       params->origin_scale, params->wind_strength, params->weather_scale, params->low_scale};
    if (!finite_n(f, (int)(sizeof(f) / sizeof(f[0])))) return MM_E_PARAM;
    if (!(params->inner_diameter > 0.f) || !(params->thickness > 0.f) || params->steps < 0 || params->steps > 4096)
        return MM_E_PARAM;
    if (params->subtexel_bits != 0 && params->subtexel_bits != 8) return MM_E_PARAM;
    if (!finite_n(camera->view, 16) || !finite_n(camera->cameraPosition, 4) || !finite_n(sun->directionBasis, 16) ||
        !finite_n(sky->wind, 4))
        return MM_E_PARAM;
    int e;
    if ((e = check_texture(weather, 0, 2, true))) return e;
    if ((e = check_texture(low_res, 2, 3, true))) return e;
    int fset;
    if (weather->layout == MM_LAYOUT_LINEAR && low_res->layout == MM_LAYOUT_LINEAR && weather->format == MM_FMT_RGBA8 &&
        low_res->format == MM_FMT_RGBA8)
        fset = mm::FSET_REF;
    else if (weather->layout == MM_LAYOUT_APRON && low_res->layout == MM_LAYOUT_APRON &&
             weather->format == MM_FMT_RG8 && low_res->format == MM_FMT_RGBA8)
        fset = mm::FSET_FAST;
    else
        return MM_E_FORMAT;
    if (n == 0) return MM_OK;
    mm::ShadowConsts c;
    memset(&c, 0, sizeof(c));
    const float* V = camera->view;
    const float* B = sun->directionBasis;
    // L = normalize((camera.view * vec4(sunDir, 0)).xyz); flipped below the horizon (model.frag:219-220)
    float lx = (V[0] * B[4] + V[4] * B[5]) + V[8] * B[6];
    float ly = (V[1] * B[4] + V[5] * B[5]) + V[9] * B[6];
    float lz = (V[2] * B[4] + V[6] * B[5]) + V[10] * B[6];
    const float dd = (lx * lx + ly * ly) + lz * lz;
    if (!(dd > 0.f)) return MM_E_PARAM;
    const float r = 1.0f / sqrtf(dd);
    lx *= r; ly *= r; lz *= r;
    if (ly < -0.05f) { lx *= -1.0f; ly *= -1.0f; lz *= -1.0f; }
    c.L[0] = lx; c.L[1] = ly; c.L[2] = lz;
    for (int k = 0; k < 3; k++) {
        c.sunW[k] = B[4 + k];
        c.camPos[k] = camera->cameraPosition[k];
        c.wind[k] = sky->wind[k];
    }
    c.time = sky->wind[3];
    c.centerY = params->center_y;
    c.innerDiam = params->inner_diameter;
    c.thickness = params->thickness;
    c.projRadius = params->proj_radius;
    c.step = params->step;
    c.originScale = params->origin_scale;
    c.windStrength = params->wind_strength;
    c.weatherScale = params->weather_scale;
    c.lowScale = params->low_scale;
    c.steps = params->steps;
    c.subtexelBits = params->subtexel_bits;
    c.fp64 = (params->flags & MM_MARCH_ISECT_FP64) ? 1 : 0;
    c.n = n;
    reset_error_state();
    return mm::launch_shadow(fset, c, view_of(*weather), view_of(*low_res), points, out, stream);
}

void mm_post_params_reference(mm_post_params* p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->size = sizeof(mm_post_params);
    p->subtexel_bits = 8;
    p->god_ray_samples = 8;                       // NUM_SAMPLES (god-ray.frag:31)
    p->god_ray_density = 0.75f;                   // DENSITY
    p->god_ray_decay = 0.99f;                     // DECAY
    p->god_ray_exposure = 0.9f;                   // EXPOSURE
    p->blur_scale = 1.5f;                         // radialBlur.frag:59
    p->blur_gain = 1.1f;
    p->exposure = 0.7f;                           // tonemap.frag:30
    p->inv_gamma = 1.0f / 2.2f;
    p->white_point = 50.2f;                       // tonemap.frag:29
    p->vignette[0] = 0.1f;                        // tonemap.frag:34
    p->vignette[1] = 0.05f;
    p->vignette[2] = 0.13f;
}

int mm_post_process(const mm_camera_ubo* camera, const mm_sun_ubo* sun, const float* image_rgba32f, int32_t width,
                    int32_t height, int32_t row_pitch_px, float* god_alpha, float* out_hdr_rgba32f, uint8_t* out_rgba8,
                    const mm_post_params* params, void* stream) {
    if (!camera || !sun || !image_rgba32f || !god_alpha) return MM_E_NULL;
    if (width <= 0 || height <= 0 || width > 16384 || height > 16384) return MM_E_DIM;
    if (row_pitch_px < width || (int64_t)row_pitch_px * height > (1ll << 28)) return MM_E_PITCH;
    mm_post_params ref;
    if (!params) {
        mm_post_params_reference(&ref);
        params = &ref;
    }
    if (params->size != sizeof(mm_post_params)) return MM_E_VERSION;
    const float f[] = {params->god_ray_density, params->god_ray_decay, params->god_ray_exposure, params->blur_scale,
                       params->blur_gain, params->exposure, params->inv_gamma, params->white_point,
                       params->vignette[0], params->vignette[1], params->vignette[2]};
    if (!finite_n(f, (int)(sizeof(f) / sizeof(f[0])))) return MM_E_PARAM;
    if (params->subtexel_bits < 1 || params->subtexel_bits > 8) return MM_E_PARAM;
    if (params->god_ray_samples < 1 || params->god_ray_samples > 256) return MM_E_PARAM;
    if (!finite_n(camera->view, 16) || !finite_n(camera->proj, 16) || !finite_n(sun->location, 4) ||
        !finite_n(sun->direction, 4) || !finite_n(sun->color, 3) || !isfinite(sun->intensity))
        return MM_E_PARAM;
    mm::PostConsts c;
    memset(&c, 0, sizeof(c));
    // the sun in NDC: (proj * view) * location, then / w (god-ray.frag:47-49, radialBlur.frag:52-53)
    const float* P = camera->proj;
    const float* V = camera->view;
    float M[16];
    for (int col = 0; col < 4; col++)
        for (int r = 0; r < 4; r++)
            M[col * 4 + r] = ((P[0 * 4 + r] * V[col * 4 + 0] + P[1 * 4 + r] * V[col * 4 + 1]) +
                              P[2 * 4 + r] * V[col * 4 + 2]) + P[3 * 4 + r] * V[col * 4 + 3];
    float sp[4];
    for (int r = 0; r < 4; r++)
        sp[r] = ((M[0 * 4 + r] * sun->location[0] + M[1 * 4 + r] * sun->location[1]) + M[2 * 4 + r] * sun->location[2]) +
                M[3 * 4 + r] * sun->location[3];
    c.sunX = sp[0] / sp[3];
    c.sunY = sp[1] / sp[3];
    c.width = width;
    c.height = height;
    c.pitch = row_pitch_px;
    c.bits = params->subtexel_bits;
    c.samples = params->god_ray_samples;
    c.below = sun->direction[1] < 0.0f ? 1 : 0;
    c.sw = 1.0f / (float)params->god_ray_samples;
    c.k = c.sw * params->god_ray_density;
    c.decay = params->god_ray_decay;
    c.godExposure = params->god_ray_exposure;
    c.blurScale = params->blur_scale;
    c.blurGain = params->blur_gain;
    for (int k = 0; k < 3; k++) {
        c.gc[k] = sun->color[k] * sun->intensity;
        c.vignette[k] = params->vignette[k];
    }
    c.exposure = params->exposure;
    c.invGamma = params->inv_gamma;
    {
        // 1 / uc2(white point) (tonemap.frag:20), the curve as the kernels evaluate it
        const float x = params->white_point;
        const float A = 0.15f, k05 = 0.1f * 0.5f, k004 = 0.2f * 0.02f, k06 = 0.2f * 0.3f, k0067 = 0.02f / 0.3f;
        c.whitemap = 1.0f / (((x * (A * x + k05) + k004) / (x * (A * x + 0.5f) + k06)) - k0067);
    }
    reset_error_state();
    return mm::launch_post(c, image_rgba32f, god_alpha, out_hdr_rgba32f, out_rgba8, stream);
}

struct mm_view_plan {
    int n_views = 0, total_tiles = 0, fset = 0, mode = 0, lanes = 1;
    bool float_weights = false;       // subtexel_bits 0: the float-sampler kernels
    bool bounded_steps = false;       // the kernels' specialised step checks (mm::bounded_steps)
    mm::ViewDesc* d_views = nullptr;  // device: ViewDesc[n_views]
    mm::MarchTextures shared;
    mm_noise_set shared_set;          // the shared textures as given (curl, low_res, hi_res)
    std::vector<int> tile_base;       // each view's first tile index (the launch's tile layout)
    int device = 0;
    // tile schedule: every launch records each global tile's cost, and every launch after the first
    // marches the tiles longest first by the previous launch's costs (the layout never changes)
    unsigned* d_sched = nullptr;  // schedule record: order state, then the global tiles' costs
    int* d_order = nullptr;
    mutable bool launched = false;  // set by mm_march_views, which takes the plan as const
};

// Every view of a plan is marched with the same kernel shape: params.lanes_per_ray (or
// kDefaultViewsLanesPerRay) and params.tile_width (or kDefaultViewsTileWidth).
static int fill_views(const mm_view* views, int32_t n, const mm_noise_set* shared, const mm_march_params* given,
                      std::vector<mm::ViewDesc>& out, mm::MarchTextures& tx, int& fset, int& total, int& lanes) {
    mm_march_params vp = *given;
    if (vp.tile_width == 0) vp.tile_width = kDefaultViewsTileWidth;
    const mm_march_params* params = &vp;
    try {  // no exception crosses the C boundary
        out.resize((size_t)n);
    } catch (...) {
        return MM_E_SIZE;
    }
    total = 0;
    fset = -1;
    for (int i = 0; i < n; i++) {
        const mm_view& v = views[i];
        mm_noise_set set = *shared;
        set.weather = v.weather;
        mm::MarchConsts c;
        mm::MarchTextures t;
        int fs = 0;
        bool empty = false;
        const int e = build_march(&v.camera, &v.sun, &v.sky, &set, nullptr, v.out_rgba32f, v.width, v.height,
                                  v.row_pitch_px, v.phase, params, nullptr, c, t, fs, empty);
        if (e) return e;
        if (fset >= 0 && fs != fset) return MM_E_FORMAT;
        fset = fs;
        lanes = params->lanes_per_ray ? lanes_for(c, params, -1) : kDefaultViewsLanesPerRay;
        set_lanes(c, lanes);
        if (i == 0) tx = t;
        mm::ViewDesc& d = out[(size_t)i];
        memset(&d, 0, sizeof(d));
        d.c = c;
        d.weather = t.weather;
        d.out = v.out_rgba32f;
        d.tileBase = total;
        const long long tiles = empty ? 0 : (long long)c.tilesX * c.tilesY;
        if ((long long)total + tiles > (1ll << 30)) return MM_E_DIM;
        total += (int)tiles;
    }
    return MM_OK;
}

int mm_view_plan_create(const mm_view* views, int32_t n_views, const mm_noise_set* shared,
                        const mm_march_params* params, mm_view_plan** plan, void* stream) {
    if (!views || !shared || !plan) return MM_E_NULL;
    *plan = nullptr;
    if (n_views < 1 || n_views > 65536) return MM_E_PARAM;
    mm_march_params ref;
    if (!params) {
        mm_march_params_reference(&ref);
        params = &ref;
    }
    std::vector<mm::ViewDesc> h;
    mm::MarchTextures tx;
    int fset = 0, total = 0, lanes = 1;
    int e = fill_views(views, n_views, shared, params, h, tx, fset, total, lanes);
    if (e) return e;
    reset_error_state();
    mm_view_plan* p = new (std::nothrow) mm_view_plan();
    if (!p) return MM_E_DEVICE;
    p->n_views = n_views;
    p->total_tiles = total;
    p->lanes = lanes;
    p->fset = fset;
    p->mode = params->cone_mode;
    p->float_weights = params->subtexel_bits == 0;
    p->bounded_steps = mm::bounded_steps(params->flags, params->max_steps, params->max_coarse_steps,
                                         params->max_fine_steps);
    p->shared = tx;
    p->shared_set = *shared;
    p->tile_base.resize((size_t)n_views);
    for (int i = 0; i < n_views; i++) p->tile_base[(size_t)i] = h[(size_t)i].tileBase;
    (void)hipGetDevice(&p->device);
    if (fset == mm::FSET_PREC && p->mode == MM_CONE_SURROGATE) {
        delete p;
        return MM_E_FORMAT;
    }
    const size_t bytes = sizeof(mm::ViewDesc) * (size_t)n_views;
    const size_t nt = (size_t)(total > 0 ? total : 1);
    if (hipMalloc((void**)&p->d_views, bytes) != hipSuccess) {
        delete p;
        return MM_E_DEVICE;
    }
    const size_t record = sizeof(unsigned) * ((size_t)mm::kSchedHeader + nt);
    if (hipMalloc((void**)&p->d_sched, record) != hipSuccess ||
        hipMalloc((void**)&p->d_order, sizeof(int) * nt) != hipSuccess) {
        (void)hipFree(p->d_views);
        (void)hipFree(p->d_sched);
        delete p;
        return MM_E_DEVICE;
    }
    hipStream_t s = (hipStream_t)stream;
    // a clear record (no counts, zero costs) until a launch records costs: a plan whose first launch
    // is captured into a graph orders its first eager launch from these
    if (hipMemcpyAsync(p->d_views, h.data(), bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemsetAsync(p->d_sched, 0, record, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipFree(p->d_views);
        (void)hipFree(p->d_sched);
        (void)hipFree(p->d_order);
        delete p;
        return MM_E_DEVICE;
    }
    *plan = p;
    return MM_OK;
}

int mm_view_plan_update(mm_view_plan* plan, const mm_view* views, int32_t n_views, const mm_march_params* params,
                        void* stream) {
    if (!plan || !views) return MM_E_NULL;
    if (n_views != plan->n_views) return MM_E_PARAM;
    mm_march_params ref;
    if (!params) {
        mm_march_params_reference(&ref);
        params = &ref;
    }
    if (params->cone_mode != plan->mode) return MM_E_PARAM;
    if ((params->subtexel_bits == 0) != plan->float_weights) return MM_E_PARAM;  // a different kernel
    if (mm::bounded_steps(params->flags, params->max_steps, params->max_coarse_steps, params->max_fine_steps) !=
        plan->bounded_steps)
        return MM_E_PARAM;  // a different kernel
    // the shared volumes stay those of the plan; views bring new uniforms, weather maps and outputs
    std::vector<mm::ViewDesc> h;
    mm::MarchTextures tx;
    int fset = 0, total = 0, lanes = 1;
    int e = fill_views(views, n_views, &plan->shared_set, params, h, tx, fset, total, lanes);
    if (e) return e;
    if (fset != plan->fset) return MM_E_FORMAT;
    // the launch shape (grid = total tiles, kernel = lanes per ray) is what a captured graph keeps:
    // an update that changes it needs a new plan, not new descriptors under an old launch
    if (total != plan->total_tiles || lanes != plan->lanes) return MM_E_PARAM;
    for (int i = 0; i < n_views; i++)
        if (h[(size_t)i].tileBase != plan->tile_base[(size_t)i]) return MM_E_PARAM;
    reset_error_state();
    hipStream_t s = (hipStream_t)stream;
    if (hipMemcpyAsync(plan->d_views, h.data(), sizeof(mm::ViewDesc) * (size_t)n_views, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return MM_E_DEVICE;
    return MM_OK;
}

int mm_view_plan_info(const mm_view_plan* plan, int32_t* total_tiles, int32_t* grid_waves) {
    if (!plan) return MM_E_NULL;
    if (total_tiles) *total_tiles = plan->total_tiles;
    if (grid_waves) *grid_waves = plan->total_tiles;
    return MM_OK;
}

int mm_march_views(const mm_view_plan* plan, unsigned long long* counts, void* stream) {
    if (!plan) return MM_E_NULL;
    if (plan->total_tiles == 0) return MM_OK;
    reset_error_state();
    const int* order = nullptr;
    if (plan->launched) {
        const int e = mm::launch_tile_order(plan->d_sched, plan->total_tiles, plan->d_order, stream);
        if (e) {
            plan->launched = false;  // the next launch marches unordered and clears the record's counts
            return e;
        }
        order = plan->d_order;
    }
    const int e = mm::launch_views(plan->fset, plan->mode, plan->lanes, plan->float_weights, plan->bounded_steps,
                                   plan->d_views, plan->n_views, plan->shared, plan->total_tiles, counts, order,
                                   plan->d_sched, stream);
    if (!e) plan->launched = true;
    return e;
}

void mm_view_plan_destroy(mm_view_plan* plan) {
    if (!plan) return;
    (void)hipFree(plan->d_views);
    (void)hipFree(plan->d_sched);
    (void)hipFree(plan->d_order);
    delete plan;
}

size_t mm_surrogate_bytes(void) { return mm::surrogate_model_bytes(); }

int mm_surrogate_init(void* model, uint32_t seed, void* stream) {
    if (!model) return MM_E_NULL;
    reset_error_state();
    return mm::surrogate_init(model, seed, stream);
}

int mm_surrogate_fit_stats(void* model, const float* samples, int32_t n, void* stream) {
    if (!model || !samples) return MM_E_NULL;
    if (n < 2) return MM_E_PARAM;
    reset_error_state();
    return mm::surrogate_fit_stats(model, samples, n, MM_SURROGATE_ROW, stream);
}

int mm_surrogate_train(void* model, const float* samples, int32_t n, int32_t batch, int32_t steps, float lr,
                       uint32_t seed, float* loss_out, void* stream) {
    if (!model || !samples) return MM_E_NULL;
    if (n < 1 || steps < 0 || batch < 64 || batch % 64 != 0 || batch > (1 << 24) || !(lr > 0.f) || !isfinite(lr))
        return MM_E_PARAM;
    reset_error_state();
    return mm::surrogate_train(model, samples, n, MM_SURROGATE_ROW, batch, steps, lr, seed, loss_out, stream);
}

int mm_surrogate_load(void* model, const float* params, const float* stats, void* stream) {
    if (!model) return MM_E_NULL;
    reset_error_state();
    return mm::surrogate_load(model, params, stats, stream);
}

int mm_surrogate_gradient(const void* model, const float* samples, int32_t n, float* grad_out, float* loss_out,
                          void* stream) {
    if (!model || !samples || !grad_out || !loss_out) return MM_E_NULL;
    if (n < 1 || n > (1 << 26)) return MM_E_PARAM;
    reset_error_state();
    return mm::surrogate_gradient(model, samples, n, MM_SURROGATE_ROW, grad_out, loss_out, stream);
}

int mm_surrogate_predict(const void* model, const float* samples, int32_t n, float* out, void* stream) {
    if (!model || !samples || !out) return MM_E_NULL;
    if (n < 0) return MM_E_PARAM;
    if (n == 0) return MM_OK;
    reset_error_state();
    return mm::surrogate_predict(model, samples, n, MM_SURROGATE_ROW, out, stream);
}

}  // extern "C"
