// CPU restatement of the cloud-shadow march of the mesh pass, SkyEngine/SkyEngine/Shaders/model.frag
// (test infrastructure: the checker of the HIP kernel, never linked into the product).
//
// model.frag:240-270 marches NUM_SHADOW_STEPS low-res density samples from each shaded point toward
// the sun and keeps the maximum density (color *= 1 - 2 * accum, model.frag:282). It carries its
// own copies of the cloud functions (model.frag:52-129) with constants that differ from
// compute-clouds.comp: ATMOSPHERE_RADIUS 1e6, centre y = -0.99 R / 2, weather scale 1e-5, low-res
// scale 5.7e-5, coverage exponent remap(0.7, 0.8 -> 1, 0.6) clamped to [0.6, 1], wind
// (0, 0.2 h, 0), and a layer blend whose second term mixes stratocumulus with stratus.
// Quirks kept: the ray starts at t = isect.t of the inner sphere along the world-space sun
// direction (not normalised), but steps along the view-space sun direction L from 4x the world
// position; L is flipped when L.y < -0.05. A miss of the sphere leaves isect.t undefined in GLSL;
// here it is 0.
#include <math.h>
#include <stdint.h>

#include "oracle.h"

namespace {

struct F3 {
    float x, y, z;
};
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline F3 operator*(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
inline float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline float length(F3 a) { return sqrtf(dot(a, a)); }
inline F3 normalize(F3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
inline float remap(float v, float a, float b, float c, float d) { return c + (((v - a) / (b - a)) * (d - c)); }
inline float remapClamped(float v, float a, float b, float c, float d) {
    return clampf(c + (((v - a) / (b - a)) * (d - c)), c, d);
}

// model.frag:88-98 (d2 mixes stratocumulus with stratus)
float layerDensity(float h, float type) {
    h = clampf(h, 0.0f, 1.0f);
    float cumulus = fmaxf(0.0f, remap(h, 0.0f, 0.2f, 0.0f, 1.0f) * remap(h, 0.7f, 1.0f, 1.0f, 0.0f));
    (void)cumulus;  // computed and unused by the reference
    float strato = fmaxf(0.0f, remap(h, 0.0f, 0.2f, 0.0f, 1.0f) * remap(h, 0.2f, 0.7f, 1.0f, 0.0f));
    float stratus = fmaxf(0.0f, remap(h, 0.0f, 0.1f, 0.0f, 1.0f) * remap(h, 0.2f, 0.3f, 1.0f, 0.0f));
    float d1 = mixf(stratus, strato, clampf(type * 2.0f, 0.0f, 1.0f));
    float d2 = mixf(strato, stratus, clampf((type - 0.5f) * 2.0f, 0.0f, 1.0f));
    return mixf(d1, d2, type);
}

// model.frag:101-103. pow(x, y) is evaluated as exp2(y * log2(x)) in double, rounded once to float
// (GLSL defines pow through exp2/log2 and leaves its precision to the implementation; the HIP kernel
// uses the same formula).
float heightBiasCoverage(float coverage, float height) {
    const float e = clampf(remap(height, 0.7f, 0.8f, 1.0f, 0.6f), 0.6f, 1.0f);
    return (float)exp2((double)e * log2((double)coverage));
}

struct Ctx {
    const mm_shadow_params* p;
    const uint8_t* weather;
    int ww, wh;
    const uint8_t* low;
    int ln;
    float camx, camz;
};

// model.frag:106-129
float cloudTest(const Ctx& c, F3 pos, float h, F3 center) {
    F3 proj = c.p->proj_radius * normalize(pos - center) + center;
    float info[4], dn[4];
    orc_sample2d(c.weather, c.ww, c.wh, c.p->weather_scale * (proj.x - c.camx), c.p->weather_scale * (proj.z - c.camz),
                 c.p->subtexel_bits, info);
    float layer = layerDensity(h, info[2]);
    orc_sample3d(c.low, c.ln, c.p->low_scale * pos.x, c.p->low_scale * pos.y, c.p->low_scale * pos.z,
                 c.p->subtexel_bits, dn);
    float density = layer * remapClamped(dn[0], 0.3f, 1.0f, 0.0f, 1.0f);
    if (density < 0.0001f) return 0.0f;
    float coverage = heightBiasCoverage(h, fminf(0.85f, info[0]));
    float erosion = ((0.625f * dn[1]) + (0.25f * dn[2])) + (0.125f * dn[3]);
    erosion = remapClamped(erosion, coverage, 1.0f, 0.0f, 1.0f);
    return remapClamped(density, erosion, 1.0f, 0.0f, 1.0f);
}

// raySphereIntersection(...).t (model.frag:132-161); 0 on a miss.
template <class I>
I sphere_t(I rox, I roy, I roz, I rdx, I rdy, I rdz, I cx, I cy, I cz, I w) {
    rox = (rox - cx) / w;
    roy = (roy - cy) / w;
    roz = (roz - cz) / w;
    I A = (rdx * rdx + rdy * rdy) + rdz * rdz;
    I B = I(2) * ((rdx * rox + rdy * roy) + rdz * roz);
    I C = ((rox * rox + roy * roy) + roz * roz) - I(0.25);
    I disc = (B * B) - ((I(4) * A) * C);
    if (disc < I(0)) return I(0);
    I t = ((-sqrt(disc) - B) / A) * I(0.5);
    if (t < I(0)) t = ((sqrt(disc) - B) / A) * I(0.5);
    if (!(t >= I(0))) return I(0);
    I px = (rox + rdx * t) * w + cx, py = (roy + rdy * t) * w + cy, pz = (roz + rdz * t) * w + cz;
    I dx = px - rox, dy = py - roy, dz = pz - roz;
    return sqrt((dx * dx + dy * dy) + dz * dz);
}

}  // namespace

extern "C" float orc_shadow_layer_density(float h, float type) { return layerDensity(h, type); }

extern "C" int orc_cloud_shadow(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                                const uint8_t* weather, int ww, int wh, const uint8_t* low, int ln,
                                const float* points, int n, float* out, const mm_shadow_params* p) {
    if (!cam || !sun || !sky || !weather || !low || !points || !out || !p) return -1;
    Ctx c{p, weather, ww, wh, low, ln, cam->cameraPosition[0], cam->cameraPosition[2]};
    const float* V = cam->view;
    const float* B = sun->directionBasis;
    const F3 sunW = f3(B[4], B[5], B[6]);
    // L = normalize((camera.view * vec4(sunDir, 0)).xyz) (model.frag:219)
    F3 L = normalize(f3((V[0] * sunW.x + V[4] * sunW.y) + V[8] * sunW.z, (V[1] * sunW.x + V[5] * sunW.y) + V[9] * sunW.z,
                        (V[2] * sunW.x + V[6] * sunW.y) + V[10] * sunW.z));
    if (L.y < -0.05f) L = L * -1.0f;
    const F3 camPos = f3(cam->cameraPosition[0], cam->cameraPosition[1], cam->cameraPosition[2]);
    F3 center = camPos;
    center.y = p->center_y;
    const float tw = sky->wind[3];
    for (int i = 0; i < n; i++) {
        const F3 wc = f3(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
        float t;
        if (p->flags & MM_MARCH_ISECT_FP64)
            t = (float)sphere_t<double>(wc.x, wc.y, wc.z, sunW.x, sunW.y, sunW.z, center.x, center.y, center.z,
                                        (double)p->inner_diameter);
        else
            t = sphere_t<float>(wc.x, wc.y, wc.z, sunW.x, sunW.y, sunW.z, center.x, center.y, center.z,
                                p->inner_diameter);
        float accum = 0.0f;
        const F3 origin = wc * p->origin_scale;
        for (int s = 0; s < p->steps; s++) {
            F3 pos = origin + L * t;
            F3 proj = p->proj_radius * normalize(pos - center) + center;
            float h = clampf(length(pos - proj) / p->thickness, 0.0f, 1.0f);
            F3 wind = p->wind_strength * (f3(sky->wind[0], sky->wind[1], sky->wind[2]) + f3(0.0f, 0.2f * h, 0.0f));
            wind = wind * (tw + h * 200.0f);
            float d = cloudTest(c, pos + wind, h, center);
            accum = fmaxf(d, accum);
            if (accum > 0.99f) {
                accum = 1.0f;
                break;
            }
            t += p->step;
        }
        out[i] = accum;
    }
    return 0;
}
